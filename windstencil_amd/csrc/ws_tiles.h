// ws_tiles.h -- tile-shape identifiers shared by the tiled kernels and the autotuner.
#pragma once

#include "../../include/windstencil_amd.h"
