"""The vector ALU throughput probe behind the measured VALU roof (ws_probe_valu)."""
import ctypes

import pytest

import windstencil_amd as W
from tests.gpu_util import log

pytestmark = pytest.mark.gpu

# wave64 instructions at one per four cycles on 4 x 16 lanes per compute unit, 256
# compute units at 2.4 GHz: 39.3 tera lane-operations/s
FOUR_CYCLE_RATE = 256 * 64 * 2.4e9 / 1e12


@pytest.mark.parametrize("prec", [W.WS_PREC_FP64, W.WS_PREC_MIXED])
@pytest.mark.parametrize("op", [W.WS_PROBE_FMA, W.WS_PROBE_MUL, W.WS_PROBE_ADD])
def test_valu_probe_reports_plausible_rates(prec, op):
    """Every operation runs at a positive rate no faster than two instructions per four
    cycles (scalar FP32 additions may issue at twice the four-cycle rate)."""
    L = W.lib()
    v = ctypes.c_double()
    W.check(L.ws_probe_valu(prec, op, ctypes.byref(v), None), "probe")
    log(f"valu probe prec {prec} op {op}: {v.value:.2f} tera lane-ops/s")
    assert 0.2 * FOUR_CYCLE_RATE < v.value < 2.1 * FOUR_CYCLE_RATE


def test_valu_probe_rejects_bad_arguments():
    L = W.lib()
    v = ctypes.c_double()
    assert L.ws_probe_valu(7, W.WS_PROBE_FMA, ctypes.byref(v), None) != 0
    assert L.ws_probe_valu(W.WS_PREC_FP64, 9, ctypes.byref(v), None) != 0
    assert L.ws_probe_valu(W.WS_PREC_FP64, W.WS_PROBE_FMA, None, None) != 0
