"""Where the mixed-precision WENO weights leave FP32's range.

The FP32 reconstruction forms its weights as products of four (eps + S_k) factors with
the indicators scaled by 1/720 (eps = 2.592e-6/720); the kernel recomputes a face in FP64
when |N| + D exceeds 2^125 (N the weighted numerator), which covers a weight sum D whose
reciprocal is not a normal FP32 number; D itself never falls below eps^4, far above
2^-126 (pinned below).  This audit evaluates D in FP64 from the oracle's indicators on every
x-window of the positive split fluxes: none of the Sod-plus-noise windows comes near the
limit (the fallback costs the benchmark nothing), while the blasts with pressure ratios of
1e3 and more have windows beyond it (the GPU blast test exercises the fallback)."""
import numpy as np
import pytest

from oracle import np_oracle as N
from windstencil_amd import initial as I

D = (0.0401954833730, 0.2493800006710, 0.4802686256260, 0.2009775476730, 0.0291783426580)
EPS = 2.592e-6
LO, HI = 2.0 ** -126, 2.0 ** 125


def _fp32_weight_sums(orc, prims):
    rho, u, v, w, T = prims
    fp, _ = N.split_sw(rho, u, v, w, np.sqrt(T), 1.0, 0.0, 0.0, 1.4, 1e-3)
    out = []
    for comp in np.asarray(fp):
        lines = comp.reshape(-1, comp.shape[-1])
        ext = np.concatenate([lines[:, -4:], lines, lines[:, :4]], axis=1)   # periodic windows
        for line in lines[:: max(1, len(lines) // 40)]:
            i = np.where((lines == line).all(axis=1))[0][0]
            for f in range(comp.shape[-1] + 1):
                S, _, _ = orc.weno_p_parts(ext[i, f:f + 8])
                e = (EPS + S) / 720.0
                out.append(sum(D[k] * np.prod([e[j] for j in range(5) if j != k]) for k in range(5)))
    return np.array(out)


def test_sod_windows_stay_in_fp32_range(orc):
    dsum = _fp32_weight_sums(orc, I.sod_noise(40, 8, 4, 5))
    # D >= sum(d_k) (eps/720)^4 = 1.68e-34 whatever the window (flat windows reach it:
    # the transverse momentum fluxes of the Sod state are exactly zero)
    assert dsum.min() >= 0.999 * (EPS / 720.0) ** 4 > 1e3 * LO
    assert dsum.max() < 1e-6 * HI


@pytest.mark.parametrize("ratio,beyond", [(1e1, False), (1e3, True), (1e4, True), (1e6, True)])
def test_blast_windows_beyond_fp32_range(orc, ratio, beyond):
    dsum = _fp32_weight_sums(orc, I.blast(48, 20, 16, 3, ratio))
    out = int(np.count_nonzero((dsum < LO) | (dsum > HI)))
    print(f"ratio {ratio:g}: {out} of {dsum.size} windows out of FP32 range, max D {dsum.max():.3e}")
    assert (out > 0) == beyond
