"""bench.py's constants and CLI contract (CPU only)."""
import json
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_algorithmic_flop_constant_matches_op_count(orc):
    import bench
    oc = orc.opcount()
    assert bench.ALGO_FLOPS_PER_CELL == oc["ALGO_FLOPS_PER_CELL"]
    assert bench.ALGO_FLOPS_PER_DIRECTION == oc["per_direction"]
    assert bench.ALGO_FLOPS_PER_CELL_SHARED_EPS == oc["ALGO_FLOPS_PER_CELL_SHARED_EPS"]
    assert bench.VISCOUS_FLOPS_PER_CELL_CD6 == oc["VISCOUS_FLOPS_PER_CELL_CD6"]


def test_viscous_op_count_by_hand(orc):
    """The viscous count's pieces, derived from the formulas: the viscosity c0 sqrt(T^3)/(tsb+T)
    is two multiplications for T^3, the root, a multiplication, an addition and a division (6);
    a CD6 derivative three differences, two multiplications by 45 and 9, two additions, the
    division by 60 and the one by h (9), CD8 four differences, four multiplications, three
    additions and the two divisions (13); one direction's fluxes are twelve metric products
    of three terms (5 each: 60), div (2), three diagonal stresses (4 each: 12), three shear
    stresses (2 each: 6), Cp/Pr mu (1), three energy fluxes of four products (7 each: 21),
    three scaled metric terms (3) and four fluxes of three products (5 each: 20): 125."""
    oc = orc.opcount()
    assert (oc["vis_amu"], oc["vis_cd6"], oc["vis_cd8"], oc["vis_flux_per_direction"]) == (6, 9, 13, 125)
    assert oc["VISCOUS_FLOPS_PER_CELL_CD6"] == 6 + 12 * 9 + 3 * 125 + 12 * (9 + 2)


def test_multi_gpu_request_exits_with_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in json.loads(p.stdout.strip().splitlines()[-1])["error"]
