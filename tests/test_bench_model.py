"""The bench's roofline model of whole thermal solves (CPU: arithmetic only)."""
import importlib.util
import os

import pytest

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_model", os.path.join(HERE, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _levels(n, faces):
    out = []
    while True:
        out.append({"shape": (n, n, n), "faces": (1, 1, 1) if faces else None})
        if n // 2 <= 4:
            return out
        n //= 2


@pytest.mark.parametrize("faces", [True, False])
def test_vcycle_bytes_sum_the_levels(bench, faces):
    lv = _levels(32, faces)  # 32, 16, 8
    f = 0 if faces else 1
    want = 32 ** 3 * bench._VC_LEVEL_B[f] + 16 ** 3 * bench._VC_LEVEL_B[f] + 8 ** 3 * bench._VC_COARSE_B[f]
    assert bench.vcycle_bytes(lv) == want


@pytest.mark.parametrize("faces", [True, False])
def test_solve_roofline_counts_iterations_network_and_hybrid_work(bench, faces):
    lv = _levels(64, faces)
    f = 0 if faces else 1
    n0 = 64.0 ** 3
    vc = bench.vcycle_bytes(lv)
    kb = bench._KRYLOV_B["gcr6"][f] * n0
    tr = 2 * bench._TRUE_RES_B[f] * n0
    plain = bench.solve_roofline(lv, 10, "gcr6", 1.0)
    assert plain["algorithmic_bytes"] == int(10 * (vc + kb) + tr)
    flops = 1e12
    hyb = bench.solve_roofline(lv, 10, "gcr6", 1.0, network_flops=flops, network_iterations=10, hybrid=True)
    hb = 10 * (kb + vc + bench._HYBRID_B[f] * n0) + tr
    assert hyb["algorithmic_bytes"] == int(hb)
    t = hb / (bench.HBM_PEAK_GBS * 1e9) + 10 * flops / bench.MFMA_BF16_PEAK_FLOPS
    assert abs(hyb["t_roof_s"] - round(t, 5)) < 1e-9 and abs(hyb["frac"] - round(t, 4)) < 1e-9
    alone = bench.solve_roofline(lv, 10, "gcr6", 1.0, network_flops=flops, network_iterations=4)
    assert alone["algorithmic_bytes"] == int(6 * (vc + kb) + 4 * kb + tr)
