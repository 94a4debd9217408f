"""The algorithmic bytes per RK4 step that bench.py divides by its
measured step time (step roofline) equal the per-step figures the
workloads are priced at: B_alg = T (b_X + b_K) + R b_K with T transforms and R
state-algebra streams per step (ns2d 21 / 17, ns3d 36 / 51, strat 52 / 68),
b_X = p n^d and b_K = 2 p n^(d-1) (n/2 + 1)."""

import pytest

import bench


@pytest.mark.parametrize("fn,n,prec,expected", [
    (bench.b_alg_ns3d, 512, 4, 66.22e9),      # ns3d 512^3 FP32
    (bench.b_alg_ns3d, 512, 8, 132.44e9),     # ns3d 512^3 FP64
    (bench.b_alg_ns3d, 1024, 4, 529.0e9),     # ns3d 1024^3 FP32
    (bench.b_alg_strat, 1024, 4, 739.7e9),    # strat 1024^3 FP32
    (bench.b_alg_strat, 256, 8, 23.21e9),     # strat 256^3 FP64
    (bench.b_alg_ns2d, 2048, 4, 990.5e6),     # ns2d 2048^2 FP32
    (bench.b_alg_ns2d, 256, 8, 31.09e6),      # ns2d 256^2 FP64
    (bench.b_alg_ns2d, 128, 4, 3.91e6),       # ns2d 128^2 FP32 (LES)
])
def test_b_alg_matches_the_priced_figures(fn, n, prec, expected):
    assert fn(n, prec) == pytest.approx(expected, rel=2e-3)


def test_ensemble_b_alg():
    total = sum(count * bench.b_alg_ns2d(n, 4) for n, count in bench.ENSEMBLE_SPEC)
    assert total == pytest.approx(7.954e9, rel=1e-3)
    assert sum(c for _, c in bench.ENSEMBLE_SPEC) == 512


def test_b_alg_counts_transforms_and_streams():
    # a planted mistake (R = 36 instead of 51 for ns3d) leaves the table
    n, p = 512, 4
    bX, bK = p * n**3, 2 * p * n * n * (n // 2 + 1)
    assert bench.b_alg_ns3d(n, p) == 36 * (bX + bK) + 51 * bK
    assert bench.b_alg_ns3d(n, p) != pytest.approx(36 * (bX + bK) + 36 * bK, rel=1e-2)
