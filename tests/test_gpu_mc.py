"""The reference's Monte Carlo loop in LU mode on one GPU (lufqhe_mc_lu_run,
MonteCarloLoop::perform, LU-fqhe/src/MonteCarloLoop.cu:121-401): each
iteration's value is the GPU share's pair sum plus the host share's, checked
against the oracle's restated LUCoreCPU::CALL on each share."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nelem,cpus", [(16, 2000, 8), (32, 1030, 16), (7, 501, 0)])
def test_mc_lu_loop_against_oracle(lc, orc, wl, dev, n, nelem, cpus):
    A = wl.gaussian_np(nelem, n, seed=n + nelem)
    ng, nc = lc.mc_split(nelem, cpus)
    res, st = lc.mc_lu_run(A, cpu_count=cpus, iterations=3)
    assert (st["gpu_matrices"], st["cpu_matrices"]) == (ng, nc)
    assert st["wall_s"] > 0 and (ng == 0 or st["gpu_s"] > 0)
    # every iteration reruns the same work: identical values
    assert res[0] == res[1] == res[2]
    # host share: the product Crout, bit for bit the oracle's pair sum
    cpu = orc.crout_ref_pairsum(A[ng:]) if nc >= 2 else 0j
    # GPU share: fused arithmetic, so against the oracle within rounding of
    # the sum of the pair products' magnitudes
    gpu = orc.crout_ref_pairsum(A[:ng])
    d = np.linalg.det(A)
    mag = np.sum(np.abs(d[0:ng:2] * d[1:ng:2]))
    assert abs(res[0] - cpu - gpu) <= 1e-12 * mag, (res[0], cpu + gpu, mag)
    if nc >= 2:
        assert res[0] - gpu == pytest.approx(cpu, rel=1e-9, abs=1e-12 * mag)
