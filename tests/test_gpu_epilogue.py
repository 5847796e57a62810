"""GPU parity of the reference's determinant epilogues and of the
reference-shaped C++ API (LUCoreGPU::CALL / globalReduction).

  * in-place write-back (WRITE_DET + PERFORM_PROD): bit-exact against the
    oracle mirror's determinants combined with the reference's product formula;
  * the wavefunction sum: the plain complex sum against a sum in exact
    rational arithmetic, the scaled (mant, exp) sum likewise, including sums
    whose terms overflow float64 (FQHE n = 48).
"""
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest
import torch

from conftest import bits

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "cpp" / "lucore_gpu_driver"


def plain_det(mant, exp):
    return np.ldexp(mant.real, exp) + 1j * np.ldexp(mant.imag, exp)


def ref_product(a, b):
    """LU-fqhe/src/LUCoreCPU.cu:88-89, in float64 without contraction."""
    x = a.real * b.real - a.imag * b.imag
    y = a.imag * b.real + a.real * b.imag
    return x + 1j * y


def exact_sum(vals):
    """Exact rational sum of float64 complex values -> (re, im) as Fractions."""
    re = sum((Fraction(float(v.real)) for v in vals), Fraction(0))
    im = sum((Fraction(float(v.imag)) for v in vals), Fraction(0))
    return re, im


def expected_inplace(orc, A):
    o = orc.mirror(A, "crout_ref", "fused", pivots=False)
    d = plain_det(o["mant"], o["exp"])
    want = d.copy()
    pairs = len(A) // 2
    want[0:2 * pairs:2] = ref_product(d[0:2 * pairs:2], d[1:2 * pairs:2])
    return want


@pytest.mark.parametrize("n,count", [(5, 9), (16, 40), (33, 14)])
def test_det_inplace_matches_reference_contract(lc, orc, wl, dev, n, count):
    A_np = wl.gaussian_np(count, n, seed=300 + n)
    A = torch.from_numpy(A_np).to(dev)
    lc.det_inplace(A, pair_product=True)
    torch.cuda.synchronize()
    got = A.cpu().numpy()
    want = expected_inplace(orc, A_np)
    assert np.array_equal(bits(got[:, 0, 0]), bits(want))
    rest = np.ones_like(got, bool)
    rest[:, 0, 0] = False
    assert np.array_equal(bits(got[rest]), bits(A_np[rest]))  # nothing else is written


def test_det_inplace_without_pair_product_and_workspace(lc, orc, wl, dev):
    A_np = wl.gaussian_np(21, 12, seed=7)
    A = torch.from_numpy(A_np).to(dev)
    ws = torch.empty(lc.load().lufqhe_det_inplace_workspace_bytes(21), dtype=torch.uint8, device=dev)
    lc.det_inplace(A, pair_product=False, workspace=ws)
    torch.cuda.synchronize()
    o = orc.mirror(A_np, "crout_ref", "fused", pivots=False)
    assert np.array_equal(bits(A.cpu().numpy()[:, 0, 0]), bits(plain_det(o["mant"], o["exp"])))


@pytest.mark.parametrize("count", [0, 1, 2, 7, 1000, 300001])
def test_global_reduction_exact_sum(lc, dev, count):
    n = 3
    rng = np.random.default_rng(count)
    A_np = rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n))
    A = torch.from_numpy(A_np).to(dev)
    s = lc.global_reduction(A)
    torch.cuda.synchronize()
    got = complex(s.cpu().numpy()[0])
    terms = A_np[0:2 * (count // 2):2, 0, 0]
    re, im = exact_sum(terms)
    scale = float(np.sum(np.abs(terms))) if len(terms) else 0.0
    assert abs(got.real - float(re)) <= 1e-15 * max(scale, 1e-300) * max(1, np.log2(max(count, 2)))
    assert abs(got.imag - float(im)) <= 1e-15 * max(scale, 1e-300) * max(1, np.log2(max(count, 2)))
    # deterministic: the same bits every time
    s2 = lc.global_reduction(A)
    torch.cuda.synchronize()
    assert np.array_equal(bits(s.cpu().numpy()), bits(s2.cpu().numpy()))


@pytest.mark.parametrize("n,count", [(16, 2000), (24, 501)])
def test_pair_sum_scaled_matches_exact(lc, orc, wl, dev, n, count):
    A_np = wl.gaussian_np(count, n, seed=11 + n)
    r = lc.det_batched(torch.from_numpy(A_np).to(dev))
    m, e = lc.pair_sum(r)
    torch.cuda.synchronize()
    mant = r.mant.cpu().numpy()
    ex = r.exp.cpu().numpy().astype(np.int64)
    pairs = count // 2
    re = Fraction(0)
    im = Fraction(0)
    tot = 0.0
    for i in range(pairs):
        a, b = mant[2 * i], mant[2 * i + 1]
        p = ref_product(a, b)
        f = Fraction(2) ** int(ex[2 * i] + ex[2 * i + 1])
        re += Fraction(float(p.real)) * f
        im += Fraction(float(p.imag)) * f
        tot += abs(p) * float(f)
    got_m = complex(m.cpu().numpy()[0])
    got_e = int(e.cpu().numpy()[0])
    g = Fraction(2) ** got_e
    assert abs(Fraction(got_m.real) * g - re) <= Fraction(1e-13) * Fraction(tot)
    assert abs(Fraction(got_m.imag) * g - im) <= Fraction(1e-13) * Fraction(tot)
    # the reference's plain form agrees where it does not overflow
    o = orc.crout_ref_pairsum(A_np)
    assert abs(complex(got_m) * 2.0 ** got_e - o) <= 1e-10 * tot


def test_pair_sum_beyond_float64_range(lc, orc, wl, dev):
    """FQHE n = 48: every determinant and pair product overflows float64; the
    scaled sum keeps it, checked in exact rational arithmetic."""
    rng = np.random.default_rng(48)
    A_np = np.stack([wl.fqhe_matrix(wl.fqhe_positions(48, rng)) for _ in range(10)])
    r = lc.det_batched(torch.from_numpy(A_np).to(dev), policy="abs")
    m, e = lc.pair_sum(r)
    torch.cuda.synchronize()
    mant = r.mant.cpu().numpy()
    ex = r.exp.cpu().numpy().astype(np.int64)
    assert ex.max() > 1024  # beyond float64
    re = im = Fraction(0)
    tot = Fraction(0)
    for i in range(5):
        p = ref_product(mant[2 * i], mant[2 * i + 1])
        f = Fraction(2) ** int(ex[2 * i] + ex[2 * i + 1])
        re += Fraction(float(p.real)) * f
        im += Fraction(float(p.imag)) * f
        tot += Fraction(float(abs(p))) * f
    got_m = complex(m.cpu().numpy()[0])
    g = Fraction(2) ** int(e.cpu().numpy()[0])
    assert abs(Fraction(got_m.real) * g - re) <= Fraction(1e-13) * tot
    assert abs(Fraction(got_m.imag) * g - im) <= Fraction(1e-13) * tot


@pytest.mark.parametrize("n,count,padded", [(16, 64, 0), (20, 30, 1)])
def test_cpp_reference_api(lc, orc, wl, tmp_path, n, count, padded):
    """LUCoreGPU::CALL + globalReduction driven from C++ like the reference's
    Monte Carlo loop (padded = the __PADDED__ layout, ld = 32 for n = 20)."""
    assert DRIVER.exists(), "tests/cpp/lucore_gpu_driver is built by __graft_entry__.build()"
    A_np = wl.gaussian_np(count, n, seed=500 + n)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    A_np.astype(np.complex128).tofile(inp)
    subprocess.run([str(DRIVER), str(count), str(n), str(padded), str(inp), str(out)],
                   check=True, timeout=60)
    res = np.fromfile(out, dtype=np.complex128)
    want = expected_inplace(orc, A_np)
    assert np.array_equal(bits(res[:count]), bits(want))
    re, im = exact_sum(want[0::2])
    scale = float(np.sum(np.abs(want[0::2])))
    assert abs(res[count].real - float(re)) <= 1e-14 * scale
    assert abs(res[count].imag - float(im)) <= 1e-14 * scale
