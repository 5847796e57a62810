"""GPU parity of the batched LUP-determinant kernels (through the C ABI).

Bar: the fused kernels equal the oracle's mirror of their operation order bit
for bit (mantissa, exponent, pivot sequence, status); the reference-arithmetic
kernels equal the restated reference CPU Crout (LU-fqhe/src/LUCoreCPU.cu:
112-267) bit for bit; log|det| and phase agree with the mantissa/exponent to a
few ulp; accuracy against the long-double truth is reported and bounded.
"""
import numpy as np
import pytest
import torch

from conftest import bits, golden

pytestmark = pytest.mark.gpu

ALL_N = list(range(1, 65))


def run(lc, A_np, dev, policy="crout_ref", arith="fused", pivots=True):
    A = torch.from_numpy(np.ascontiguousarray(A_np)).to(dev)
    r = lc.det_batched(A, policy=policy, arith=arith, pivots=pivots)
    torch.cuda.synchronize()
    return dict(mant=r.mant.cpu().numpy(), exp=r.exp.cpu().numpy(), logabs=r.logabs.cpu().numpy(),
                phase=r.phase.cpu().numpy(), info=r.info.cpu().numpy(),
                piv=None if r.piv is None else r.piv.cpu().numpy().astype(np.int32))


def assert_same(g, o, n, what):
    assert np.array_equal(g["info"], o["info"]), (what, n)
    assert np.array_equal(bits(g["mant"]), bits(o["mant"])), (what, n)
    assert np.array_equal(g["exp"], o["exp"]), (what, n)
    if o.get("piv") is not None:
        assert np.array_equal(g["piv"], o["piv"]), (what, n)


def check_log_outputs(orc, g):
    la, ph = orc.log_form(g["mant"], g["exp"])
    ok = g["info"] >= 0
    fin = np.isfinite(la) & ok
    assert np.all(np.abs(g["logabs"][fin] - la[fin]) <= 4e-16 * np.maximum(1.0, np.abs(la[fin])))
    assert np.all(np.abs(np.angle(np.exp(1j * (g["phase"][fin] - ph[fin])))) <= 1e-15)


@pytest.mark.parametrize("n", ALL_N)
def test_fused_bit_exact_vs_mirror(lc, orc, wl, dev, n):
    count = 37 if n > 8 else 61  # not a multiple of the matrices per wave
    A = wl.gaussian_np(count, n, seed=1000 + n)
    for policy in ("crout_ref", "abs"):
        g = run(lc, A, dev, policy, "fused")
        o = orc.mirror(A, policy, "fused")
        assert_same(g, o, n, policy)
        check_log_outputs(orc, g)


@pytest.mark.parametrize("n", ALL_N)
def test_reference_arithmetic_bit_exact_vs_crout(lc, orc, wl, dev, n):
    A = wl.gaussian_np(29, n, seed=2000 + n)
    g = run(lc, A, dev, "crout_ref", "ref")
    o = orc.crout_ref(A, pivots=True)
    assert_same(g, o, n, "ref")
    # the plain reference det is the same number whenever it is representable
    det = np.ldexp(g["mant"].real, g["exp"]) + 1j * np.ldexp(g["mant"].imag, g["exp"])
    assert np.array_equal(bits(det), bits(o["det"]))


@pytest.mark.parametrize("fixture", ["gauss", "near_singular", "fqhe", "ref_fixture16"])
def test_golden_fixtures(lc, orc, dev, fixture):
    gd = golden(fixture)
    for key in [k[2:] for k in gd if k.startswith("A_")]:
        A = gd[f"A_{key}"]
        for policy in ("crout_ref", "abs"):
            g = run(lc, A, dev, policy, "fused")
            assert_same(g, orc.mirror(A, policy, "fused"), key, policy)
            d = (g["logabs"] - gd[f"logabs_{key}"]) + 1j * (g["phase"] - gd[f"phase_{key}"])
            err = np.abs(np.expm1(d))
            tol = {"gauss": 1e-12, "ref_fixture16": 1e-12, "near_singular": 1e-5, "fqhe": 1e-9}[fixture]
            assert err.max() < tol, (fixture, key, policy, err.max())


def test_reference_fixture_bit_exact(lc, orc, dev):
    """The reference's own input file: GPU (reference arithmetic) == restated
    LUCoreCPU::determinant at FP64, including the pivot sequences."""
    gd = golden("ref_fixture16")
    A = gd["A32_16"].astype(np.complex128)
    g = run(lc, A, dev, "crout_ref", "ref")
    assert_same(g, orc.crout_ref(A, pivots=True), 16, "ref-fixture")


def test_strided_and_padded_layout(lc, orc, wl, dev):
    """ld > n (padded rows, LU-fqhe/src/RunInfo.cpp:72-76) and stride > n*ld."""
    n, ld, count = 20, 32, 50
    stride = n * ld + 17
    A = wl.gaussian_np(count, n, seed=5)
    buf = torch.full((count * stride + ld,), complex(np.nan, np.nan), dtype=torch.complex128)
    for m in range(count):
        for r in range(n):
            buf[m * stride + r * ld: m * stride + r * ld + n] = torch.from_numpy(A[m, r])
    buf = buf.to(dev)
    view = torch.as_strided(buf, (count, n, n), (stride, ld, 1))
    res = lc.det_batched(view, pivots=True)
    torch.cuda.synchronize()
    o = orc.mirror(A, "crout_ref", "fused")
    assert np.array_equal(bits(res.mant.cpu().numpy()), bits(o["mant"]))
    assert np.array_equal(res.piv.cpu().numpy().astype(np.int32), o["piv"])


def test_status_codes_on_gpu(lc, orc, dev):
    n = 6
    A = np.tile(np.eye(n, dtype=complex), (4, 1, 1))
    A[1, 3] = 1j  # row 3 has no real part -> reference exit(0) case
    A[2] = 1.0
    A[2, :, 0] = 0  # zero column (rows keep a real part) -> zero pivot at step 0
    A[3] = 0  # all zero: rejected under CROUT_REF (row 0), zero pivot under ABS
    for policy in ("crout_ref", "abs"):
        g = run(lc, A, dev, policy, "fused")
        o = orc.mirror(A, policy, "fused")
        assert np.array_equal(g["info"], o["info"]), policy
        assert np.array_equal(g["exp"], o["exp"])
    g = run(lc, A, dev, "crout_ref", "fused")
    assert list(g["info"]) == [0, -4, 1, -1]
    assert np.isnan(g["mant"][1]) and np.isnan(g["logabs"][1])
    assert g["mant"][2] == 0 and g["logabs"][2] == -np.inf and g["phase"][2] == 0


def test_nullable_outputs(lc, dev, wl):
    lib = lc.load()
    A = torch.from_numpy(wl.gaussian_np(10, 12, seed=1)).to(dev)
    mant = torch.empty(10, dtype=torch.complex128, device=dev)
    st = lib.lufqhe_det_batched(A.data_ptr(), 10, 12, 12, 144, mant.data_ptr(), None, None,
                                None, None, None, 0, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert st == 0
    full = lc.det_batched(A)
    assert torch.equal(mant, full.mant)


def same_or_both_nan(g, o):
    """Element-wise: equal bits, or NaN on both sides (a NaN the hardware
    generates carries its own sign bit: +NaN on the GPU, -NaN on x86)."""
    g = np.ascontiguousarray(g).view(np.float64)
    o = np.ascontiguousarray(o).view(np.float64)
    return (g.view(np.uint64) == o.view(np.uint64)) | (np.isnan(g) & np.isnan(o))


@pytest.mark.parametrize("n", [5, 16, 32, 48, 64])
def test_extreme_and_non_finite_inputs(lc, orc, wl, dev, n):
    """Entries near the ends of the FP64 range and non-finite entries take
    the mirror's path: rows below the row-scale floor are rejected, pivots
    are scaled before squaring (no |p|^2 overflow), NaN criteria count as
    DBL_MAX, subnormals are kept. Status, exponent and pivots match exactly,
    the mantissa bit for bit or as NaN on both sides."""
    base = wl.gaussian_np(12, n, seed=300 + n)
    A = np.empty_like(base)
    # exact power-of-two scalings: tiny (rejected), large, |p|^2 near and past overflow
    for m, k in enumerate((-90, 60, 400, 508, 515, 600)):
        A[m] = base[m] * 2.0 ** k
    A[6] = base[6]
    A[6][1 % n, 2 % n] = np.inf                       # one infinite real part
    A[7] = base[7]
    A[7][0, 0] = complex(np.nan, 0.0)                  # NaN on the first pivot's row
    A[8] = base[8]
    A[8][n - 1, :] = complex(np.nan, np.nan)           # a whole NaN row
    A[9] = base[9]
    A[9][n // 2, 0] = complex(1.0, np.inf)             # infinite imaginary part
    A[10] = base[10]
    A[10][::2, :] *= 1e-310 / 2 ** -1022 * 2.0 ** -1030  # rows of subnormal magnitude
    A[11] = base[11] * 1e-12                           # small but above the row floor
    for policy in ("crout_ref", "abs"):
        g = run(lc, A, dev, policy, "fused")
        o = orc.mirror(A, policy, "fused")
        assert np.array_equal(g["info"], o["info"]), (policy, g["info"], o["info"])
        assert np.array_equal(g["exp"], o["exp"]), (policy, g["exp"], o["exp"])
        assert same_or_both_nan(g["mant"], o["mant"]).all(), policy
        ok = g["info"] >= 0
        assert np.array_equal(g["piv"][ok], o["piv"][ok]), policy


def test_pivot_with_underflowing_square(lc, orc, dev):
    """A pivot p = 1e-170 whose |p|^2 underflows is an ordinary pivot: the
    kernel scales p by a power of two before squaring, so the multiplier
    1e-21/p = 1e149 is formed and the determinant is the true 1e129 (the
    first row's 1 times 1e-21 times the 1e150 block), checked against the
    long-double truth; status 0, bits equal to the mirror."""
    A = np.zeros((2, 3, 3), np.complex128)
    A[0] = [[1e-170, 0.0, 1.0],
            [1e-21, 1e150, 1e-180],
            [0.0, 1e150, 3e-180]]
    A[1] = A[0] * (0.6 + 0.8j)
    for policy in ("crout_ref", "abs"):
        g = run(lc, A, dev, policy, "fused")
        o = orc.mirror(A, policy, "fused")
        assert_same(g, o, 3, policy)
        assert g["info"].tolist() == [0, 0]
        t = orc.truth(A)
        err = np.abs(np.expm1((g["logabs"] - t["logabs"])
                              + 1j * np.angle(np.exp(1j * (g["phase"] - t["phase"])))))
        assert err.max() < 1e-12, (policy, err)
    assert abs(g["logabs"][0] - np.log(1e129)) < 1e-12 * np.log(1e129)


@pytest.mark.parametrize("n", [3, 16, 33, 64])
def test_power_of_two_scaling_is_exact(lc, orc, wl, dev, n):
    """det(2^k A) = 2^(k n) det(A) exactly on the GPU: same mantissa bits and
    pivots, exponent moved by k*n -- past |p|^2 overflow (k >= 512) and
    underflow (k <= -512, ABS only: CROUT_REF rejects rows below 1e-24) and,
    at k = 1000, with every CROUT_REF criterion taken by division."""
    base = wl.gaussian_np(24, n, seed=700 + n)
    for policy, ks in (("crout_ref", (515, 600, 1000)), ("abs", (-900, -515, 515, 600, 1000))):
        b = run(lc, base, dev, policy)
        assert_same(b, orc.mirror(base, policy, "fused"), n, policy)
        for k in ks:
            g = run(lc, base * 2.0 ** k, dev, policy)
            assert np.all(g["info"] == 0), (policy, k)
            assert np.array_equal(bits(g["mant"]), bits(b["mant"])), (policy, k)
            assert np.array_equal(g["exp"], b["exp"] + k * n), (policy, k)
            assert np.array_equal(g["piv"], b["piv"]), (policy, k)
            assert np.allclose(g["logabs"], b["logabs"] + k * n * np.log(2.0), rtol=1e-15, atol=0)


def test_binding_refuses_short_buffers(lc, dev, wl):
    """The Python binding checks every buffer a launch would touch before the
    launch: short or misplaced outputs and workspaces raise, nothing runs."""
    A = torch.from_numpy(wl.gaussian_np(8, 5, seed=2)).to(dev)
    short = lc.alloc_result(7, 5, dev)
    with pytest.raises(ValueError):
        lc.det_batched(A, out=short)
    no_piv_room = lc.alloc_result(8, 4, dev, pivots=True)
    with pytest.raises(ValueError):
        lc.det_batched(A, out=no_piv_room)
    on_host = lc.alloc_result(8, 5, "cpu")
    with pytest.raises(ValueError):
        lc.det_batched(A, out=on_host)
    ws = torch.empty(8 * 20 - 1, dtype=torch.uint8, device=dev)
    with pytest.raises(ValueError):
        lc.det_inplace(A.clone(), workspace=ws)
    # outputs left as None are skipped by the library
    r = lc.alloc_result(8, 5, dev)
    r.logabs = r.phase = r.info = None
    lc.det_batched(A, out=r)
    torch.cuda.synchronize()
    assert torch.equal(r.mant, lc.det_batched(A).mant)


def test_mixed_plan_refuses_short_flat_buffer(lc, dev):
    ns = np.array([3, 4], np.int32)
    offs = np.array([0, 9], np.int64)
    plan = lc.MixedPlan(ns, offs, nstreams=1)
    try:
        res = lc.alloc_result(2, 1, dev)
        with pytest.raises(ValueError):
            plan.execute(torch.zeros(24, dtype=torch.complex128, device=dev), res, mode="plain")
        with pytest.raises(ValueError):
            plan.execute(torch.zeros(25, dtype=torch.complex128, device=dev),
                         lc.alloc_result(1, 1, dev), mode="plain")
        # positions need only offs + n elements
        plan.execute(torch.ones(13, dtype=torch.complex128, device=dev), res, mode="plain",
                     source="fqhe")
        flat = torch.zeros(25, dtype=torch.complex128, device=dev)
        flat[0] = flat[4] = flat[8] = 1.0
        flat[9] = flat[14] = flat[19] = flat[24] = 2.0
        plan.execute(flat, res, mode="plain")
        torch.cuda.synchronize()
        assert res.det().cpu().numpy().tolist() == [1.0, 16.0]
    finally:
        plan.close()


@pytest.mark.parametrize("n,count,seed", [(16, 4096, 11), (32, 65536, 20250101)])
def test_full_size_pivots_equal_reference_crout(lc, orc, wl, dev, n, count, seed):
    """Pivot parity with the reference CPU Crout (LU-fqhe/src/LUCoreCPU.cu:
    188-216) on the 4096 x 16x16 and 65536 x 32x32 batches the benchmark
    times (same seeds), in the production mode: every pivot of every matrix
    equals the restated Crout's, exact ties included, and none of them is an
    excluded near-tie (a first difference at a step whose criteria lie
    within 1e-12 of each other).  |det/det_crout - 1| stays within 1e-12
    for all but a handful of ill-conditioned matrices."""
    A = wl.gaussian(count, n, seed=seed, device=dev)
    r = lc.det_batched(A, pivots=True)
    torch.cuda.synchronize()
    A_h = A.cpu().numpy()
    ref = orc.crout_ref_gaps(A_h)
    par = orc.pivot_parity(r.piv.cpu().numpy(), ref, 1e-12)
    print(f"n={n}: {par}")
    assert par["mismatches"] == 0 and par["differing"] == 0, par
    assert par["matrices_with_exact_ties"] > count // 8
    cr = orc.crout_ref(A_h)
    la, ph = orc.log_form(cr["mant"], cr["exp"])
    e = np.abs(np.expm1((r.logabs.cpu().numpy() - la)
                        + 1j * np.angle(np.exp(1j * (r.phase.cpu().numpy() - ph)))))
    assert np.median(e) < 1e-14 and int((e > 1e-12).sum()) <= count // 10000 + 2, (
        np.median(e), e.max(), int((e > 1e-12).sum()))


@pytest.mark.parametrize("n,count", [(8, 2048), (16, 2048), (32, 1024), (48, 256), (64, 256)])
def test_abs_pivots_equal_lapack(lc, wl, dev, n, count):
    """Under the ABS policy the kernels interchange the rows LAPACK's zgetrf
    does (scipy's lu_factor: izamax over cabs1, first index on ties) at every
    step of N(0,1) matrices."""
    from scipy.linalg import lu_factor
    A = wl.gaussian(count, n, seed=300 + n, device=dev)
    r = lc.det_batched(A, pivots=True, policy="abs")
    torch.cuda.synchronize()
    piv = r.piv.cpu().numpy().astype(np.int64)
    lap = np.stack([lu_factor(a, check_finite=False)[1] for a in A.cpu().numpy()])
    bad = np.nonzero(np.any(piv != lap, axis=1))[0]
    assert len(bad) == 0, (n, bad[:8])


def test_mixed_sample_pivots_equal_reference_crout(lc, orc, wl, dev):
    """The same pivot parity on a sample of the mixed n = 8..64 batch with 5%
    near-singular matrices (each size through its tuned kernel, which the
    mixed plans launch)."""
    flat, ns, offs, _ = wl.mixed_batch(6000, 2024)
    flat = flat.numpy() if hasattr(flat, "numpy") else flat
    tot = 0
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([flat[offs[i]:offs[i] + n * n].reshape(n, n) for i in idx])
        g = run(lc, A, dev)
        par = orc.pivot_parity(g["piv"], orc.crout_ref_gaps(A), 1e-12)
        assert par["mismatches"] == 0, (int(n), par)
        tot += par["compared"]
    assert tot == 6000


@pytest.mark.parametrize("n,count", [(16, 4096), (32, 65536)])
def test_full_size_uniform_configs(lc, orc, wl, dev, n, count):
    """The 4096 x 16x16 and 65536 x 32x32 batches at full size: every status
    clean, a random sample bit-exact against the mirror, all of them within tolerance of the
    independent rocSOLVER/MAGMA slogdet, and the accuracy against the
    long-double truth reported on a sample."""
    A = wl.gaussian(count, n, seed=77 + n, device=dev)
    r = lc.det_batched(A, pivots=False)
    torch.cuda.synchronize()
    assert int((r.info != 0).sum()) == 0
    sample = torch.randperm(count, generator=torch.Generator().manual_seed(1))[:512]
    As = A[sample.to(dev)].cpu().numpy()
    o = orc.mirror(As, "crout_ref", "fused", pivots=False)
    assert np.array_equal(bits(r.mant[sample.to(dev)].cpu().numpy()), bits(o["mant"]))
    sign, ld = torch.linalg.slogdet(A)
    dla = (r.logabs - ld).abs()
    dph = torch.angle(torch.exp(1j * (r.phase - torch.angle(sign))))
    err = (dla + dph.abs())
    assert float(torch.quantile(err, 0.5)) < 1e-13
    assert float(err.max()) < 1e-8
    t = orc.truth(As)
    e = np.abs(np.expm1((r.logabs[sample.to(dev)].cpu().numpy() - t["logabs"])
                        + 1j * np.angle(np.exp(1j * (r.phase[sample.to(dev)].cpu().numpy() - t["phase"])))))
    assert np.median(e) < 1e-13 and np.quantile(e, 0.99) < 1e-11


@pytest.mark.parametrize("n", ALL_N)
def test_every_launch_shape_bit_exact(lc, orc, wl, dev, n):
    """Every compiled candidate shape (the autotuner may pick any) gives the
    oracle's determinants and pivots bit for bit."""
    A = wl.gaussian_np(45, n, seed=3000 + n)
    for policy in ("crout_ref", "abs"):
        o = orc.mirror(A, policy, "fused")
        keep = lc.get_shape(policy, "fused", n)
        try:
            for shape in range(lc.shape_count(n)):
                assert lc.set_shape(policy, "fused", n, shape)
                assert_same(run(lc, A, dev, policy, "fused"), o, n, (policy, shape))
        finally:
            lc.set_shape(policy, "fused", n, keep)
