"""GPU parity of the on-device FQHE builder.

The kernels build A_ij = z_i^j exp(-|z_i|^2/4) from particle positions in
their load phase.  Checks:
  * the fused build gives bit for bit the determinant of the materialised
    matrices (lufqhe_fqhe_build), for every launch shape of each n;
  * the materialised matrices equal the sequential-power host build to
    rounding, and the mirror oracle's determinant bit for bit;
  * the FQHE batch at full size (32768 matrices, n ~ U{10..40}) through the mixed plan
    against the analytic log|det| + i arg det (bounds in that test);
  * row permutations (the symmetrisation terms) change the phase by the
    permutation's sign only.
"""
import numpy as np
import pytest
import torch

from conftest import bits

pytestmark = pytest.mark.gpu


def positions(wl, count, n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([wl.fqhe_positions(n, rng) for _ in range(count)])


def same(a, b):
    assert np.array_equal(a.info.cpu().numpy(), b.info.cpu().numpy())
    assert np.array_equal(bits(a.mant.cpu().numpy()), bits(b.mant.cpu().numpy()))
    assert np.array_equal(a.exp.cpu().numpy(), b.exp.cpu().numpy())


def log_err(la, ph, la_ref, ph_ref):
    """|log det - log det_ref| / |log det_ref| with the phase wrapped."""
    dph = np.angle(np.exp(1j * (ph - ph_ref)))
    return np.abs((la - la_ref) + 1j * dph) / np.abs(la_ref + 1j * ph_ref)


@pytest.mark.parametrize("n", [1, 5, 10, 16, 24, 32, 33, 40, 48, 64])
def test_fused_build_equals_materialised(lc, orc, wl, dev, n):
    z = torch.from_numpy(positions(wl, 301, n, 900 + n)).to(dev)
    A = lc.fqhe_build(z)
    host = np.stack([wl.fqhe_matrix_seq(r) for r in z.cpu().numpy()])
    Ad = A.cpu().numpy()
    assert np.allclose(Ad, host, rtol=1e-14, atol=0)
    ref = lc.det_batched(A)
    o = orc.mirror(Ad, "crout_ref", "fused", pivots=False)
    assert np.array_equal(bits(ref.mant.cpu().numpy()), bits(o["mant"]))
    old = lc.get_shape("crout_ref", "fused", n)
    try:
        for sh in range(lc.shape_count(n)):
            lc.set_shape("crout_ref", "fused", n, sh)
            same(lc.fqhe_det(z), ref)
    finally:
        lc.set_shape("crout_ref", "fused", n, old)
    torch.cuda.synchronize()


@pytest.mark.parametrize("policy", ["crout_ref", "abs"])
def test_config3_analytic_log_form(lc, orc, wl, dev, policy):
    """The FQHE batch: 32768 FQHE matrices, n ~ U{10..40}, seeded; one mixed-plan graph
    launch with the FQHE source.

    The monomial matrices of 10..40 random points in a disk are ill
    conditioned: rounding the entries to FP64 alone moves log det by more
    than 1e-10 relative for a few matrices in a thousand (a long-double LU of
    the same FP64 matrices misses 1e-10 against the analytic form on 58 of
    these 32768, LAPACK's zgetrf on 97).  So the checks are: 99.5% within
    1e-10 of the analytic form; and, against the long-double LU of the same
    device-built matrices, the engine's error stays within the conditioning
    that LAPACK's own error on the same matrix shows -- at most 50x it (or
    1e-13) for 99.9% of the matrices and 700x (or 1e-12) for all (measured
    with the oracle's mirror: 36x at the 99.9th percentile, 351x at most).
    Parity is the bit-exact check against the mirror oracle on a sample of
    device-built matrices."""
    z, ns, offs = wl.fqhe_configs(32768, 10, 40, seed=3)
    plan = lc.MixedPlan(ns, offs, policy=policy, nstreams=4)
    res = lc.alloc_result(len(ns), 64, dev)
    zd = torch.from_numpy(z).to(dev)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        plan.execute(zd, res, mode="graph", stream=side, source="fqhe")
    side.synchronize()
    assert int((res.info != 0).sum()) == 0
    la = res.logabs.cpu().numpy()
    ph = res.phase.cpu().numpy()
    mant = res.mant.cpu().numpy()
    la_ref = np.empty(len(ns))
    ph_ref = np.empty(len(ns))
    for m in range(len(ns)):
        la_ref[m], ph_ref[m] = wl.fqhe_analytic(z[offs[m]:offs[m] + ns[m]])
    err = log_err(la, ph, la_ref, ph_ref)
    assert (err <= 1e-10).mean() >= 0.995, (err <= 1e-10).mean()
    assert np.median(err) < 1e-14
    # per n: device-built matrices -> error against their long-double LU,
    # bounded by LAPACK's error on the same matrices, and a bit-exact sample
    rng = np.random.default_rng(1)
    e_ld = np.empty(len(ns))
    e_lap = np.empty(len(ns))
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        zz = torch.from_numpy(np.stack([z[offs[m]:offs[m] + n] for m in idx])).to(dev)
        A = lc.fqhe_build(zz).cpu().numpy()
        t = orc.truth(A)
        s, l = np.linalg.slogdet(A)
        e_ld[idx] = log_err(la[idx], ph[idx], t["logabs"], t["phase"])
        e_lap[idx] = log_err(l, np.angle(s), t["logabs"], t["phase"])
        pick = rng.choice(len(idx), min(len(idx), 40), replace=False)
        o = orc.mirror(A[pick], policy, "fused", pivots=False)
        assert np.array_equal(bits(mant[idx[pick]]), bits(o["mant"])), n
    assert np.mean(e_ld <= np.maximum(50.0 * e_lap, 1e-13)) >= 0.999
    assert np.all(e_ld <= np.maximum(700.0 * e_lap, 1e-12)), float(np.max(e_ld / np.maximum(e_lap, 1e-16)))
    plan.close()


def test_mixed_fqhe_equals_uniform(lc, wl, dev):
    """A mixed FQHE batch gives, bit for bit, what per-size uniform fqhe_det
    calls give for the same configurations."""
    z, ns, offs = wl.fqhe_configs(3000, 1, 64, seed=11)
    zd = torch.from_numpy(z).to(dev)
    plan = lc.MixedPlan(ns, offs, nstreams=3)
    res = lc.alloc_result(len(ns), 64, dev)
    side = torch.cuda.Stream()
    for mode in ("plain", "graph", "persistent"):
        res.mant.zero_()
        with torch.cuda.stream(side):
            plan.execute(zd, res, mode=mode, stream=side, source="fqhe")
        side.synchronize()
        for n in np.unique(ns):
            idx = np.nonzero(ns == n)[0]
            zz = torch.from_numpy(np.stack([z[offs[m]:offs[m] + n] for m in idx])).to(dev)
            u = lc.fqhe_det(zz)
            sel = torch.from_numpy(idx).to(dev)
            assert torch.equal(res.mant[sel].view(torch.float64), u.mant.view(torch.float64))
            assert torch.equal(res.exp[sel], u.exp)
    plan.close()


def perm_sign(p):
    p = list(p)
    s = 1
    seen = [False] * len(p)
    for i in range(len(p)):
        if not seen[i]:
            j, c = i, 0
            while not seen[j]:
                seen[j] = True
                j = p[j]
                c += 1
            if c % 2 == 0:
                s = -s
    return s


@pytest.mark.parametrize("n", [8, 24, 48, 64])
def test_row_permuted_terms(lc, wl, dev, n):
    """k permutation terms of one shared configuration (z_stride 0): equal to
    the materialised permuted matrices bit for bit, and det(P A) = sgn(P) det(A)
    exactly -- without pivot ties partial pivoting picks the same pivot rows in
    any row order, so the same products are formed and only the sign moves."""
    rng = np.random.default_rng(40 + n)
    z = wl.fqhe_positions(n, rng)
    k = 700
    perms = np.stack([rng.permutation(n) for _ in range(k)]).astype(np.uint8)
    perms[0] = np.arange(n)
    zd = torch.from_numpy(z[None]).to(dev)
    pd = torch.from_numpy(perms).to(dev)
    fused = lc.fqhe_det(zd, row_perm=pd, count=k)
    A = lc.fqhe_build(zd, row_perm=pd, count=k)
    same(fused, lc.det_batched(A))
    signs = np.array([perm_sign(p) for p in perms])
    mant = fused.mant.cpu().numpy()
    assert np.array_equal(bits(mant), bits(np.where(signs < 0, -mant[0], mant[0])))
    assert np.all(fused.exp.cpu().numpy() == int(fused.exp[0]))
    la_ref, ph_ref = wl.fqhe_analytic(z)
    s, l = np.linalg.slogdet(A[0].cpu().numpy())
    e_lap = log_err(l, np.angle(s), la_ref, ph_ref)
    e = log_err(float(fused.logabs[0]), float(fused.phase[0]), la_ref, ph_ref)
    assert e <= max(1000.0 * e_lap, 1e-12)


def test_fqhe_wrapper_validation(lc, dev):
    z = torch.zeros(1, 8, dtype=torch.complex128, device=dev)
    with pytest.raises(ValueError):
        lc.fqhe_det(z, row_perm=torch.zeros(3, 8, dtype=torch.int64, device=dev), count=3)
    with pytest.raises(ValueError):
        lc.fqhe_det(torch.zeros(4, 8, dtype=torch.complex128, device=dev), count=3)


def test_out_of_range_permutation_entry_is_reported(lc, wl, dev):
    """A row-permutation entry >= n is an error of that matrix, not a
    clamped row: its determinant outputs are NaN with info -1000, the other
    matrices of the launch are untouched, and the materialised build writes
    that row as NaN."""
    n, count = 12, 40
    z = torch.from_numpy(positions(wl, 1, n, 4)).to(dev)
    rng = np.random.default_rng(8)
    perm = np.stack([rng.permutation(n) for _ in range(count)]).astype(np.uint8)
    good = lc.fqhe_det(z, row_perm=torch.from_numpy(perm).to(dev), count=count)
    bad_perm = perm.copy()
    bad_perm[3, 5] = n          # one past the end
    bad_perm[17, 0] = 255
    res = lc.fqhe_det(z, row_perm=torch.from_numpy(bad_perm).to(dev), count=count, pivots=True)
    torch.cuda.synchronize()
    info = res.info.cpu().numpy()
    assert info[3] == -1000 and info[17] == -1000
    keep = np.ones(count, bool)
    keep[[3, 17]] = False
    assert np.all(info[keep] == 0)
    mant = res.mant.cpu().numpy()
    assert np.isnan(mant[[3, 17]]).all() and np.isnan(res.logabs.cpu().numpy()[[3, 17]]).all()
    assert np.array_equal(bits(mant[keep]), bits(good.mant.cpu().numpy()[keep]))
    assert np.all(res.piv.cpu().numpy()[[3, 17]] == -1)
    A = lc.fqhe_build(z.expand(count, n).contiguous(), row_perm=torch.from_numpy(bad_perm).to(dev))
    A = A.cpu().numpy()
    assert np.isnan(A[3, 5]).all() and np.isnan(A[17, 0]).all()
    assert np.isfinite(A[keep]).all() and np.isfinite(np.delete(A[3], 5, axis=0)).all()
