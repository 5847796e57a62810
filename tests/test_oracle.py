"""Pins the CPU oracle before it is trusted as the checker.

Each restatement is checked against values fixed independently of it:
mpmath determinants of the committed fixtures (60 digits), closed forms,
brute-force Leibniz expansion, LAPACK (numpy) and the analytic FQHE
(Vandermonde) form.  A dropped term, a wrong sign, a transposed operand or a
wrong pivot rule fails at least one of these.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import bits, golden


def rel_err_log(la, ph, la_ref, ph_ref):
    """|det/det_ref - 1| from log forms."""
    d = (np.asarray(la) - la_ref) + 1j * np.angle(np.exp(1j * (np.asarray(ph) - ph_ref)))
    return np.abs(np.expm1(d))


# ---------------------------------------------------------------- mpmath goldens
@pytest.mark.parametrize("fixture,tol", [("gauss", 1e-12), ("ref_fixture16", 1e-12), ("fqhe", 1e-9)])
def test_truth_and_crout_vs_mpmath(orc, fixture, tol):
    g = golden(fixture)
    for key in [k[2:] for k in g if k.startswith("A_")]:
        A = g[f"A_{key}"]
        t = orc.truth(A)
        e_truth = rel_err_log(t["logabs"], t["phase"], g[f"logabs_{key}"], g[f"phase_{key}"])
        assert np.all(t["info"] == 0)
        assert e_truth.max() < 1e-15 if fixture != "fqhe" else e_truth.max() < 1e-12, (key, e_truth)
        c = orc.crout_ref(A)
        la, ph = orc.log_form(c["mant"], c["exp"])
        e = rel_err_log(la, ph, g[f"logabs_{key}"], g[f"phase_{key}"])
        assert e.max() < tol, (fixture, key, e.max())


def test_near_singular_vs_mpmath(orc):
    g = golden("near_singular")
    for key in ("8", "16", "32"):
        t = orc.truth(g[f"A_{key}"])
        e = rel_err_log(t["logabs"], t["phase"], g[f"logabs_{key}"], g[f"phase_{key}"])
        # cond ~ 1e8: long double keeps ~1e-19 * cond
        assert e.max() < 1e-9, (key, e)


def test_fqhe_analytic_form_matches_mpmath(wl):
    g = golden("fqhe")
    for key in ("10", "16", "24", "32", "40"):
        for z, la_ref, ph_ref in zip(g[f"z_{key}"], g[f"logabs_{key}"], g[f"phase_{key}"]):
            la, ph = wl.fqhe_analytic(z)
            assert abs(la - la_ref) < 1e-11 * max(1.0, abs(la_ref))
            assert abs(np.angle(np.exp(1j * (ph - ph_ref)))) < 1e-9


def test_reference_fixture_complex64(orc):
    """The reference's own input file at the reference's own precision."""
    g = golden("ref_fixture16")
    c = orc.crout_ref_c64(g["A32_16"])
    assert np.all(c["info"] == 0)
    ref = np.exp(g["logabs_16"] + 1j * g["phase_16"])
    err = np.abs(c["det"].astype(np.complex128) / ref - 1)
    assert err.max() < 1e-3  # FP32 Crout on 16x16 inputs in (0,1)
    # the FP64 restatement on the same (exactly widened) inputs
    d = orc.crout_ref(g["A32_16"].astype(np.complex128))
    assert np.all(np.abs(d["det"] / ref - 1) < 1e-12)


# ---------------------------------------------------------------- closed forms
def _leibniz(A):
    n = A.shape[0]
    tot = 0j
    for p in itertools.permutations(range(n)):
        sgn = 1
        for i in range(n):
            for j in range(i + 1, n):
                if p[i] > p[j]:
                    sgn = -sgn
        prod = 1 + 0j
        for i in range(n):
            prod *= complex(A[i, p[i]])
        tot += sgn * prod
    return tot


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_leibniz_brute_force(orc, n):
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((8, n, n)) + 1j * rng.standard_normal((8, n, n))
    ref = np.array([_leibniz(a) for a in A])
    for out in (orc.crout_ref(A), orc.mirror(A, "crout_ref", "fused"), orc.mirror(A, "abs", "fused")):
        d = out["mant"] * np.exp2(out["exp"].astype(float))
        assert np.all(np.abs(d / ref - 1) < 1e-12), n


def test_triangular_diagonal_and_permutation(orc):
    rng = np.random.default_rng(7)
    n = 12
    U = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    diag = np.prod(np.diag(U))
    P = np.eye(n)[rng.permutation(n)]
    sgnP = np.linalg.det(P).real
    for M, want in ((U, diag), (U.T.copy(), diag), (P, sgnP), (P @ U, sgnP * diag)):
        for out in (orc.crout_ref(M), orc.mirror(M, "abs", "fused"), orc.mirror(M, "crout_ref", "fused")):
            d = out["mant"][0] * 2.0 ** out["exp"][0]
            assert abs(d / want - 1) < 1e-13


def test_row_permutation_sign_and_scaling(orc):
    """det(PA) = sgn(P) det(A); det(alpha A) = alpha^n det(A)."""
    rng = np.random.default_rng(8)
    n = 20
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    base = orc.truth(A)
    perm = rng.permutation(n)
    sgn = np.linalg.det(np.eye(n)[perm]).real
    alpha = 0.3 - 1.7j
    for M, dla, dph in ((A[perm], 0.0, 0.0 if sgn > 0 else np.pi),
                        (alpha * A, n * np.log(abs(alpha)), n * np.angle(alpha))):
        for out in (orc.crout_ref(M), orc.mirror(M, "abs", "fused")):
            la, ph = orc.log_form(out["mant"], out["exp"])
            e = rel_err_log(la, ph, base["logabs"] + dla, base["phase"] + dph)
            assert e.max() < 1e-12


def test_block_diagonal_product(orc):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    B = rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))
    M = np.zeros((16, 16), complex)
    M[:7, :7] = A
    M[7:, 7:] = B
    want = np.linalg.det(A) * np.linalg.det(B)
    out = orc.crout_ref(M)
    assert abs(out["mant"][0] * 2.0 ** out["exp"][0] / want - 1) < 1e-13


def test_lapack_cross_check_and_mirror_identity(orc, wl):
    for n in (1, 2, 3, 7, 16, 33, 64):
        A = wl.gaussian_np(64, n, seed=n)
        ref = np.linalg.det(A)
        cr = orc.crout_ref(A, pivots=True)
        assert np.all(cr["info"] == 0)
        assert np.all(np.abs(cr["det"] / ref - 1) < 1e-9)
        # the plain product equals mant * 2^exp bit for bit when it is normal
        assert np.array_equal(bits(cr["det"]), bits(np.ldexp(cr["mant"].real, cr["exp"]) + 1j * np.ldexp(cr["mant"].imag, cr["exp"])))
        # right-looking mirror in the reference's arithmetic == restated Crout
        mr = orc.mirror(A, "crout_ref", "ref")
        assert np.array_equal(bits(mr["mant"]), bits(cr["mant"]))
        assert np.array_equal(mr["exp"], cr["exp"])
        assert np.array_equal(mr["piv"], cr["piv"])


def test_pivot_rule_is_scaled_real_part_with_last_tie(orc):
    """Hand-checked pivots: |Re|/rowscale decides, exact ties go to the last row."""
    # row scales: row0 max|Re| = 4, row1 = 1, row2 = 2
    A = np.array([[2 + 0j, 4, 0], [1 + 9j, 0, 0], [1, 2, 0.5]], complex)
    # column 0 criteria: 2/4 = 0.5, 1/1 = 1 (large imaginary part ignored), 1/2 = 0.5
    out = orc.crout_ref(A, pivots=True)
    assert out["piv"][0, 0] == 1
    # exact tie: both criteria 1 -> the last row wins
    B = np.array([[1 + 0j, 0.5], [2, 1]], complex)
    out = orc.crout_ref(B, pivots=True)
    assert out["piv"][0, 0] == 1
    # ABS policy uses |Re| + |Im| (LAPACK's cabs1): row 1 (1 + 9) wins column 0 of A
    m = orc.mirror(A, "abs", "fused")
    assert m["piv"][0, 0] == 1
    m2 = orc.mirror(np.array([[3 + 0j, 1], [0 + 3j, 1]]), "abs", "fused")
    assert m2["piv"][0, 0] == 0  # |3| + 0 == 0 + |3|: tie to the first row, as izamax
    m3 = orc.mirror(np.array([[2 + 2j, 1], [0 + 3j, 1]]), "abs", "fused")
    assert m3["piv"][0, 0] == 0  # 4 > 3 although |2+2i| < |3i|: cabs1, not the modulus


def test_abs_pivots_equal_lapack_zgetrf(orc, wl):
    """The ABS policy is LAPACK's partial pivoting: scipy's lu_factor (zgetrf,
    izamax over cabs1, first index on ties) interchanges the same rows at every
    step.  N(0,1) matrices agree at every step; small-integer matrices, whose
    first column is full of exact cabs1 ties, agree at step 0 every time (the
    later steps of a few of them differ because zgetrf's blocked elimination
    rounds the updates differently at ties)."""
    from scipy.linalg import lu_factor
    for n, count in ((5, 600), (16, 400), (32, 200), (64, 60)):
        A = wl.gaussian_np(count, n, seed=900 + n)
        piv = orc.mirror(A, "abs", "fused")["piv"]
        lap = np.stack([lu_factor(a, check_finite=False)[1] for a in A])
        assert np.array_equal(piv, lap), n
    rng = np.random.default_rng(5)
    A = (rng.integers(-2, 3, (1000, 8, 8)) + 1j * rng.integers(-2, 3, (1000, 8, 8))).astype(complex)
    piv = orc.mirror(A, "abs", "fused")["piv"]
    lap = np.stack([lu_factor(a, check_finite=False)[1] for a in A])
    assert np.array_equal(piv[:, 0], lap[:, 0])
    assert np.mean(np.all(piv == lap, axis=1)) > 0.98


def test_fused_criterion_keeps_non_power_of_two_ties(orc):
    """Both rows of column 0 have the criterion 1 exactly (their max |Re| is in
    column 0), so the reference's `>=` picks the last row.  49 * RN(1/49)
    rounds to 1 - 2^-53, so a product criterion would pick row 0; the fused
    mirror's corrected quotient keeps the tie."""
    assert 49.0 * (1.0 / 49.0) < 1.0
    A = np.array([[3 + 0j, 2], [49 + 1j, 1]], complex)
    assert orc.crout_ref(A, pivots=True)["piv"][0, 0] == 1
    assert orc.mirror(A, "crout_ref", "fused")["piv"][0, 0] == 1
    # the same tie inside a larger matrix, at several non-power-of-two scales
    rng = np.random.default_rng(12)
    for x in (49.0, 0.1, 3.7e-5, 98.3, 1e20 / 3):
        B = rng.standard_normal((9, 9)) * 0.01 + 0j
        B[:, 1] = 1.0     # every other row's largest real part is off column 0
        B[2, 0] = x
        B[2, 1:] = x / 7
        B[6, 0] = 7.0
        B[6, 1:] = 1.0
        cr = orc.crout_ref(B, pivots=True)
        assert cr["piv"][0, 0] == 6
        assert np.array_equal(orc.mirror(B, "crout_ref", "fused")["piv"], cr["piv"])


def _hard_quotients(count, seed):
    """(x, s) in [1, 2) whose quotient x/s lies within 6 * 2^-106 of a
    rounding midpoint: B*M = A*2^54 - d for small d, M odd (the hardest cases
    for correct rounding of a division)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        M = int(rng.integers(2 ** 53, 2 ** 54)) | 1
        d = int(rng.integers(1, 7)) * (1 if rng.random() < 0.5 else -1)
        B = (-d * pow(M, -1, 2 ** 54)) % 2 ** 54
        if not 2 ** 52 <= B < 2 ** 53:
            continue
        A, rem = divmod(B * M + d, 2 ** 54)
        assert rem == 0
        if A >= 2 ** 53:
            continue
        out.append((A * 2.0 ** -52, B * 2.0 ** -52))
    return out


def test_corrected_quotient_is_the_division(orc):
    """The fused criterion's corrected product equals the IEEE division on
    the hardest-to-round quotients (where the uncorrected product is wrong
    about half the time), on random quotients over many binades, at the
    edges of its fast range and on its fallback inputs."""
    hard = _hard_quotients(4000, 3)
    wrong_product = 0
    for x, s in hard:
        assert orc.quot_rn(x, s) == x / s, (x.hex(), s.hex())
        wrong_product += (x * (1.0 / s)) != x / s
    assert wrong_product > 1000
    rng = np.random.default_rng(4)
    xs = np.ldexp(rng.random(20000) + 0.5, rng.integers(-300, 300, 20000))
    ss = np.ldexp(rng.random(20000) + 0.5, rng.integers(-80, 300, 20000))
    for x, s in zip(xs.tolist(), ss.tolist()):
        assert orc.quot_rn(x, s) == x / s
    for x, s in ((0.0, 3.0), (5e-324, 1e-24), (2.0 ** -900, 1.0), (2.0 ** 1000, 2.0 ** -70),
                 (1.5 * 2.0 ** 1023, 2.0 ** 1000), (3.0, 2.0 ** 1020), (math.inf, 2.0),
                 (2.0 ** -879, 1.0), (2.0 ** -881, 1.0), (2.0 ** 959, 1.0), (2.0 ** 961, 1.0)):
        assert orc.quot_rn(x, s) == x / s, (x, s)
    assert math.isnan(orc.quot_rn(math.nan, 2.0))


@pytest.mark.parametrize("n,count", [(16, 4096), (32, 65536)])
def test_fused_pivots_equal_reference_crout(orc, wl, n, count):
    """The fused mirror (the GPU kernel's operation order) picks the
    restated reference Crout's pivot at every step of every matrix of the
    4096 x 16x16 and 65536 x 32x32 batches, exact criterion ties included:
    only the arithmetic of the eliminations differs, and no step comes within
    1e-12 of a tie between different values."""
    A = wl.gaussian_np(count, n, seed=77 + n)
    r = orc.pivot_parity(orc.mirror(A, "crout_ref", "fused")["piv"], orc.crout_ref_gaps(A), 1e-12)
    assert r["mismatches"] == 0 and r["differing"] == 0, r
    assert r["matrices_with_exact_ties"] > count // 8   # ties at 1.0 in column 0 are common


def test_fused_pivots_equal_reference_crout_mixed_and_fqhe(orc, wl):
    """The same on a mixed n = 8..64 sample with 5% near-singular matrices
    and on FQHE matrices (whose column 0 is exp(-|z|^2/4), often the row's
    largest real part)."""
    flat, ns, offs, _ = wl.mixed_batch(3000, 2024)
    flat = flat.numpy() if hasattr(flat, "numpy") else flat
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        A = np.stack([flat[offs[i]:offs[i] + n * n].reshape(n, n) for i in idx])
        r = orc.pivot_parity(orc.mirror(A, "crout_ref", "fused")["piv"], orc.crout_ref_gaps(A),
                             1e-12)
        assert r["mismatches"] == 0, (int(n), r)
    rng = np.random.default_rng(31)
    for n in (10, 20, 33, 40):
        A = np.stack([wl.fqhe_matrix(wl.fqhe_positions(n, rng)) for _ in range(200)])
        r = orc.pivot_parity(orc.mirror(A, "crout_ref", "fused")["piv"], orc.crout_ref_gaps(A),
                             1e-12)
        assert r["mismatches"] == 0, (n, r)


@pytest.mark.parametrize("policy,ks", [("crout_ref", (60, 400, 515, 600, 900, 1000)),
                                       ("abs", (-900, -600, -515, 60, 515, 600, 900, 1000))])
def test_power_of_two_scaling_is_exact(orc, wl, policy, ks):
    """det(2^k A) = 2^(k n) det(A) exactly: the fused mirror returns the same
    mantissa bits and pivots, and the exponent moves by k*n, including where
    |p|^2 of an unscaled pivot would overflow (k >= 512) or underflow
    (k <= -512) and where the row scale leaves the corrected quotient's fast
    range (k = 1000).  (CROUT_REF rejects rows of max |Re| < 1e-24, so it
    takes no negative k.)"""
    for n in (3, 16, 33, 64):
        base = wl.gaussian_np(16, n, seed=500 + n)
        b = orc.mirror(base, policy, "fused")
        assert np.all(b["info"] == 0)
        for k in ks:
            m = orc.mirror(base * 2.0 ** k, policy, "fused")
            assert np.all(m["info"] == 0), (n, k)
            assert np.array_equal(bits(m["mant"]), bits(b["mant"])), (n, k)
            assert np.array_equal(m["exp"], b["exp"] + k * n), (n, k)
            assert np.array_equal(m["piv"], b["piv"]), (n, k)


def test_status_codes(orc):
    A = np.eye(4, dtype=complex)
    A[2] = 1j * np.ones(4)  # row 2 has no real part: the reference exit(0)s
    out = orc.crout_ref(A)
    assert out["info"][0] == -3 and np.isnan(out["mant"][0])
    m = orc.mirror(A, "crout_ref", "fused")
    assert m["info"][0] == -3
    m = orc.mirror(A, "abs", "fused")
    assert m["info"][0] == 0
    Z = np.ones((5, 5), complex)
    Z[:, 0] = 0  # zero first column: zero pivot at step 0
    for out in (orc.crout_ref(Z + 0), orc.mirror(Z, "abs", "fused"), orc.mirror(Z, "crout_ref", "ref")):
        assert out["info"][0] == 1 and out["mant"][0] == 0


def test_fp64_overflow_carried_in_exponent(orc, wl):
    """FQHE determinants pass the float64 range near n = 33; mant/exp keep them."""
    rng = np.random.default_rng(5)
    z = wl.fqhe_positions(64, rng)
    A = wl.fqhe_matrix(z)
    la, ph = wl.fqhe_analytic(z)
    assert la > 710  # beyond float64
    t = orc.truth(A)
    assert abs(t["logabs"][0] - la) < 1e-9 * la
    for pol in ("abs", "crout_ref"):
        m = orc.mirror(A, pol, "fused")
        l2, p2 = orc.log_form(m["mant"], m["exp"])
        assert abs(l2[0] - la) < 1e-8 * la, pol


def test_pairsum_matches_products(orc, wl):
    A = wl.gaussian_np(10, 6, seed=3)
    d = np.linalg.det(A)
    want = np.sum(d[0::2] * d[1::2])
    assert abs(orc.crout_ref_pairsum(A) / want - 1) < 1e-12
