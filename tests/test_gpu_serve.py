"""GPU tests of the walker server (lufqhe_server_*).

* every request's result is bit-identical served alone (one request per
  batch, one batch in flight) and in the mix (large batches, several in
  flight, Poisson arrivals, a tight HBM budget);
* each request's psi' equals the sum of its k permutation-term determinants
  recomputed through lufqhe_fqhe_det from the walker's reconstructed
  trajectory, to 1e-12 of sum |det_t|;
* the Metropolis bookkeeping leaves each walker at its last accepted
  configuration.
"""
import math

import numpy as np
import pytest
import torch

from conftest import bits

pytestmark = pytest.mark.gpu

SIZES = [8, 16, 24, 32, 48, 64]


def make_server(lc, walkers=24, perms=300, seed=5):
    wn = [SIZES[w % len(SIZES)] for w in range(walkers)]
    return lc.Server(wn, perms_per_walker=perms, seed=seed)


def key(rec):
    return (rec["walker"], rec["particle"], rec["k"], rec["perm_start"], rec["accepted"],
            rec["sum_exp"], bits(rec["sum_mant"]), bits(rec["delta"]))


def assert_same(a, b):
    for f in ("walker", "particle", "k", "perm_start", "accepted", "sum_exp"):
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(bits(a["sum_mant"]), bits(b["sum_mant"]))
    assert np.array_equal(bits(a["delta"]), bits(b["delta"]))


def test_alone_equals_in_the_mix(lc, dev):
    srv = make_server(lc)
    z_init = [srv.walker(w) for w in range(srv.walkers)]
    common = dict(requests=480, k_min=1, k_max=300, seed=9)
    st_a, alone = srv.run(rate=0.0, max_batch_requests=1, max_inflight=1, **common)
    z_a = [srv.walker(w) for w in range(srv.walkers)]
    assert st_a["batches"] == 480
    srv.reset()
    assert all(np.array_equal(srv.walker(w), z_init[w]) for w in range(srv.walkers))
    st_b, mix = srv.run(rate=0.0, max_batch_requests=64, max_inflight=4, **common)
    assert st_b["batches"] < 480 / 4
    assert_same(alone, mix)
    assert all(np.array_equal(bits(srv.walker(w)), bits(z_a[w])) for w in range(srv.walkers))
    # Poisson arrivals under a budget that holds only a few requests at once
    tight_cfg = dict(rate=20000.0, max_batch_requests=16, max_inflight=3,
                     max_batch_dets=16 * 300)
    srv.reset()
    st_c, _ = srv.run(**tight_cfg, **common)   # resident + batch buffers of this shape
    srv.reset()
    budget = st_c["resident_bytes"] + 3 * (300 * 44 + 1108)  # about three requests in flight
    st_c, tight = srv.run(hbm_budget=budget, **tight_cfg, **common)
    assert st_c["budget_waits"] > 0
    assert_same(alone, tight)
    assert 0.0 < st_c["p50_s"] <= st_c["p99_s"] <= st_c["max_s"]
    srv.close()


def scaled_sum(mant, exp, e0):
    """sum mant_t * 2^(exp_t - e0) exactly (Fractions), and sum |det_t| * 2^-e0."""
    from fractions import Fraction
    re = sum(Fraction(float(m.real)) * Fraction(2) ** (int(e) - e0) for m, e in zip(mant, exp))
    im = sum(Fraction(float(m.imag)) * Fraction(2) ** (int(e) - e0) for m, e in zip(mant, exp))
    mag = sum(abs(complex(m)) * 2.0 ** (int(e) - e0) for m, e in zip(mant, exp))
    return re, im, mag


def test_sums_match_recomputed_terms(lc, dev):
    srv = make_server(lc, walkers=12, perms=200, seed=21)
    z = {w: srv.walker(w) for w in range(srv.walkers)}
    tables = {w: srv.perms(w) for w in range(srv.walkers)}
    _, rec = srv.run(requests=96, rate=0.0, k_min=1, k_max=200, seed=4,
                     max_batch_requests=32, max_inflight=2)
    from fractions import Fraction
    for r in rec:
        w = int(r["walker"])
        zp = z[w].copy()
        zp[r["particle"]] = complex(zp[r["particle"]].real + r["delta"].real,
                                    zp[r["particle"]].imag + r["delta"].imag)
        idx = (int(r["perm_start"]) + np.arange(int(r["k"]))) % srv.perms_per_walker
        perm = torch.from_numpy(np.ascontiguousarray(tables[w][idx])).to(dev)
        res = lc.fqhe_det(torch.from_numpy(zp[None]).to(dev), row_perm=perm, count=int(r["k"]))
        ex = res.exp.cpu().numpy()
        e0 = int(ex.max())   # everything scaled by 2^-e0 (the determinants overflow doubles)
        re, im, mag = scaled_sum(res.mant.cpu().numpy(), ex, e0)
        got_re = Fraction(float(r["sum_mant"].real)) * Fraction(2) ** (int(r["sum_exp"]) - e0)
        got_im = Fraction(float(r["sum_mant"].imag)) * Fraction(2) ** (int(r["sum_exp"]) - e0)
        err = math.hypot(float(got_re - re), float(got_im - im))
        assert err <= 1e-12 * mag, (int(r["walker"]), err, mag)
        if r["accepted"]:
            z[w] = zp
    for w in range(srv.walkers):
        assert np.array_equal(bits(srv.walker(w)), bits(z[w]))
    srv.close()


def test_server_validation(lc, dev):
    srv = make_server(lc, walkers=6, perms=10)
    with pytest.raises(lc.LufqheError) as e:
        srv.run(requests=4, k_max=100, hbm_budget=1 << 16)
    assert e.value.status == 5  # out of budget
    with pytest.raises(lc.LufqheError):
        srv.run(requests=4, k_max=100, max_batch_dets=50)   # a request cannot fit a batch
    with pytest.raises(lc.LufqheError):
        srv.run(requests=4, max_inflight=9)
    st, rec = srv.run(requests=0)
    assert st["requests"] == 0
    srv.close()


def exact_sum(mant, exp, e0):
    """sum mant_t * 2^(exp_t - e0) as exact Fractions (re, im), and sum |det_t| * 2^-e0."""
    from fractions import Fraction
    re = im = Fraction(0)
    for m, e in zip(mant, exp):
        s = Fraction(2) ** (int(e) - e0)
        re += Fraction(float(m.real)) * s
        im += Fraction(float(m.imag)) * s
    mag = float(np.sum(np.abs(mant) * np.exp2(exp.astype(np.float64) - e0)))
    return re, im, mag


def test_serving_config_full_size_against_oracle(lc, orc, wl, dev):
    """The serving configuration at its real size: 512 walkers (n = 8, 16, 24,
    32, 48, 64 in turn), 2000 permutations per walker, k ~ U[1, 2000] terms
    per request, as in the reference's Monte Carlo iteration
    (LU-fqhe/src/MonteCarloLoop.cu:280-400).
      * every record is bit-identical served alone (one request per batch, one
        batch in flight) and in the mix (Poisson arrivals near capacity, up to
        128 requests per batch, 4 batches in flight);
      * for 3 requests of each size (among them the largest k of each n),
        replaying the walker's trajectory: psi' equals the exact sum of the
        oracle mirror's determinants of the permuted FQHE matrices (as the
        device builds them) to 1e-12 of sum |det_t|, and the long-double
        truth of the host-built matrices to 1e-9 of it."""
    from fractions import Fraction
    walkers, perms = 512, 2000
    wn = [SIZES[w % len(SIZES)] for w in range(walkers)]
    srv = lc.Server(wn, perms_per_walker=perms, seed=7)
    z = {w: srv.walker(w) for w in range(walkers)}
    tables = {w: srv.perms(w) for w in range(walkers)}
    common = dict(requests=1536, k_min=1, k_max=2000, seed=13, hbm_budget=200 * 10 ** 9)
    st_a, alone = srv.run(rate=0.0, max_batch_requests=1, max_inflight=1, **common)
    srv.reset()
    st_m, mix = srv.run(rate=40000.0, max_batch_requests=128, max_inflight=4, **common)
    print(f"alone: {st_a['dets_per_s']:.3e} det/s; mix: {st_m['dets_per_s']:.3e} det/s, "
          f"p50 {st_m['p50_s'] * 1e3:.3f} ms, p99 {st_m['p99_s'] * 1e3:.3f} ms, "
          f"{st_m['batches']} batches", flush=True)
    assert st_a["batches"] == 1536 and st_m["batches"] < 1536   # requests shared batches
    assert_same(alone, mix)
    assert int(alone["k"].max()) > 1900 and int(alone["k"].min()) < 100

    # requests to check: per n, the largest k and two others
    pick = set()
    for n in SIZES:
        rs = [r for r in range(len(mix)) if wn[int(mix["walker"][r])] == n]
        rs.sort(key=lambda r: -int(mix["k"][r]))
        pick.update([rs[0], rs[len(rs) // 2], rs[-1]])
    worst_mirror = worst_truth = 0.0
    for r in range(len(mix)):
        rec = mix[r]
        w = int(rec["walker"])
        zp = z[w].copy()
        zp[rec["particle"]] += rec["delta"]
        if r in pick:
            n, k = wn[w], int(rec["k"])
            idx = (int(rec["perm_start"]) + np.arange(k)) % perms
            perm = np.ascontiguousarray(tables[w][idx])
            A = lc.fqhe_build(torch.from_numpy(zp[None]).to(dev),
                              row_perm=torch.from_numpy(perm).to(dev), count=k).cpu().numpy()
            o = orc.mirror(A, "crout_ref", "fused", pivots=False)
            assert np.all(o["info"] == 0)
            e0 = int(o["exp"].max())
            re, im, mag = exact_sum(o["mant"], o["exp"], e0)
            got_re = Fraction(float(rec["sum_mant"].real)) * Fraction(2) ** (int(rec["sum_exp"]) - e0)
            got_im = Fraction(float(rec["sum_mant"].imag)) * Fraction(2) ** (int(rec["sum_exp"]) - e0)
            err = math.hypot(float(got_re - re), float(got_im - im)) / mag
            worst_mirror = max(worst_mirror, err)
            assert err <= 1e-12, (r, n, k, err)
            # accuracy: long-double truth of the host-built matrices
            base = wl.fqhe_matrix_seq(zp)
            t = orc.truth(base[perm])
            d = np.exp((t["logabs"] - t["logabs"].max()) + 1j * t["phase"])
            scale = math.log(2.0) * int(rec["sum_exp"]) - float(t["logabs"].max())
            psi = complex(rec["sum_mant"]) * math.exp(scale)
            errt = abs(psi - d.sum()) / np.abs(d).sum()
            worst_truth = max(worst_truth, errt)
            assert errt <= 1e-9, (r, n, k, errt)
        if rec["accepted"]:
            z[w] = zp
    print(f"checked {len(pick)} requests: worst vs mirror {worst_mirror:.2e}, "
          f"vs long-double truth {worst_truth:.2e} (relative to sum |det_t|)", flush=True)
    for w in range(walkers):
        assert np.array_equal(bits(srv.walker(w)), bits(z[w]))
    srv.close()
