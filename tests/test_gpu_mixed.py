"""GPU parity of mixed-size batches (one flat buffer, n = 1..64 per matrix).

Every mode -- one hipGraph replayed, forked streams, plain per-size launches,
persistent class kernels draining device work queues -- must give the
oracle's determinants bit for bit at each matrix's own index. The graphs are
replayed several times with eager work and synchronisation in between; every
replay is checked.
"""
import numpy as np
import pytest
import torch

from conftest import bits

pytestmark = pytest.mark.gpu


def mixed_input(wl, count, seed, lo=1, hi=64):
    flat, ns, offs, sing = wl.mixed_batch(count, seed, n_lo=lo, n_hi=hi, near_singular=0.05)
    return flat, ns, offs


def check(orc, flat_np, ns, offs, res, policy):
    o = orc.mirror_mixed(flat_np, ns, offs, policy, "fused")
    assert np.array_equal(res.info.cpu().numpy(), o["info"])
    assert np.array_equal(bits(res.mant.cpu().numpy()), bits(o["mant"]))
    assert np.array_equal(res.exp.cpu().numpy(), o["exp"])


@pytest.mark.parametrize("policy", ["crout_ref", "abs"])
def test_mixed_modes_bit_exact(lc, orc, wl, dev, policy):
    flat, ns, offs = mixed_input(wl, 1500, seed=71)
    flat_np = flat.numpy()
    d = flat.to(dev)
    plan = lc.MixedPlan(ns, offs, policy=policy, nstreams=4)
    assert sorted(b[0] for b in plan.buckets()) == sorted(set(ns.tolist()))
    side = torch.cuda.Stream()
    for mode in ("plain", "streams", "graph", "persistent"):
        res = lc.alloc_result(len(ns), 64, dev)
        with torch.cuda.stream(side):
            plan.execute(d, res, mode=mode, stream=side)
        side.synchronize()
        check(orc, flat_np, ns, offs, res, policy)
    plan.close()


@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_graph_replays_with_eager_work_between(lc, orc, wl, dev, mode):
    flat, ns, offs = mixed_input(wl, 900, seed=5, lo=8, hi=64)
    flat_np = flat.numpy()
    d = flat.to(dev)
    plan = lc.MixedPlan(ns, offs, nstreams=3)
    res = lc.alloc_result(len(ns), 64, dev)
    side = torch.cuda.Stream()
    o = orc.mirror_mixed(flat_np, ns, offs, "crout_ref", "fused")
    for rep in range(4):
        res.mant.zero_()
        res.exp.fill_(-7)
        with torch.cuda.stream(side):
            plan.execute(d, res, mode=mode, stream=side)
        side.synchronize()
        # eager work of the program between replays
        x = torch.randn(1 << 16, device=dev).sum()
        torch.cuda.synchronize()
        assert np.array_equal(bits(res.mant.cpu().numpy()), bits(o["mant"])), rep
        assert np.array_equal(res.exp.cpu().numpy(), o["exp"]), rep
        del x
    plan.close()


def test_persistent_new_pointers_and_single_class(lc, orc, wl, dev):
    """Changing the output buffers re-captures the persistent graph with new
    per-bucket arguments, and switching back works; a batch whose sizes all
    fall in one class (n = 3..9) uses one class kernel; modes alternate on
    one plan."""
    side = torch.cuda.Stream()
    for lo, hi, count in ((3, 9, 700), (1, 64, 1200)):
        flat, ns, offs = mixed_input(wl, count, seed=lo + hi, lo=lo, hi=hi)
        flat_np = flat.numpy()
        d = flat.to(dev)
        o = orc.mirror_mixed(flat_np, ns, offs, "crout_ref", "fused")
        plan = lc.MixedPlan(ns, offs, nstreams=2)
        r1 = lc.alloc_result(len(ns), 64, dev)
        r2 = lc.alloc_result(len(ns), 64, dev)
        for res, mode in ((r1, "persistent"), (r2, "persistent"), (r1, "graph"),
                          (r1, "persistent"), (r2, "streams"), (r2, "persistent")):
            res.mant.zero_()
            with torch.cuda.stream(side):
                plan.execute(d, res, mode=mode, stream=side)
            side.synchronize()
            assert np.array_equal(bits(res.mant.cpu().numpy()), bits(o["mant"])), (lo, mode)
            assert np.array_equal(res.exp.cpu().numpy(), o["exp"]), (lo, mode)
        plan.close()


def test_mixed_graph_rejects_default_stream(lc, wl, dev):
    flat, ns, offs = mixed_input(wl, 10, seed=1, lo=4, hi=8)
    plan = lc.MixedPlan(ns, offs)
    res = lc.alloc_result(len(ns), 8, dev)
    with pytest.raises(lc.LufqheError):
        plan.execute(flat.to(dev), res, mode="graph", stream=torch.cuda.default_stream())
    plan.close()


def test_config4_full_size(lc, orc, wl, dev):
    """The mixed-size batch at full size: 200000 matrices, n ~ U{8..64}, 5% near-singular.
    Statuses clean, a sample bit-exact against the mirror, the rest against
    the long-double truth by matrix class."""
    flat, ns, offs, sing = wl.mixed_batch(200000, seed=2024, device=dev)
    plan = lc.MixedPlan(ns, offs, nstreams=4)
    res = lc.alloc_result(len(ns), 64, dev)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        plan.execute(flat, res, mode="graph", stream=side)
    side.synchronize()
    assert int((res.info != 0).sum()) == 0
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(ns), 600, replace=False))
    sub_ns = ns[idx]
    sub_offs = np.zeros(len(idx), np.int64)
    sub_offs[1:] = np.cumsum(sub_ns.astype(np.int64) ** 2)[:-1]
    flat_cpu = flat.cpu().numpy()
    sub = np.concatenate([flat_cpu[offs[i]:offs[i] + ns[i] ** 2] for i in idx])
    o = orc.mirror_mixed(sub, sub_ns, sub_offs, "crout_ref", "fused")
    assert np.array_equal(bits(res.mant.cpu().numpy()[idx]), bits(o["mant"]))
    # the persistent class kernels give the same bits for every matrix
    rp = lc.alloc_result(len(ns), 64, dev)
    with torch.cuda.stream(side):
        plan.execute(flat, rp, mode="persistent", stream=side)
    side.synchronize()
    assert torch.equal(rp.mant.view(torch.float64), res.mant.view(torch.float64))
    assert torch.equal(rp.exp, res.exp) and torch.equal(rp.info, res.info)
    t = orc.truth_mixed(sub, sub_ns, sub_offs)
    la = res.logabs.cpu().numpy()[idx]
    ph = res.phase.cpu().numpy()[idx]
    err = np.abs(np.expm1((la - t["logabs"]) + 1j * np.angle(np.exp(1j * (ph - t["phase"])))))
    well = ~sing[idx]
    assert np.median(err[well]) < 1e-13 and np.quantile(err[well], 0.99) < 1e-10
    assert np.median(err[~well]) < 1e-6  # cond ~ 1e8: error ~ cond * eps
    plan.close()
