"""lufqhe_det_inplace_static: the in-place determinant call with its
workspace in a fixed device array (for callers that cannot allocate).
Checked through the C ABI against lufqhe_det_inplace, the allocating entry
point it mirrors, bit for bit, on two calls in a row (the second smaller, so
it runs on what the first left in the array), plus the size limit."""
import numpy as np
import pytest

from conftest import bits

MAX = 1 << 20   # LUFQHE_STATIC_MAX_MATRICES


def test_static_entry_point_validation(lc):
    """Argument checks answer before any device work (runs without a GPU)."""
    lib = lc.load()
    assert lib.lufqhe_det_inplace_static(None, MAX + 1, 8, 8, 64, 1, 0, None) == 7
    assert lib.lufqhe_det_inplace_static(None, 10, 0, 8, 64, 1, 0, None) == 1
    assert lib.lufqhe_det_inplace_static(None, 10, 8, 7, 64, 1, 0, None) == 1
    assert lib.lufqhe_det_inplace_static(None, 10, 8, 8, 64, 1, 5, None) == 2
    assert lib.lufqhe_det_inplace_static(None, 0, 8, 8, 64, 1, 0, None) == 0
    assert b"fixed workspace" in lib.lufqhe_status_string(7)


def _static_call(lc, A, pair, stream):
    B, n, _ = A.shape
    return lc.load().lufqhe_det_inplace_static(A.data_ptr(), B, n, A.stride(1), A.stride(0),
                                               int(pair), 0, stream.cuda_stream)


@pytest.mark.gpu
def test_static_entry_point_matches_allocating_one(lc, wl, dev):
    import torch
    stream = torch.cuda.current_stream()
    cases = [(5000, 24, True, 1), (777, 9, False, 2)]   # second call smaller
    for count, n, pair, seed in cases:
        A = wl.gaussian(count, n, seed=seed, device=dev)
        ref = A.clone()
        lc.det_inplace(ref, pair_product=pair)
        got = A.clone()
        assert _static_call(lc, got, pair, stream) == 0
        torch.cuda.synchronize()
        assert np.array_equal(bits(got.cpu().numpy()), bits(ref.cpu().numpy())), (count, n)
        # only the [0][0] entries changed
        keep = torch.ones_like(A, dtype=torch.bool)
        keep[:, 0, 0] = False
        assert torch.equal(got[keep], A[keep])
    # a padded layout (ld > n) on the same array
    base = wl.gaussian(300, 16, seed=3, device=dev)
    buf = torch.zeros(300, 16, 20, dtype=torch.complex128, device=dev)
    buf[:, :, :16] = base
    view = buf[:, :, :16]
    ref = buf.clone()
    lc.det_inplace(ref[:, :, :16], pair_product=True)
    assert _static_call(lc, view, True, stream) == 0
    torch.cuda.synchronize()
    assert np.array_equal(bits(buf.cpu().numpy()), bits(ref.cpu().numpy()))
    # over the limit: refused, nothing written
    small = wl.gaussian(4, 5, seed=4, device=dev)
    before = small.clone()
    st = lc.load().lufqhe_det_inplace_static(small.data_ptr(), MAX + 1, 5, 5, 25, 1, 0,
                                             stream.cuda_stream)
    torch.cuda.synchronize()
    assert st == 7 and torch.equal(small, before)
