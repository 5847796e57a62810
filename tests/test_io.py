"""The reference's raw matrix file (LU-fqhe/src/PermHandler.cpp:91-170).

CPU: the per-rank run bounds (known answers worked out by hand from the
reference's float arithmetic, and the tiling invariant).  GPU: files in
the reference's layout -- the first 32 matrices of its own input file
LU-fqhe/inc/1000_16x16.bin, kept as a fixture -- read back bit for bit,
for one rank and split over ranks, in both precisions, and the
determinants of what was read equal those of the fixture.
"""
import numpy as np
import pytest
import torch

from conftest import bits, golden


@pytest.mark.parametrize("count,n,nodes,expect", [
    # total = 2*count*n*n scalars; rank r starts at (int)((float)(r*total)/nodes)
    (1000, 16, 1, [(0, 512000)]),
    (1000, 16, 3, [(0, 170666), (170666, 170667), (341333, 170667)]),
    (1000, 16, 4, [(0, 128000), (128000, 128000), (256000, 128000), (384000, 128000)]),
    # 2*3*5*5 = 150 over 4 ranks: 37.5 -> 37, 75, 112.5 -> 112
    (3, 5, 4, [(0, 37), (37, 38), (75, 37), (112, 38)]),
])
def test_file_chunk_known_answers(lc, count, n, nodes, expect):
    got = [lc.file_chunk(count, n, r, nodes) for r in range(nodes)]
    assert got == expect


def test_file_chunk_float_rounding(lc):
    # total = 2*4097*64*64 = 33562624 > 2^24: (float)total rounds to 33562624
    # (representable: a multiple of 4); with 3 ranks the bounds round in float
    count, n, nodes = 4097, 64, 3
    total = 2 * count * n * n
    got = [lc.file_chunk(count, n, r, nodes) for r in range(nodes)]
    f32 = np.float32
    starts = [int(f32(f32(r * total) / f32(nodes))) for r in range(nodes + 1)]
    assert got == [(starts[r], starts[r + 1] - starts[r]) for r in range(nodes)]
    # runs tile [0, last) with no gap or overlap
    assert all(got[r][0] + got[r][1] == got[r + 1][0] for r in range(nodes - 1))


def test_file_chunk_validation(lc):
    with pytest.raises(lc.LufqheError):
        lc.file_chunk(10, 4, 2, 2)    # rank >= nodes
    with pytest.raises(lc.LufqheError):
        lc.file_chunk(10, 0, 0, 1)


def test_read_missing_file(lc, tmp_path):
    lib = lc.load()
    assert lib.lufqhe_read_matrices(str(tmp_path / "nope.bin").encode(), 0, 4, 4, 0, 1,
                                    1, 1 << 20, None, None) == 6


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["complex64", "complex128"])
def test_read_reference_file_layout(lc, orc, dev, tmp_path, fmt):
    g = golden("ref_fixture16")
    A32 = g["A32_16"]                       # complex64 [32, 16, 16], the reference's bytes
    data = A32 if fmt == "complex64" else A32.astype(np.complex128)
    path = tmp_path / f"m_{fmt}.bin"
    data.tofile(path)                       # raw interleaved (re, im), row major, no header
    A = lc.read_matrices(path, 32, 16, fmt=fmt, device=dev)
    assert tuple(A.shape) == (32, 16, 16)
    assert np.array_equal(bits(A.cpu().numpy()), bits(A32.astype(np.complex128)))
    res = lc.det_batched(A)
    o = orc.mirror(A32.astype(np.complex128), "crout_ref", "fused", pivots=False)
    assert np.array_equal(bits(res.mant.cpu().numpy()), bits(o["mant"]))
    # the same file split over 3 ranks: the runs concatenate to the whole
    flat = np.concatenate([
        torch.view_as_real(lc.read_matrices(path, 32, 16, fmt=fmt, rank=r, nodes=3,
                                            device=dev)).reshape(-1).cpu().numpy()
        [:lc.file_chunk(32, 16, r, 3)[1]]
        for r in range(3)])
    assert np.array_equal(bits(flat), bits(A32.astype(np.complex128).view(np.float64).ravel()))


@pytest.mark.gpu
def test_read_large_file_in_blocks(lc, dev, tmp_path):
    """More than two staging blocks (32 MiB each): 9000 matrices of 32x32
    complex64 = 73.7 MB."""
    rng = np.random.default_rng(8)
    data = (rng.standard_normal((9000, 32, 32)) + 1j * rng.standard_normal((9000, 32, 32))).astype(np.complex64)
    path = tmp_path / "big.bin"
    data.tofile(path)
    A = lc.read_matrices(path, 9000, 32, device=dev)
    assert np.array_equal(bits(A.cpu().numpy()), bits(data.astype(np.complex128)))
    # a count larger than the file: the run is cut at the end of the file
    B = lc.read_matrices(path, 9001, 32, device=dev)
    assert B.numel() == 9000 * 32 * 32
