"""Seeded synthetic workloads of the benchmarks and tests.

All generators are deterministic functions of their seed.  Gaussian entries
use torch's counter-based (Philox) generator, so a batch generated on the GPU
is reproducible; numpy variants exist for small CPU-side cases.

  small batch     4096 x n=16, re,im ~ N(0,1)
  uniform batch   65536 x n=32, re,im ~ N(0,1)
  FQHE batch      32768 FQHE matrices, n ~ U{10..40}: z_i uniform in the
                  disk of radius sqrt(2n), A_ij = z_i^j exp(-|z_i|^2/4);
                  log|det| is known analytically
  mixed batch     200000 x n ~ U{8..64}, N(0,1), 5% near-singular (one row
                  replaced by alpha*row_a + beta*row_b + 1e-8*N(0,1))
  walker serving  FQHE walkers of n in {8,16,24,32,48,64} particles; the
                  walk itself runs on the device (lucore.Server)
"""
from __future__ import annotations

import math

import numpy as np
import torch


def gaussian(count: int, n: int, seed: int, device="cpu") -> torch.Tensor:
    """complex128 [count, n, n], re and im i.i.d. N(0,1)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    re = torch.randn(count, n, n, generator=g, dtype=torch.float64, device=device)
    im = torch.randn(count, n, n, generator=g, dtype=torch.float64, device=device)
    return torch.complex(re, im)


def gaussian_np(count: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n)))


# --------------------------------------------------------------------------- FQHE
def fqhe_positions(n: int, rng: np.random.Generator) -> np.ndarray:
    """n particle positions uniform in the disk of radius sqrt(2n)."""
    R = math.sqrt(2.0 * n)
    u = rng.random(n)
    v = rng.random(n)
    return R * np.sqrt(u) * np.exp(2j * np.pi * v)


def fqhe_matrix(z: np.ndarray) -> np.ndarray:
    """A_ij = z_i^j exp(-|z_i|^2/4), j = 0..n-1 (complex128)."""
    n = len(z)
    powers = z[:, None] ** np.arange(n)[None, :]
    return powers * np.exp(-np.abs(z) ** 2 / 4.0)[:, None]


def fqhe_analytic(z: np.ndarray) -> tuple[float, float]:
    """Exact log|det| and arg det of fqhe_matrix(z) (Vandermonde times the
    Gaussian factors): sum_{i<j} log|z_j - z_i| - sum |z_i|^2/4 and
    sum_{i<j} arg(z_j - z_i) mod 2pi."""
    n = len(z)
    iu, ju = np.triu_indices(n, 1)
    d = z[ju] - z[iu]
    la = float(np.sum(np.log(np.abs(d))) - np.sum(np.abs(z) ** 2) / 4.0)
    ph = float(np.angle(np.prod(d / np.abs(d))))
    return la, ph


def fqhe_batch(count: int, n_lo: int, n_hi: int, seed: int):
    """FQHE batch with n ~ U{n_lo..n_hi}: (flat complex128, ns, offs, z list)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_lo, n_hi + 1, size=count).astype(np.int32)
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    flat = np.empty(int(np.sum(ns.astype(np.int64) ** 2)), np.complex128)
    zs = []
    for m in range(count):
        z = fqhe_positions(int(ns[m]), rng)
        zs.append(z)
        flat[offs[m]:offs[m] + ns[m] ** 2] = fqhe_matrix(z).ravel()
    return flat, ns, offs, zs


def fqhe_configs(count: int, n_lo: int, n_hi: int, seed: int):
    """Packed particle configurations, n ~ U{n_lo..n_hi}:
    (z_flat complex128, ns int32, offs int64) with matrix m's positions at
    z_flat[offs[m] : offs[m] + ns[m]] -- the input of the on-device builder."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_lo, n_hi + 1, size=count).astype(np.int32)
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    total = int(offs[-1] + ns[-1]) if count else 0
    R = np.sqrt(2.0 * np.repeat(ns, ns).astype(np.float64))
    u = rng.random(total)
    v = rng.random(total)
    z = R * np.sqrt(u) * np.exp(2j * np.pi * v)
    return z, ns, offs


def fqhe_matrix_seq(z: np.ndarray) -> np.ndarray:
    """fqhe_matrix with powers as sequential products z^j = z^(j-1) * z (the
    device builder's order)."""
    n = len(z)
    A = np.empty((n, n), np.complex128)
    g = np.exp(-0.25 * (z.real * z.real + z.imag * z.imag))
    pw = np.ones(n, np.complex128)
    for j in range(n):
        A[:, j] = pw * g
        pw = pw * z
    return A


# --------------------------------------------------------------------------- mixed
def mixed_sizes(count: int, n_lo: int, n_hi: int, seed: int):
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_lo, n_hi + 1, size=count).astype(np.int32)
    offs = np.zeros(count, np.int64)
    offs[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    return ns, offs


def mixed_batch(count: int, seed: int, n_lo: int = 8, n_hi: int = 64,
                near_singular: float = 0.05, device="cpu"):
    """Packed mixed-size batch: (flat complex128 tensor, ns, offs, singular mask).

    Matrix m occupies flat[offs[m] : offs[m] + ns[m]**2] (row major, ld = n).
    A fraction `near_singular` of the matrices get one row replaced by
    alpha*row_a + beta*row_b + 1e-8*N(0,1) (complex alpha, beta ~ N(0,1))."""
    ns, offs = mixed_sizes(count, n_lo, n_hi, seed)
    total = int(offs[-1] + int(ns[-1]) ** 2)
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1)
    flat = torch.complex(torch.randn(total, generator=g, dtype=torch.float64, device=device),
                         torch.randn(total, generator=g, dtype=torch.float64, device=device))
    rng = np.random.default_rng(seed + 2)
    sing = (rng.random(count) < near_singular) & (ns >= 3)  # needs 3 distinct rows
    idx = np.nonzero(sing)[0]
    for m in idx:
        n = int(ns[m])
        r, a, b = rng.choice(n, size=3, replace=False)
        alpha = complex(rng.standard_normal(), rng.standard_normal())
        beta = complex(rng.standard_normal(), rng.standard_normal())
        base = int(offs[m])
        M = flat[base:base + n * n].view(n, n)
        noise = torch.complex(torch.tensor(rng.standard_normal(n)), torch.tensor(rng.standard_normal(n)))
        M[r] = alpha * M[a] + beta * M[b] + 1e-8 * noise.to(device)
    return flat, ns, offs, sing
