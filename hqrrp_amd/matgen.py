"""Seeded test matrices for the workload families HQRRP is measured on.

* ``gaussian``      -- i.i.d. N(0,1) entries;
* ``graded``        -- U diag(sigma) V' with Haar-orthogonal U, V (QR of seeded
                       Gaussians, sign-fixed) and sigma_i = 10^(-decades*i/k);
* ``lowrank_noise`` -- rank-k product of Gaussian factors with sigma_i = 1/i plus
                       a small i.i.d. Gaussian perturbation.

numpy versions are used at test sizes; ``graded_device`` builds the same
family on the GPU with torch for the large bench sizes (input generation is
outside every timed region).
"""
from __future__ import annotations

import numpy as np


def gaussian(m: int, n: int, seed: int = 0) -> np.ndarray:
    return np.asfortranarray(np.random.default_rng(seed).standard_normal((m, n)))


def _haar(rng, m: int, k: int) -> np.ndarray:
    Q, R = np.linalg.qr(rng.standard_normal((m, k)))
    return Q * np.sign(np.diag(R))


def graded(m: int, n: int, seed: int = 0, decades: float = 12.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = min(m, n)
    U = _haar(rng, m, k)
    V = _haar(rng, n, k)
    sig = 10.0 ** (-decades * np.arange(k) / k)
    return np.asfortranarray((U * sig) @ V.T)


def lowrank_noise(m: int, n: int, rank: int = 512, seed: int = 0,
                  noise: float = 1e-8) -> np.ndarray:
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((m, rank)) / np.sqrt(m)
    R = rng.standard_normal((n, rank)) / np.sqrt(n)
    sig = 1.0 / np.arange(1, rank + 1)
    return np.asfortranarray((L * sig) @ R.T + noise * rng.standard_normal((m, n)))


def graded_device(m: int, n: int, seed: int = 0, decades: float = 12.0,
                  device: str = "cuda"):
    """Same family as :func:`graded`, built with torch on the GPU.

    Returns a contiguous (n, m) tensor, i.e. the column-major m x n matrix.
    """
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    k = min(m, n)

    def haar(rows):
        Z = torch.randn(rows, k, generator=g, device=device, dtype=torch.float64)
        Q, R = torch.linalg.qr(Z)
        return Q * torch.sign(torch.diagonal(R))

    U = haar(m)
    V = haar(n)
    sig = 10.0 ** (-decades * torch.arange(k, device=device,
                                           dtype=torch.float64) / k)
    # (U diag(sig) V')' = V diag(sig) U'  -> row-major (n, m) buffer.
    return ((V * sig) @ U.T).contiguous()


def gaussian_device(m: int, n: int, seed: int = 0, device: str = "cuda"):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return torch.randn(n, m, generator=g, device=device, dtype=torch.float64)


def lowrank_device(m: int, n: int, rank: int = 512, seed: int = 0,
                   noise: float = 1e-8, device: str = "cuda",
                   core_sigma: bool = False):
    """Same family as :func:`lowrank_noise` (rank `rank`, sigma_i = 1/i,
    plus `noise` N(0,1)), built with torch on the GPU; returned column-major
    as a (n, m) tensor like :func:`gaussian_device`.

    core_sigma=True also returns the singular values of the noise-free core
    L diag(1/i) R' (exactly, from the SVD of R_L diag(1/i) R_R' with
    L = Q_L R_L, R = Q_R R_R); by Weyl's inequality those of A differ from
    them by at most ||noise||_2 <= noise (sqrt(m) + sqrt(n)) (1 + o(1))."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = torch.float64
    L = torch.randn(m, rank, generator=g, device=device, dtype=f64) / m ** 0.5
    R = torch.randn(n, rank, generator=g, device=device, dtype=f64) / n ** 0.5
    sig = 1.0 / torch.arange(1, rank + 1, device=device, dtype=f64)
    At = (R * sig) @ L.t()  # (n, m): A' = R diag(sig) L'
    At += noise * torch.randn(n, m, generator=g, device=device, dtype=f64)
    if not core_sigma:
        return At
    RL = torch.linalg.qr(L, mode="r")[1]
    RR = torch.linalg.qr(R, mode="r")[1]
    s = torch.linalg.svdvals((RL * sig) @ RR.t())
    return At, s


# ---------------------------------------------------------------- batches
def batch_shapes(count: int, seed: int = 0, m_range=(64, 512),
                 n_min: int = 32):
    """Shapes of the batched workload: m ~ U[m_range], n ~ U[n_min, m]."""
    rng = np.random.default_rng(seed)
    m = rng.integers(m_range[0], m_range[1] + 1, size=count)
    u = rng.random(count)
    n = n_min + np.floor(u * (m - n_min + 1)).astype(np.int64)
    return m.astype(np.int32), np.minimum(n, m).astype(np.int32)


def _batch_one(args):
    m, n, seed, decades = args
    return graded(int(m), int(n), int(seed), decades)


def _one_thread():
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:
        pass


def batch_matrices(m, n, seed: int = 0, decades: float = 8.0,
                   workers: int = 1):
    """Matrix i of a batch: N(0,1)-distributed singular vectors (Haar U, V)
    with singular values rescaled to 10^(-decades*i/k), seeded by (seed, i).
    `workers` > 1 generates them in a process pool."""
    jobs = [(m[i], n[i], seed * 1000003 + i, decades) for i in range(len(m))]
    if workers <= 1 or len(jobs) < 64:
        return [_batch_one(j) for j in jobs]
    import multiprocessing as mp

    with mp.get_context("fork").Pool(workers, initializer=_one_thread) as p:
        return p.map(_batch_one, jobs, chunksize=32)
