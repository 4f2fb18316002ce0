"""Learned subgrid closure: training data from box-filtered fine-grid runs that the
solver itself produces, normalisation, training on the GPU (laps_mlp_train) and the
parameter vector for Laps.closure_set.

No counterpart in the reference. The coarse-grid state is the box filter (mean over
r^3 fine cells) of the fine state. The network learns, per coarse point, the
subgrid terms of the filtered momentum and induction equations:

    R_ij = bar(m_i u_j) - M_i U_j          (Reynolds stress, M = bar(m), U = M / bar(rho))
    M_ij = bar(B_i B_j) - B_i B_j          (Maxwell stress, B = bar(B))
    EMF  = bar(E) - E(bar state),  E = -u x B + (d_i / rho) J x B

with the features the library computes on the coarse state (rho, u, B and their
central-difference gradients). The energy flux is left without a closure term.
"""
import math

import numpy as np

from .runtime import Laps, mlp_train, MLP_NPARAM, CLOSURE_NPARAM

NIN, NH, NOUT = 28, 64, 15
PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (MLP_NPARAM floats)."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in ((NIN, NH), (NH, NH), (NH, NOUT)):
        lim = math.sqrt(6.0 / (a + b))
        parts.append(rng.uniform(-lim, lim, a * b))
        parts.append(np.zeros(b))
    out = np.concatenate(parts).astype(np.float32)
    assert out.size == MLP_NPARAM
    return out


def box_filter(f, r):
    """Mean over r^3 blocks of cells of a Fortran-ordered (nx, ny, nz, ...) array."""
    nx, ny, nz = f.shape[:3]
    rest = f.shape[3:]
    g = np.asarray(f).reshape((r, nx // r, r, ny // r, r, nz // r) + rest, order="F")
    return np.asfortranarray(g.mean(axis=(0, 2, 4)))


def _k(n, L):
    return 2.0 * math.pi * np.fft.fftfreq(n, d=L / n)


def curl(B, L):
    """Spectral curl of a real (nx, ny, nz, 3) field on the box L = (Lx, Ly, Lz)."""
    nx, ny, nz = B.shape[:3]
    kx, ky, kz = np.meshgrid(_k(nx, L[0]), _k(ny, L[1]), _k(nz, L[2]), indexing="ij")
    Bh = [np.fft.fftn(B[..., a]) for a in range(3)]
    J = np.empty_like(B)
    J[..., 0] = np.real(np.fft.ifftn(1j * (ky * Bh[2] - kz * Bh[1])))
    J[..., 1] = np.real(np.fft.ifftn(1j * (kz * Bh[0] - kx * Bh[2])))
    J[..., 2] = np.real(np.fft.ifftn(1j * (kx * Bh[1] - ky * Bh[0])))
    return J


def emf(rho, u, B, J, di):
    E = -np.cross(u, B)
    if di:
        E = E + (di / rho)[..., None] * np.cross(J, B)
    return E


def filtered_sample(uu, prim, r, gamma, di, L):
    """Coarse state (uu_c, prim_c) of a fine state and the (Nc, 15) subgrid targets,
    points in Fortran order."""
    rho, m, B = uu[..., 0], uu[..., 1:4], uu[..., 4:7]
    u = prim[..., 0:3]
    E = emf(rho, u, B, curl(B, L), di)
    uu_c = box_filter(uu, r)
    rho_c, m_c, B_c, e_c = uu_c[..., 0], uu_c[..., 1:4], uu_c[..., 4:7], uu_c[..., 7]
    u_c = m_c / rho_c[..., None]
    p_c = (gamma - 1.0) * (e_c - 0.5 * (m_c * u_c).sum(-1) - 0.5 * (B_c * B_c).sum(-1))
    prim_c = np.asfortranarray(np.concatenate([u_c, p_c[..., None]], axis=-1))
    E_c = emf(rho_c, u_c, B_c, curl(B_c, L), di)
    T = []
    for i, j in PAIRS:
        T.append(box_filter(m[..., i] * u[..., j], r) - m_c[..., i] * u_c[..., j])
    for i, j in PAIRS:
        T.append(box_filter(B[..., i] * B[..., j], r) - B_c[..., i] * B_c[..., j])
    Ef = box_filter(E, r)
    for a in range(3):
        T.append(Ef[..., a] - E_c[..., a])
    T = np.stack([t.reshape(-1, order="F") for t in T], axis=1)
    return uu_c, prim_c, T


def coarse_params(fine, r):
    """Parameter set of the coarse box for a fine one (grid / r, same physics)."""
    from .params import LapsParams
    import ctypes
    p = LapsParams()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(fine), ctypes.sizeof(LapsParams))
    p.nx, p.ny, p.nz = fine.nx // r, fine.ny // r, fine.nz // r
    return p


def features_of(p_coarse, uu_c, prim_c):
    """Raw features of a coarse state, computed by the library's own feature kernel."""
    with Laps(p_coarse) as g:
        g.set_state(uu_c, prim_c)
        return g.closure_features()


def normalise(X, T):
    xmean = X.mean(axis=0)
    xstd = X.std(axis=0)
    xstd = np.where(xstd > 0, xstd, 1.0)
    ystd = T.std(axis=0)
    ystd = np.where(ystd > 0, ystd, 1.0)
    return xmean.astype(np.float32), xstd.astype(np.float32), ystd.astype(np.float32)


def closure_vector(P, xmean, xstd, ystd):
    out = np.concatenate([P, xmean, xstd, ystd]).astype(np.float32)
    assert out.size == CLOSURE_NPARAM
    return out


def train(X, T, steps=3000, batch=4096, lr=2e-3, seed=0, device=0):
    """Trains the network on raw features X (n, 28) and targets T (n, 15): returns the
    closure parameter vector and the loss history (on normalised targets)."""
    xmean, xstd, ystd = normalise(X, T)
    Xn = (X - xmean) / xstd
    Tn = T / ystd
    perm = np.random.default_rng(seed).permutation(X.shape[0])
    P, hist = mlp_train(Xn[perm], Tn[perm], init_params(seed), steps, batch, lr, device=device)
    return closure_vector(P, xmean, xstd, ystd), hist


def shell_spectrum(uhat, n, kmax):
    """(rho u)^2/2 + B^2/2 per integer shell of a cubic spectrum (nx/2+1, n, n, 8)."""
    nxh = uhat.shape[0]
    mx = np.arange(nxh)[:, None, None]
    m = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(mx ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2)
    w = np.where(mx == 0, 1.0, 2.0) * np.ones_like(kk)
    e = sum(np.abs(uhat[..., v]) ** 2 for v in (1, 2, 3, 4, 5, 6)) * 0.5 * w
    shell = np.rint(kk).astype(np.int64).ravel()
    return np.bincount(shell, weights=e.ravel(), minlength=kmax + 1)[: kmax + 1]
