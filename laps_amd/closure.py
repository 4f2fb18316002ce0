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
    """Spectral curl of a real (nx, ny, nz, 3) field on the box L = (Lx, Ly, Lz) (real
    transforms over (z, y, x), multithreaded)."""
    import os
    import scipy.fft as sfft
    nx, ny, nz = B.shape[:3]
    w = min(16, os.cpu_count() or 1)
    kx = 2.0 * math.pi * np.arange(nx // 2 + 1) / L[0]
    kx, ky, kz = np.meshgrid(kx, _k(ny, L[1]), _k(nz, L[2]), indexing="ij")
    Bh = [sfft.rfftn(B[..., a], axes=(2, 1, 0), workers=w) for a in range(3)]
    J = np.empty_like(B)

    def inv(F):
        return sfft.irfftn(F, s=(nz, ny, nx), axes=(2, 1, 0), workers=w)

    J[..., 0] = inv(1j * (ky * Bh[2] - kz * Bh[1]))
    J[..., 1] = inv(1j * (kz * Bh[0] - kx * Bh[2]))
    J[..., 2] = inv(1j * (kx * Bh[1] - ky * Bh[0]))
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


def fine_run(p_fine, seed, r, nsteps, every, amp, dt):
    """Snapshots (step, uu, prim) of a fine run at a fixed dt, every `every` steps."""
    from .ic import cons_from_prim, seeded_primitive
    n = (p_fine.nx, p_fine.ny, p_fine.nz)
    uu, pr = cons_from_prim(seeded_primitive(n, seed, kmin=1, kmax=8, du=amp, db=amp), p_fine.adiabatic_index)
    out = []
    with Laps(p_fine) as g:
        g.set_state(uu, pr)
        g.set_fixed_dt(dt)
        for s in range(0, nsteps + 1, every):
            if s:
                g.step(every)
            u, q = g.state()
            out.append((s, u, q))
    return out


def experiment(n_fine=128, r=4, train_seeds=(11, 12), test_seed=21, nsteps=384, every=32, amp=0.4,
               di=0.05, visc=5e-3, train_steps=4000, log=print):
    """Trains the closure on box-filtered fine runs and compares coarse runs with and
    without it against the filtered fine run of another seed: energy spectra per shell up
    to the coarse grid's dealiased range, and the cost of a coarse step with the closure
    against the fine steps it stands in for."""
    import time
    from .params import default_params
    L = 2.0 * math.pi
    pf = default_params(nx=n_fine, ny=n_fine, nz=n_fine, Lx=L, Ly=L, Lz=L, adiabatic_index=5.0 / 3.0, cfl=0.4,
                        dealias_option=1, if_visc=1, viscosity=visc, if_resis=1, resistivity=visc,
                        if_hall=1 if di else 0, ion_inertial_length=di)
    pc = coarse_params(pf, r)
    nc = pc.nx
    # the fine dt: 0.8 of the CFL dt of the first training state, kept for every run
    from .ic import cons_from_prim, seeded_primitive
    uu0, pr0 = cons_from_prim(seeded_primitive((n_fine,) * 3, train_seeds[0], kmin=1, kmax=8, du=amp, db=amp),
                              pf.adiabatic_index)
    with Laps(pf) as g:
        g.set_state(uu0, pr0)
        dt = 0.8 * g.cfl()[1]
    t0 = time.perf_counter()
    X, T = [], []
    for sd in train_seeds:
        for s, u, q in fine_run(pf, sd, r, nsteps, every, amp, dt):
            uc, qc, t = filtered_sample(u, q, r, pf.adiabatic_index, di, (L, L, L))
            X.append(features_of(pc, uc, qc))
            T.append(t)
    X = np.concatenate(X)
    T = np.concatenate(T).astype(np.float32)
    log(f"closure: {X.shape[0]} training samples from {len(train_seeds)} fine {n_fine}^3 runs "
        f"({time.perf_counter() - t0:.1f} s)")
    t1 = time.perf_counter()
    vec, hist = train(X, T, steps=train_steps)
    log(f"closure: trained {train_steps} Adam steps in {time.perf_counter() - t1:.1f} s, "
        f"loss {hist[:50].mean():.3f} -> {hist[-50:].mean():.3f}")
    # evaluation on an unseen seed
    snaps = fine_run(pf, test_seed, r, nsteps, nsteps, amp, dt)
    (_, u0, q0), (_, u1, q1) = snaps[0], snaps[-1]
    uc0, qc0, _ = filtered_sample(u0, q0, r, pf.adiabatic_index, di, (L, L, L))
    uc1, _, T1 = filtered_sample(u1, q1, r, pf.adiabatic_index, di, (L, L, L))
    kmax = nc // 3
    with Laps(pc) as ref_g:
        ref_g.set_state(uc1, filtered_sample(u1, q1, r, pf.adiabatic_index, di, (L, L, L))[1])
        e_ref = shell_spectrum(ref_g.spectral(), nc, kmax)
    res = {}
    for tag, use in (("plain", False), ("closure", True)):
        with Laps(pc) as g:
            g.set_state(uc0, qc0)
            g.set_fixed_dt(r * dt)
            if use:
                g.closure_set(vec)
            g.step(nsteps // r)
            g.sync()
            t2 = time.perf_counter()
            g.step(nsteps // r)
            g.sync()
            tc = (time.perf_counter() - t2) / (nsteps // r)
            g.closure_set(None)
            g.set_state(uc0, qc0)
            g.set_fixed_dt(r * dt)
            if use:
                g.closure_set(vec)
            g.step(nsteps // r)
            e = shell_spectrum(g.spectral(), nc, kmax)
        dev = np.abs(e[1:] - e_ref[1:]) / e_ref[1:]
        res[tag] = dict(max_shell_dev=float(dev.max()), mean_shell_dev=float(dev.mean()),
                        shell_dev=[round(float(x), 4) for x in dev], s_per_step=tc)
        log(f"closure: coarse {nc}^3 {tag}: max shell deviation {dev.max():.4f}, mean {dev.mean():.4f}")
    with Laps(pf) as g:
        g.set_state(u0, q0)
        g.set_fixed_dt(dt)
        g.step(4)
        g.sync()
        t3 = time.perf_counter()
        g.step(nsteps // 4)
        g.sync()
        tf = (time.perf_counter() - t3) / (nsteps // 4)
    cost = (tf * r) / res["closure"]["s_per_step"]
    return dict(workload=f"fine {n_fine}^3 -> coarse {nc}^3 (box filter of {r}^3 cells), Hall d_i={di}, "
                         f"nu=eta={visc}, seeded k^-5/3 shells 1-8 rms {amp}; trained on seeds {list(train_seeds)}, "
                         f"evaluated on seed {test_seed} after {nsteps} fine steps",
                samples=int(X.shape[0]), train_steps=train_steps,
                loss_first=float(hist[:50].mean()), loss_last=float(hist[-50:].mean()),
                shells=f"1..{kmax}", plain=res["plain"], closure=res["closure"],
                cost_ratio_fine_over_coarse_closure=round(float(cost), 2))


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train the closure on box-filtered fine runs and compare")
    ap.add_argument("--n-fine", type=int, default=128)
    ap.add_argument("--ratio", type=int, default=4)
    ap.add_argument("--steps", type=int, default=384, help="fine steps (the coarse run takes steps / ratio)")
    ap.add_argument("--every", type=int, default=32, help="fine steps between training snapshots")
    ap.add_argument("--visc", type=float, default=5e-3)
    ap.add_argument("--train-steps", type=int, default=4000)
    a = ap.parse_args()
    print(json.dumps(experiment(n_fine=a.n_fine, r=a.ratio, nsteps=a.steps, every=a.every, visc=a.visc,
                                train_steps=a.train_steps, log=lambda m: print(m, file=sys.stderr, flush=True))))
