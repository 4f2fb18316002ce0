"""Learned subgrid closure: training data from box-filtered fine-grid runs that the
solver itself produces, normalisation, training on the GPU (laps_mlp_train) and the
parameter vector for Laps.closure_set.

No counterpart in the reference. The coarse-grid state is the box filter (mean over
r^3 fine cells) of the fine state. The network learns, per coarse point, the
subgrid terms of the filtered momentum, induction and energy equations:

    R_ij = bar(m_i u_j) - M_i U_j          (Reynolds stress, M = bar(m), U = M / bar(rho))
    M_ij = bar(B_i B_j) - B_i B_j          (Maxwell stress, B = bar(B))
    EMF  = bar(E) - E(bar state),  E = -u x B + (d_i / rho) J x B
    F_e  = bar(F) - F(bar state),  F = (e + p + B^2/2) u - (u.B) B   (mhdrhs.f90:87-89)

with the features the library computes on the coarse state (rho, u, B and their
central-difference gradients).
"""
import math

import numpy as np

from .runtime import Laps, mlp_train, MLP_NPARAM, CLOSURE_NPARAM

NIN, NH, NOUT = 28, 64, 18
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


def _cut_modes(n, nc):
    """signed modes |m| < nc/2 (the coarse grid's Nyquist row dropped) as (fine, coarse) indices"""
    h = nc // 2
    m = np.concatenate([np.arange(0, h), np.arange(-(h - 1), 0)]) if h > 1 else np.zeros(1, dtype=np.int64)
    return m % n, m % nc


def spectral_cut(f, nc, fine_grid=False):
    """Sharp spectral cut-off of a real (nx, ny, nz) field: the modes kx < ncx/2, |ky| < ncy/2,
    |kz| < ncz/2, taken to the points of the (ncx, ncy, ncz) grid, or with fine_grid back to the
    points of the fine grid (numpy restatement of laps_closure_filter_spectral's cut)."""
    nx, ny, nz = f.shape
    ncx, ncy, ncz = nc
    F = np.fft.rfftn(f, axes=(1, 2, 0)) / (nx * ny * nz)
    kx = np.arange(ncx // 2)
    yf, yc = _cut_modes(ny, ncy)
    zf, zc = _cut_modes(nz, ncz)
    if fine_grid:
        G = np.zeros_like(F)
        G[np.ix_(kx, yf, zf)] = F[np.ix_(kx, yf, zf)]
        return np.fft.irfftn(G, s=(ny, nz, nx), axes=(1, 2, 0)) * (nx * ny * nz)
    G = np.zeros((ncx // 2 + 1, ncy, ncz), dtype=complex)
    G[np.ix_(kx, yc, zc)] = F[np.ix_(kx, yf, zf)]
    return np.fft.irfftn(G, s=(ncy, ncz, ncx), axes=(1, 2, 0)) * (ncx * ncy * ncz)


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


def energy_flux(e, u, B, p):
    """(e + p + B^2/2) u - (u.B) B (mhdrhs.f90:87-89)"""
    ptot = p + 0.5 * (B * B).sum(-1)
    return (e + ptot)[..., None] * u - (u * B).sum(-1)[..., None] * B


def filtered_sample(uu, prim, r, gamma, di, L):
    """Coarse state (uu_c, prim_c) of a fine state and the (Nc, 18) subgrid targets,
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
    Ff = box_filter(energy_flux(uu[..., 7], u, B, prim[..., 3]), r)
    F_c = energy_flux(e_c, u_c, B_c, p_c)
    for a in range(3):
        T.append(Ff[..., a] - F_c[..., a])
    T = np.stack([t.reshape(-1, order="F") for t in T], axis=1)
    return uu_c, prim_c, T


def coarse_state(f, gamma):
    """Conserved and primitive coarse fields from the first 8 filtered fields (rho, m, B, e)"""
    uu_c = np.asfortranarray(f[..., 0:8])
    rho_c, m_c, B_c, e_c = uu_c[..., 0], uu_c[..., 1:4], uu_c[..., 4:7], uu_c[..., 7]
    u_c = m_c / rho_c[..., None]
    p_c = (gamma - 1.0) * (e_c - 0.5 * (m_c * u_c).sum(-1) - 0.5 * (B_c * B_c).sum(-1))
    return uu_c, np.asfortranarray(np.concatenate([u_c, p_c[..., None]], axis=-1))


def filtered_sample_device(g_fine, g_coarse, r, spectral=False):
    """filtered_sample formed with the library on the device: the box filter of the fine state
    and of its flux products (laps_closure_filter on the fine context, with J = curl B from its
    own transforms), and the same products of the filtered state (the filter of width 1 on the
    coarse context, loaded with the filtered state). Targets are their differences: filter of
    the products minus products of the filter. Returns the coarse state, the raw features of
    the coarse context and the (Nc, 18) targets; the coarse context keeps the filtered state.
    With spectral, the filter is the sharp cut-off of the coarse grid instead
    (laps_closure_filter_spectral), and the products of the cut state are formed on the fine
    grid, free of aliasing."""
    if spectral:
        C, D = g_fine.closure_filter_spectral(g_coarse)
        uu_c, prim_c = coarse_state(C, g_fine.params.adiabatic_index)
        g_coarse.set_state(uu_c, prim_c)
        T = (C[..., 8:26] - D[..., 8:26]).reshape(-1, 18, order="F")
        return uu_c, prim_c, g_coarse.closure_features(), T
    f = g_fine.closure_filter(r)
    uu_c, prim_c = coarse_state(f, g_fine.params.adiabatic_index)
    g_coarse.set_state(uu_c, prim_c)
    f1 = g_coarse.closure_filter(1)
    T = (f[..., 8:26] - f1[..., 8:26]).reshape(-1, 18, order="F")
    return uu_c, prim_c, g_coarse.closure_features(), T


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
    """Trains the network on raw features X (n, 28) and targets T (n, 18): returns the
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


def initial_state(p_fine, seed, amp, ic):
    """Seeded fine initial state: 'kolmogorov' (k^-5/3 in shells 1-8, du = db = amp) or
    'alfvenic' (the expanding-box workload's kind without expansion: outward Alfven waves,
    amplitude k^-1 in shells 1..n/16, db/B0 = amp, generated on a quarter-size grid and brought
    to the fine grid by the library's inverse transform)."""
    from .ic import band_limited_state, cons_from_prim, seeded_primitive
    n = p_fine.nx
    if ic == "alfvenic":
        kmax = n // 16
        with Laps(p_fine) as g:
            return band_limited_state(n, n // 4, p_fine.adiabatic_index, g.fft_c2r, seed=seed, kmin=1, kmax=kmax,
                                      slope=-1.0, du=amp, db=amp, alfvenic=True)
    prim = seeded_primitive((n, p_fine.ny, p_fine.nz), seed, kmin=1, kmax=8, du=amp, db=amp)
    return cons_from_prim(prim, p_fine.adiabatic_index)


def check_finite(what, *arrays):
    for a in arrays:
        if not np.all(np.isfinite(a)):
            raise FloatingPointError(f"closure experiment: non-finite values in the {what} data "
                                     f"(the fixed dt or the viscosity does not hold this run)")


def fine_samples(p_fine, p_coarse, state, r, nsteps, every, dt, spectral=False):
    """Training samples of a fine run at a fixed dt, every `every` steps: (step, uu_c, prim_c,
    X, T) with the filtered state, its raw features and the subgrid targets, all formed on the
    device (filtered_sample_device); the fine state never leaves the GPU."""
    with Laps(p_fine) as g, Laps(p_coarse) as gc:
        g.set_state(*state)
        g.set_fixed_dt(dt)
        for s in range(0, nsteps + 1, every):
            if s:
                g.step(every)
            yield (s,) + filtered_sample_device(g, gc, r, spectral)


def experiment(n_fine=128, r=4, train_seeds=(11, 12), test_seed=21, nsteps=384, every=32, amp=0.4,
               di=0.05, visc=5e-3, train_steps=4000, ic="kolmogorov", dt_frac=0.8, filt="box", log=print):
    """Trains the closure on filtered fine runs (box filter, or with filt="spectral" the sharp
    cut-off of the coarse grid) and compares coarse runs with and
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
    # the fine dt: dt_frac times the CFL dt of the first training state, kept for every run
    t0 = time.perf_counter()
    state0 = initial_state(pf, train_seeds[0], amp, ic)
    with Laps(pf) as g:
        g.set_state(*state0)
        dt = dt_frac * g.cfl()[1]
    X, T = [], []
    for k, sd in enumerate(train_seeds):
        st = state0 if k == 0 else initial_state(pf, sd, amp, ic)
        for s, uc, qc, x, t in fine_samples(pf, pc, st, r, nsteps, every, dt, filt == "spectral"):
            check_finite(f"seed {sd}, fine step {s}", uc, x, t)
            X.append(x)
            T.append(t.astype(np.float32))
            log(f"closure: seed {sd}, fine step {s}: {x.shape[0]} samples, filtered energy "
                f"{0.5 * float(np.mean(uc[..., 1:7] ** 2)):.5e} ({time.perf_counter() - t0:.1f} s)")
        del st
    del state0
    X = np.concatenate(X)
    T = np.concatenate(T)
    log(f"closure: {X.shape[0]} training samples from {len(train_seeds)} fine {n_fine}^3 runs "
        f"({time.perf_counter() - t0:.1f} s, filtered on the device)")
    t1 = time.perf_counter()
    vec, hist = train(X, T, steps=train_steps)
    log(f"closure: trained {train_steps} Adam steps in {time.perf_counter() - t1:.1f} s, "
        f"loss {hist[:50].mean():.3f} -> {hist[-50:].mean():.3f}")
    del X, T
    # evaluation on an unseen seed: the filtered fine states at the start and after nsteps
    log(f"closure: evaluation run, seed {test_seed}, {nsteps} fine steps")
    snaps = []
    for snap in fine_samples(pf, pc, initial_state(pf, test_seed, amp, ic), r, nsteps, max(1, nsteps // 4), dt,
                             filt == "spectral"):
        check_finite(f"evaluation seed {test_seed}, fine step {snap[0]}", snap[1])
        snaps.append(snap[:3])
        log(f"closure: evaluation run at fine step {snap[0]} ({time.perf_counter() - t0:.1f} s)")
    (_, uc0, qc0), (_, uc1, qc1) = snaps[0], snaps[-1]
    kmax = nc // 3
    with Laps(pc) as ref_g:
        ref_g.set_state(uc1, qc1)
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
        check_finite(f"coarse {tag} run", e)
        dev = np.abs(e[1:] - e_ref[1:]) / e_ref[1:]
        off = ~(dev <= 0.05)
        within = int(np.argmax(off)) if np.any(off) else len(dev)
        res[tag] = dict(max_shell_dev=float(dev.max()), mean_shell_dev=float(dev.mean()),
                        shells_within_5pct=within, shell_dev=[round(float(x), 4) for x in dev], s_per_step=tc)
        log(f"closure: coarse {nc}^3 {tag}: max shell deviation {dev.max():.4f}, mean {dev.mean():.4f}, "
            f"within 5 % up to shell {within}")
    with Laps(pf) as g:
        g.set_state(*initial_state(pf, test_seed, amp, ic))
        g.set_fixed_dt(dt)
        g.step(4)
        g.sync()
        t3 = time.perf_counter()
        nf = max(8, min(nsteps // 4, 100))
        g.step(nf)
        g.sync()
        tf = (time.perf_counter() - t3) / nf
    cost = (tf * r) / res["closure"]["s_per_step"]
    icname = ("seeded k^-5/3 shells 1-8 rms" if ic != "alfvenic" else
              f"outward Alfven waves k^-1 shells 1-{n_fine // 16}, db/B0 =")
    fname = f"box filter of {r}^3 cells" if filt != "spectral" else "sharp spectral cut-off of the coarse grid"
    return dict(workload=f"fine {n_fine}^3 -> coarse {nc}^3 ({fname}), Hall d_i={di}, "
                         f"nu=eta={visc}, {icname} {amp}; trained on seeds {list(train_seeds)}, "
                         f"evaluated on seed {test_seed} after {nsteps} fine steps ({nsteps // r} coarse)",
                samples=int(sum(1 for _ in range(0, nsteps + 1, every)) * len(train_seeds) * nc ** 3),
                dt_fine=dt, dt_coarse=r * dt,
                train_steps=train_steps, loss_first=float(hist[:50].mean()), loss_last=float(hist[-50:].mean()),
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
    ap.add_argument("--amp", type=float, default=0.4)
    ap.add_argument("--di", type=float, default=0.05)
    ap.add_argument("--ic", choices=["kolmogorov", "alfvenic"], default="kolmogorov")
    ap.add_argument("--dt-frac", type=float, default=0.8, help="fixed fine dt as a fraction of the first CFL dt")
    ap.add_argument("--filter", choices=["box", "spectral"], default="box")
    a = ap.parse_args()
    print(json.dumps(experiment(n_fine=a.n_fine, r=a.ratio, nsteps=a.steps, every=a.every, visc=a.visc,
                                train_steps=a.train_steps, amp=a.amp, di=a.di, ic=a.ic, dt_frac=a.dt_frac,
                                filt=a.filter,
                                log=lambda m: print(m, file=sys.stderr, flush=True))))
