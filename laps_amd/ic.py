"""Seeded initial conditions for the benchmark and parity workloads.

The reference's own perturbation hooks (mhdinit.f90:302-1036) draw phases
from the compiler's RNG; these workloads instead use a counter-based hash
(splitmix64 of seed, mode and component), so every run and every platform
builds the same field. Closest reference hook: ipert = 7
(mhdinit.f90:695-829, random-phase modes in shells).

Box [0, L)^3, uniform background rho0, P0, B0 (ifield = 3 style,
mhdinit.f90:251-256) plus random-phase Fourier modes with integer
|k| in [kmin, kmax] and amplitude |k|^slope per mode; db is made
solenoidal by projection, and du, db are rescaled to the requested rms.
Returns primitive fields (rho, ux, uy, uz, Bx, By, Bz, p) in the reference's
(nx,ny,nz,8) Fortran layout, before initial_calc_conserve_variable.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def _uniform(seed, mx, my, mz, comp):
    """uniform [0,1) from hash(seed, mode, component)"""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) & _M64
        key = _splitmix(key ^ (mx.astype(np.int64).astype(np.uint64) & np.uint64(0xFFFF)))
        key = _splitmix(key ^ ((my.astype(np.int64).astype(np.uint64) & np.uint64(0xFFFF)) << np.uint64(16)))
        key = _splitmix(key ^ ((mz.astype(np.int64).astype(np.uint64) & np.uint64(0xFFFF)) << np.uint64(32)))
        key = _splitmix(key ^ (np.uint64(comp) << np.uint64(48)))
    return (key >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def _random_vector_field(n, seed, kmin, kmax, slope, solenoidal, salt):
    nx, ny, nz = n
    mx = np.arange(nx // 2 + 1)[:, None, None] * np.ones((1, ny, nz), dtype=np.int64)
    my = np.fft.fftfreq(ny, 1.0 / ny).astype(np.int64)[None, :, None] * np.ones((nx // 2 + 1, 1, nz), dtype=np.int64)
    mz = np.fft.fftfreq(nz, 1.0 / nz).astype(np.int64)[None, None, :] * np.ones((nx // 2 + 1, ny, 1), dtype=np.int64)
    kk = np.sqrt(mx.astype(float) ** 2 + my ** 2 + mz ** 2)
    shell = np.rint(kk)
    sel = (shell >= kmin) & (shell <= kmax)
    amp = np.where(sel, np.maximum(kk, 1e-30) ** slope, 0.0)
    comps = []
    for c in range(3):
        ph = 2 * np.pi * _uniform(seed, mx, my, mz, 2 * c + salt)
        mag = np.sqrt(-2 * np.log(1.0 - _uniform(seed, mx, my, mz, 2 * c + 1 + salt) * 0.999999))
        comps.append(amp * mag * np.exp(1j * ph))
    v = np.stack(comps, -1)
    if solenoidal:
        kvec = np.stack([mx, my, mz], -1).astype(float)
        k2 = np.maximum((kvec ** 2).sum(-1, keepdims=True), 1e-30)
        v = v - kvec * ((kvec * v).sum(-1, keepdims=True) / k2)
    out = np.empty((nx, ny, nz, 3))
    for c in range(3):
        # numpy's irfftn over (z, y, x) of the transposed half spectrum
        out[..., c] = np.fft.irfftn(v[..., c].T, s=(nz, ny, nx), axes=(0, 1, 2)).T
    return out


def seeded_primitive(n, seed, kmin=1, kmax=8, slope=-5.0 / 3.0, du=0.1, db=0.1, rho0=1.0, p0=1.0,
                     b0=(1.0, 0.0, 0.0), alfvenic=False):
    """Primitive (rho,u,B,p) fields of shape (nx,ny,nz,8), Fortran order."""
    nx, ny, nz = n
    prim = np.zeros((nx, ny, nz, 8), order="F")
    prim[..., 0] = rho0
    prim[..., 7] = p0
    b = _random_vector_field(n, seed, kmin, kmax, slope, True, salt=0)
    brms = np.sqrt((b ** 2).sum(-1).mean())
    if brms > 0:
        b *= db / brms
    if alfvenic:
        u = -b / np.sqrt(rho0)
    else:
        u = _random_vector_field(n, seed, kmin, kmax, slope, False, salt=16)
        urms = np.sqrt((u ** 2).sum(-1).mean())
        if urms > 0:
            u *= du / urms
    prim[..., 1:4] = u
    for c in range(3):
        prim[..., 4 + c] = b0[c] + b[..., c]
    return prim


def cons_from_prim(prim, gamma):
    """initial_calc_conserve_variable (mhdinit.f90:1038-1056)."""
    prim = np.asfortranarray(prim, dtype=np.float64)
    uu = np.array(prim, order="F", copy=True)
    pr = np.zeros(prim.shape[:3] + (4,), order="F")
    pr[..., 0:3] = prim[..., 1:4]
    pr[..., 3] = prim[..., 7]
    uu[..., 1] = uu[..., 0] * pr[..., 0]
    uu[..., 2] = uu[..., 0] * pr[..., 1]
    uu[..., 3] = uu[..., 0] * pr[..., 2]
    uu[..., 7] = pr[..., 3] / (gamma - 1) + 0.5 * (
        uu[..., 0] * (pr[..., 0] ** 2 + pr[..., 1] ** 2 + pr[..., 2] ** 2)
        + uu[..., 4] ** 2 + uu[..., 5] ** 2 + uu[..., 6] ** 2)
    return uu, pr


def band_limited_state(n, m, gamma, fft_c2r, **seeded_kw):
    """Conserved and primitive fields (uu, uu_prim) on the n^3 grid of a seeded state whose
    modes all lie in shells up to kmax <= m/2 - 1: seeded_primitive on m^3 (same hash, same
    rms), its spectrum zero-padded to n^3 and brought to real space by `fft_c2r` (a context's
    laps_fft_c2r on n^3, seconds instead of minutes of numpy at 512^3). rho = p = 1 and B0 =
    (1, 0, 0), as seeded_primitive's defaults; the fluctuations of u and B are transformed."""
    pc = seeded_primitive((m, m, m), **seeded_kw)
    idx = np.arange(m)
    idx = np.where(idx <= m // 2, idx, idx - m) % n
    spec = np.zeros((n // 2 + 1, n, n, 6), dtype=np.complex128, order="F")
    for f, (col, off) in enumerate(((1, 0.0), (2, 0.0), (3, 0.0), (4, 1.0), (5, 0.0), (6, 0.0))):
        F = np.fft.rfftn(pc[..., col] - off, axes=(2, 1, 0)) / m ** 3
        spec[np.ix_(np.arange(m // 2 + 1), idx, idx, [f])] = F[..., None]
    del pc
    real = fft_c2r(spec)
    del spec
    prim = np.zeros((n, n, n, 8), order="F")
    prim[..., 0] = 1.0
    prim[..., 7] = 1.0
    prim[..., 1:4] = real[..., 0:3]
    prim[..., 4:7] = real[..., 3:6]
    prim[..., 4] += 1.0
    del real
    return cons_from_prim(prim, gamma)
