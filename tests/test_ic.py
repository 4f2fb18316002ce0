"""Host-side initial conditions (laps_amd/ic.py)."""
import numpy as np

from laps_amd.ic import band_limited_state, cons_from_prim, seeded_primitive


def _c2r_numpy(spec):
    """laps_fft_c2r's convention with numpy: unnormalised inverse of (nx/2+1, ny, nz, nf)
    coefficients (FFT / N) to (nx, ny, nz, nf) real fields"""
    nxh, ny, nz, nf = spec.shape
    nx = 2 * (nxh - 1)
    out = np.empty((nx, ny, nz, nf), order="F")
    for f in range(nf):
        out[..., f] = np.fft.irfftn(spec[..., f].T, s=(nz, ny, nx), axes=(0, 1, 2)).T * (nx * ny * nz)
    return out


def test_band_limited_state_equals_direct_generation():
    # the same seeded modes generated on 16^3 and zero-padded to 32^3 give the state generated
    # on 32^3 directly (the rms of a band-limited field is the same on both grids)
    kw = dict(seed=5, kmin=1, kmax=7, slope=-1.0, du=0.5, db=0.5, alfvenic=True)
    uu, pr = band_limited_state(32, 16, 5.0 / 3.0, _c2r_numpy, **kw)
    uu0, pr0 = cons_from_prim(seeded_primitive((32, 32, 32), **kw), 5.0 / 3.0)
    assert np.abs(uu - uu0).max() < 1e-13
    assert np.abs(pr - pr0).max() < 1e-13
