"""ctypes binding of laps_amd/liblaps_amd.so (the C ABI of include/laps.h).

Arrays are numpy arrays in the reference's Fortran layouts: real fields
(nx,ny,nz,nvar) float64 and spectral fields (nx/2+1,ny,nz,nvar) complex128.
The binding fails loudly when the HIP library is missing: there is no CPU
fallback behind it.
"""
import ctypes
import os

import numpy as np

from .params import LapsParams, default_params, LAPS_FP64, LAPS_MIXED, LAPS_RUNNING, LAPS_DT_STOP  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblaps_amd.so")
_lib = None
_D = ctypes.POINTER(ctypes.c_double)

# every entry point declared in include/laps.h
EXPORTS = (
    "laps_default_params", "laps_create", "laps_destroy", "laps_last_error", "laps_set_state",
    "laps_set_time", "laps_set_fixed_dt", "laps_cfl", "laps_step", "laps_get_scalars", "laps_get_state",
    "laps_stage_output", "laps_staged_output",
    "laps_get_spectral", "laps_moments", "laps_budgets", "laps_max_divb", "laps_fft_r2c", "laps_fft_c2r", "laps_sync",
    "laps_timing_enable", "laps_timing_read", "laps_get_stream",
    "laps_ens_create", "laps_ens_destroy", "laps_ens_size", "laps_ens_groups", "laps_ens_box", "laps_ens_step", "laps_ens_sync",
    "laps_closure_set", "laps_closure_features", "laps_closure_eval", "laps_closure_filter",
    "laps_closure_filter_spectral", "laps_mlp_forward", "laps_mlp_grad",
    "laps_mlp_train", "laps_mlp_grad_noalloc",
)

MLP_NPARAM = 7186          # LAPS_MLP_NPARAM
MLP_NIN, MLP_NOUT = 28, 18  # network inputs and outputs per sample
CLOSURE_NPARAM = 7186 + 74  # LAPS_CLOSURE_NPARAM
_F = ctypes.POINTER(ctypes.c_float)


class LapsError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LapsError(f"{LIB_PATH} is not built: run `python -m laps_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.laps_default_params.argtypes = [ctypes.POINTER(LapsParams)]
        L.laps_create.argtypes = [ctypes.POINTER(LapsParams), ctypes.POINTER(vp)]
        L.laps_destroy.argtypes = [vp]
        L.laps_destroy.restype = None
        L.laps_last_error.restype = ctypes.c_char_p
        L.laps_set_state.argtypes = [vp, _D, _D]
        L.laps_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.laps_set_fixed_dt.argtypes = [vp, ctypes.c_double]
        L.laps_cfl.argtypes = [vp, _D, _D]
        L.laps_step.argtypes = [vp, ctypes.c_int]
        L.laps_get_scalars.argtypes = [vp, _D, _D, _D, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]
        L.laps_get_state.argtypes = [vp, _D, _D]
        L.laps_get_spectral.argtypes = [vp, _D]
        L.laps_stage_output.argtypes = [vp, ctypes.c_int]
        L.laps_staged_output.argtypes = [vp, ctypes.POINTER(_D), _D, _D, ctypes.POINTER(ctypes.c_int64)]
        L.laps_moments.argtypes = [vp, _D, _D, _D]
        L.laps_max_divb.argtypes = [vp, _D]
        L.laps_budgets.argtypes = [vp, _D]
        L.laps_fft_r2c.argtypes = [vp, ctypes.c_int, _D, _D]
        L.laps_fft_c2r.argtypes = [vp, ctypes.c_int, _D, _D]
        L.laps_sync.argtypes = [vp]
        L.laps_timing_enable.argtypes = [vp, ctypes.c_int]
        L.laps_timing_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _D, _D,
                                       ctypes.POINTER(ctypes.c_int64)]
        L.laps_get_stream.argtypes = [vp, ctypes.POINTER(ctypes.c_void_p)]
        L.laps_ens_create.argtypes = [ctypes.c_int, ctypes.POINTER(LapsParams), ctypes.POINTER(vp)]
        L.laps_ens_destroy.argtypes = [vp]
        L.laps_ens_destroy.restype = None
        L.laps_ens_size.argtypes = [vp]
        L.laps_ens_groups.argtypes = [vp]
        L.laps_ens_box.argtypes = [vp, ctypes.c_int]
        L.laps_ens_box.restype = vp
        L.laps_ens_step.argtypes = [vp, ctypes.c_int]
        L.laps_ens_sync.argtypes = [vp]
        L.laps_closure_set.argtypes = [vp, _F, ctypes.c_int]
        L.laps_closure_features.argtypes = [vp, _F]
        L.laps_closure_eval.argtypes = [vp, _F, _D]
        L.laps_closure_filter.argtypes = [vp, ctypes.c_int, _D]
        L.laps_closure_filter_spectral.argtypes = [vp, vp, _D]
        L.laps_mlp_forward.argtypes = [ctypes.c_int, ctypes.c_int64, _F, _F, _F]
        L.laps_mlp_grad.argtypes = [ctypes.c_int, ctypes.c_int64, _F, _F, _F, _F, _F]
        L.laps_mlp_train.argtypes = [ctypes.c_int, ctypes.c_int64, _F, _F, _F, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, _F]
        L.laps_mlp_grad_noalloc.argtypes = [ctypes.c_int, ctypes.c_int64] + [vp] * 6
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().laps_last_error().decode(errors="replace")
        raise LapsError(f"laps error {rc}: {msg}")


def _fptr(a):
    return a.ctypes.data_as(_D)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def _sptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _samples(X, T=None):
    """the network's samples as contiguous FP32 arrays, shapes checked: X (n, 28), T (n, 18)"""
    X = _f32(X)
    if X.ndim != 2 or X.shape[1] != MLP_NIN:
        raise LapsError(f"network inputs must be (n, {MLP_NIN}), got {X.shape}")
    if T is None:
        return X
    T = _f32(T)
    if T.shape != (X.shape[0], MLP_NOUT):
        raise LapsError(f"network targets must be ({X.shape[0]}, {MLP_NOUT}), got {T.shape}")
    return X, T


def mlp_forward(X, params, device=0):
    """Closure network on samples X (n, 28) -> (n, 18)."""
    X = _samples(X)
    P = _f32(params, (MLP_NPARAM,))
    Y = np.zeros((X.shape[0], MLP_NOUT), np.float32)
    _check(lib().laps_mlp_forward(device, X.shape[0], _sptr(X), _sptr(P), _sptr(Y)))
    return Y


def mlp_grad(X, T, params, device=0):
    """Gradient of the mean squared error over all samples, and the loss."""
    X, T = _samples(X, T)
    P = _f32(params, (MLP_NPARAM,))
    G = np.zeros(MLP_NPARAM, np.float32)
    loss = ctypes.c_float()
    _check(lib().laps_mlp_grad(device, X.shape[0], _sptr(X), _sptr(T), _sptr(P), _sptr(G), ctypes.byref(loss)))
    return G, float(loss.value)


def mlp_train(X, T, params, steps, batch, lr, device=0):
    """Adam steps over consecutive batches of the (already shuffled) samples; returns the new
    parameters and the loss of every batch."""
    X, T = _samples(X, T)
    P = _f32(params, (MLP_NPARAM,)).copy()
    hist = np.zeros(max(steps, 1), np.float32)
    _check(lib().laps_mlp_train(device, X.shape[0], _sptr(X), _sptr(T), _sptr(P), int(steps), int(batch),
                                float(lr), _sptr(hist)))
    return P, hist[:steps]


class Laps:
    """One periodic box advanced on the GPU through the C ABI."""

    def __init__(self, params, _handle=None, _owner=None):
        self.params = params
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz
        self.nxh = self.nx // 2 + 1
        self._owner = _owner       # an Ensemble owns its boxes' contexts
        if _handle is not None:
            self._h = ctypes.c_void_p(_handle)
            return
        h = ctypes.c_void_p()
        _check(lib().laps_create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            if self._owner is None:
                lib().laps_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _real(self, nv):
        return np.zeros((self.nx, self.ny, self.nz, nv), dtype=np.float64, order="F")

    def _spec(self, nv):
        return np.zeros((self.nxh, self.ny, self.nz, nv), dtype=np.complex128, order="F")

    def set_state(self, uu, prim):
        uu = np.asfortranarray(uu, dtype=np.float64)
        prim = np.asfortranarray(prim, dtype=np.float64)
        assert uu.shape == (self.nx, self.ny, self.nz, 8) and prim.shape == (self.nx, self.ny, self.nz, 4)
        _check(lib().laps_set_state(self._h, _fptr(uu), _fptr(prim)))

    def set_time(self, t, dt_prev):
        _check(lib().laps_set_time(self._h, float(t), float(dt_prev)))

    def set_fixed_dt(self, dt):
        _check(lib().laps_set_fixed_dt(self._h, float(dt)))

    def step(self, n=1):
        _check(lib().laps_step(self._h, int(n)))

    def sync(self):
        _check(lib().laps_sync(self._h))

    def scalars(self):
        t, dt, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ist, st = ctypes.c_int64(), ctypes.c_int32()
        _check(lib().laps_get_scalars(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(r),
                                      ctypes.byref(ist), ctypes.byref(st)))
        return dict(time=t.value, dt=dt.value, radius=r.value, istep=ist.value, status=st.value)

    def state(self):
        uu, pr = self._real(8), self._real(4)
        _check(lib().laps_get_state(self._h, _fptr(uu), _fptr(pr)))
        return uu, pr

    def stage_output(self, primitive=True):
        """queue the output payload of the current step for an asynchronous copy to pinned memory
        (laps_stage_output); later steps run while it is copied"""
        _check(lib().laps_stage_output(self._h, 1 if primitive else 0))

    def staged_output(self, copy=True):
        """the staged payload (nx, ny, nz, 8) and the time, dt and istep it belongs to; copy=False
        gives a view of the pinned buffer, valid until the next stage_output or close"""
        ptr = _D()
        t, dt, ist = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _check(lib().laps_staged_output(self._h, ctypes.byref(ptr), ctypes.byref(t), ctypes.byref(dt),
                                        ctypes.byref(ist)))
        buf = np.ctypeslib.as_array(ptr, shape=(self.nx * self.ny * self.nz * 8,))
        w = buf.reshape((self.nx, self.ny, self.nz, 8), order="F")
        return (np.array(w, order="F") if copy else w,
                t.value, dt.value, int(ist.value))

    def spectral(self):
        u = self._spec(8)
        _check(lib().laps_get_spectral(self._h, _fptr(u)))
        return u

    def cfl(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().laps_cfl(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def moments(self):
        ave, var, r = np.zeros(8), np.zeros(8), np.zeros(3)
        _check(lib().laps_moments(self._h, _fptr(ave), _fptr(var), _fptr(r)))
        return ave, var, r

    BUDGETS = ("e", "kinetic", "magnetic", "internal", "cross_helicity", "magnetic_helicity",
               "current_helicity")

    def budgets(self):
        """Volume averages: <e>, <rho u.u>/2, <B.B>/2, <p>/(gamma-1), <u.B>, <A.B>, <J.B>."""
        out = np.zeros(len(self.BUDGETS))
        _check(lib().laps_budgets(self._h, _fptr(out)))
        return dict(zip(self.BUDGETS, out.tolist()))

    def max_divb(self):
        d = ctypes.c_double()
        _check(lib().laps_max_divb(self._h, ctypes.byref(d)))
        return d.value

    def timing_enable(self, on=True):
        _check(lib().laps_timing_enable(self._h, 1 if on else 0))

    def timing(self):
        """{kernel class: (total_ms, algorithmic_bytes, launches)} since timing_enable."""
        out = {}
        i = 0
        while True:
            name = ctypes.c_char_p()
            ms, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
            _check(lib().laps_timing_read(self._h, i, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(by),
                                          ctypes.byref(n)))
            if not name.value:
                return out
            out[name.value.decode()] = (ms.value, by.value, n.value)
            i += 1

    def closure_set(self, params):
        """Learned closure on (params: CLOSURE_NPARAM floats) or off (None)."""
        if params is None:
            _check(lib().laps_closure_set(self._h, None, 0))
            return
        P = _f32(params, (CLOSURE_NPARAM,))
        _check(lib().laps_closure_set(self._h, _sptr(P), P.size))

    def closure_features(self):
        """Raw features (nx*ny*nz, 28) of the current state, points in Fortran order."""
        out = np.zeros((self.nx * self.ny * self.nz, 28), np.float32)
        _check(lib().laps_closure_features(self._h, _sptr(out)))
        return out

    def closure_eval(self):
        """Normalised network outputs (nx*ny*nz, 18) and the closure RHS term (nxh, ny, nz, 7)."""
        out = np.zeros((self.nx * self.ny * self.nz, 18), np.float32)
        rhs = self._spec(7)
        _check(lib().laps_closure_eval(self._h, _sptr(out), _fptr(rhs)))
        return out, rhs

    def closure_filter(self, r):
        """Box filter (r^3 cells) of the state and of its flux products, on the device
        (laps_closure_filter): (nx/r, ny/r, nz/r, 26) FP64"""
        out = np.zeros((self.nx // r, self.ny // r, self.nz // r, 26), order="F")
        _check(lib().laps_closure_filter(self._h, int(r), _fptr(out)))
        return out

    def closure_filter_spectral(self, coarse):
        """Sharp spectral cut-off of the state and its flux products onto the grid of the context
        `coarse` (laps_closure_filter_spectral): two (ncx, ncy, ncz, 26) FP64 arrays, the cut of
        the fine fields and the same fields formed from the cut state"""
        out = np.zeros((coarse.nx, coarse.ny, coarse.nz, 52), order="F")
        _check(lib().laps_closure_filter_spectral(self._h, coarse._h, _fptr(out)))
        return out[..., :26], out[..., 26:]

    def fft_r2c(self, real):
        real = np.asfortranarray(real, dtype=np.float64)
        nf = real.shape[3] if real.ndim == 4 else 1
        out = self._spec(nf)
        _check(lib().laps_fft_r2c(self._h, nf, _fptr(real), _fptr(out)))
        return out if real.ndim == 4 else out[..., 0]

    def fft_c2r(self, spec):
        spec = np.asfortranarray(spec, dtype=np.complex128)
        nf = spec.shape[3] if spec.ndim == 4 else 1
        out = self._real(nf)
        _check(lib().laps_fft_c2r(self._h, nf, _fptr(spec), _fptr(out)))
        return out if spec.ndim == 4 else out[..., 0]


class Ensemble:
    """Independent boxes advanced together (laps_ens_*): box(i) is a Laps view of
    box i; step(n) advances all of them: after the first step, hipGraph launches of 8
    steps (and of 1 step for the rest), boxes of identical parameters batched into
    shared launches."""

    def __init__(self, params_list):
        arr = (LapsParams * len(params_list))(*params_list)
        h = ctypes.c_void_p()
        _check(lib().laps_ens_create(len(params_list), arr, ctypes.byref(h)))
        self._h = h
        self.boxes = [Laps(p, _handle=lib().laps_ens_box(h, i), _owner=self) for i, p in enumerate(params_list)]

    def __len__(self):
        return len(self.boxes)

    def groups(self):
        """number of launch groups (boxes of identical parameters step together)"""
        return int(lib().laps_ens_groups(self._h))

    def box(self, i):
        return self.boxes[i]

    def step(self, n=1):
        _check(lib().laps_ens_step(self._h, int(n)))

    def sync(self):
        _check(lib().laps_ens_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            for b in self.boxes:
                b._h = None
            lib().laps_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
