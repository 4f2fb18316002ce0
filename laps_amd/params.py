"""ctypes mirror of ``laps_params`` (include/laps.h) and the reference defaults.

Field meanings follow the reference namelists (src_compressible/mhd.f90:30-41);
defaults follow the module declarations (mhdinit.f90:5-27, dealiasing.f90:8-10,
AEBmod.f90:10-12, mhd.f90:21).
"""
import ctypes
import math

LAPS_FP64 = 0
LAPS_MIXED = 1

LAPS_OK = 0
LAPS_EINVAL = -1
LAPS_ENOMEM = -2
LAPS_EDEVICE = -3
LAPS_ESTATE = -4

LAPS_RUNNING = 0
LAPS_DT_STOP = 1


class LapsParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("adiabatic_index", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("dealias_option", ctypes.c_int32),
        ("if_resis", ctypes.c_int32), ("if_resis_exp", ctypes.c_int32),
        ("if_visc", ctypes.c_int32), ("if_visc_exp", ctypes.c_int32),
        ("if_conserve_background", ctypes.c_int32),
        ("afx", ctypes.c_double), ("afy", ctypes.c_double), ("afz", ctypes.c_double),
        ("resistivity", ctypes.c_double), ("viscosity", ctypes.c_double),
        ("if_hall", ctypes.c_int32), ("if_aeb", ctypes.c_int32),
        ("if_corotating", ctypes.c_int32), ("device", ctypes.c_int32),
        ("ion_inertial_length", ctypes.c_double),
        ("radius0", ctypes.c_double), ("Ur0", ctypes.c_double),
        ("corotating_angle", ctypes.c_double),
    ]


def default_params(**kw):
    """Reference module defaults, overridden by keyword arguments."""
    p = LapsParams()
    p.nx, p.ny, p.nz = 128, 128, 64          # mhdinit.f90:5
    p.precision = LAPS_FP64
    p.Lx = p.Ly = p.Lz = 2.0 * math.pi
    p.adiabatic_index = 5.0 / 3.0            # mhdinit.f90:21
    p.cfl = 0.5                              # mhd.f90:21
    p.dealias_option = 2                     # dealiasing.f90:9
    p.afx = p.afy = p.afz = 0.495            # dealiasing.f90:10
    p.radius0 = 30.0                         # AEBmod.f90:12
    p.Ur0 = 0.0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"laps_params has no field {k!r}")
        setattr(p, k, v)
    return p


def as_dict(p):
    return {name: getattr(p, name) for name, _ in LapsParams._fields_}
