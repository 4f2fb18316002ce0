"""Reader/writer of the reference's output and restart file format.

outNNN.dat (mhdoutput.f90:72-131, restart.f90:17-63): one Fortran
unformatted record holding the time as float32 ([int32 4][float32 t]
[int32 4]), then the raw (nx,ny,nz,8) float64 array in Fortran order
starting at byte 12, primitive variables (rho,ux,uy,uz,Bx,By,Bz,p) unless
output_primitive is false. Format as read by
data_process/3D_Python/read_output.py:76-91.
"""
import numpy as np


def write_out(path, t, fields):
    """fields: (nx,ny,nz,8) float64."""
    a = np.asfortranarray(fields, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(np.array([4], dtype="<i4").tobytes())
        f.write(np.array([t], dtype="<f4").tobytes())
        f.write(np.array([4], dtype="<i4").tobytes())
        f.write(a.tobytes(order="F"))


def read_out(path, shape):
    """-> (time as float, (nx,ny,nz,8) float64 array)."""
    nx, ny, nz = shape
    with open(path, "rb") as f:
        head = f.read(12)
        m1, = np.frombuffer(head[:4], dtype="<i4")
        t, = np.frombuffer(head[4:8], dtype="<f4")
        m2, = np.frombuffer(head[8:12], dtype="<i4")
        if m1 != 4 or m2 != 4:
            raise ValueError(f"{path}: not an outNNN.dat file (record markers {m1}, {m2})")
        data = np.frombuffer(f.read(), dtype="<f8")
    if data.size != nx * ny * nz * 8:
        raise ValueError(f"{path}: expected {nx * ny * nz * 8} values, found {data.size}")
    return float(t), data.reshape((nx, ny, nz, 8), order="F")


def read_rms(path):
    """rms.dat rows (mhdrms.f90:24-25, :48): t, 8 means, 8 variances, 3 <rho (u-<u>)^2>."""
    return np.loadtxt(path, ndmin=2)
