"""numpy restatement of the learned closure (the test reference for the HIP kernels in
laps_amd/csrc/closure.hip): network forward and mean-squared-error gradient in
float64, the 28 features, the correction fields and their spectral RHS term."""
import numpy as np

from .helpers import wave_numbers

NIN, NH, NOUT = 28, 64, 18
PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    out, o = [], 0
    for a, b in ((NIN, NH), (NH, NH), (NH, NOUT)):
        out.append(P[o:o + a * b].reshape(a, b))
        o += a * b
        out.append(P[o:o + b])
        o += b
    return out


def forward(X, P):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ W3 + b3, h1, h2


def loss_grad(X, T, P):
    """loss = mean (Y - T)^2 over samples and outputs, and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    Y, h1, h2 = forward(X, P)
    n = X.shape[0]
    dY = 2.0 * (Y - T) / (n * NOUT)
    gW3, gb3 = h2.T @ dY, dY.sum(0)
    d2 = (dY @ W3.T) * (1 - h2 ** 2)
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * (1 - h1 ** 2)
    gW1, gb1 = X.T @ d1, d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(np.mean((Y - T) ** 2)), g


def features(uu, prim, L):
    """(N, 28) raw features, points in Fortran order."""
    n = uu.shape[:3]
    d = [L[a] / n[a] for a in range(3)]
    fields = [uu[..., 0], prim[..., 0], prim[..., 1], prim[..., 2], uu[..., 4], uu[..., 5], uu[..., 6]]
    cols = list(fields)
    grads = []
    for f in fields:
        for a in range(3):
            grads.append((np.roll(f, -1, axis=a) - np.roll(f, 1, axis=a)) / (2 * d[a]))
    cols += grads
    return np.stack([c.reshape(-1, order="F") for c in cols], axis=1)


def terms(Yn, ystd, gamma):
    """(N, 12): momentum-flux correction xx, xy, xz, yy, yz, zz, the EMF correction and the
    energy-flux correction."""
    o = np.asarray(Yn, np.float64) * np.asarray(ystd, np.float64)
    trR = o[:, 0] + o[:, 3] + o[:, 5]
    trM = o[:, 6] + o[:, 9] + o[:, 11]
    q = -0.5 * (gamma - 1.0) * (trR + trM) + 0.5 * trM
    C = np.empty((o.shape[0], 12))
    for c in range(6):
        C[:, c] = o[:, c] - o[:, 6 + c] + (q if c in (0, 3, 5) else 0.0)
    C[:, 6:12] = o[:, 12:18]
    return C


def rhs(C, shape, L):
    """Spectral RHS term (nx/2+1, ny, nz, 7): -i k.tau for momentum, -i k x E for B and
    -i k.F for the energy, with spectra normalised by 1/N like the solver's."""
    nx, ny, nz = shape
    N = nx * ny * nz
    F = [np.fft.rfftn(C[:, c].reshape(shape, order="F"), axes=(2, 1, 0)) / N for c in range(12)]
    Kx, Ky, Kz = np.meshgrid(wave_numbers(nx, L[0])[: nx // 2 + 1], wave_numbers(ny, L[1]),
                             wave_numbers(nz, L[2]), indexing="ij")
    K = (Kx, Ky, Kz)
    idx = ((0, 1, 2), (1, 3, 4), (2, 4, 5))
    out = np.empty(F[0].shape + (7,), complex)
    for a in range(3):
        out[..., a] = -1j * sum(K[b] * F[idx[a][b]] for b in range(3))
    E = F[6:9]
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        out[..., 3 + a] = -1j * (K[b] * E[c] - K[c] * E[b])
    out[..., 6] = -1j * sum(K[b] * F[9 + b] for b in range(3))
    return out
