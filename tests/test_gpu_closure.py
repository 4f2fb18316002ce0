"""Learned closure on the GPU: the HIP network (forward, gradient, Adam training),
the feature kernel and the closure RHS term against the numpy restatement
(tests/closure_np.py), and the step with a closure switched on."""
import numpy as np
import pytest

from laps_amd import closure
from tests import closure_np as ref
from tests.helpers import params, rel_l2, seeded_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


def random_params(seed):
    rng = np.random.default_rng(seed)
    P = closure.init_params(seed)
    return (P + 0.05 * rng.standard_normal(P.size)).astype(np.float32)


def test_network_forward_and_gradient_vs_numpy(rt):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3000, 28)).astype(np.float32)
    T = rng.standard_normal((3000, 18)).astype(np.float32)
    P = random_params(2)
    Y = rt.mlp_forward(X, P)
    Yr = ref.forward(X.astype(np.float64), P)[0]
    print(f"forward rel l2 {rel_l2(Y, Yr):.2e}")
    assert rel_l2(Y, Yr) < 1e-5
    G, loss = rt.mlp_grad(X, T, P)
    lr, Gr = ref.loss_grad(X.astype(np.float64), T, P)
    print(f"gradient rel l2 {rel_l2(G, Gr):.2e}, loss {loss:.6f} vs {lr:.6f}")
    assert rel_l2(G, Gr) < 1e-4
    assert loss == pytest.approx(lr, rel=1e-5)


def test_training_fits_a_teacher_network(rt):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((16384, 28)).astype(np.float32)
    teacher = random_params(4)
    T = ref.forward(X.astype(np.float64), teacher)[0].astype(np.float32)
    T /= T.std(axis=0)
    P, hist = rt.mlp_train(X, T, closure.init_params(5), steps=1500, batch=2048, lr=3e-3)
    first, last = hist[:8].mean(), hist[-8:].mean()
    print(f"training loss {first:.4f} -> {last:.5f}")
    assert last < 0.05 * first
    # the returned parameters are the trained ones
    assert ref.loss_grad(X[:2048].astype(np.float64), T[:2048], P)[0] < 0.1 * first


def test_features_vs_numpy(rt):
    p = params((32, 16, 64), Ly=3.0)
    uu, pr = seeded_state(p, seed=5, kmax=5, du=0.2, db=0.2)
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.step(2)
        X = g.closure_features()
        ug, pg = g.state()
    Xr = ref.features(ug, pg, (p.Lx, p.Ly, p.Lz))
    for q in range(28):
        assert rel_l2(X[:, q], Xr[:, q]) < 1e-6, q


def test_closure_outputs_and_rhs_vs_numpy(rt):
    p = params((32, 16, 64), Ly=3.0, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=6, kmax=5, du=0.2, db=0.2)
    rng = np.random.default_rng(7)
    P = random_params(8)
    xmean = rng.standard_normal(28).astype(np.float32) * 0.1
    xstd = (0.5 + rng.random(28)).astype(np.float32)
    ystd = (1e-3 * (0.5 + rng.random(18))).astype(np.float32)
    vec = closure.closure_vector(P, xmean, xstd, ystd)
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.closure_set(vec)
        Yn, rhs = g.closure_eval()
        ug, pg = g.state()
    Xn = (ref.features(ug, pg, (p.Lx, p.Ly, p.Lz)) - xmean) / xstd
    Yr = ref.forward(Xn, P)[0]
    assert rel_l2(Yn, Yr) < 1e-5
    # the spectral part from the device outputs themselves
    C = ref.terms(Yn, ystd, p.adiabatic_index)
    R = ref.rhs(C, (32, 16, 64), (p.Lx, p.Ly, p.Lz))
    print(f"closure rhs rel l2 {rel_l2(rhs, R):.2e}")
    assert rel_l2(rhs, R) < 1e-12


def test_zero_network_leaves_the_step_unchanged(rt):
    p = params(32, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=9, kmax=6, du=0.2, db=0.2)
    vec = closure.closure_vector(np.zeros(rt.MLP_NPARAM, np.float32), np.zeros(28, np.float32),
                                 np.ones(28, np.float32), np.ones(18, np.float32))
    with rt.Laps(p) as a, rt.Laps(p) as b:
        for g in (a, b):
            g.set_state(uu, pr)
        b.closure_set(vec)
        a.step(3)
        b.step(3)
        assert np.array_equal(a.spectral(), b.spectral())
        b.closure_set(None)
        a.step(1)
        b.step(1)
        assert np.array_equal(a.spectral(), b.spectral())


def test_device_filter_vs_numpy(rt):
    """laps_closure_filter (box filter of the state and its flux products, J from the library's
    own transforms) against the same definitions in numpy (laps_amd.closure.filtered_sample's
    products, J = curl B by scipy; the state has no mode near Nyquist, where the two transforms'
    derivative conventions differ)."""
    p = params((32, 16, 64), Ly=3.0, Lz=5.0, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=12, kmax=4, du=0.2, db=0.2)
    L = (p.Lx, p.Ly, p.Lz)
    r = 2
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.step(2)
        f = g.closure_filter(r)
        u2, q2 = g.state()
    rho, m, B = u2[..., 0], u2[..., 1:4], u2[..., 4:7]
    u = q2[..., 0:3]
    E = closure.emf(rho, u, B, closure.curl(B, L), 0.05)
    F = closure.energy_flux(u2[..., 7], u, B, q2[..., 3])
    cols = [u2[..., c] for c in range(8)]
    cols += [m[..., i] * u[..., j] for i, j in closure.PAIRS]
    cols += [B[..., i] * B[..., j] for i, j in closure.PAIRS]
    cols += [E[..., a] for a in range(3)] + [F[..., a] for a in range(3)]
    for q, c in enumerate(cols):
        want = closure.box_filter(c, r)
        err = np.abs(f[..., q] - want).max() / max(np.abs(want).max(), 1e-30)
        assert err < 1e-11, (q, err)


def test_device_spectral_filter_vs_numpy(rt):
    """laps_closure_filter_spectral against numpy: the sharp cut of the 26 fields of a fine Hall
    state onto a coarser grid, and of the same fields formed on the fine grid from the cut
    state (u = cut(rho u)/cut(rho), p from cut(e), J = cut(curl B)); numpy transforms, the
    definitions of laps_amd.closure (spectral_cut)."""
    p = params((32, 16, 64), Ly=3.0, Lz=5.0, if_hall=1, ion_inertial_length=0.05)
    pc = params((16, 16, 16), Ly=3.0, Lz=5.0, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=13, kmax=5, du=0.2, db=0.2)
    L = (p.Lx, p.Ly, p.Lz)
    nc = (pc.nx, pc.ny, pc.nz)
    gam = p.adiabatic_index
    with rt.Laps(p) as g, rt.Laps(pc) as gc:
        g.set_state(uu, pr)
        g.step(2)
        gc.set_state(*seeded_state(pc, seed=1))
        C, D = g.closure_filter_spectral(gc)
        u2, q2 = g.state()

    def fields(st, prim, J):
        rho, m, B = st[..., 0], st[..., 1:4], st[..., 4:7]
        u = prim[..., 0:3]
        E = closure.emf(rho, u, B, J, 0.05)
        F = closure.energy_flux(st[..., 7], u, B, prim[..., 3])
        cols = [st[..., c] for c in range(8)]
        cols += [m[..., i] * u[..., j] for i, j in closure.PAIRS]
        cols += [B[..., i] * B[..., j] for i, j in closure.PAIRS]
        return cols + [E[..., a] for a in range(3)] + [F[..., a] for a in range(3)]

    J = closure.curl(u2[..., 4:7], L)
    want_c = [closure.spectral_cut(c, nc) for c in fields(u2, q2, J)]
    sb = np.stack([closure.spectral_cut(u2[..., v], nc, fine_grid=True) for v in range(8)], axis=-1)
    mb = sb[..., 1:4]
    ub = mb / sb[..., 0:1]
    pb = (gam - 1.0) * (sb[..., 7] - 0.5 * (mb * ub).sum(-1) - 0.5 * (sb[..., 4:7] ** 2).sum(-1))
    qb = np.concatenate([ub, pb[..., None]], axis=-1)
    Jb = np.stack([closure.spectral_cut(J[..., a], nc, fine_grid=True) for a in range(3)], axis=-1)
    want_d = [closure.spectral_cut(c, nc) for c in fields(sb, qb, Jb)]
    for q in range(26):
        for got, want, tag in ((C[..., q], want_c[q], "C"), (D[..., q], want_d[q], "D")):
            err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-30)
            assert err < 1e-11, (tag, q, err)
