"""Pins of the learned-reconstruction oracle (oracle/learn_oracle.py), CPU only.

Each check is chosen so that a plausible mistake fails it: the backward pass
against central finite differences, GELU and Adam closed forms, bf16 rounding
known answers, the plain face values against the C oracle's reconstruction,
conservation and exactness of the restriction and of the face targets on
linear fields, invariance of the training rows under a symmetry the field
has, and the sensor's closed forms.
"""
import numpy as np
import pytest

from oracle import learn_oracle as LO


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, LO.NPARAMS)


def test_layout_size():
    assert LO.NPARAMS == 11664
    p = np.arange(LO.NPARAMS, dtype=np.float64)
    assert np.array_equal(LO.pack(LO.unpack(p)), p)
    P = LO.unpack(p)
    assert P["W1"][1, 0] == 100 and P["b1"][0] == 6400 and P["b3"][-1] == LO.NPARAMS - 1


def test_gelu_closed_forms():
    from scipy.special import erf
    z = np.linspace(-4, 4, 41)
    assert LO.gelu(np.array(0.0)) == 0.0
    assert LO.gelu_d(np.array(0.0)) == 0.5
    # z sigmoid(v(z)) with v odd: sigmoid(v) + sigmoid(-v) = 1
    assert np.allclose(LO.gelu(z) - LO.gelu(-z), z, rtol=0, atol=1e-15)
    h = 1e-6
    assert np.allclose(LO.gelu_d(z), (LO.gelu(z + h) - LO.gelu(z - h)) / (2 * h), atol=1e-8)
    # the tanh form stays within 4.8e-4 of the exact z Phi(z) (a dropped cubic
    # term or a wrong factor in v is off by 1e-2 or more)
    w = np.linspace(-10, 10, 20001)
    exact = 0.5 * w * (1.0 + erf(w / np.sqrt(2.0)))
    err = np.max(np.abs(LO.gelu(w) - exact))
    assert 4e-4 < err < 4.8e-4


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(4)
    X = rng.normal(size=(7, LO.NIN))
    T = rng.normal(size=(7, LO.NOUT))
    w = np.array([1, 1, 0, 1, 0.5, 1, 1.0])
    loss, g = LO.loss_grad(p, X, T, w)
    idx = np.concatenate([rng.choice(6400, 40, replace=False), 6400 + rng.choice(64, 5, replace=False),
                          6464 + rng.choice(4096, 30, replace=False), 10560 + np.arange(3),
                          10624 + rng.choice(1024, 20, replace=False), 11648 + np.arange(16)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (LO.loss_grad(pp, X, T, w)[0] - LO.loss_grad(pm, X, T, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    # loss definition: sum w (yhat - t)^2 / (16 B), yhat = P + sigmoid(z) (Q - P)
    yh, z, _ = LO.model(p, X)
    Pc, Qc = LO.candidates(X)
    assert np.allclose(yh, Pc + (Qc - Pc) / (1.0 + np.exp(-z)))
    assert np.isclose(loss, np.sum(w[:, None] * (yh - T) ** 2) / (16 * 7))


def test_candidates_closed_forms():
    """Q is exact for cell averages of a quartic (fifth-order face value), P is
    the minmod value; both exact on linear data; both faces and both axes."""
    rng = np.random.default_rng(12)
    coef = rng.normal(size=5)
    def avg(i):  # average of the quartic over [i - 1/2, i + 1/2]
        F = lambda x: sum(coef[k] * x ** (k + 1) / (k + 1) for k in range(5))
        return F(i + 0.5) - F(i - 0.5)
    poly = lambda x: sum(coef[k] * x ** k for k in range(5))
    line = np.array([avg(i) for i in range(-2, 3)])
    x = np.zeros((4, 5, 5))
    x[0, 2, :] = line          # rho along x
    x[1, :, 2] = line          # rho u along y
    x[2, 2, :] = line[::-1]    # rho v along x, reversed
    P, Q = LO.candidates(x.reshape(1, 100))
    assert np.isclose(Q[0, 4 * 1 + 0], poly(0.5), rtol=0, atol=1e-12)    # x+ face, rho
    assert np.isclose(Q[0, 4 * 0 + 0], poly(-0.5), rtol=0, atol=1e-12)   # x- face, rho
    assert np.isclose(Q[0, 4 * 3 + 1], poly(0.5), rtol=0, atol=1e-12)    # y+ face, rho u
    assert np.isclose(Q[0, 4 * 0 + 2], poly(0.5), rtol=0, atol=1e-12)    # x- face of the reversed line
    mm = LO.minmod(line[3] - line[2], line[2] - line[1])
    assert P[0, 4 * 1 + 0] == line[2] + 0.5 * mm
    lin = np.arange(-2, 3) * 0.3 + 1.0
    x = np.zeros((4, 5, 5))
    x[3, 2, :] = lin
    P, Q = LO.candidates(x.reshape(1, 100))
    assert np.isclose(P[0, 4 * 1 + 3], 1.15) and np.isclose(Q[0, 4 * 1 + 3], 1.15)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(5)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = LO.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bf16_known_answers():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e-3, 65504.0], dtype=np.float32)
    b = LO.bf16(x)
    assert b[0] == 1.0
    assert b[1] == 1.0                  # tie -> even
    assert b[2] == 1 + 2.0 ** -6        # tie -> even (upper)
    assert b[3] == -2.5
    assert abs(b[4] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8
    assert b[5] == 65536.0


def random_field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    prim = np.empty((ny, nx, 5))
    prim[..., 0] = rng.uniform(0.3, 2.0, (ny, nx))
    prim[..., 1:3] = rng.uniform(-1, 1, (ny, nx, 2))
    prim[..., 3] = 0.0
    prim[..., 4] = rng.uniform(0.5, 2.0, (ny, nx))
    return prim_to_cons(prim)


def prim_to_cons(prim):
    rho = prim[..., 0] / prim[..., 4]
    U = np.empty_like(prim)
    U[..., 0] = rho
    U[..., 1:4] = rho[..., None] * prim[..., 1:4]
    U[..., 4] = rho * prim[..., 4] / 0.4 + 0.5 * rho * (prim[..., 1:4] ** 2).sum(-1)
    return U


@pytest.mark.parametrize("periodic", [(True, True), (False, False), (True, False)])
def test_plain_face_states_equal_c_oracle(orc, periodic):
    ny, nx = 6, 7
    U = random_field(ny, nx, 11)
    # one cell with a steep jump so that the positivity fallback triggers
    U[3, 3, 4] = 0.05 * U[3, 3, 0]
    F = LO.plain_face_states(U, periodic)
    bc = [4 if periodic[0] else 0] * 2 + [4 if periodic[1] else 0] * 2 + [0, 0]
    o = orc.Solver(orc.problem(orc.RADSOD, 2), 2, [nx, ny], [0.0, 0.0, 0.0], 0.1, side_bc=bc)
    Fo = o.face_states(U.reshape(-1, 5))
    assert np.array_equal(F.reshape(-1, 4, 5), Fo)


def test_restrict_conserves_and_is_exact_on_constants():
    U = random_field(8, 12, 2)
    C4 = LO.restrict(U, 4)
    assert C4.shape == (2, 3, 5)
    assert np.allclose(C4.sum((0, 1)) * 16, U.sum((0, 1)), rtol=1e-14)
    # dyadic values: the 16-term sums are exact
    const = np.ones((8, 8, 5)) * np.array([1.0, 0.25, -0.375, 0.0, 2.5])
    assert np.array_equal(LO.restrict(const, 4), np.ones((2, 2, 5)) * np.array([1.0, 0.25, -0.375, 0.0, 2.5]))


def test_face_targets_exact_on_linear_field():
    """A field linear in x: the minmod slopes are exact, so the coarse-face
    targets are the exact face values, and the y faces carry the coarse mean."""
    r, nx, ny = 4, 32, 16
    H = 1.0
    xf = (np.arange(nx) + 0.5) * (H / r)
    base = np.array([1.0, 0.3, 0.1, 0.0, 3.0])
    slope = np.array([0.01, 0.002, -0.001, 0.0, 0.02])
    U = base[None, None, :] + slope[None, None, :] * xf[None, :, None] * np.ones((ny, 1, 1))
    T = LO.face_targets(U, r, (False, True))
    xc = (np.arange(nx // r) + 0.5) * H
    for I in range(1, nx // r - 1):
        lo = base + slope * (xc[I] - 0.5 * H)
        hi = base + slope * (xc[I] + 0.5 * H)
        mid = base + slope * xc[I]
        assert np.allclose(T[1, I, 0], lo, rtol=0, atol=1e-13)
        assert np.allclose(T[1, I, 1], hi, rtol=0, atol=1e-13)
        assert np.allclose(T[1, I, 2], mid, rtol=0, atol=1e-13)
        assert np.allclose(T[1, I, 3], mid, rtol=0, atol=1e-13)


def test_symmetry_group_laws():
    rng = np.random.default_rng(8)
    x = rng.normal(size=100).astype(np.float32)
    t = rng.normal(size=16).astype(np.float32)
    for sg in (1, 2, 4):
        x2, t2 = LO.symmetry(*LO.symmetry(x, t, sg), sg)
        assert np.array_equal(x2, x) and np.array_equal(t2, t)
    # rotation by 90 degrees (transpose then reflect x) has order 4
    xr, tr = x, t
    for _ in range(4):
        xr, tr = LO.symmetry(xr, tr, 3)
    assert np.array_equal(xr, x) and np.array_equal(tr, t)
    xr, tr = LO.symmetry(x, t, 3)
    assert not np.array_equal(xr, x)


def test_rows_invariant_under_a_symmetry_of_the_field():
    """A fine field mirror-symmetric about x = x_c of one coarse cell (rho, rho v,
    rho E even, rho u odd): that cell's row equals its x-reflected image."""
    r, n = 4, 40
    rng = np.random.default_rng(21)
    half = rng.uniform(0, 1, (n, n // 2, 5))
    U = np.empty((n, n, 5))
    U[:, n // 2:] = half
    U[:, :n // 2] = half[:, ::-1]
    U[:, :n // 2, 1] *= -1.0
    U[..., 0] = 0.8 + 0.3 * U[..., 0]
    U[..., 3] = 0.0
    U[..., 4] = 3.0 + U[..., 4]
    # the mirror plane x = n/2 falls between coarse cells 4 and 5; shift the
    # field by r/2 fine cells so that it passes through the centre of coarse cell 4
    U = np.roll(U, -(n // 2 - (4 * r + r // 2)), axis=1)
    X, T, w = LO.collect(U, r, (True, True), [4 + (n // r) * 5], theta=1e9)
    assert np.allclose(X[2], X[0], atol=1e-6) and np.allclose(T[2], T[0], atol=1e-6)
    assert not np.allclose(X[1], X[0], atol=1e-3)


def test_sensor_closed_forms():
    ny, nx = 5, 6
    lin = prim_to_cons(np.stack([1.0 + 0.1 * np.arange(nx)[None, :] * np.ones((ny, 1)),
                                 np.zeros((ny, nx)), np.zeros((ny, nx)), np.zeros((ny, nx)),
                                 np.ones((ny, nx))], axis=-1))
    phi = LO.sensor(lin, (False, False))
    # rho = 1 + 0.1 i, p = 1: the largest jump is to the lower neighbour, 0.1 / (rho_{i-1} + rho_i)
    i = np.arange(1, nx - 1)
    assert np.allclose(phi[1:-1, 1:-1], (0.1 / (2.0 + 0.1 * (2 * i - 1)))[None, :], rtol=1e-12)
    assert np.all(np.isinf(phi[0])) and np.all(np.isinf(phi[:, 0]))
    step = np.ones((ny, nx, 5))
    step[..., 1:4] = 0.0
    step[..., 4] = 2.5
    step[:, 3:, 0] = 4.0
    step[:, 3:, 4] = 10.0
    phi = LO.sensor(step, (True, True))
    # the largest jump at the last low cell: rho |4 - 1| / (4 + 1) = 3/5 (p jumps 1 -> 2.4, less)
    assert phi[2, 2] == 3.0 / 5.0
    # two cells away from the step the field is flat: nothing fires
    assert phi[2, 1] == 0.0
    # p = rho T with T = 1 (rho E = 2.5 rho at rest): the pressure jump equals the density jump
    wave = np.ones((ny, nx, 5))
    wave[..., 1:4] = 0.0
    wave[..., 0] = 1.0 + 0.01 * np.arange(nx)[None, :]
    wave[..., 4] = 2.5 * wave[..., 0]
    phi = LO.sensor(wave, (False, False))
    assert np.allclose(phi[1:-1, 1:-1], (0.01 / (2.0 + 0.01 * (2 * i - 1)))[None, :], rtol=1e-12)


def test_collect_weights_and_zero_network():
    n, r = 32, 4
    U = random_field(n, n, 31)
    U[..., 0] = 1.0 + 0.01 * U[..., 0]
    U[..., 4] = 3.0 + 0.01 * U[..., 4]
    X, T, w = LO.collect(U, r, (False, False), [0, 1, 2 + 8 * 3, 8 * 8 - 1], theta=1e9)
    assert w.reshape(4, 8)[:, 0].tolist() == [0.0, 0.0, 1.0, 0.0]
    # a network that outputs zero weighs the two candidates equally
    Uc = LO.restrict(U, r)
    F, fb = LO.learned_face_states(Uc, np.zeros(LO.NPARAMS), (True, True), 1e9)
    assert not fb.any()
    Pc, Qc = LO.candidates_physical(Uc, (True, True))
    half = (Pc + 0.5 * (Qc - Pc)).reshape(F.shape)
    for k in LO.VARS:
        assert np.array_equal(F[..., k], half[..., k])
    for f in range(4):
        assert np.array_equal(F[:, :, f, 3], Uc[..., 3])
