"""Learned face reconstruction: numpy float64 restatement of the network, its
training step, the data it is trained on and the face values it produces.

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels of the learned form are
compared against.  Only tests/ import this module; nothing in the product
path calls it.

The learned form has no counterpart in the reference.  What it replaces is
the face-value evaluation of the limited linear reconstruction
(reconstruction.cpp:67-97: mean +/- slope/2, fallback to the mean when the
face density or pressure is not positive), which stays the fallback and is
restated here exactly as the oracle and the HIP kernels evaluate it.

Definitions shared with the HIP kernels (every choice is this package's own, with no reference counterpart):
  * stencil: the 5x5 coarse cells around a cell, conservative components
    rho, rho u, rho v, rho E (VARS); input index v*25 + (dj+2)*5 + (di+2);
  * normalisation per cell and component: c = centre value, s = max(max|S-c|,
    1e-3 (|c_rho| + |c_rhoE|)); inputs x = (S - c) / s (the floor keeps
    variations far below the solution's scale from being magnified);
  * outputs z[4 f + v]: face f (0 x-, 1 x+, 2 y-, 3 y+), component v; the face
    value is F = P + sigmoid(z) (Q - P) between two upwind-biased candidates,
    P the minmod value and Q the fifth-order upwind-biased linear value along
    the face's axis (a learned limiter); rho w is copied from the cell;
    training targets are (F - c) / s;
  * network 100 -> 64 -> 64 -> 16, GELU on both hidden layers in its tanh
    form z sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)) (Hendrycks & Gimpel 2016);
  * loss sum_s w_s sum_o (yhat - t)^2 / (16 B) over a batch of B rows, with
    yhat = P + sigmoid(z) (Q - P) in normalised units;
  * Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected);
  * shock sensor: the largest relative jump |q_n - q0| / (q_n + q0) of
    q = rho and q = p to the four face neighbours; the cell falls back to the
    plain face values when it exceeds the threshold, when its stencil crosses
    a non-periodic border, or when a learned face state has non-positive
    density or pressure;
  * safeguards on learned values: clipped between the means of the two cells
    sharing the face, optionally blended towards the plain value;
  * targets: per coarse face, the mean of the fine grid's plain face values
    over the fine faces that tile it (fine run restricted by the factor r).
"""
import numpy as np

GAMMA = 1.4
NIN, NH, NOUT = 100, 64, 16
LAYOUT = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NOUT, NH)), ("b3", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)
VARS = (0, 1, 2, 4)
TAU_REL = 1e-3  # floor of the per-cell scale, relative to |rho| + |rho E|
GELU_K = 2.0 * np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


# ---- network ---------------------------------------------------------------------

def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def gelu(z):
    return z * sigmoid(GELU_K * (z + GELU_C * z ** 3))


def gelu_d(z):
    sg = sigmoid(GELU_K * (z + GELU_C * z ** 3))
    return sg + z * sg * (1.0 - sg) * GELU_K * (1.0 + 3.0 * GELU_C * z * z)


def forward(p, X):
    P = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ P["W1"].T + P["b1"]
    a1 = gelu(z1)
    z2 = a1 @ P["W2"].T + P["b2"]
    a2 = gelu(z2)
    y = a2 @ P["W3"].T + P["b3"]
    return y, (X, z1, a1, z2, a2)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def candidates(X):
    """Normalised limiter candidates of every output from the normalised
    stencil rows X (N, 100): P = minmod face value (reconstruction.cpp:67-97
    without the positivity fallback), Q = fifth-order upwind-biased linear
    face value (2 u-2 - 13 u-1 + 47 u0 + 27 u+1 - 3 u+2) / 60, both along the
    face's axis with offsets counted towards the face."""
    A = np.asarray(X, dtype=np.float64).reshape(-1, 4, 5, 5)
    P = np.empty((A.shape[0], NOUT))
    Q = np.empty((A.shape[0], NOUT))
    for f in range(4):
        axis, side = f // 2, f % 2
        for v in range(4):
            line = A[:, v, 2, :] if axis == 0 else A[:, v, :, 2]
            if side == 0:
                line = line[:, ::-1]
            m2, m1, c0, p1, p2 = (line[:, i] for i in range(5))
            P[:, 4 * f + v] = c0 + 0.5 * minmod(p1 - c0, c0 - m1)
            Q[:, 4 * f + v] = (2.0 * m2 - 13.0 * m1 + 47.0 * c0 + 27.0 * p1 - 3.0 * p2) / 60.0
    return P, Q


def model(p, X):
    """Normalised face values: P + sigmoid(z) (Q - P), z the network output."""
    z, cache = forward(p, X)
    P, Q = candidates(X)
    return P + sigmoid(z) * (Q - P), z, cache


def loss_grad(p, X, T, w):
    P = unpack(p)
    B = X.shape[0]
    yh, z, (X, z1, a1, z2, a2) = model(p, X)
    Pc, Qc = candidates(X)
    r = yh - np.asarray(T, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)[:, None]
    loss = float(np.sum(w * r * r) / (NOUT * B))
    sg = sigmoid(z)
    dy = 2.0 * w * r / (NOUT * B) * sg * (1.0 - sg) * (Qc - Pc)
    g = {"W3": dy.T @ a2, "b3": dy.sum(0)}
    dz2 = (dy @ P["W3"]) * gelu_d(z2)
    g["W2"], g["b2"] = dz2.T @ a1, dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * gelu_d(z1)
    g["W1"], g["b1"] = dz1.T @ X, dz1.sum(0)
    return loss, pack(g)


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(a):
    """float32 -> bfloat16, round to nearest even, returned as float32 values."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def forward_bf16(p, X):
    """The inference path: inputs, weights and hidden activations rounded to
    bf16 (biases kept in fp32), products summed in float64; the device sums
    the same products in fp32 on the matrix cores."""
    P = {k: np.asarray(v, dtype=np.float32) for k, v in unpack(p).items()}
    xb = bf16(np.asarray(X, dtype=np.float32)).astype(np.float64)
    W = {k: bf16(P[k]).astype(np.float64) for k in ("W1", "W2", "W3")}
    h1 = bf16(gelu(xb @ W["W1"].T + P["b1"].astype(np.float64)).astype(np.float32)).astype(np.float64)
    h2 = bf16(gelu(h1 @ W["W2"].T + P["b2"].astype(np.float64)).astype(np.float32)).astype(np.float64)
    return h2 @ W["W3"].T + P["b3"].astype(np.float64)


# ---- fields on a 2D grid: arrays (ny, nx, 5), x fastest ------------------------

def pressure(U):
    """p of cons2prim (utils.cpp:48-63) in the device's operation order."""
    K = (U[..., 1] * U[..., 1] + U[..., 2] * U[..., 2] + U[..., 3] * U[..., 3]) / (U[..., 0] * U[..., 0])
    T = (2.0 * U[..., 4] / U[..., 0] - K) / (2.0 / (GAMMA - 1.0))
    return U[..., 0] * T


def minmod(a, b):
    return np.where((a > 0) & (b > 0), np.minimum(a, b), np.where((a < 0) & (b < 0), np.maximum(a, b), 0.0))


def neighbour(A, d, axis, periodic):
    """Values of the neighbour at offset d along array axis 'axis' (1: x, 0: y)
    and whether it exists (periodic wrap, else inside)."""
    n = A.shape[axis]
    idx = np.arange(n) + d
    exists = (idx >= 0) & (idx < n)
    if periodic:
        exists[:] = True
    B = np.take(A, idx % n, axis=axis)
    shape = [1] * A.ndim
    shape[axis] = n
    return B, exists.reshape(shape[:2])


def plain_face_states(U, periodic):
    """Face values of the limited linear reconstruction for every cell:
    F[j, i, f, k], f = 0 x-, 1 x+, 2 y-, 3 y+ (oracle/mmf_oracle.c
    orc_compute_polynomials + orc_reconstruct, the HIP gather kernel)."""
    U = np.asarray(U, dtype=np.float64)
    F = np.empty(U.shape[:2] + (4, 5))
    for a, axis in ((0, 1), (1, 0)):
        Xm, em = neighbour(U, -1, axis, periodic[a])
        Xp, ep = neighbour(U, +1, axis, periodic[a])
        dm = np.where(em[..., None], U - Xm, 0.0)
        dp = np.where(ep[..., None], Xp - U, 0.0)
        s = minmod(dp, dm)
        for side, v in ((0, U - 0.5 * s), (1, U + 0.5 * s)):
            ok = (v[..., 0] > 0.0) & (pressure(v) > 0.0)
            F[:, :, 2 * a + side] = np.where(ok[..., None], v, U)
    return F


def restrict(U, r):
    """Conservative average over r x r blocks, summed row by row in order."""
    U = np.asarray(U, dtype=np.float64)
    ny, nx = U.shape[0] // r, U.shape[1] // r
    acc = np.zeros((ny, nx, U.shape[2]))
    for b in range(r):
        for a in range(r):
            acc = acc + U[b::r, a::r][:ny, :nx]
    return acc * (1.0 / (r * r))


def sensor(U, periodic):
    """Largest relative jump |q_n - q0| / (q_n + q0) of rho and p to the four
    face neighbours; +inf where a neighbour is missing."""
    q = (U[..., 0], pressure(U))
    phi = np.zeros(U.shape[:2])
    for a, axis in ((0, 1), (1, 0)):
        for Q in q:
            for d in (-1, 1):
                Qn, en = neighbour(Q[..., None], d, axis, periodic[a])
                Qn = Qn[..., 0]
                val = np.where(en, np.abs(Qn - Q) / (Qn + Q), np.inf)
                phi = np.maximum(phi, val)
    return phi


def features(U, periodic):
    """Normalised 5x5 stencils of every cell: X (N, 100) float32, centre c and
    scale s (N, 4) float64, and whether the whole stencil exists."""
    U = np.asarray(U, dtype=np.float64)
    ny, nx = U.shape[:2]
    S = np.empty((ny, nx, 4, 5, 5))
    valid = np.ones((ny, nx), dtype=bool)
    for dj in range(-2, 3):
        Bj, ej = neighbour(U, dj, 0, periodic[1])
        for di in range(-2, 3):
            Bij, ei = neighbour(Bj, di, 1, periodic[0])
            for v, k in enumerate(VARS):
                S[:, :, v, dj + 2, di + 2] = Bij[..., k]
            valid &= ej & ei
    c = S[:, :, :, 2, 2]
    dev = np.abs(S - c[..., None, None]).reshape(ny, nx, 4, 25).max(axis=3)
    tau = TAU_REL * (np.abs(c[..., 0]) + np.abs(c[..., 3]))
    s = np.maximum(dev, tau[..., None])
    X = ((S - c[..., None, None]) / s[..., None, None]).astype(np.float32)
    return X.reshape(ny * nx, 100), c.reshape(-1, 4), s.reshape(-1, 4), valid.ravel()


def face_targets(Uf, r, periodic):
    """Per coarse cell and face, the mean of the fine plain face values over
    the r fine faces that tile the coarse face: (nyc, nxc, 4, 5)."""
    Ff = plain_face_states(Uf, periodic)
    ny, nx = Uf.shape[0] // r, Uf.shape[1] // r
    T = np.zeros((ny, nx, 4, 5))
    for b in range(r):
        T[:, :, 0] = T[:, :, 0] + Ff[b::r, 0::r][:ny, :nx, 0]
        T[:, :, 1] = T[:, :, 1] + Ff[b::r, r - 1::r][:ny, :nx, 1]
        T[:, :, 2] = T[:, :, 2] + Ff[0::r, b::r][:ny, :nx, 2]
        T[:, :, 3] = T[:, :, 3] + Ff[r - 1::r, b::r][:ny, :nx, 3]
    return T * (1.0 / r)


SWAP = (0, 2, 1, 3)  # rho u <-> rho v


def symmetry(x, t, sigma):
    """Sample (x (100,), t (16,)) under the dihedral symmetry sigma of the 2D
    Euler equations: bit 0 transposes x and y, then bit 1 reflects x, then
    bit 2 reflects y (velocity components change sign or swap with them)."""
    X = np.asarray(x, dtype=np.float32).reshape(4, 5, 5).copy()
    T = np.asarray(t, dtype=np.float32).reshape(4, 4).copy()
    if sigma & 1:
        X = np.stack([X[SWAP[v]].T for v in range(4)])
        T = np.stack([T[(2, 3, 0, 1)[f]][list(SWAP)] for f in range(4)])
    if sigma & 2:
        sg = np.array([1, -1, 1, 1], dtype=np.float32)
        X = X[:, :, ::-1] * sg[:, None, None]
        T = T[[1, 0, 2, 3]] * sg[None, :]
    if sigma & 4:
        sg = np.array([1, 1, -1, 1], dtype=np.float32)
        X = X[:, ::-1, :] * sg[:, None, None]
        T = T[[0, 1, 3, 2]] * sg[None, :]
    return np.ascontiguousarray(X).ravel(), np.ascontiguousarray(T).ravel()


def collect(Uf, r, periodic, cells, theta):
    """Training rows of the coarse cells 'cells' (flat indices of the coarse
    grid), eight symmetric rows each: X (8n, 100), T (8n, 16), w (8n,)."""
    Uc = restrict(Uf, r)
    X, c, s, valid = features(Uc, periodic)
    phi = sensor(Uc, periodic).ravel()
    Tf = face_targets(Uf, r, periodic).reshape(-1, 4, 5)
    n = len(cells)
    Xo = np.empty((8 * n, 100), dtype=np.float32)
    To = np.empty((8 * n, 16), dtype=np.float32)
    wo = np.empty(8 * n, dtype=np.float32)
    for q, cell in enumerate(cells):
        t = ((Tf[cell][:, list(VARS)] - c[cell][None, :]) / s[cell][None, :]).astype(np.float32).ravel()
        w = 1.0 if (valid[cell] and phi[cell] <= theta) else 0.0
        for sg in range(8):
            xs, ts = symmetry(X[cell], t, sg)
            Xo[8 * q + sg], To[8 * q + sg], wo[8 * q + sg] = xs, ts, w
    return Xo, To, wo


def candidates_physical(U, periodic):
    """P (minmod, no positivity fallback) and Q (fifth-order upwind-biased)
    face values in physical units, (ny*nx, 4, 5) each, for the components in
    VARS; evaluated in the device's operation order."""
    U = np.asarray(U, dtype=np.float64)
    ny, nx = U.shape[:2]
    P = np.zeros((ny * nx, 4, 5))
    Q = np.zeros((ny * nx, 4, 5))
    for f in range(4):
        axis_arr, side = (1 if f < 2 else 0), f % 2
        per = periodic[0] if f < 2 else periodic[1]
        sgn = 1 if side else -1
        vals = [neighbour(U, sgn * d, axis_arr, per)[0].reshape(-1, 5) for d in (-2, -1, 0, 1, 2)]
        m2, m1, c0, p1, p2 = vals
        for k in VARS:
            P[:, f, k] = c0[:, k] + 0.5 * minmod(p1[:, k] - c0[:, k], c0[:, k] - m1[:, k])
            Q[:, f, k] = ((((2.0 * m2[:, k] - 13.0 * m1[:, k]) + 47.0 * c0[:, k]) + 27.0 * p1[:, k])
                          - 3.0 * p2[:, k]) / 60.0
    return P, Q


def learned_face_states(U, p, periodic, theta, clip=False, blend=1.0):
    """Face values of the learned reconstruction with its safeguards and
    fallback: F (ny, nx, 4, 5) and the per-cell fallback flag.  The network
    output z weighs the two limiter candidates, F = P + sigmoid(z) (Q - P);
    clip bounds a component between the means of the two cells sharing the
    face; blend returns plain + blend (learned - plain)."""
    U = np.asarray(U, dtype=np.float64)
    ny, nx = U.shape[:2]
    X, c, s, valid = features(U, periodic)
    phi = sensor(U, periodic).ravel()
    z = forward_bf16(p, X).astype(np.float32)
    sg = (1.0 / (1.0 + np.exp(-z.astype(np.float32)))).astype(np.float32).astype(np.float64)
    Pc, Qc = candidates_physical(U, periodic)
    P = plain_face_states(U, periodic).reshape(-1, 4, 5)
    FL = np.empty((ny * nx, 4, 5))
    for f in range(4):
        a, side = f // 2, f % 2
        Un, _ = neighbour(U, 1 if side else -1, 1 if a == 0 else 0, True)
        Un = Un.reshape(-1, 5)
        for v, k in enumerate(VARS):
            val = Pc[:, f, k] + sg[:, 4 * f + v] * (Qc[:, f, k] - Pc[:, f, k])
            if clip:
                u0, u1 = U.reshape(-1, 5)[:, k], Un[:, k]
                lo, hi = np.minimum(u0, u1), np.maximum(u0, u1)
                val = np.where(val < lo, lo, np.where(val > hi, hi, val))
            if blend != 1.0:
                val = P[:, f, k] + blend * (val - P[:, f, k])
            FL[:, f, k] = val
        FL[:, f, 3] = U.reshape(-1, 5)[:, 3]
    pos = np.all((FL[..., 0] > 0.0) & (pressure(FL) > 0.0), axis=1)
    ok = valid & (phi <= theta) & pos
    F = np.where(ok[:, None, None], FL, P)
    return F.reshape(ny, nx, 4, 5), ~ok.reshape(ny, nx)
