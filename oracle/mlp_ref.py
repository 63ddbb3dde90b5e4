"""NumPy restatement of the plate-vote surrogate network (forward, backward,
Adam) and of its training targets.

TEST INFRASTRUCTURE ONLY: the checker for the library's MFMA kernels. The
network has no counterpart in the reference; its targets are the reference's
plate integral (tensorvote.h:1651-1695), computed here with the FP64 oracle
(oracle_plate3_sum_f64).

Conventions mirror the library: parameters FP32 row-major [out][in]
(W1 64x32, W2 64x64, W3 64x64, W4 32x64, flat 12288); BF16 weights (round
to nearest even); hidden activations carried as BF16 high part + BF16
residual (see act); input slots 0-4 feature high parts,
5-9 residuals (sharing the high part's weight), 10 constant 1; hidden unit
63 pinned to 1 (bias carrier); loss = sum over samples and the 6 plate-frame
components of w_c (Y - T)^2 / n with w_c = 2 on off-diagonals (Frobenius).
"""
from __future__ import annotations

import numpy as np

NPARAM = 12288
LAYERS = ((0, 64, 32), (2048, 64, 64), (6144, 64, 64), (10240, 32, 64))
COMP_W = np.array([1, 2, 2, 1, 2, 1], np.float64)
NFEAT = 5
CONST_IN = 10
BIAS_UNIT = 63


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = ((u + r) & 0xFFFF0000).astype(np.uint32).view(np.float32)
    return out.reshape(np.shape(x))


def bf16_bits_to_f32(bits):
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def layers(p):
    return [np.asarray(p[o:o + r * c]).reshape(r, c) for o, r, c in LAYERS]


def mask():
    m = np.zeros(NPARAM, bool)
    W1 = m[0:2048].reshape(64, 32)
    W1[:63, :NFEAT] = True
    W1[:63, CONST_IN] = True
    m[2048:6144].reshape(64, 64)[:63] = True
    m[6144:10240].reshape(64, 64)[:63] = True
    m[10240:].reshape(32, 64)[:6] = True
    return m


def act(a, exact=False):
    """ReLU activation as the library carries it into the next layer: the
    FP32 value split into a BF16 high part and a BF16 residual (hi + lo, about
    16 mantissa bits). Returns (hi + lo, hi); exact: (relu(a), relu(a))."""
    a = np.maximum(np.asarray(a, np.float64), 0.0)
    if exact:
        return a, a
    af = a.astype(np.float32)
    hi = bf16(af)
    lo = bf16(af - hi)
    return hi.astype(np.float64) + lo.astype(np.float64), hi.astype(np.float64)


def _forward(p, X, exact=False):
    W = [w.astype(np.float64) if exact else bf16(w.astype(np.float32)).astype(np.float64) for w in layers(p)]
    X = np.asarray(X, np.float64)
    H1, H1h = act(X @ W[0].T, exact)
    H2, H2h = act(H1 @ W[1].T, exact)
    H3, H3h = act(H2 @ W[2].T, exact)
    Y = H3 @ W[3].T
    return Y, W, (X, H1, H2, H3), (H1h, H2h, H3h)


def forward(p, X, exact=False, affine=None):
    """X: n x 32 input values (already BF16 values unless exact). Returns
    Y (n x 32; columns 0-5 mapped through mu + sd * y when affine = (mu, sd))
    and the activations (hi + lo pairs, see act). exact=True: FP64, no
    rounding."""
    Y, _, acts, _ = _forward(p, X, exact)
    if affine is not None:
        mu, sd = (np.asarray(a, np.float64) for a in affine)
        Y = Y.copy()
        Y[:, :6] = mu + sd * Y[:, :6]
    return Y, acts


def loss(Y, T):
    d = Y[:, :6] - np.asarray(T, np.float64)[:, :6]
    return float((COMP_W * d * d).sum() / Y.shape[0])


def gradients(p, X, T, exact=False, affine=None):
    """Raw parameter gradients in the flat layout (residual columns of W1 get
    their own gradient; Adam adds them to the shared weight) of the loss of
    mu + sd * Y against T. As the library: the forward pass carries hi + lo
    activations, the weight gradients take the high parts, the ReLU masks
    test the high parts, and dY and the back-propagated dZ are BF16."""
    r = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)).astype(np.float64))
    mu, sd = (np.zeros(6), np.ones(6)) if affine is None else (np.asarray(a, np.float64) for a in affine)
    Y, W, (X, H1, H2, H3), (H1h, H2h, H3h) = _forward(p, X, exact)
    n = Y.shape[0]
    dY = np.zeros_like(Y)
    dY[:, :6] = 2.0 * COMP_W * (mu + sd * Y[:, :6] - np.asarray(T, np.float64)[:, :6]) * sd / n
    dY = r(dY)
    g = np.zeros(NPARAM)
    g[10240:] = (dY.T @ H3h).ravel()
    dZ3 = r(dY @ W[3]) * (H3h > 0)
    g[6144:10240] = (dZ3.T @ H2h).ravel()
    dZ2 = r(dZ3 @ W[2]) * (H2h > 0)
    g[2048:6144] = (dZ2.T @ H1h).ravel()
    dZ1 = r(dZ2 @ W[1]) * (H1h > 0)
    g[0:2048] = (dZ1.T @ X).ravel()
    return g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step as the library takes it (t counts from 1): the residual
    columns of W1 share the high part's weight and gradient."""
    p, g, m, v = (np.array(a, np.float64) for a in (p, g, m, v))
    g1 = g[0:2048].reshape(64, 32)
    g1[:, :NFEAT] += g1[:, NFEAT:2 * NFEAT]
    msk = mask()
    c1 = 1 - b1 ** t
    c2 = 1 - b2 ** t
    lr_t = lr * np.sqrt(c2) / c1
    eps_t = eps * np.sqrt(c2)
    m = np.where(msk, b1 * m + (1 - b1) * g, m)
    v = np.where(msk, b2 * v + (1 - b2) * g * g, v)
    p = np.where(msk, p - lr_t * m / (np.sqrt(v) + eps_t), p)
    p1 = p[0:2048].reshape(64, 32)
    p1[:, NFEAT:2 * NFEAT] = p1[:, :NFEAT]
    return p, m, v


def plate_frame(n):
    """(u, v) of tensorvote.h:1667-1673 for a unit normal n."""
    n = np.asarray(n, np.float64)
    a = np.array([0.0, 0.0, 1.0]) if abs(n[2]) < 0.999 else np.array([1.0, 0.0, 0.0])
    u = np.cross(a, n)
    u /= np.linalg.norm(u)
    return u, np.cross(n, u)


def aligned_frame(d, n):
    """Plate frame turned about n so that the offset d lies in the (u, n)
    half-plane (the reference's basis when d runs along n). The N-sample
    plate sum is the same in every orthonormal plate basis once N >= p + 2."""
    n = np.asarray(n, np.float64)
    dh = np.asarray(d, np.float64) / np.linalg.norm(d)
    w = dh - (dh @ n) * n
    if w @ w > 1e-12:
        u = w / np.linalg.norm(w)
    else:
        u, _ = plate_frame(n)
    return u, np.cross(n, u)


def features(d, n, sigma1):
    """The 5 features of offset vector d from a plate voter with unit normal n:
    plate-frame direction of the offset (3), |d| / sigma1 clamped to 3, and
    2 (d.n)^2 - 1."""
    u, v = aligned_frame(d, n)
    r = np.linalg.norm(d)
    dh = np.asarray(d, np.float64) / r
    f2 = dh @ n
    return np.array([dh @ u, dh @ v, f2, min(r / sigma1, 3.0), 2 * f2 * f2 - 1])


def input_row(f):
    """BF16 input slots of a feature vector (high parts, residuals, constant)."""
    f = np.asarray(f, np.float32)
    row = np.zeros(32, np.float32)
    hi = bf16(f)
    row[:NFEAT] = hi
    row[NFEAT:2 * NFEAT] = bf16(f - hi)
    row[CONST_IN] = 1.0
    return row


def target(d, n, sigma1, sigma2, power, samples, plate_sum):
    """Plate-frame components (uu, uv, un, vv, vn, nn) of the plate vote / c1
    in the offset-aligned frame, `plate_sum` being the FP64 oracle of the
    N-sample integral."""
    d = np.asarray(d, np.float64)
    r2 = d @ d
    c1 = np.exp(-r2 / sigma1 ** 2)
    c2 = np.exp(-r2 / sigma2 ** 2) if sigma2 > 0 else 0.0
    V = plate_sum(d / np.sqrt(r2), 1.0, c2 / c1 if sigma2 > 0 else 0.0, power, n, samples)
    M = np.array([[V[0], V[1], V[2]], [V[1], V[3], V[4]], [V[2], V[4], V[5]]])
    u, v = aligned_frame(d, n)
    E = np.stack([u, v, np.asarray(n, np.float64)])
    L = E @ M @ E.T
    return np.array([L[0, 0], L[0, 1], L[0, 2], L[1, 1], L[1, 2], L[2, 2]]), c1


def to_field(o, n, d):
    """Field-frame symmetric tensor (xx, xy, xz, yy, yz, zz) of the aligned-frame components o."""
    u, v = aligned_frame(d, n)
    R = np.stack([u, v, np.asarray(n, np.float64)], axis=1)
    L = np.array([[o[0], o[1], o[2]], [o[1], o[3], o[4]], [o[2], o[4], o[5]]])
    M = R @ L @ R.T
    return np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])


# ------------------------------------------------------------------ tabulated network
# The library evaluates the network once per weight change on a grid and
# interpolates it per pair (TIRA_MLP_TABLE). In the offset-aligned frame the
# features depend on s = n.d and rho = |d| / sigma1 only, and the plate vote
# has the invariant form V / c1 = A I + B n n^T + s g (n d^T + d n^T) + D d d^T
# with A, B, g, D functions of x = 1 - s^2. These helpers restate that
# tabulation: grid, inputs, conversion, lookup and the vote it gives.
TAB_M = 64


def table_grid(M=TAB_M):
    """Cell centres x_j = (j + 0.5) / M."""
    return (np.arange(M) + 0.5) / M


def table_features(rho, M=TAB_M):
    """Feature rows (2M x 5) of one distance row: cell j at s = +sqrt(1 - x_j)
    (row 2j) and s = -sqrt(1 - x_j) (row 2j + 1)."""
    x = table_grid(M)
    f = np.zeros((2 * M, NFEAT))
    for j in range(M):
        for k, sg in enumerate((1.0, -1.0)):
            s = sg * np.sqrt(1.0 - x[j])
            f[2 * j + k] = [np.sqrt(x[j]), 0.0, s, rho, 2.0 * s * s - 1.0]
    return f


def table_row(Y, M=TAB_M):
    """(A, B, g, D/2) per cell (M x 4) from plate-frame outputs Y (2M x 6,
    uu, uv, un, vv, vn, nn; rows as table_features), using the even parts
    of uu, vv, nn and the odd part of un under s -> -s."""
    Y = np.asarray(Y, np.float64)
    Yp, Ym = Y[0::2], Y[1::2]
    x = table_grid(M)
    s, sn = np.sqrt(1.0 - x), np.sqrt(x)
    uu, un = 0.5 * (Yp[:, 0] + Ym[:, 0]), 0.5 * (Yp[:, 2] - Ym[:, 2])
    vv, nn = 0.5 * (Yp[:, 3] + Ym[:, 3]), 0.5 * (Yp[:, 5] + Ym[:, 5])
    D = (uu - vv) / x
    g = -D + un / (s * sn)
    B = (nn - vv) - s * s * D - 2.0 * s * s * g
    return np.stack([vv, B, g, 0.5 * D], axis=1)


def table_lookup(row, x, M=TAB_M):
    """Linear interpolation of a row (M x 4) at x, extrapolated over the half
    cells at both ends (cells j and j + 1 around x M - 1/2)."""
    t = x * M - 0.5
    j = min(int(t), M - 2)  # int() truncates: t in [-0.5, 0) gives cell 0
    f = t - j
    return row[j] + f * (row[j + 1] - row[j])


def table_vote(coef, n, d):
    """Plate vote / c1 (xx, xy, xz, yy, yz, zz) of unit normal n at unit
    offset d from (A, B, g, D/2)."""
    A, B, g, Dh = coef
    n = np.asarray(n, np.float64)
    d = np.asarray(d, np.float64)
    s = n @ d
    M = A * np.eye(3) + B * np.outer(n, n) + s * g * (np.outer(n, d) + np.outer(d, n)) + 2.0 * Dh * np.outer(d, d)
    return np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]])


def shells(sigma1):
    """Squared lattice distances of the table's rows 1..K: every l2 > 0 that
    is a sum of three squares with sqrt(l2) / sigma1 < 3 (float32), ascending,
    by enumeration of the lattice."""
    R = int(np.ceil(3.0 * sigma1))
    a = np.arange(-R, R + 1)
    l2 = np.unique((a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2).ravel())
    l2 = l2[l2 > 0]
    rho = np.sqrt(l2.astype(np.float32)) / np.float32(sigma1)
    return l2[np.minimum(rho, np.float32(3.0)) < np.float32(3.0)]
