"""NumPy restatement of the plate-vote surrogate network (forward, backward,
Adam) and of its training targets.

TEST INFRASTRUCTURE ONLY: the checker for the library's MFMA kernels. The
network has no counterpart in the reference; its targets are the reference's
plate integral (tensorvote.h:1651-1695), computed here with the FP64 oracle
(oracle_plate3_sum_f64).

Conventions mirror the library: parameters FP32 row-major [out][in]
(W1 64x32, W2 64x64, W3 64x64, W4 32x64, flat 12288); BF16 weights and
activations (round to nearest even); input slots 0-4 feature high parts,
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


def forward(p, X, exact=False, affine=None):
    """X: n x 32 input values (already BF16 values unless exact). Returns
    Y (n x 32; columns 0-5 mapped through mu + sd * y when affine = (mu, sd))
    and the activations. exact=True: FP64, no rounding."""
    W = [w.astype(np.float64) if exact else bf16(w.astype(np.float32)).astype(np.float64) for w in layers(p)]
    r = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)).astype(np.float64))
    X = np.asarray(X, np.float64)
    H1 = r(np.maximum(X @ W[0].T, 0))
    H2 = r(np.maximum(H1 @ W[1].T, 0))
    H3 = r(np.maximum(H2 @ W[2].T, 0))
    Y = H3 @ W[3].T
    if affine is not None:
        mu, sd = (np.asarray(a, np.float64) for a in affine)
        Y = Y.copy()
        Y[:, :6] = mu + sd * Y[:, :6]
    return Y, (X, H1, H2, H3)


def loss(Y, T):
    d = Y[:, :6] - np.asarray(T, np.float64)[:, :6]
    return float((COMP_W * d * d).sum() / Y.shape[0])


def gradients(p, X, T, exact=False, affine=None):
    """Raw parameter gradients in the flat layout (residual columns of W1 get
    their own gradient; Adam adds them to the shared weight) of the loss of
    mu + sd * Y against T."""
    W = [w.astype(np.float64) if exact else bf16(w.astype(np.float32)).astype(np.float64) for w in layers(p)]
    r = (lambda a: a) if exact else (lambda a: bf16(a.astype(np.float32)).astype(np.float64))
    mu, sd = (np.zeros(6), np.ones(6)) if affine is None else (np.asarray(a, np.float64) for a in affine)
    Y, (X, H1, H2, H3) = forward(p, X, exact)
    n = Y.shape[0]
    dY = np.zeros_like(Y)
    dY[:, :6] = 2.0 * COMP_W * (mu + sd * Y[:, :6] - np.asarray(T, np.float64)[:, :6]) * sd / n
    dY = r(dY)
    g = np.zeros(NPARAM)
    g[10240:] = (dY.T @ H3).ravel()
    dZ3 = r(dY @ W[3]) * (H3 > 0)
    g[6144:10240] = (dZ3.T @ H2).ravel()
    dZ2 = r(dZ3 @ W[2]) * (H2 > 0)
    g[2048:6144] = (dZ2.T @ H1).ravel()
    dZ1 = r(dZ2 @ W[1]) * (H1 > 0)
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
