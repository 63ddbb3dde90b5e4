"""Seeded synthetic tensor fields for the benchmark configurations.

All generators use numpy.random.default_rng(seed) and return float32 arrays:
  2D fields (H, W, 4) = [a, b, b, c], tensor axis 0 along the fast image axis;
  3D fields (s0, s1, s2, 9) column-major 3x3 (symmetric), tensor axis 0 along
  the fastest volume axis s2 (the reference's convention,
  tensorvote.h:1148, 1426).
Seeds: 1001..1005 for the five workloads, 2000 + i for batch volume i.
"""
from __future__ import annotations

import numpy as np

SEED_CFG1, SEED_CFG2, SEED_CFG3, SEED_CFG4, SEED_CFG5 = 1001, 1002, 1003, 1004, 1005


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ------------------------------------------------------------------ 2D
def bezier_field(H=256, W=256, n_curves=8, noise_frac=0.02, seed=SEED_CFG1):
    """t t^T along seeded cubic Bezier curves (later curves overwrite) plus
    noise sticks m q q^T (m ~ U(0, 0.5)) on `noise_frac` of the pixels."""
    rng = np.random.default_rng(seed)
    f = np.zeros((H, W, 4), np.float32)
    for _ in range(n_curves):
        P = rng.uniform(0, [W, H], size=(4, 2))  # (x, y) control points
        # arc length bound of the control polygon -> <= 0.25 px sample spacing
        L = np.sum(np.linalg.norm(np.diff(P, axis=0), axis=1))
        n = max(8, int(np.ceil(L / 0.25)) + 1)
        t = np.linspace(0.0, 1.0, n)[:, None]
        pts = ((1 - t) ** 3) * P[0] + 3 * ((1 - t) ** 2) * t * P[1] + 3 * (1 - t) * t ** 2 * P[2] + t ** 3 * P[3]
        der = 3 * ((1 - t) ** 2) * (P[1] - P[0]) + 6 * (1 - t) * t * (P[2] - P[1]) + 3 * t ** 2 * (P[3] - P[2])
        nrm = np.linalg.norm(der, axis=1, keepdims=True)
        ok = nrm[:, 0] > 0
        tan = der[ok] / nrm[ok]
        ix = np.rint(pts[ok, 0]).astype(int)
        iy = np.rint(pts[ok, 1]).astype(int)
        inb = (ix >= 0) & (ix < W) & (iy >= 0) & (iy < H)
        ix, iy, tan = ix[inb], iy[inb], tan[inb]
        f[iy, ix, 0] = tan[:, 0] * tan[:, 0]
        f[iy, ix, 1] = tan[:, 0] * tan[:, 1]
        f[iy, ix, 2] = tan[:, 0] * tan[:, 1]
        f[iy, ix, 3] = tan[:, 1] * tan[:, 1]
    nn = int(round(noise_frac * H * W))
    idx = rng.choice(H * W, size=nn, replace=False)
    th = rng.uniform(0, np.pi, nn)
    m = rng.uniform(0, 0.5, nn)
    q = np.stack([np.cos(th), np.sin(th)], 1)
    ff = f.reshape(-1, 4)
    ff[idx, 0] = m * q[:, 0] ** 2
    ff[idx, 1] = m * q[:, 0] * q[:, 1]
    ff[idx, 2] = m * q[:, 0] * q[:, 1]
    ff[idx, 3] = m * q[:, 1] ** 2
    return f


# ------------------------------------------------------------------ 3D helpers
def _ball_offsets(radius):
    r = int(np.ceil(radius))
    g = np.mgrid[-r:r + 1, -r:r + 1, -r:r + 1].reshape(3, -1).T
    return g[np.sum(g * g, axis=1) <= radius * radius + 1e-9]


def _stamp_tube(vol9, pts, tans, radius):
    """Set t t^T on every voxel within `radius` of the sampled centreline.
    pts/tans in (x, y, z) = (s2, s1, s0) axes."""
    s0, s1, s2 = vol9.shape[:3]
    off = _ball_offsets(radius)
    c = np.rint(pts).astype(np.int64)
    allp = (c[:, None, :] + off[None, :, :]).reshape(-1, 3)
    allt = np.repeat(tans, off.shape[0], axis=0)
    x, y, z = allp[:, 0], allp[:, 1], allp[:, 2]
    inb = (x >= 0) & (x < s2) & (y >= 0) & (y < s1) & (z >= 0) & (z < s0)
    x, y, z, t = x[inb], y[inb], z[inb], allt[inb]
    T = (t[:, :, None] * t[:, None, :]).reshape(-1, 9).astype(np.float32)
    vol9[z, y, x] = T


def tubes_sheets_volume(shape=(128, 128, 128), n_straight=10, n_helix=10, tube_radius=2.0, n_sheets=4,
                        noise_frac=0.01, seed=SEED_CFG2):
    """3D stick + plate test volume: straight and helical tubes (t t^T inside
    radius `tube_radius`), planar sheets of thickness 1 (I - n n^T) and
    isotropic noise m I (m ~ U(0, 0.5)) on `noise_frac` of the voxels."""
    rng = np.random.default_rng(seed)
    s0, s1, s2 = shape
    vol = np.zeros((s0, s1, s2, 9), np.float32)
    ext = np.array([s2, s1, s0], np.float64)
    # sheets
    zz, yy, xx = np.meshgrid(np.arange(s0), np.arange(s1), np.arange(s2), indexing="ij")
    for _ in range(n_sheets):
        n = _unit(rng.normal(size=3))
        p = rng.uniform(0.2, 0.8, 3) * ext
        dist = (xx - p[0]) * n[0] + (yy - p[1]) * n[1] + (zz - p[2]) * n[2]
        m = np.abs(dist) <= 0.5
        T = (np.eye(3) - np.outer(n, n)).astype(np.float32).reshape(9)
        vol[m] = T
    del zz, yy, xx
    L = float(np.linalg.norm(ext))
    # straight tubes
    for _ in range(n_straight):
        d = _unit(rng.normal(size=3))
        c = rng.uniform(0.1, 0.9, 3) * ext
        s = np.arange(-L, L, 0.25)
        pts = c[None, :] + s[:, None] * d[None, :]
        _stamp_tube(vol, pts, np.repeat(d[None, :], len(s), 0), tube_radius)
    # helices
    for _ in range(n_helix):
        a = _unit(rng.normal(size=3))
        u = _unit(np.cross(a, [1.0, 0.0, 0.0] if abs(a[0]) < 0.9 else [0.0, 1.0, 0.0]))
        v = np.cross(a, u)
        R = rng.uniform(4.0, 12.0)
        pitch = rng.uniform(8.0, 24.0)  # axial advance per turn
        c = rng.uniform(0.2, 0.8, 3) * ext
        k = pitch / (2 * np.pi)
        ds = 0.25 / np.sqrt(R * R + k * k)  # parameter step for 0.25 voxel arc spacing
        phi = np.arange(-L / k, L / k, ds) if k > 0 else np.arange(0, 2 * np.pi, ds)
        pts = c + np.outer(k * phi, a) + R * (np.outer(np.cos(phi), u) + np.outer(np.sin(phi), v))
        der = np.outer(np.full_like(phi, k), a) + R * (np.outer(-np.sin(phi), u) + np.outer(np.cos(phi), v))
        _stamp_tube(vol, pts, _unit(der), tube_radius)
    # isotropic noise
    nv = s0 * s1 * s2
    nn = int(round(noise_frac * nv))
    idx = rng.choice(nv, size=nn, replace=False)
    m = rng.uniform(0, 0.5, nn).astype(np.float32)
    flat = vol.reshape(-1, 9)
    flat[idx] = 0.0
    flat[idx, 0] = m
    flat[idx, 4] = m
    flat[idx, 8] = m
    return vol


def sheet_params(shape=(256, 256, 256), n_sheets=12, seed=SEED_CFG4):
    """Parameters of the curved plate sheets (15 floats each: R row-major with
    the local axes as columns, centre (x, y, z) in the middle 60% of the
    volume, curvatures (a, b, c) in U(-0.01, 0.01): radii of 50+ voxels) and
    the patch radius (0.35 of the smallest edge)."""
    rng = np.random.default_rng(seed)
    s0, s1, s2 = shape
    ext = np.array([s2, s1, s0], np.float64)
    R = random_rotations(n_sheets, rng)
    C = rng.uniform(0.2, 0.8, (n_sheets, 3)) * ext
    K = rng.uniform(-0.01, 0.01, (n_sheets, 3))
    p = np.concatenate([R.reshape(n_sheets, 9), C, K], axis=1).astype(np.float32)
    return np.ascontiguousarray(p), float(0.35 * min(shape))


def curved_sheets_volume_device(shape=(256, 256, 256), n_sheets=12, noise_frac=0.02, seed=SEED_CFG4):
    """Plate-dominant volume generated on the device (tira_synth_sheets_async):
    `n_sheets` curved sheets (see sheet_params), thickness 1, tensor I - n n^T
    with n the surface normal, plus plate noise m (I - q q^T) (q uniform on the
    sphere, m ~ U(0, 0.5)) on `noise_frac` of the voxels. Returns a
    (s0, s1, s2, 6) packed DeviceArray."""
    from ._lib import DeviceArray, call
    p, patch = sheet_params(shape, n_sheets, seed)
    out = DeviceArray(tuple(shape) + (6,))
    call("tira_synth_sheets_async", out.ptr, *[int(s) for s in shape], n_sheets, p.ctypes.data, patch,
         float(noise_frac), seed)
    call("tira_hip_synchronize")
    return out


def vessel_tree(shape=(384, 512, 512), seed=SEED_CFG5, n_roots=4, r_root=8.0, r_min=2.0, max_segments=600):
    """Seeded branching vascular tree: n x 8 float32 segments (p0 xyz, p1 xyz,
    r0, r1; x along the fastest axis). Roots start on random faces pointing
    inward with radius r_root; each segment runs 4-8 radii, tapers by
    0.85-0.97, bends by up to ~20 degrees, and bifurcates with probability
    0.45 into two children of radius 0.8 r (Murray's law) at 25-45 degrees;
    branches end below r_min or outside the volume."""
    rng = np.random.default_rng(seed)
    s0, s1, s2 = shape
    ext = np.array([s2, s1, s0], np.float64)
    stack = []
    for _ in range(n_roots):
        ax = rng.integers(0, 3)
        side = rng.integers(0, 2)
        p = rng.uniform(0.25, 0.75, 3) * ext
        p[ax] = 1.0 if side == 0 else ext[ax] - 2.0
        d = np.zeros(3)
        d[ax] = 1.0 if side == 0 else -1.0
        d = _unit(d + 0.3 * rng.normal(size=3))
        stack.append((p, d, r_root))
    segs = []
    while stack and len(segs) < max_segments:
        p, d, r = stack.pop()
        L = rng.uniform(4.0, 8.0) * r
        p1 = p + L * d
        r1 = max(r * rng.uniform(0.85, 0.97), 0.5 * r_min)
        segs.append(np.concatenate([p, p1, [r, r1]]))
        if r1 < r_min or np.any(p1 < -r1) or np.any(p1 > ext + r1):
            continue
        if rng.random() < 0.45:
            perp = _unit(np.cross(d, rng.normal(size=3)))
            for sgn in (1.0, -1.0):
                ang = np.deg2rad(rng.uniform(25, 45))
                stack.append((p1, _unit(np.cos(ang) * d + sgn * np.sin(ang) * perp), 0.8 * r1))
        else:
            stack.append((p1, _unit(d + 0.35 * _unit(rng.normal(size=3))), r1))
    return np.ascontiguousarray(segs, np.float32)


def vessel_volume_device(shape=(384, 512, 512), seed=SEED_CFG5, noise_sigma=0.1):
    """Intensity volume of vessel_tree (1 inside, 0 outside, + N(0, noise^2))
    generated on the device; returns (DeviceArray (s0, s1, s2), segments)."""
    from ._lib import DeviceArray, call
    segs = vessel_tree(shape, seed)
    dseg = DeviceArray.from_host(segs)
    out = DeviceArray(tuple(shape))
    call("tira_synth_vessels_async", out.ptr, *[int(s) for s in shape], dseg.ptr, len(segs), float(noise_sigma),
         seed)
    call("tira_hip_synchronize")
    return out, segs


def random_rotations(n, rng):
    """Haar-uniform rotations from normalised Gaussian quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - w * z); R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y); R[:, 2, 1] = 2 * (y * z + w * x); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def eig_tensors(n, seed, lo=1e-3, hi=1e2, degenerate_frac=0.10, gap_lo=1e-7, gap_hi=1e-4):
    """n packed-6 (xx, xy, xz, yy, yz, zz) tensors with log-uniform
    eigenvalues in [lo, hi], Haar rotations, and `degenerate_frac` of them
    with one adjacent eigenvalue pair at relative gap 10^U(log gap_lo, log gap_hi).
    Returns (t6 float32, eigenvalues float64 sorted ascending)."""
    rng = np.random.default_rng(seed)
    lam = np.sort(10 ** rng.uniform(np.log10(lo), np.log10(hi), size=(n, 3)), axis=1)
    nd = int(round(degenerate_frac * n))
    if nd:
        idx = rng.choice(n, nd, replace=False)
        gap = 10 ** rng.uniform(np.log10(gap_lo), np.log10(gap_hi), nd)
        which = rng.integers(0, 2, nd)  # pair (0,1) or (1,2)
        lo_i = which
        lam[idx, lo_i + 1] = lam[idx, lo_i] * (1 + gap)
        lam[idx] = np.sort(lam[idx], axis=1)
    R = random_rotations(n, rng)
    A = np.einsum("nij,nj,nkj->nik", R, lam, R)
    t6 = np.stack([A[:, 0, 0], A[:, 0, 1], A[:, 0, 2], A[:, 1, 1], A[:, 1, 2], A[:, 2, 2]], 1).astype(np.float32)
    return t6, lam


def mat9_from_packed6(t6):
    t = np.asarray(t6, np.float32).reshape(-1, 6)
    m = np.empty((t.shape[0], 9), np.float32)
    m[:, 0] = t[:, 0]; m[:, 1] = t[:, 1]; m[:, 2] = t[:, 2]
    m[:, 3] = t[:, 1]; m[:, 4] = t[:, 3]; m[:, 5] = t[:, 4]
    m[:, 6] = t[:, 2]; m[:, 7] = t[:, 4]; m[:, 8] = t[:, 5]
    return m


def packed6_from_mat9(m9):
    m = np.asarray(m9, np.float32).reshape(-1, 9)
    return np.ascontiguousarray(m[:, [0, 1, 2, 4, 5, 8]])


def batch_shapes(n_volumes=48, seed=SEED_CFG3, edges=tuple(range(32, 257, 16))):
    rng = np.random.default_rng(seed)
    return [tuple(int(e) for e in rng.choice(edges, 3)) for _ in range(n_volumes)]
