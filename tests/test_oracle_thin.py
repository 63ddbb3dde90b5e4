"""Pins the oracle's saliency threshold (reconstruct.py:14-63) and Lee 3D
thinning (thinning/lee_thinning.h:112-706) without running the reference
(lee_thinning.h needs CImg through tira/volume.h, which this image lacks, so
it cannot be compiled here).

Neighbourhood tests, exhaustively over all 2^26 neighbourhoods:
  * the octant-LUT Euler test (the reference's table, restated) is zero exactly
    when the Euler characteristic V - E + F - C of the union of closed voxel
    cubes does not change on removing the centre, counted independently;
  * the octree-labelling simple-point test equals breadth-first counting of
    26-connected components among the 26 neighbours.
Thinning: an independent sequential implementation (numpy loops over the
reference's border directions and raster order, using the BFS and Euler-count
tests above) gives the same volume bit for bit; the result is a subset of the
input, never splits an object component, and is a fixed point of thinning;
lines and single voxels are fixed points; bars and balls keep their topology.
(The reference's re-check of candidates, lee_thinning.h:651-661, tests only
simplicity, and an empty neighbourhood counts as simple, so the last voxels of
a small component can all be deleted: components may vanish, which the
restatement reproduces.)
Otsu: the threshold equals the one found by an independent route (numpy's own
histogram, then the split of least within-class variance, which maximises the
between-class variance) on data with a unique optimum; degenerate inputs.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def test_neighbourhood_tests_exhaustive(oracle):
    euler_mismatch, simple_mismatch = oracle.thin_exhaustive(nthreads=8)
    assert euler_mismatch == 0
    assert simple_mismatch == 0


def test_neighbourhood_known_codes(oracle):
    L = oracle.lib()
    c = 1 << 13
    assert L.oracle_thin_delta_euler(c) == -1 + 8 - 12 + 6 and L.oracle_thin_euler_lut_sum(c) != 0  # isolated
    line = c | (1 << 12) | (1 << 14)  # x-1 and x+1: removing the centre splits them
    assert not L.oracle_thin_simple_octree(line) and not L.oracle_thin_simple_bfs(line)
    corner = c | (1 << 0) | (1 << 1)  # two touching neighbours: one component
    assert L.oracle_thin_simple_octree(corner)
    full = (1 << 27) - 1  # interior voxel: removing it opens a cavity
    assert L.oracle_thin_delta_euler(full) != 0 and L.oracle_thin_euler_lut_sum(full) != 0


def _euler(v):
    """V - E + F - C of the union of closed unit cubes at the set voxels."""
    v = np.pad(v.astype(bool), 1)
    C = int(v.sum())
    # a lattice cell is present when any voxel of the 2^k voxels sharing it is set
    def count(axes):
        m = v.copy()
        for a in axes:
            m = m | np.roll(m, 1, axis=a)
        return int(m.sum())
    F = sum(count([a]) for a in range(3))
    E = sum(count([a, b]) for a, b in ((0, 1), (0, 2), (1, 2)))
    V = count([0, 1, 2])
    return V - E + F - C


def _topology(v):
    obj = ndimage.label(v, structure=np.ones((3, 3, 3)))[1]
    bg = ndimage.label(np.pad(v == 0, 1), structure=ndimage.generate_binary_structure(3, 1))[1]
    return obj, bg, _euler(v)


def _blob(shape, seed, q=0.5, sigma=1.5):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.random(shape), sigma)
    return (f > np.quantile(f, q)).astype(np.uint8)


def _code(v, x, y, z):
    s0, s1, s2 = v.shape
    c = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                X, Y, Z = x + dx, y + dy, z + dz
                if 0 <= X < s2 and 0 <= Y < s1 and 0 <= Z < s0 and v[Z, Y, X] == 1:
                    c |= 1 << (9 * (dz + 1) + 3 * (dy + 1) + dx + 1)
    return c


def _lee_sequential(vol, L):
    """Independent sequential Lee thinning: directions N S E W U B
    (y-1, y+1, x+1, x-1, z+1, z-1), candidates on the unchanged volume,
    re-check in raster order; Euler test by the cell count, simplicity by BFS."""
    v = (np.asarray(vol) != 0).astype(np.uint8)
    s0, s1, s2 = v.shape
    dirs = [(0, -1, 0), (0, 1, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, -1)]
    unchanged = 0
    while unchanged < 6:
        unchanged = 0
        for dx, dy, dz in dirs:
            cands = []
            for z, y, x in zip(*np.nonzero(v)):  # raster order
                X, Y, Z = x + dx, y + dy, z + dz
                if 0 <= X < s2 and 0 <= Y < s1 and 0 <= Z < s0 and v[Z, Y, X]:
                    continue
                c = _code(v, x, y, z)
                if bin(c).count("1") == 2 or L.oracle_thin_delta_euler(c) != 0 or not L.oracle_thin_simple_bfs(c):
                    continue
                cands.append((z, y, x))
            deleted = 0
            for z, y, x in cands:
                if L.oracle_thin_simple_bfs(_code(v, x, y, z)):
                    v[z, y, x] = 0
                    deleted += 1
            if not deleted:
                unchanged += 1
    return v


@pytest.mark.parametrize("shape,seed,q", [((14, 13, 12), 1, 0.5), ((9, 11, 16), 2, 0.6), ((1, 20, 25), 3, 0.5),
                                          ((6, 24, 7), 5, 0.4)])
def test_lee_matches_independent_sequential(oracle, shape, seed, q):
    v = _blob(shape, seed, q)
    t, it = oracle.lee_thin3(v)
    assert it >= 1
    assert np.array_equal(t, _lee_sequential(v, oracle.lib()))


@pytest.mark.parametrize("shape,seed,q", [((14, 13, 12), 1, 0.5), ((20, 9, 31), 2, 0.6), ((1, 30, 40), 3, 0.5),
                                          ((25, 25, 25), 4, 0.7), ((6, 40, 7), 5, 0.4)])
def test_lee_invariants(oracle, shape, seed, q):
    v = _blob(shape, seed, q)
    t, it = oracle.lee_thin3(v)
    assert np.all(t <= v)
    lab, n = ndimage.label(v, structure=np.ones((3, 3, 3)))
    for i in range(1, n + 1):  # no component splits
        assert ndimage.label((lab == i) & (t == 1), structure=np.ones((3, 3, 3)))[1] <= 1
    again, it2 = oracle.lee_thin3(t)
    assert np.array_equal(again, t) and it2 == 1


def test_lee_fixed_points_and_shapes(oracle):
    v = np.zeros((9, 9, 20), np.uint8)
    v[4, 4, 2:18] = 1  # a line along x
    t, _ = oracle.lee_thin3(v)
    assert np.array_equal(t, v)
    p = np.zeros((5, 5, 5), np.uint8)
    p[2, 2, 2] = 1
    assert np.array_equal(oracle.lee_thin3(p)[0], p)
    bar = np.zeros((10, 10, 30), np.uint8)
    bar[3:6, 3:6, 3:27] = 1  # 3x3 bar: thins to a curve through it
    t, _ = oracle.lee_thin3(bar)
    assert _topology(t) == _topology(bar)
    nb = ndimage.convolve(t.astype(int), np.ones((3, 3, 3), int), mode="constant") - 1
    assert np.all(nb[t == 1] <= 2) and t.sum() >= 20
    ball = np.zeros((15, 15, 15), np.uint8)
    z, y, x = np.mgrid[:15, :15, :15]
    ball[(x - 7) ** 2 + (y - 7) ** 2 + (z - 7) ** 2 <= 25] = 1
    t, _ = oracle.lee_thin3(ball)
    assert _topology(t) == _topology(ball) and t.sum() <= 3
    for empty in (np.zeros((4, 5, 6), np.uint8),):
        assert not oracle.lee_thin3(empty)[0].any()
    # nonzero values of any size count as object (lee(): binarise)
    w = bar.astype(np.int32) * 7
    assert np.array_equal(oracle.lee_thin3(w)[0], oracle.lee_thin3(bar)[0])


def _otsu_by_within_variance(values):
    v = np.asarray(values, np.float32)
    vmax = float(v.max())
    hist, edges = np.histogram(v, bins=256, range=(0.0, vmax))
    c = ((edges[:-1] + edges[1:]) / 2.0).astype(np.float64)
    h = hist.astype(np.float64)
    best, arg = np.inf, 0
    for k in range(256):
        a, b = h[: k + 1], h[k + 1:]
        wa, wb = a.sum(), b.sum()
        sw = 0.0
        if wa > 0:
            ma = (a * c[: k + 1]).sum() / wa
            sw += (a * (c[: k + 1] - ma) ** 2).sum()
        if wb > 0:
            mb = (b * c[k + 1:]).sum() / wb
            sw += (b * (c[k + 1:] - mb) ** 2).sum()
        if sw < best * (1 - 1e-9):
            best, arg = sw, k
    return np.float32(((edges[:-1] + edges[1:]) / 2.0)[arg])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_otsu_matches_within_class_route(oracle, seed):
    rng = np.random.default_rng(seed)
    a = np.concatenate([rng.normal(1.0, 0.2, 20000), rng.normal(4.0 + seed, 0.5, 7000)])
    v = np.abs(a).astype(np.float32)
    t = oracle.otsu_threshold(v)
    assert np.float32(t) == _otsu_by_within_variance(v)
    assert 1.5 < t < 4.0 + seed


def test_otsu_degenerate(oracle):
    assert oracle.otsu_threshold(np.zeros(100, np.float32)) == 0.0
    # a constant positive field: one class is always empty, every between-class
    # variance is 0 and the first bin's centre wins (argmax of a zero vector)
    t = oracle.otsu_threshold(np.full(50, 2.5, np.float32))
    assert t == np.float32(np.float32(2.5 / 256) / np.float32(2.0))
    # values on bin edges land where numpy puts them
    vmax = np.float32(3.0)
    e = np.linspace(0.0, 3.0, 257).astype(np.float32)
    v = np.concatenate([e, e[::7], [vmax]]).astype(np.float32)
    h_np = np.histogram(v, bins=256, range=(0.0, float(vmax)))[0]
    assert h_np.sum() == v.size
    assert oracle.otsu_threshold(v) == _otsu_by_within_variance(v)


@pytest.mark.parametrize("seed,r", [(0, 0), (1, 1), (2, 2), (3, 0), (4, 1)])
def test_fill_enclosed_matches_definition(oracle, seed, r):
    """The scipy route of saliency_threshold(fill=True) (reconstruct.py:98-116)
    against a flood of the background from the border written from the
    definitions, on small random masks."""
    rng = np.random.default_rng(seed)
    b = rng.random((13 + seed, 17)) < (0.35 + 0.05 * seed)
    np.testing.assert_array_equal(oracle.fill_enclosed(b, r), oracle.fill_enclosed_bruteforce(b, r))


def test_fill_enclosed_known_shapes(oracle):
    ring = np.zeros((9, 9), bool)
    ring[2:7, 2:7] = True
    ring[3:6, 3:6] = False
    full = np.zeros((9, 9), bool)
    full[2:7, 2:7] = True
    np.testing.assert_array_equal(oracle.fill_enclosed(ring, 0), full)
    # a one-pixel gap in the ring lets the flood in; closing with r = 1 shuts it
    gap = ring.copy()
    gap[2, 4] = False
    np.testing.assert_array_equal(oracle.fill_enclosed(gap, 0), gap)
    np.testing.assert_array_equal(oracle.fill_enclosed(gap, 1), full)
    # background is 4-connected: a pixel whose only background neighbours are
    # diagonal is enclosed
    diag = np.zeros((5, 5), bool)
    diag[1, 2] = diag[2, 1] = diag[2, 3] = diag[3, 2] = True
    out = diag.copy()
    out[2, 2] = True
    np.testing.assert_array_equal(oracle.fill_enclosed(diag, 0), out)
    # closing with border_value 0 clears objects touching the image edge
    edge = np.ones((6, 6), bool)
    assert not oracle.fill_enclosed(edge, 1)[0].any()
