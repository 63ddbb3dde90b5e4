"""GPU parity of the plate-vote surrogate: training targets against the FP64
integrator oracle, MFMA forward / backward / Adam against the NumPy
restatement (same BF16 roundings), and the vote kernel's in-register
network rounds against the host evaluation of the same network."""
import numpy as np
import pytest

from oracle import mlp_ref as M
from tiralib_amd import _lib, api
from tiralib_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def _x(X):
    return M.bf16_bits_to_f32(X.to_host()).reshape(X.shape)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_samples_match_fp64_integrator(hip, oracle):
    net = api.PlateMLP(2.0, 0.0, 1, 64)
    X, T, R = net.samples(2048, seed=11, raw=True)
    x, t, r = _x(X), T.to_host(), R.to_host()
    radii = np.linalg.norm(r[:, :3], axis=1)
    assert radii.max() <= 6.0 * (1 + 1e-6) and radii.min() > 0
    assert np.allclose(np.linalg.norm(r[:, 3:6], axis=1), 1.0, atol=1e-6)
    for i in range(0, 2048, 41):
        d = r[i, :3].astype(np.float64)
        n = r[i, 3:6].astype(np.float64)
        n /= np.linalg.norm(n)
        want, c1 = M.target(d, n, 2.0, 0.0, 1, 64, oracle.plate3_sum_f64)
        assert np.allclose(t[i, :6], want, rtol=1e-4, atol=2e-5 * np.abs(want).max()), i
        assert abs(r[i, 6] - c1) <= 1e-5 * max(c1, 1e-3)
        f = M.features(d, n, 2.0)
        got = x[i, :5].astype(np.float64) + x[i, 5:10]
        assert np.allclose(got, f, atol=5e-5), i
        assert x[i, 10] == 1.0 and not x[i, 11:].any()
    assert not t[:, 6:].any()


@pytest.mark.parametrize("with_affine", [False, True])
def test_forward_and_gradients_match_reference(hip, with_affine):
    net = api.PlateMLP(2.0, 0.0, 1, 64, seed=21)
    p = net.params()
    aff = None
    if with_affine:
        rng = np.random.default_rng(3)
        aff = (rng.normal(0, 1, 6).astype(np.float32), rng.uniform(0.1, 2.0, 6).astype(np.float32))
        net.set_affine(*aff)
        mu, sd = net.affine()
        assert np.array_equal(mu, aff[0]) and np.array_equal(sd, aff[1])
    X, T, _ = net.samples(512, seed=12)
    x, t = _x(X), T.to_host()
    Y = net.forward(X).to_host()
    Yr, _ = M.forward(p, x, affine=aff)
    assert _rel(Y[:, :6], Yr[:, :6]) <= 2e-3
    assert not Y[:, 6:].any()
    g, sums = net.gradients(X, T)
    gr = M.gradients(p, x, t, affine=aff)
    for name, off, rows, cols in api.PlateMLP.LAYERS:
        a, b = g[off:off + rows * cols], gr[off:off + rows * cols]
        assert _rel(a, b) <= 1e-2, (name, _rel(a, b))
    assert abs(sums[0] / 512 - M.loss(Yr, t)) <= 5e-3 * M.loss(Yr, t)
    net.close()


def test_adam_step_matches_reference(hip):
    net = api.PlateMLP(2.0, 0.0, 1, 64, seed=22)
    p0 = net.params()
    X, T, _ = net.samples(1024, seed=13)
    g, _ = net.gradients(X, T)
    net.step(X, T, 1e-3)
    p1 = net.params()
    pr, _, _ = M.adam_step(p0, g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, 1e-3)
    # each step is ~lr * sign(g); entries whose gradient is within float noise of 0 may differ
    close = np.isclose(p1, pr, rtol=0, atol=2e-6)
    assert close.mean() > 0.999, close.mean()
    assert np.array_equal(p1.reshape(-1)[~M.mask() & ~_residual()], p0[~M.mask() & ~_residual()])
    W1 = p1[:2048].reshape(64, 32)
    assert np.array_equal(W1[:, 5:10], W1[:, :5])


def _residual():
    r = np.zeros(M.NPARAM, bool)
    r[0:2048].reshape(64, 32)[:, 5:10] = True
    return r


@pytest.mark.parametrize("seed", [0, 1])
def test_vote_kernel_rounds_equal_host_network(hip, seed):
    """One plate voter (eigenvalues 0, 1, 1) in a 21^3 volume: every receiver
    but the voter itself holds c1 / pi * R net(features) R^T."""
    rng = np.random.default_rng(seed)
    net = api.PlateMLP(2.0, 0.0, 1, 64, seed=30 + seed)
    p = net.params()
    aff = (rng.normal(0, 1, 6).astype(np.float32), rng.uniform(0.2, 2.0, 6).astype(np.float32))
    net.set_affine(*aff)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    n *= np.sign(n[np.argmax(np.abs(n))])  # canonical sign: largest component positive
    S = 21
    c = S // 2
    field = np.zeros((S, S, S, 3, 3), np.float32)
    field[c, c, c] = np.eye(3) - np.outer(n, n)
    d_in = DeviceArray.from_host(field.reshape(S, S, S, 9))
    d_out = DeviceArray((S, S, S, 6))
    w = 13
    net.vote(d_in, d_out, (S, S, S), w, stick=False, tol=1e30)
    got = d_out.to_host()
    r = w // 2
    rows, want, idx = [], [], []
    for z in range(c - r, c + r + 1):
        for y in range(c - r, c + r + 1):
            for x in range(c - r, c + r + 1):
                d = np.array([c - x, c - y, c - z], np.float64)  # receiver -> voter, (du, dv, dw)
                if not d.any():
                    continue
                f = M.features(d, n, 2.0).astype(np.float32)
                rows.append(M.input_row(f))
                idx.append((z, y, x))
                want.append(np.exp(-(d @ d) / 4.0) / np.pi)
    Y, _ = M.forward(p, np.array(rows), affine=aff)
    ref = np.zeros_like(got, dtype=np.float64)
    for k, (z, y, x) in enumerate(idx):
        ref[z, y, x] = want[k] * M.to_field(Y[k, :6], n, np.array([c - x, c - y, c - z], np.float64))
    mask = np.ones((S, S, S), bool)
    mask[c, c, c] = False
    assert _rel(got[mask], ref[mask]) <= 2e-3, _rel(got[mask], ref[mask])
    outside = np.ones((S, S, S), bool)
    outside[c - r:c + r + 1, c - r:c + r + 1, c - r:c + r + 1] = False
    assert not got[outside].any()
