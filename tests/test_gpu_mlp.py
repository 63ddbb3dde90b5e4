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
    """Per-pair evaluation: one plate voter (eigenvalues 0, 1, 1) in a 21^3
    volume; every receiver but the voter itself holds
    c1 / pi * R net(features) R^T."""
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
    net.vote(d_in, d_out, (S, S, S), w, stick=False, tol=1e30, method=_lib.MLP_PER_PAIR)
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


def _host_table(net, shells):
    """The vote table restated on the host: the NumPy network at the grid's
    features of every distance row, converted by mlp_ref.table_row."""
    p = net.params()
    aff = net.affine()
    rows = [np.zeros((M.TAB_M, 4))]
    rhos = [np.sqrt(np.float32(l2)) / np.float32(net.sigma1) for l2 in shells] + [3.0]
    for rho in rhos:
        X = np.array([M.input_row(f.astype(np.float32)) for f in M.table_features(rho)])
        Y, _ = M.forward(p, X, affine=aff)
        rows.append(M.table_row(Y[:, :6]))
    return np.array(rows)


def _table_values(tab):
    """Cell values (A, B, g, D/2) of a library table (rows, cells, 8)."""
    return tab[:, :, :4].astype(np.float64)


def test_vote_table_matches_host_network(hip):
    """The library's tabulated network (MFMA forward on the grid, conversion to
    A, B, g, D/2) against the NumPy network and conversion, compared through
    the plate-frame components it encodes at every cell (uu = A + D x,
    un = sqrt(x) s (g + D), vv = A, nn = A + B + 2 s^2 (g + D/2)), which do
    not amplify rounding by 1/x; and the distance rows against the lattice."""
    rng = np.random.default_rng(8)
    net = api.PlateMLP(2.0, 0.0, 1, 64, seed=41)
    net.set_affine(rng.normal(0, 1, 6).astype(np.float32), rng.uniform(0.2, 2.0, 6).astype(np.float32))
    tab, shells = net.vote_table()
    assert np.array_equal(shells, M.shells(2.0))
    assert tab.shape == (len(shells) + 2, M.TAB_M, 8)
    got, want = _table_values(tab), _host_table(net, shells)
    # self row: the integrator's (pi / 2)(I - n n^T)
    assert np.allclose(got[0], [np.pi / 2, -np.pi / 2, 0, 0], atol=1e-6)
    # the stored differences are those of consecutive cells
    assert np.allclose(tab[1:, :-1, 4:], np.diff(got[1:], axis=1), atol=1e-5 * np.abs(got[1:]).max())
    x = M.table_grid()
    s = np.sqrt(1 - x)

    def comps(v):
        A, B, g, Dh = v[..., 0], v[..., 1], v[..., 2], v[..., 3]
        return np.stack([A + 2 * Dh * x, np.sqrt(x) * s * (g + 2 * Dh), A, A + B + 2 * s * s * (g + Dh)], -1)

    cg, cw = comps(got[1:]), comps(want[1:])
    # BF16 activations may round differently under the two accumulation
    # orders: Frobenius per distance row
    per_row = np.linalg.norm((cg - cw).reshape(len(cg), -1), axis=1) / np.linalg.norm(cw.reshape(len(cw), -1), axis=1)
    assert per_row.max() <= 3e-3, per_row.max()
    net.close()


def test_table_vote_single_voter(hip):
    """One plate voter in a 21^3 volume: every receiver holds
    c1 / pi * scale * (A I + B nn + s g (nd + dn) + D dd) with (A, B, g, D/2)
    looked up in the fetched table (row of the offset's distance, linear in
    x = 1 - s^2): pins the kernel's row choice, lookup and tensor assembly."""
    rng = np.random.default_rng(2)
    net = api.PlateMLP(2.0, 0.0, 1, 64, seed=43)
    net.set_affine(rng.normal(0, 1, 6).astype(np.float32), rng.uniform(0.2, 2.0, 6).astype(np.float32))
    tab, shells = net.vote_table()
    vals = _table_values(tab)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    n *= np.sign(n[np.argmax(np.abs(n))])
    S, w = 21, 13
    c, r = S // 2, w // 2
    field = np.zeros((S, S, S, 3, 3), np.float32)
    field[c, c, c] = np.eye(3) - np.outer(n, n)
    d_in = DeviceArray.from_host(field.reshape(S, S, S, 9))
    d_out = DeviceArray((S, S, S, 6))
    net.vote(d_in, d_out, (S, S, S), w, stick=False, tol=1e30)
    got = d_out.to_host()
    ref = np.zeros_like(got, dtype=np.float64)
    for z in range(c - r, c + r + 1):
        for y in range(c - r, c + r + 1):
            for x in range(c - r, c + r + 1):
                d = np.array([c - x, c - y, c - z], np.float64)
                l2 = int(d @ d)
                if l2 == 0:
                    row = 0
                elif l2 in shells:
                    row = 1 + int(np.searchsorted(shells, l2))
                else:
                    row = len(shells) + 1
                dh = d / np.sqrt(l2) if l2 else d
                coef = M.table_lookup(vals[row], 1.0 - (n @ dh) ** 2)
                ref[z, y, x] = np.exp(-l2 / 4.0) / np.pi * M.table_vote(coef, n, dh)
    assert _rel(got, ref) <= 1e-5, _rel(got, ref)
    outside = np.ones((S, S, S), bool)
    outside[c - r:c + r + 1, c - r:c + r + 1, c - r:c + r + 1] = False
    assert not got[outside].any()
    net.close()


def test_training_is_deterministic(hip):
    """Every reduction of training (weight gradients, loss, validation, the
    per-bin error table) runs in a fixed order: the same seeds give
    bit-identical parameters, errors and vote tables."""
    out = []
    for _ in range(2):
        net = api.PlateMLP(2.0, 0.0, 1, 64, seed=9)
        tr, va = net.train(nsamples=1 << 17, epochs=3, batch=8192, lr=1e-2, lr_final=1e-4, seed=10)
        out.append((net.params(), tr, va, net.bin_errors(), net.table_errors(), net.vote_table()[0]))
        net.close()
    a, b = out
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2]
    for i in (3, 4, 5):
        assert np.array_equal(a[i], b[i]), i


def test_sheet_volume_generator(hip):
    from tiralib_amd import synth
    v = synth.curved_sheets_volume_device((40, 36, 48), 3, 0.02, seed=9).to_host().reshape(-1, 6)
    v2 = synth.curved_sheets_volume_device((40, 36, 48), 3, 0.02, seed=9).to_host().reshape(-1, 6)
    assert np.array_equal(v, v2)
    nz = np.abs(v).sum(1) > 0
    assert 0.02 <= nz.mean() <= 0.2
    M9 = synth.mat9_from_packed6(v[nz]).reshape(-1, 3, 3).astype(np.float64)
    ev = np.linalg.eigvalsh(M9)
    # every tensor is a plate m (I - n n^T): eigenvalues 0, m, m
    assert np.allclose(ev[:, 0], 0, atol=1e-6) and np.allclose(ev[:, 1], ev[:, 2], rtol=1e-5, atol=1e-6)


def test_trained_network_gain_and_fallback(hip, oracle):
    """Training takes the network from useless to a vote within 1e-3 of the
    FP64 integrator (both evaluations: the tabulated network and the network
    per pair); the per-voter check routes every voter to the integrator at
    tolerance 0 (closed form, the integrator kernel's result to FP32
    summation order); and the trained network votes faster than the
    integrator kernel (the tabulated network than its 20-sample sum, the
    per-pair network than its 64-sample sum)."""
    from tiralib_amd import synth
    sigma, w, shape = 2.0, 13, (32, 40, 48)
    net = api.PlateMLP(sigma, 0.0, 1, 64, seed=5)
    X, T, _ = net.samples(65536, seed=123)
    t = T.to_host()[:, :6]

    def rel_err():
        d = net.forward(X).to_host()[:, :6] - t
        return float(np.sqrt((M.COMP_W * d * d).sum() / (M.COMP_W * t * t).sum()))

    e0 = rel_err()
    tr, va = net.train(nsamples=1 << 22, epochs=100, batch=4096, lr=3e-3, lr_final=1e-6, seed=6)
    e1 = rel_err()
    print("untrained", e0, "trained", e1, "validation", va)
    assert e1 < 0.01 * e0 and e1 < 2e-3
    vol = synth.curved_sheets_volume_device(shape, 4, 0.02, seed=10)
    n = int(np.prod(shape))
    host = vol.to_host()
    o_c = DeviceArray((n, 6))
    _lib.call("tira_tensorvote3_async", vol.ptr, _lib.LAYOUT_PACKED6, o_c.ptr, _lib.LAYOUT_PACKED6, *shape, sigma,
              0.0, w, 1, 1, 1, 64, _lib.PLATE_CLOSED)
    _lib.call("tira_hip_synchronize")
    cl = o_c.to_host()
    rng = np.random.default_rng(1)
    recv = rng.choice(n, 96, replace=False)
    ref = oracle.tensorvote3_f64(synth.mat9_from_packed6(host).reshape(shape + (9,)), sigma, 0.0, 1, w, samples=64,
                                recv=recv, nthreads=8)
    ref6 = ref[:, [0, 1, 2, 4, 5, 8]]
    wgt = np.array([1, 2, 2, 1, 2, 1.0])
    votes = {}
    for method in (_lib.MLP_TABLE, _lib.MLP_PER_PAIR):
        o_m, o_f = DeviceArray((n, 6)), DeviceArray((n, 6))
        nf = net.vote(vol, o_m, shape, w, stick=True, tol=1e30, in_layout=_lib.LAYOUT_PACKED6, count_fallback=True,
                      method=method)
        assert nf == 0
        nf0 = net.vote(vol, o_f, shape, w, stick=True, tol=0.0, in_layout=_lib.LAYOUT_PACKED6, count_fallback=True,
                       method=method)
        assert nf0 == int((np.abs(host).sum(-1) > 0).sum())
        assert _rel(o_f.to_host(), cl) <= 1e-6  # same closed-form plate, different summation order
        got = o_m.to_host().reshape(-1, 6)
        votes[method] = got
        err = np.sqrt((wgt * (got[recv] - ref6) ** 2).sum() / (wgt * ref6 ** 2).sum())
        print("method", method, "MLP vote vs FP64 oracle", err)
        assert err <= 1e-3
    assert _rel(votes[_lib.MLP_TABLE], votes[_lib.MLP_PER_PAIR]) <= 1e-3
    st = _lib.Stream().use()
    e_0, e_1 = _lib.Event(), _lib.Event()

    def ms(fn):
        fn()
        st.synchronize()
        e_0.record(st)
        fn()
        e_1.record(st)
        st.synchronize()
        return e_0.elapsed_ms(e_1)

    o_m = DeviceArray((n, 6))
    t_tab = ms(lambda: net.vote(vol, o_m, shape, w, stick=True, tol=1e30, in_layout=_lib.LAYOUT_PACKED6))
    t_pp = ms(lambda: net.vote(vol, o_m, shape, w, stick=True, tol=1e30, in_layout=_lib.LAYOUT_PACKED6,
                               method=_lib.MLP_PER_PAIR))
    t_sum = {N: ms(lambda: _lib.call("tira_tensorvote3_async", vol.ptr, _lib.LAYOUT_PACKED6, o_c.ptr,
                                     _lib.LAYOUT_PACKED6, *shape, sigma, 0.0, w, 1, 1, 1, N, _lib.PLATE_SUM))
             for N in (20, 64)}
    print("ms mlp table", t_tab, "mlp per pair", t_pp, "integrated N=20", t_sum[20], "N=64", t_sum[64])
    assert t_tab < t_sum[20] and t_pp < t_sum[64]
    net.close()
