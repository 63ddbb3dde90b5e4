"""GPU parity of the 3D stick + plate vote against the CPU oracle.

Bar: relative Frobenius error <= 1e-4 against the
reference-faithful FP32 restatement of cpu::tensorvote3 on identical seeded
inputs, compared on the flat slots {0,1,2,4,5,8} (the reference CPU path
accumulates only the upper triangle of the plate, tensorvote.h:1747-1751, so
slots 3, 6, 7 differ from the GPU's symmetric output by design)."""
import numpy as np
import pytest

from tiralib_amd import _lib, api, synth
from tiralib_amd._lib import DeviceArray
from tests.conftest import rel_frob

pytestmark = pytest.mark.gpu

SLOTS = [0, 1, 2, 4, 5, 8]
NT = 8  # oracle threads


def _random_field(shape, seed, sparse=0.0, lo=0.1, hi=10.0):
    """Random SPD tensors, 5% near-degenerate. The default eigenvalue range keeps
    off-diagonal energy far above the reference's absolute 1e-6 "diagonal"
    threshold (eigen.h:399-406); small-tensor fields are covered separately."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    t6, _ = synth.eig_tensors(n, seed=seed, lo=lo, hi=hi, degenerate_frac=0.05)
    m9 = synth.mat9_from_packed6(t6)
    if sparse:
        m9[rng.random(n) < sparse] = 0.0
    return m9.reshape(*shape, 9)


def _check(gpu, ref, tol=1e-4):
    g = np.asarray(gpu).reshape(-1, 9)[:, SLOTS]
    r = np.asarray(ref).reshape(-1, 9)[:, SLOTS]
    e = rel_frob(g, r)
    print(f"rel Frobenius error {e:.3e}")
    assert e <= tol, e
    return e


@pytest.mark.parametrize("shape,sigma,w", [((20, 22, 24), 1.5, 7), ((13, 70, 65), 1.0, 5)])
def test_vote3_random_field(hip, oracle, shape, sigma, w):
    f = _random_field(shape, seed=21)
    gpu = api.tensorvote3(f, sigma, 0.0, w, 1, samples=20)
    ref = oracle.tensorvote3(f, sigma, 0.0, 1, w, samples=20, nthreads=NT)
    _check(gpu, ref)
    # output is symmetric
    g = gpu.reshape(-1, 3, 3)
    assert np.array_equal(g, np.transpose(g, (0, 2, 1)))


def test_vote3_tubes_sheets(hip, oracle):
    f = synth.tubes_sheets_volume((32, 32, 32), n_straight=3, n_helix=2, n_sheets=2, seed=5)
    gpu = api.tensorvote3(f, 2.0, 0.0, 13, 1, samples=20)
    ref = oracle.tensorvote3(f, 2.0, 0.0, 1, 13, samples=20, nthreads=NT)
    _check(gpu, ref)


@pytest.mark.parametrize("power,samples,stick,plate,sigma2", [
    (1, 2, True, True, 0.0),     # N < p+2: literal sum kernel
    (1, 1, False, True, 0.0),
    (1, 0, True, True, 0.0),     # samples == 0: analytic plate == exact integral
    (2, 20, True, True, 0.7),    # closed form P=2, sigma2 > 0
    (3, 20, True, True, 0.0),
    (4, 7, True, True, 0.5),
    (5, 10, True, True, 0.0),    # p > 4: literal sum
    (2, 3, True, True, 0.0),     # N < p+2 at p=2
    (1, 20, True, False, 0.0),
    (1, 20, False, True, 1.2),
])
def test_vote3_modes(hip, oracle, power, samples, stick, plate, sigma2):
    f = _random_field((12, 14, 16), seed=31 + power + samples)
    gpu = api.tensorvote3(f, 1.3, sigma2, 7, power, stick=stick, plate=plate, samples=samples)
    ref = oracle.tensorvote3(f, 1.3, sigma2, power, 7, stick=stick, plate=plate, samples=samples, nthreads=NT)
    ref64 = oracle.tensorvote3_f64(f, 1.3, sigma2, power, 7, stick=stick, plate=plate, samples=samples,
                                   nthreads=NT)
    r32 = ref.reshape(-1, 9)[:, SLOTS]
    r64 = ref64.reshape(-1, 9)[:, SLOTS]
    inherent = rel_frob(r32, r64)  # the reference FP32 path's own distance from FP64
    print(f"reference FP32 vs FP64: {inherent:.3e}")
    _check(gpu, ref, tol=max(1e-4, 3 * inherent))
    _check(gpu, ref64, tol=max(1e-4, 3 * inherent))


def test_vote3_none(hip):
    f = _random_field((5, 6, 7), seed=3)
    out = api.tensorvote3(f, 1.0, 0.0, 3, 1, stick=False, plate=False)
    assert np.all(out == 0)


def test_vote3_literal_vs_closed(hip, oracle):
    """The literal N-sample kernel and the closed form agree with the FP64
    oracle (and with each other) for N >= p + 2."""
    f = _random_field((16, 16, 16), seed=41)
    d = DeviceArray.from_host(f)
    outs = {}
    for mode in (_lib.PLATE_SUM, _lib.PLATE_CLOSED):
        o = DeviceArray((16, 16, 16, 9))
        api.tensorvote3_async(d, o, (16, 16, 16), 2.0, 0.0, 13, 1, samples=20, plate_mode=mode)
        _lib.call("tira_hip_synchronize")
        outs[mode] = o.to_host().reshape(-1, 9)[:, SLOTS]
    ref64 = oracle.tensorvote3_f64(f, 2.0, 0.0, 1, 13, samples=20, nthreads=NT).reshape(-1, 9)[:, SLOTS]
    e_sum, e_cl = rel_frob(outs[_lib.PLATE_SUM], ref64), rel_frob(outs[_lib.PLATE_CLOSED], ref64)
    e_sc = rel_frob(outs[_lib.PLATE_SUM], outs[_lib.PLATE_CLOSED])
    print(f"sum vs FP64 {e_sum:.3e}, closed vs FP64 {e_cl:.3e}, sum vs closed {e_sc:.3e}")
    assert e_sum <= 1e-4 and e_cl <= 1e-4 and e_sc <= 1e-4


def test_vote3_small_tensors_scale_invariant(hip, oracle):
    """Tensors with norm ~1e-3: the reference's absolute 1e-6 off-diagonal
    threshold (eigen.h:399-406) treats many as diagonal and votes with the
    wrong eigenvectors; this implementation's thresholds are relative, so it
    tracks the FP64 oracle. Documented deviation from the reference."""
    f = _random_field((12, 12, 12), seed=77, lo=1e-4, hi=1e-2)
    gpu = api.tensorvote3(f, 1.5, 0.0, 7, 1, samples=20).reshape(-1, 9)[:, SLOTS]
    ref = oracle.tensorvote3(f, 1.5, 0.0, 1, 7, samples=20, nthreads=NT).reshape(-1, 9)[:, SLOTS]
    ref64 = oracle.tensorvote3_f64(f, 1.5, 0.0, 1, 7, samples=20, nthreads=NT).reshape(-1, 9)[:, SLOTS]
    e_gpu, e_ref = rel_frob(gpu, ref64), rel_frob(ref, ref64)
    print(f"GPU vs FP64 {e_gpu:.3e}; reference FP32 vs FP64 {e_ref:.3e}")
    assert e_gpu <= 1e-4
    assert e_gpu <= e_ref


def test_vote3_tk_entry(hip, oracle):
    f = _random_field((10, 12, 14), seed=51)
    m9 = f.reshape(-1, 9)
    lam = oracle.evals3(m9)
    ang = oracle.evecs3spherical(m9, lam)
    th, ph = ang[:, 0::2].astype(np.float64), ang[:, 1::2].astype(np.float64)
    cart = np.stack([np.cos(th) * np.sin(ph), np.sin(th) * np.sin(ph), np.cos(ph)], -1).astype(np.float32)
    e0, e2 = np.ascontiguousarray(cart[:, 0]), np.ascontiguousarray(cart[:, 2])
    gpu = api.tensorvote3_tk(lam, e2, e0, (10, 12, 14), 1.5, 0.0, 1, 9, samples=20)
    ref = oracle.tensorvote3_tk(lam, e2, e0, (10, 12, 14), 1.5, 0.0, 1, 9, samples=20, nthreads=NT)
    _check(gpu, ref)


def test_vote3_device_pointers(hip):
    f = _random_field((9, 10, 11), seed=61)
    host = api.tensorvote3(f, 1.0, 0.0, 5, 1)
    dev = api.tensorvote3(DeviceArray.from_host(f), 1.0, 0.0, 5, 1).to_host()
    np.testing.assert_array_equal(host, dev)


def test_vote3_config2_sampled(hip, oracle):
    """128^3 tubes + sheets, sigma=2, w=13, N=20: GPU vs oracle on 1500
    sampled receivers (FP32 faithful) and against the FP64 oracle."""
    f = synth.tubes_sheets_volume((128, 128, 128), seed=synth.SEED_CFG2)
    gpu = api.tensorvote3(f, 2.0, 0.0, 13, 1, samples=20).reshape(-1, 9)
    rng = np.random.default_rng(0)
    mag = np.abs(f.reshape(-1, 9)).sum(1)
    near = np.flatnonzero(mag > 0)
    recv = np.unique(np.concatenate([rng.choice(near, 1000, replace=False), rng.choice(f.shape[0] ** 3, 500)]))
    ref = oracle.tensorvote3(f, 2.0, 0.0, 1, 13, samples=20, recv=recv, nthreads=NT)
    _check(gpu[recv], ref)
    ref64 = oracle.tensorvote3_f64(f, 2.0, 0.0, 1, 13, samples=20, recv=recv[:300], nthreads=NT)
    e64 = rel_frob(gpu[recv[:300]][:, SLOTS], ref64[:, SLOTS])
    print(f"config-2 sampled error vs FP64 oracle: {e64:.3e}")
    assert e64 <= 1e-4
