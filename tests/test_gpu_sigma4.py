"""GPU parity at the vote parameters of the 256^3 learned-plate workload and
the 512 x 512 x 384 pipeline: sigma 4, window 25 (radius 12), power 1.

A 48^3 curved-sheet volume (4 sheets, 2% plate noise) is voted by every 3D
path of the library and compared with the FP64 oracle (cpu::tensorvote3 of
tensorvote.h:1776-1859 restated in double) on sampled receivers: 48 on the
sheets and 16 uniform, relative Frobenius error over the upper-triangle
slots. Bars:
  * spectral vote (FFT convolution, exact plate): <= 1e-5;
  * closed-form gather and the literal 20- and 64-sample sums
    (tensorvote.h:2457-2485, the reference's integrated plate): <= 1e-4;
  * learned plate (tabulated network and network per pair, trained on the
    device for sigma 4): <= 1e-3, the tolerance of the learned surrogate.
The fused vote + eig3 call is checked too: its eigenvalues against FP64
eigenvalues of its own voted tensors (<= 1e-6 normwise: the solver) and
against FP64 eigenvalues of the FP64 oracle's vote (<= 1e-5 over the
sampled receivers: solver and vote)."""
import numpy as np
import pytest

from tiralib_amd import _lib, api, synth
from tiralib_amd._lib import DeviceArray, call

pytestmark = pytest.mark.gpu

SHAPE = (48, 48, 48)
SIGMA, W = 4.0, 25
SLOTS = [0, 1, 2, 4, 5, 8]
CW = np.array([1, 2, 2, 1, 2, 1.0])


def _frob(a, b):
    d = np.asarray(a, np.float64) - b
    return float(np.sqrt((CW * d * d).sum() / (CW * np.asarray(b, np.float64) ** 2).sum()))


@pytest.fixture(scope="module")
def case(hip, oracle):
    vol = synth.curved_sheets_volume_device(SHAPE, 4, 0.02, seed=1041)
    host = vol.to_host().reshape(-1, 6)
    n = host.shape[0]
    rng = np.random.default_rng(11)
    occ = np.flatnonzero(np.abs(host).sum(1) > 0)
    recv = np.sort(np.concatenate([rng.choice(occ, 48, replace=False), rng.choice(n, 16, replace=False)]))
    vol9 = synth.mat9_from_packed6(host).reshape(SHAPE + (9,))
    ref = {N: oracle.tensorvote3_f64(vol9, SIGMA, 0.0, 1, W, samples=N, recv=recv, nthreads=8)[:, SLOTS]
           for N in (20, 64)}
    yield vol, host, recv, ref
    vol.free()


def _vote(vol, mode, samples):
    n = int(np.prod(SHAPE))
    out = DeviceArray((n, 6))
    try:
        call("tira_tensorvote3_async", vol.ptr, _lib.LAYOUT_PACKED6, out.ptr, _lib.LAYOUT_PACKED6, *SHAPE, SIGMA,
             0.0, W, 1, 1, 1, samples, mode)
        call("tira_hip_synchronize")
        return out.to_host()
    finally:
        out.free()


@pytest.mark.parametrize("name,mode,samples,bar", [("spectral", _lib.VOTE_SPECTRAL, 20, 1e-5),
                                                     ("closed", _lib.PLATE_CLOSED, 20, 1e-4),
                                                     ("sum20", _lib.PLATE_SUM, 20, 1e-4),
                                                     ("sum64", _lib.PLATE_SUM, 64, 1e-4)])
def test_vote_sigma4_w25_vs_fp64(case, name, mode, samples, bar):
    vol, host, recv, ref = case
    print(f"sigma 4 w 25 vote: {name}", flush=True)
    got = _vote(vol, mode, samples)[recv]
    err = _frob(got, ref[samples])
    print(name, f"{err:.2e}")
    assert err <= bar, err


def test_learned_plate_sigma4_w25_vs_fp64(case):
    vol, host, recv, ref = case
    print("sigma 4 w 25 vote: learned plate (training)", flush=True)
    net = api.PlateMLP(SIGMA, 0.0, 1, 64, seed=1004)
    try:
        net.train(nsamples=1 << 22, epochs=100, batch=4096, lr=3e-3, lr_final=1e-6, seed=1004)
        n = int(np.prod(SHAPE))
        out = DeviceArray((n, 6))
        try:
            for method in (_lib.MLP_TABLE, _lib.MLP_PER_PAIR):
                nf = net.vote(vol, out, SHAPE, W, stick=True, tol=4e-3, in_layout=_lib.LAYOUT_PACKED6,
                              count_fallback=True, method=method)
                got = out.to_host()[recv]
                err = _frob(got, ref[64])
                print("method", method, f"error {err:.2e}, fallback voters {nf}")
                assert err <= 1e-3, (method, err)
        finally:
            out.free()
    finally:
        net.close()


def test_fused_eigenvalues_sigma4_vs_fp64(case, oracle):
    vol, host, recv, ref = case
    print("sigma 4 w 25 vote + eig3 (fused)", flush=True)
    n = int(np.prod(SHAPE))
    f, ev = DeviceArray((n, 6)), DeviceArray((n, 3))
    try:
        call("tira_tensorvote3_eig_async", vol.ptr, _lib.LAYOUT_PACKED6, f.ptr, _lib.LAYOUT_PACKED6, *SHAPE, SIGMA,
             0.0, W, 1, 1, 1, 20, _lib.PLATE_AUTO, ev.ptr, _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT, None)
        call("tira_hip_synchronize")
        field, evals = f.to_host(), ev.to_host().astype(np.float64)
    finally:
        f.free()
        ev.free()
    # the solver: FP64 eigenvalues of the same FP32 tensors, every voxel
    own = oracle.eig3_packed6_f64(field)
    scale = np.maximum(np.abs(own).max(axis=1), 1e-30)
    live = np.abs(field).max(axis=1) > 0
    solver = (np.abs(evals - own).max(axis=1) / scale)[live]
    print(f"solver normwise max {solver.max():.2e}")
    assert solver.max() <= 1e-6
    # solver and vote: FP64 eigenvalues (LAPACK) of the FP64 oracle vote at the sampled receivers
    r = np.asarray(ref[20], np.float64)
    M = np.stack([r[:, [0, 1, 2]], r[:, [1, 3, 4]], r[:, [2, 4, 5]]], axis=1)
    want = np.linalg.eigvalsh(M)
    got = evals[recv]
    e = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    per = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    print(f"eigenvalues vs FP64 oracle vote: {e:.2e} (Frobenius over receivers); per-eigenvalue relative "
          f"median {np.median(per):.2e} max {per.max():.2e}")
    assert e <= 1e-5
