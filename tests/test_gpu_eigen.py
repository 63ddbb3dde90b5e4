"""GPU parity of the batched eigen-solvers against the CPU oracle.

Bars: eigenvalues within 1e-6 of the FP64 Jacobi oracle normwise
(max_i |dl_i| / max_j |l_j|), within 2e-5 normwise of the reference-faithful
FP32 restatement (FP32 rounding + the FP64 re-solve of near-degenerate
tensors), eigenvectors satisfying ||A v - l v|| <= 1e-5 ||A|| with the
reference's canonical sign."""
import numpy as np
import pytest

from tiralib_amd import api, synth
from tiralib_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def _special_mats():
    rng = np.random.default_rng(7)
    mats = [np.zeros(9), np.eye(3).ravel() * 0.3, np.diag([1.0, 2.0, 3.0]).ravel(),
            np.diag([3.0, 1.0, 2.0]).ravel()]
    for _ in range(8):
        t = rng.normal(size=3); t /= np.linalg.norm(t)
        mats.append(np.outer(t, t).ravel())                  # stick
        mats.append((np.eye(3) - np.outer(t, t)).ravel())    # plate
    return np.array(mats, np.float32)


def _normwise(ev, ref):
    scale = np.max(np.abs(ref), axis=1)
    scale[scale == 0] = 1
    return np.max(np.abs(ev - ref), axis=1) / scale


def test_evals3_random_vs_fp64(hip, oracle):
    t6, lam = synth.eig_tensors(200_000, seed=11)
    m9 = synth.mat9_from_packed6(t6)
    ev = api.evals3_symmetric(m9)
    ref64 = oracle.eig3_packed6_f64(t6)
    err = _normwise(ev.astype(np.float64), ref64)
    print("max normwise eig3 error vs FP64:", err.max())
    assert err.max() <= 1e-6


def test_evals3_vs_reference_fp32(hip, oracle):
    """Where the reference's own FP32 result is accurate (within 1e-6 of FP64)
    ours matches it within 2e-6; elsewhere the reference is the one that is off
    (its absolute 1e-6 diagonal threshold and the trigonometric formula's
    1/gap error growth) and ours stays within 1e-6 of FP64."""
    t6, _ = synth.eig_tensors(100_000, seed=12)
    m9 = np.concatenate([synth.mat9_from_packed6(t6), _special_mats()])
    ev = api.evals3_symmetric(m9).astype(np.float64)
    ref = oracle.evals3(m9).astype(np.float64)
    t6all = synth.packed6_from_mat9(m9)
    ref64 = oracle.eig3_packed6_f64(t6all)
    ref_ok = _normwise(ref, ref64) <= 1e-6
    print(f"reference FP32 accurate on {ref_ok.mean():.3f} of the tensors")
    assert _normwise(ev, ref)[ref_ok].max() <= 2e-6
    assert _normwise(ev, ref64).max() <= 1e-6
    # exactly diagonal / zero tensors are exact
    sp = _special_mats()
    np.testing.assert_array_equal(api.evals3_symmetric(sp[:4]), oracle.evals3(sp[:4]))


def test_evals3_edge_sizes(hip, oracle):
    for n in [1, 3, 1023, 1024, 1025, 4097]:
        t6, _ = synth.eig_tensors(n, seed=100 + n, degenerate_frac=0.3)
        m9 = synth.mat9_from_packed6(t6)
        ev = api.evals3_symmetric(m9)
        assert ev.shape == (n, 3)
        assert _normwise(ev.astype(np.float64), oracle.eig3_packed6_f64(t6)).max() <= 1e-6


def test_eig3_packed6_device(hip, oracle):
    t6, _ = synth.eig_tensors(300_001, seed=13)
    d = DeviceArray.from_host(t6)
    ev, e0, e2 = api.eig3(d, vectors=True)
    ev, e0, e2 = ev.to_host().astype(np.float64), e0.to_host(), e2.to_host()
    ref = oracle.eig3_packed6_f64(t6)
    assert _normwise(ev, ref).max() <= 1e-6
    # eigen-equation residual for the well-separated tensors
    A = synth.mat9_from_packed6(t6).reshape(-1, 3, 3).astype(np.float64)
    nrm = np.linalg.norm(A, axis=(1, 2))
    for v, lam in ((e0, ev[:, 0]), (e2, ev[:, 2])):
        v = v.astype(np.float64)
        res = np.linalg.norm(np.einsum("nij,nj->ni", A, v) - lam[:, None] * v, axis=1) / nrm
        gaps = np.minimum(ev[:, 1] - ev[:, 0], ev[:, 2] - ev[:, 1]) / np.max(np.abs(ev), axis=1)
        ok = gaps > 1e-3
        assert np.quantile(res[ok], 0.999) <= 1e-5
        assert np.allclose(np.linalg.norm(v, axis=1), 1, atol=1e-5)


def test_evecs3spherical_vs_reference(hip, oracle):
    rng = np.random.default_rng(3)
    t6, _ = synth.eig_tensors(50_000, seed=14, degenerate_frac=0.0, lo=0.1, hi=10)
    m9 = synth.mat9_from_packed6(t6)
    lam = oracle.evals3(m9)
    ang = api.evecs3spherical_symmetric(m9, lam)
    ref = oracle.evecs3spherical(m9, lam)

    def cart(a):
        th, ph = a[:, 0::2], a[:, 1::2]
        return np.stack([np.cos(th) * np.sin(ph), np.sin(th) * np.sin(ph), np.cos(ph)], -1)
    c, cr = cart(ang.astype(np.float64)), cart(ref.astype(np.float64))
    gaps = np.minimum(lam[:, 1] - lam[:, 0], lam[:, 2] - lam[:, 1]) / lam[:, 2]
    ok = gaps > 1e-2
    assert np.max(np.linalg.norm(c - cr, axis=2)[ok]) <= 1e-4


def test_evals2_evecs2_vs_reference(hip, oracle):
    rng = np.random.default_rng(5)
    m = rng.normal(size=(70_001, 4)).astype(np.float32)
    m[:, 2] = m[:, 1]
    ev = api.evals2_symmetric(m).astype(np.float64)
    ref = oracle.evals2(m).astype(np.float64)
    # FP64 truth: mean -+ hypot
    a, b, c = m[:, 0].astype(np.float64), m[:, 1].astype(np.float64), m[:, 3].astype(np.float64)
    h = np.hypot(0.5 * (a - c), b)
    t64 = np.stack([0.5 * (a + c) - h, 0.5 * (a + c) + h], 1)
    scale = np.maximum(np.abs(t64).max(1), 1e-30)
    e_gpu = np.abs(ev - t64).max(1) / scale
    e_ref = np.abs(ref - t64).max(1) / scale
    print(f"evals2 normwise: GPU {e_gpu.max():.2e}, reference FP32 {e_ref.max():.2e}")
    assert e_gpu.max() <= 1e-6
    assert (np.abs(ev - ref).max(1) / scale).max() <= 1e-5
    th = api.evecs2polar_symmetric(m, ref.astype(np.float32))
    thr = oracle.evecs2polar(m, ref.astype(np.float32))
    d = np.abs(np.angle(np.exp(1j * (th.astype(np.float64) - thr))))
    gap = (ref[:, 1] - ref[:, 0]) / np.maximum(np.abs(ref).max(1), 1e-30)
    assert np.max(d[gap > 1e-3]) <= 1e-4
