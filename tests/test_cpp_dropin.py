"""The C++ drop-in header (include/tira/hip/tira.hpp): reference-style call
sites compile against it with g++ (CPU), and on the GPU their results match
the oracle (eig3, 3D Gaussian, 3D vote, 2D neighbour table) and invalid
arguments throw std::invalid_argument."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "tiralib_amd")


def _build(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libtira_hip.so")):
        pytest.skip("libtira_hip.so not built")
    exe = str(tmp_path / "dropin")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{ROOT}/include",
                        os.path.join(ROOT, "tests", "cpp", "dropin.cpp"), "-o", exe, f"-L{LIBDIR}", "-ltira_hip",
                        f"-Wl,-rpath,{LIBDIR}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_dropin_header_compiles(tmp_path):
    _build(tmp_path)


def _run(exe, case):
    r = subprocess.run([exe, case], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.split()


@pytest.mark.gpu
def test_dropin_results_match_oracle(tmp_path, hip, oracle):
    exe = _build(tmp_path)
    ev = np.array(_run(exe, "eig3"), np.float64).reshape(2, 3)
    m = np.array([4, 1, 0, 1, 3, 0, 0, 0, 2, 2, 0, 0, 0, 2, 0, 0, 0, 5], np.float64).reshape(2, 3, 3)
    assert np.allclose(ev, np.linalg.eigvalsh(m), rtol=1e-6)

    g = _run(exe, "gauss3")
    ow, oh, od = (int(v) for v in g[:3])
    got = np.array(g[3:], np.float32).reshape(od, oh, ow)
    W, H, D = 20, 18, 16
    src = ((np.arange(W * H * D, dtype=np.uint64) * 2654435761) % 1000).astype(np.float32) / 1000  # size_t product
    ref = oracle.gaussian_convolve3(src.reshape(D, H, W), 1.0, 1.0, 1.0)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-5

    S = 12
    vol = np.zeros((S * S * S, 9), np.float32)
    vol[::7, 0], vol[::7, 4], vol[::7, 8] = 1.0, 0.5, 0.25
    out = np.array(_run(exe, "vote3"), np.float32).reshape(-1, 9)
    ref = oracle.tensorvote3_f64(vol.reshape(S, S, S, 9), 1.5, 0.0, 1, 7, samples=20).reshape(-1, 9)
    up = [0, 1, 2, 4, 5, 8]
    assert np.linalg.norm(out[:, up] - ref[:, up]) <= 1e-4 * np.linalg.norm(ref[:, up])

    vol = np.zeros((S * S * S, 9), np.float32)
    vol[::5, 0], vol[::5, 4], vol[::5, 8] = 1.0, 0.5, 0.25
    out = np.array(_run(exe, "atv3"), np.float32).reshape(-1, 9)
    ref = oracle.tensorvote3_atv(vol.reshape(S, S, S, 9), 1.5, 7).reshape(-1, 9)
    assert np.linalg.norm(out - ref) <= 1e-5 * np.linalg.norm(ref)

    a2 = np.array(_run(exe, "atv2"), np.float32)
    fin, fout = a2[:10 * 13 * 4].reshape(10, 13, 4), a2[10 * 13 * 4:].reshape(10, 13, 4)
    ref = oracle.tensorvote2_atv(fin, 2.0, 9)
    assert np.abs(fout - ref).max() <= 1e-5 * np.abs(ref).max()

    nb = _run(exe, "nb2")
    count = int(nb[0])
    rows = np.array(nb[1:], np.float64).reshape(count, 4)
    ref_nb = oracle.build_neighbors2d(7, 2.0, 0.0, 1, 10, True, True)
    assert count == len(ref_nb)
    assert np.array_equal(rows[:, 0], ref_nb["du"]) and np.array_equal(rows[:, 1], ref_nb["dv"])
    assert np.array_equal(rows[:, 2].astype(np.float32), ref_nb["c1"])
    assert np.allclose(rows[:, 3], ref_nb["Va"], rtol=1e-6, atol=1e-7)

    assert _run(exe, "bad") == ["invalid_argument"]
