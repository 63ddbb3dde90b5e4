"""Writes tests/golden/eigen_filter_ref.npz: inputs and outputs of the
reference's own CPU eigen-solvers and Gaussian filter (eigen.h, filter.h),
built from the TIRA sources by `make -C oracle ref` into oracle/_ref.

Run in a container that has the TIRA tree:  python tests/golden/make_golden.py
The fixture is data only (inputs + expected outputs); the tests compare the
oracle restatement against it bit for bit.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def sym9(rng, n, scale_lo=-3, scale_hi=2):
    Q, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    lam = (10 ** rng.uniform(scale_lo, scale_hi, size=(n, 3))) * rng.choice([-1, 1], size=(n, 3), p=[0.2, 0.8])
    A = np.einsum("nij,nj,nkj->nik", Q, lam, Q)
    return A.reshape(n, 9).astype(np.float32)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libtira_ref.so missing: run `make -C oracle ref` with the TIRA tree present")
    rng = np.random.default_rng(20260)
    m9 = np.concatenate([
        sym9(rng, 600),
        np.zeros((1, 9), np.float32),
        np.diag([1.0, 2.0, 3.0]).reshape(1, 9).astype(np.float32),
        np.diag([3.0, 3.0, 3.0]).reshape(1, 9).astype(np.float32),
        np.eye(3, dtype=np.float32).reshape(1, 9) - np.outer([0.6, 0.8, 0.0], [0.6, 0.8, 0.0]).reshape(1, 9).astype(np.float32),
        np.outer([0.0, 0.6, 0.8], [0.0, 0.6, 0.8]).reshape(1, 9).astype(np.float32),
    ])
    # near-degenerate pairs
    Q, _ = np.linalg.qr(rng.normal(size=(60, 3, 3)))
    lam = np.stack([np.ones(60), 1 + 10 ** rng.uniform(-7, -3, 60), np.full(60, 5.0)], 1)
    m9 = np.concatenate([m9, np.einsum("nij,nj,nkj->nik", Q, lam, Q).reshape(60, 9).astype(np.float32)])
    ev3 = O.evals3(m9, impl="ref")
    vec3 = O.evecs3spherical(m9, ev3, impl="ref")
    m4 = rng.normal(size=(500, 4)).astype(np.float32)
    m4[:, 2] = m4[:, 1]
    m4 = np.concatenate([m4, np.array([[0, 0, 0, 0], [1, 0, 0, 1], [2, 1e-4, 1e-4, 2]], np.float32)])
    ev2 = O.evals2(m4, impl="ref")
    vec2 = O.evecs2polar(m4, ev2, impl="ref")
    sigmas = np.array([0.5, 1.0, 1.5, 2.0, 3.3], np.float32)
    g1d = {f"g1d_{i}": O.gaussian1d(O.gaussian_ksize(s), 0.0, float(s), 1.0, impl="ref") for i, s in enumerate(sigmas)}
    vol = rng.normal(size=(14, 15, 16)).astype(np.float32)
    gconv = O.gaussian_convolve3(vol, 1.0, 0.7, 0.5, impl="ref")
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "eigen_filter_ref.npz"),
                        m9=m9, evals3=ev3, evecs3=vec3, m4=m4, evals2=ev2, evecs2=vec2,
                        sigmas=sigmas, vol=vol, gconv=gconv, gconv_sigmas=np.array([1.0, 0.7, 0.5], np.float32),
                        **g1d)
    print("wrote eigen_filter_ref.npz")


if __name__ == "__main__":
    main()
