"""Writes tests/golden/lapack_qr.npz: third-party LAPACK vectors that pin the
oracle's Householder conventions.

The reference ships no test vectors, so these come from scipy's LAPACK
(OpenBLAS) in this container: dgeqrf (R, reflectors V, tau), dgeqrt with one
block (the compact-WY T of the whole reflector sequence) and dorgqr (thin Q).
The reference follows LAPACK's dlarfg sign convention except where the part of
a column below the diagonal is exactly zero (it still reflects, tau = 2); the
tests handle that case separately.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.linalg.lapack as lapack

SHAPES = [(1, 1), (2, 1), (3, 2), (4, 4), (5, 3), (6, 6), (7, 2), (8, 8), (8, 5),
          (17, 9), (17, 17), (33, 17), (33, 33), (100, 64), (257, 100),
          (3, 5), (5, 8), (17, 33)]


def main() -> None:
    rng = np.random.default_rng(20261203)
    out = {"scipy_version": np.array(scipy.__version__)}
    for (m, n) in SHAPES:
        A = np.asfortranarray(rng.standard_normal((m, n)))
        VR, tau, _, info = lapack.dgeqrf(A)
        assert info == 0
        k = min(m, n)
        _, T, info = lapack.dgeqrt(k, A)
        assert info == 0
        Q, _, info = lapack.dorgqr(np.asfortranarray(VR[:, :k].copy()), tau[:k])
        assert info == 0
        key = f"{m}x{n}"
        out[f"A_{key}"] = A
        out[f"VR_{key}"] = VR
        out[f"tau_{key}"] = tau[:k]
        out[f"T_{key}"] = T[:k, :k]
        out[f"Q_{key}"] = Q
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "lapack_qr.npz"), **out)


if __name__ == "__main__":
    main()
