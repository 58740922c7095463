"""Writes tests/golden/lapack_qr.npz: third-party LAPACK vectors that pin the
oracle's Householder conventions.

The reference ships no test vectors, so these come from scipy's LAPACK
(OpenBLAS) in this container: dgeqrf (R, reflectors V, tau), dgeqrt with one
block (the compact-WY T of the whole reflector sequence) and dorgqr (thin Q).
The reference follows LAPACK's dlarfg sign convention except where the part of
a column below the diagonal is exactly zero (it still reflects, tau = 2); the
tests handle that case separately.

It also writes three of these matrices as binary MAT files (A_17x9.mat,
VR_17x9.mat, A_17x9_f32.mat) for the file-format tests.

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
    write_mat_files(out, here)


def write_mat(path: str, A: np.ndarray) -> None:
    """The reference's binary MAT layout (Mat-5.2/_extra/__File.h:56-66):
    int32 rows, int32 cols, then the elements in storage order (column-major
    for MAT<T>), written here with numpy alone so that the library's reader
    is checked against an independent writer."""
    A = np.asfortranarray(A)
    with open(path, "wb") as f:
        f.write(np.array(A.shape, dtype="<i4").tobytes())
        f.write(A.tobytes(order="F"))


def write_mat_files(out: dict, here: str) -> None:
    """A few of the LAPACK fixtures also as MAT files: the 17x9 input and
    its dgeqrf output (FP64), and the input rounded to FP32."""
    write_mat(os.path.join(here, "A_17x9.mat"), out["A_17x9"])
    write_mat(os.path.join(here, "VR_17x9.mat"), out["VR_17x9"])
    write_mat(os.path.join(here, "A_17x9_f32.mat"), out["A_17x9"].astype(np.float32))


if __name__ == "__main__":
    main()
