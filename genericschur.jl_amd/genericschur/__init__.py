"""genericschur — GenericSchur.jl's Schur hot path on AMD Instinct MI355X.

Complex and real Schur decompositions (Householder Hessenberg reduction,
Francis QR with deflation and Schur-vector accumulation), eigenvectors of a
Schur factorization, an eigenvalue vector-Jacobian product, and the
Hermitian / real symmetric eigenproblem by QR iteration, computed by
hand-written HIP kernels for gfx950 behind the C ABI in
include/gschur_mi355x.h.
"""
from .api import (ArgumentError, DimensionMismatch, Schur, UnconvergedException, eigvals_, eigvecs, geigvecs,
                  gschur, gschur_, gschur_batched, gschur_batched_device, gschur_hessenberg_, pack, schur_,
                  synchronize, unpack, eig_offsets, set_stream, gschur_device, dgemm_device, batched_max_n,
                  Balancer, Eigen, balance_, lmul_, ldiv_, triangularize, enormalize_, eigen_, ordschur_, eigcond,
                  eigvalscond, subspacesep, heev_max_n, hessenberg_hermitian, heev_batched, geigen_hermitian,
                  geigvals_hermitian, steqr_, gschur_symtridiag, geigen_symtridiag, geigvals_symtridiag,
                  IllConditionException)
from ._lib import GSError, last_stats, last_launches

__all__ = [
    "ArgumentError", "DimensionMismatch", "Schur", "UnconvergedException", "GSError",
    "gschur", "gschur_", "schur_", "gschur_hessenberg_", "eigvals_", "eigvecs", "geigvecs",
    "gschur_batched", "gschur_batched_device", "pack", "unpack", "eig_offsets", "synchronize", "set_stream", "last_stats",
    "gschur_device", "dgemm_device", "batched_max_n", "last_launches",
    "Balancer", "Eigen", "balance_", "lmul_", "ldiv_", "triangularize", "enormalize_", "eigen_",
    "ordschur_", "eigcond", "eigvalscond", "subspacesep",
    "heev_max_n", "hessenberg_hermitian", "heev_batched", "geigen_hermitian", "geigvals_hermitian", "steqr_",
    "gschur_symtridiag", "geigen_symtridiag", "geigvals_symtridiag", "IllConditionException",
]
