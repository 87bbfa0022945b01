"""Differentiation of eigenvalues through the batched Schur decomposition.

``eig_vjp_batched`` is the reverse-mode rule (the rrule of the eigenvalue map
A -> lambda(A)) for real matrices: given the Schur factors of a batch and a
cotangent gbar_k = dL/dRe(lambda_k) + i dL/dIm(lambda_k) per eigenvalue slot,
it returns

    Abar = Re( sum_k gbar_k / conj(y_k^H x_k) y_k x_k^H )

with right and left eigenvectors recomputed on the GPU from T and Z (they are
not stored by the forward pass).  ``SchurEigvals`` wraps forward and backward
as a torch.autograd.Function over device tensors.
"""
import ctypes

import numpy as np

from . import _lib
from .api import eig_offsets, pack


def eig_vjp_batched(ns, off, T, Z, gbar, device=0):
    """Host arrays: T, Z packed real (as from gschur_batched), gbar complex per
    eigenvalue slot.  Returns (Abar packed real, info)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    off = np.ascontiguousarray(off, dtype=np.int64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    g = np.ascontiguousarray(gbar, dtype=np.complex128)
    Abar = np.zeros_like(T)
    info = np.zeros(len(ns), dtype=np.int32)
    h = _lib.handle(device)
    p = lambda x: ctypes.c_void_p(x.ctypes.data)
    rc = _lib.lib().gs_eig_vjp_batched_d(h, len(ns), p(ns), p(off), p(T), p(Z), p(g), p(Abar), p(info), 0)
    _lib.check(rc, h, "gs_eig_vjp_batched_d")
    return Abar, info


def _addr(t):
    return ctypes.c_void_p(t.data_ptr())


def make_schur_eigvals(torch):
    """Build the autograd Function (torch is imported by the caller)."""

    class SchurEigvals(torch.autograd.Function):
        """lambda = eigvals(A_i) for a packed batch of real matrices on the GPU.

        forward(A_packed, ns, off) -> complex eigenvalues packed per matrix
        (Schur diagonal order); backward computes Abar with the VJP kernel.
        """

        @staticmethod
        def forward(ctx, A, ns, off):
            ns = np.ascontiguousarray(ns, dtype=np.int32)
            off = np.ascontiguousarray(off, dtype=np.int64)
            # the library works on its own stream: A must be complete first
            torch.cuda.current_stream(A.device).synchronize()
            T = torch.zeros_like(A)
            Z = torch.zeros_like(A)
            W = torch.zeros(2 * int(ns.sum()), dtype=torch.float64, device=A.device)
            info = torch.zeros(len(ns), dtype=torch.int32, device=A.device)
            h = _lib.handle(A.device.index or 0)
            rc = _lib.lib().gs_schur_batched_d(h, len(ns), ctypes.c_void_p(ns.ctypes.data),
                                               ctypes.c_void_p(off.ctypes.data), _addr(A), _addr(T), _addr(Z),
                                               _addr(W), _addr(info), _lib.GS_WANTZ | _lib.GS_SCALE |
                                               _lib.GS_DEVICE_PTRS)
            _lib.check(rc, h, "gs_schur_batched_d")
            ctx.save_for_backward(T, Z)
            ctx.ns, ctx.off = ns, off
            return torch.view_as_complex(W.view(-1, 2))

        @staticmethod
        def backward(ctx, gw):
            T, Z = ctx.saved_tensors
            ns, off = ctx.ns, ctx.off
            torch.cuda.current_stream(T.device).synchronize()
            # torch's gradient of a real loss w.r.t. a complex output is
            # dL/dRe + i dL/dIm, the cotangent convention of the kernel
            g = torch.view_as_real(gw.contiguous()).reshape(-1).contiguous()
            Abar = torch.zeros_like(T)
            info = torch.zeros(len(ns), dtype=torch.int32, device=T.device)
            h = _lib.handle(T.device.index or 0)
            rc = _lib.lib().gs_eig_vjp_batched_d(h, len(ns), ctypes.c_void_p(ns.ctypes.data),
                                                 ctypes.c_void_p(off.ctypes.data), _addr(T), _addr(Z), _addr(g),
                                                 _addr(Abar), _addr(info), _lib.GS_DEVICE_PTRS)
            _lib.check(rc, h, "gs_eig_vjp_batched_d")
            return Abar, None, None

    return SchurEigvals


__all__ = ["eig_vjp_batched", "make_schur_eigvals", "pack", "eig_offsets"]
