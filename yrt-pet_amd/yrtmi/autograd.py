"""Differentiable projector: a torch.autograd.Function over the yrtmi
kernels (the reference's autograd wrapper around pyyrtpet's
ProjectionOper: scripts/deep_image_prior/run_dip.py:121-143,
python/pyyrtpet/_pyyrtpet.py:82-157).

``ProjectionOperator`` holds a context, a projector, a LOR set and an image
geometry and applies A (forward projection) and A^T (the hand-written
backprojection kernel).  Backprojection is A's exact adjoint, so the
Function's backward is one backprojection of the incoming gradient; nothing
per voxel segment is kept between forward and backward: the backward walks
the LORs again (chunk by chunk when ``chunk`` is set).

Dtypes: float32 tensors run the FP32 kernels (tile-planned backprojection
when a plan is given); float64 tensors run the FP64 projector
(yrtmi_forward_f64 / yrtmi_back_f64), for double-precision gradient checks.

Batches: B images with one LOR set each are one B-frame image whose LORs
carry their image index as the dynamic frame, so a batch is one launch.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import capi
from .core import Context, ImageParams, LorSet, Plan, Projector, _ptr


class ProjectionOperator:
    """A and A^T for one projector, image geometry and LOR set."""

    def __init__(self, ctx: Context, proj: Projector, im: ImageParams, lors: LorSet,
                 plan: Plan | None = None, chunk: int | None = None):
        if chunk is not None and lors.kind == capi.SRC_HISTO_SUBSET:
            raise ValueError("a histogram subset iterator cannot be chunked")
        self.ctx, self.proj, self.im, self.lors, self.plan = ctx, proj, im, lors, plan
        self.chunk = chunk
        self.n = lors.n
        self.image_shape = im.shape

    def _pieces(self):
        if self.chunk is None or self.chunk >= self.n:
            yield 0, self.n, self.lors
            return
        for a in range(0, self.n, self.chunk):
            b = min(self.n, a + self.chunk)
            yield a, b, self.lors.slice(a, b)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """y = A x (one value per LOR)."""
        x = x.contiguous()
        if x.numel() != self.im.nvox * self.im.nt:
            raise ValueError("image size does not match the operator's geometry")
        y = torch.empty(self.n, device=x.device, dtype=x.dtype)
        ci = self.im.c()
        lib = capi.lib()
        for a, b, L in self._pieces():
            cl = L.c()
            if x.dtype == torch.float32:
                capi.check(lib.yrtmi_forward(self.ctx.h, self.proj.h, C.byref(ci), _ptr(x),
                                             C.byref(cl), _ptr(y[a:b])))
            elif x.dtype == torch.float64:
                capi.check(lib.yrtmi_forward_f64(self.ctx.h, self.proj.h, C.byref(ci),
                                                 _ptr(x), C.byref(cl), _ptr(y[a:b])))
            else:
                raise TypeError("float32 or float64 images")
        return y

    def adjoint(self, v: torch.Tensor) -> torch.Tensor:
        """A^T v (a zero image plus the backprojection of v)."""
        v = v.contiguous()
        if v.numel() != self.n:
            raise ValueError("one value per LOR expected")
        x = torch.zeros(self.image_shape, device=v.device, dtype=v.dtype)
        ci = self.im.c()
        lib = capi.lib()
        if v.dtype == torch.float32 and self.plan is not None:
            cl = self.lors.c()
            capi.check(lib.yrtmi_back_planned(self.ctx.h, self.proj.h, C.byref(ci), _ptr(x),
                                              C.byref(cl), _ptr(v), self.plan.h))
            return x
        for a, b, L in self._pieces():
            cl = L.c()
            if v.dtype == torch.float32:
                capi.check(lib.yrtmi_back(self.ctx.h, self.proj.h, C.byref(ci), _ptr(x),
                                          C.byref(cl), _ptr(v[a:b])))
            elif v.dtype == torch.float64:
                capi.check(lib.yrtmi_back_f64(self.ctx.h, self.proj.h, C.byref(ci), _ptr(x),
                                              C.byref(cl), _ptr(v[a:b])))
            else:
                raise TypeError("float32 or float64 values")
        return x


class Project(torch.autograd.Function):
    """y = A x with backward x_grad = A^T y_grad."""

    @staticmethod
    def forward(fctx, x, op: ProjectionOperator):
        fctx.op = op
        return op.forward(x)

    @staticmethod
    def backward(fctx, gy):
        return fctx.op.adjoint(gy), None


class BackProject(torch.autograd.Function):
    """x = A^T v with backward v_grad = A x_grad."""

    @staticmethod
    def forward(fctx, v, op: ProjectionOperator):
        fctx.op = op
        return op.adjoint(v)

    @staticmethod
    def backward(fctx, gx):
        return fctx.op.forward(gx), None


def project(x: torch.Tensor, op: ProjectionOperator) -> torch.Tensor:
    return Project.apply(x, op)


def backproject(v: torch.Tensor, op: ProjectionOperator) -> torch.Tensor:
    return BackProject.apply(v, op)


def poisson_nll(ybar: torch.Tensor, m: torch.Tensor, b) -> torch.Tensor:
    """Poisson negative log-likelihood sum(ybar + b - m log(ybar + b)) of
    measured counts m given the expected counts ybar and background b
    (randoms + scatter)."""
    q = ybar + b
    return (q - m * torch.log(q)).sum()


def batch_lors(d1s, d2s, tofs=None) -> LorSet:
    """One LOR set for a batch of images: the LORs of image i get frame i."""
    frames = [torch.full_like(d, i, dtype=torch.int32) for i, d in enumerate(d1s)]
    tof = torch.cat(tofs).contiguous() if tofs is not None else None
    return LorSet.pairs(torch.cat(d1s).contiguous(), torch.cat(d2s).contiguous(), tof=tof,
                        frame=torch.cat(frames).contiguous())
