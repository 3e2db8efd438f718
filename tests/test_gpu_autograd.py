"""The differentiable projector (yrtmi.autograd) on the GPU.

- adjoint (dot-product) test <A x, y> = <x, A^T y> in FP32, direct and
  tile-planned backprojection: <= 1e-6 relative (inner products summed in
  FP64);
- autograd's backward is the reference backprojector: x.grad of
  sum(w * A x) against the oracle's backprojection of w, <= 1e-5 relative L2;
- the FP64 projector against the FP64 oracle (<= 1e-10 relative L2) and
  torch.autograd.gradcheck in FP64 on a 16 x 16 x 8 image with 5000 events,
  of A alone and of the Poisson log-likelihood through A (the FP64
  backprojection sums with atomics, so two backward calls agree to rounding
  only: nondet_tol 1e-12, against values of order 1e-2 to 1e1);
- a batch of images as frames of one launch equals projecting them one by
  one.
"""
import numpy as np
import pytest
import torch

import oracle as O
import yrtmi
from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
from yrtmi import autograd as AG
from yrtmi import synth

pytestmark = pytest.mark.gpu


def setup(gpu, shape, vox, n, tof=350.0, rays=4, nt=1, seed=1):
    sc = yrtmi.Scanner("ag", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    im = yrtmi.ImageParams(*shape, *vox, nt=nt)
    pp = yrtmi.ProjectorParams(sc, numRays=rays)
    pp.addTOF(tof, 5)
    ctx = yrtmi.Context(sc)
    proj = ctx.projector(pp)
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=tof)
    if nt > 1:
        ev["frame"] = rng.integers(0, nt, n).astype(np.int32)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    L = yrtmi.LorSet.pairs(T(ev["d1"], np.int32), T(ev["d2"], np.int32),
                           tof=T(ev["tof"], np.float32),
                           frame=T(ev["frame"], np.int32) if nt > 1 else None)
    return sc, im, pp, ctx, proj, ev, L, rng


def oracle_lors(sc, ev):
    pos, ori = O.lut(oracle_scanner(sc))
    return O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev["tof"], frame=ev.get("frame"))


@pytest.mark.parametrize("planned", [False, True])
def test_adjoint_dot_fp32(gpu, planned):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, (40, 40, 20), (4.0, 4.0, 3.0), 20000, nt=2)
    plan = ctx.plan(proj, im, L) if planned else None
    op = AG.ProjectionOperator(ctx, proj, im, L, plan=plan)
    x = torch.from_numpy(rng.uniform(0, 1, im.shape).astype(np.float32)).to(gpu)
    y = torch.from_numpy(rng.uniform(0, 1, L.n).astype(np.float32)).to(gpu)
    lhs = float(torch.dot(op.forward(x).double(), y.double()))
    rhs = float(torch.dot(x.reshape(-1).double(), op.adjoint(y).reshape(-1).double()))
    r = abs(lhs - rhs) / abs(lhs)
    print(f"dot test planned={planned}: <Ax,y>={lhs:.9e} <x,ATy>={rhs:.9e} rel {r:.2e}")
    assert r <= 1e-6


def test_backward_is_reference_backprojector(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, (32, 32, 16), (4.5, 4.5, 3.5), 8000)
    op = AG.ProjectionOperator(ctx, proj, im, L, chunk=3000)  # chunked walks
    x = torch.from_numpy(rng.uniform(0, 1, im.shape).astype(np.float32)).to(gpu)
    x.requires_grad_(True)
    w = rng.uniform(0.1, 1.0, L.n).astype(np.float32)
    (AG.project(x, op) * torch.from_numpy(w).to(gpu)).sum().backward()
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), w)
    r = rel_l2(x.grad.cpu().numpy(), ref)
    print(f"backward vs oracle backprojection: rel L2 {r:.2e}")
    assert r <= 1e-5
    # and the forward of the chunked operator is the oracle's forward
    fo = O.forward(oracle_image(im), x.detach().cpu().numpy(), oracle_proj(pp), oracle_lors(sc, ev))
    assert rel_l2(AG.project(x.detach(), op).cpu().numpy(), fo) <= 1e-5


def test_fp64_projector_matches_fp64_oracle(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, (24, 24, 12), (5.0, 5.0, 4.0), 4000, nt=2)
    op = AG.ProjectionOperator(ctx, proj, im, L)
    x = rng.uniform(0, 1, im.shape)
    v = rng.uniform(0.1, 1.0, L.n)
    OL = oracle_lors(sc, ev)
    po = oracle_proj(pp)
    f = op.forward(torch.from_numpy(x).to(gpu)).cpu().numpy()
    b = op.adjoint(torch.from_numpy(v).to(gpu)).cpu().numpy()
    f64 = O.forward(oracle_image(im), x, po, OL, dtype=np.float64)
    b64 = O.back(oracle_image(im), po, OL, v, dtype=np.float64)
    print(f"FP64 projector vs FP64 oracle: fwd {rel_l2(f, f64):.2e} back {rel_l2(b, b64):.2e}")
    assert rel_l2(f, f64) <= 1e-10
    assert rel_l2(b, b64) <= 1e-10


NONDET = 1e-12


def test_gradcheck_fp64(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, (16, 16, 8), (8.0, 8.0, 6.0), 5000)
    op = AG.ProjectionOperator(ctx, proj, im, L)
    x = torch.from_numpy(rng.uniform(0.5, 1.5, im.shape)).to(gpu).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: AG.project(t, op), (x,), eps=1e-6, atol=1e-8,
                                    rtol=1e-6, fast_mode=True, nondet_tol=NONDET)
    with torch.no_grad():
        ybar = op.forward(x)
        b = 0.1 * ybar.mean()
        m = torch.poisson(ybar + b)
    f = lambda t: AG.poisson_nll(AG.project(t, op), m, b)
    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-6, rtol=1e-5, nondet_tol=NONDET)
    v = torch.from_numpy(rng.uniform(0.1, 1.0, L.n)).to(gpu).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: AG.backproject(t, op), (v,), eps=1e-6,
                                    atol=1e-8, rtol=1e-6, fast_mode=True, nondet_tol=NONDET)


def test_batch_is_frames_of_one_launch(gpu):
    sc = yrtmi.Scanner("ag", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(350.0, 5)
    ctx = yrtmi.Context(sc)
    proj = ctx.projector(pp)
    B = 3
    im1 = yrtmi.ImageParams(24, 24, 12, 5.0, 5.0, 4.0)
    imB = yrtmi.ImageParams(24, 24, 12, 5.0, 5.0, 4.0, nt=B)
    rng = np.random.default_rng(9)
    d1s, d2s, tofs, ops = [], [], [], []
    for i in range(B):
        ev = synth.random_events(sc, 3000 + 1000 * i, rng, tof_fwhm_ps=350.0)
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
        d1s.append(T(ev["d1"], np.int32)); d2s.append(T(ev["d2"], np.int32))
        tofs.append(T(ev["tof"], np.float32))
        Li = yrtmi.LorSet.pairs(d1s[-1], d2s[-1], tof=tofs[-1])
        ops.append(AG.ProjectionOperator(ctx, proj, im1, Li))
    LB = AG.batch_lors(d1s, d2s, tofs)
    opB = AG.ProjectionOperator(ctx, proj, imB, LB)
    x = torch.from_numpy(rng.uniform(0, 1, imB.shape).astype(np.float32)).to(gpu)
    yB = opB.forward(x)
    y1 = torch.cat([ops[i].forward(x[i:i + 1].contiguous()) for i in range(B)])
    assert rel_l2(yB.cpu().numpy(), y1.cpu().numpy()) <= 1e-6
    gB = opB.adjoint(yB)
    g1 = torch.cat([ops[i].adjoint(y1[sum(d.numel() for d in d1s[:i]):
                                      sum(d.numel() for d in d1s[:i + 1])]) for i in range(B)])
    assert rel_l2(gB.cpu().numpy(), g1.cpu().numpy()) <= 1e-6
