"""yrtmi_em_subset_noalloc, the fused EM subset that allocates nothing (its
per-XCD partial images live in the library's static scratch array), against
the allocating yrtmi_em_subset on the same inputs, through the C ABI: two
calls in a row, the second on a smaller image and fewer events (the scratch
keeps the first call's stride and contents), and with direct accumulation
(no scratch).  Sizes beyond the scratch are refused with a clear error.

Bar: <= 1e-5 relative L2 (both accumulate with FP32 atomics in a
nondeterministic order)."""
import ctypes as C

import numpy as np
import pytest
import torch

import yrtmi
from helpers import rel_l2
from yrtmi import capi, synth

pytestmark = pytest.mark.gpu

TOL = 1e-5


def noalloc(ctx, proj, im, x, u, lors, meas=None):
    corr = capi.Corrections(None if meas is None else C.c_void_p(meas.data_ptr()), None, None,
                            None, None, 1.0, 1e-35)
    ci, cl = im.c(), lors.c()
    return capi.lib().yrtmi_em_subset_noalloc(ctx.h, proj.h, C.byref(ci),
                                              C.c_void_p(x.data_ptr()),
                                              C.c_void_p(u.data_ptr()), C.byref(cl),
                                              C.byref(corr))


def case(gpu, ctx, sc, proj, dims, n, seed):
    im = yrtmi.ImageParams(*dims)
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=300.0)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    L = yrtmi.LorSet.pairs(T(ev["d1"], np.int32), T(ev["d2"], np.int32),
                           tof=T(ev["tof"], np.float32))
    x = T(rng.uniform(0.2, 2.0, im.shape), np.float32)
    meas = T(rng.uniform(0.5, 2.0, n), np.float32)
    ref = torch.zeros_like(x)
    ctx.em_subset(proj, im, x, ref, L, meas=meas)
    u = torch.zeros_like(x)
    capi.check(noalloc(ctx, proj, im, x, u, L, meas))
    ctx.synchronize()
    return rel_l2(u.cpu().numpy(), ref.cpu().numpy())


def test_em_subset_noalloc_matches_em_subset(gpu):
    sc = yrtmi.Scanner("na", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    ctx = yrtmi.Context(sc)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(300.0, 5)
    proj = ctx.projector(pp)
    proj.set_accumulation(True)  # per-XCD partial images: the scratch
    r1 = case(gpu, ctx, sc, proj, (48, 44, 20, 5.0, 5.5, 4.0), 30000, 1)
    r2 = case(gpu, ctx, sc, proj, (24, 20, 10, 8.0, 9.0, 7.0), 7000, 2)
    print(f"noalloc vs em_subset: rel L2 {r1:.2e}, then (smaller) {r2:.2e}", flush=True)
    assert r1 <= TOL and r2 <= TOL
    proj.set_accumulation(False)  # direct atomics: no scratch used
    r3 = case(gpu, ctx, sc, proj, (32, 32, 16, 6.0, 6.0, 4.0), 9000, 3)
    print(f"direct accumulation: rel L2 {r3:.2e}", flush=True)
    assert r3 <= TOL


def test_em_subset_noalloc_refuses_images_beyond_the_scratch(gpu):
    sc = yrtmi.Scanner("na", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    ctx = yrtmi.Context(sc)
    proj = ctx.projector(yrtmi.ProjectorParams(sc))
    proj.set_accumulation(True)
    # 8 partial images of 2.6 M voxels exceed 16 Mi floats of scratch; the
    # call is refused on the host before anything is launched
    im = yrtmi.ImageParams(200, 200, 64, 2.0, 2.0, 2.0)
    x = torch.ones(im.shape, device=gpu)
    u = torch.zeros_like(x)
    d = torch.zeros(4, dtype=torch.int32, device=gpu)
    L = yrtmi.LorSet.pairs(d, d + 1)
    rc = noalloc(ctx, proj, im, x, u, L)
    ctx.synchronize()
    if rc == 0:
        pytest.skip("a device of one XCD: no partial images, no scratch")
    assert rc == -1  # YRTMI_EINVAL
    assert "scratch" in capi.lib().yrtmi_last_error().decode()
