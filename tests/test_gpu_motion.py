"""Time-averaged moved images on the GPU (yrtmi_time_average_move through
the C ABI; timeAverageMoveImage / timeAverageMoveImageDynamic /
getFullTimeRange through pyyrtmi; --move_sens / -m in the reconstruct
command) against the oracle's restatement of timeAverageMoveImage_kernel
and its drivers (ReconstructionUtilsDevice.cu:78-300).

Bar: <= 1e-5 relative L2 against the oracle, and the reference's own
CPU-vs-GPU check (allclose, rtol = atol = 1e-3, test_TimeAverageMove.cpp:
96).  The oracle emulates fmaf through float64 (one ulp apart, rarely) and
a voxel centre on a cell boundary may floor either way; both move values
continuously, by ulps."""
import numpy as np
import pytest
import torch

import oracle as O
import pyyrtmi as Y
import yrtmi
from helpers import oracle_image, rel_l2
from test_oracle_motion import rigid

pytestmark = pytest.mark.gpu


def ctx_py():
    return yrtmi.Context(yrtmi.Scanner("m", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24))


def check(got, want, what):
    r = rel_l2(got, want)
    print(f"{what}: rel L2 {r:.2e}", flush=True)
    assert r <= 1e-5
    np.testing.assert_allclose(got, want, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("shape,nframes,seed", [((17, 23, 29), 12, 1), ((40, 31, 31), 40, 2),
                                                ((5, 64, 70), 3, 3)])
def test_time_average_move_matches_oracle(gpu, shape, nframes, seed):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    im = yrtmi.ImageParams(nx, ny, nz, *rng.uniform(0.5, 3.0, 3))
    oi = oracle_image(im)
    x = np.zeros(shape, np.float32)
    # a random prism of positive values (test_utils' makeImageWithRandomPrism)
    lo = [rng.integers(0, s // 3 + 1) for s in shape]
    hi = [rng.integers(2 * s // 3, s + 1) for s in shape]
    x[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = rng.uniform(0.5, 2.0, [h - l for l, h in zip(lo, hi)])
    xf = np.stack([rigid(rng, 0.4, 3.0) for _ in range(nframes)])
    w = rng.uniform(0.0, 1.0, nframes).astype(np.float32)
    base = rng.uniform(0.0, 0.1, shape).astype(np.float32)  # accumulated into
    ctx = ctx_py()
    xg = torch.from_numpy(x).to(gpu)
    og = torch.from_numpy(base).to(gpu)
    ctx.time_average_move(im, xg, og, xf, w)
    ctx.synchronize()
    want = O.move_accumulate(x, oi, O.invert_transform(xf), w, base)
    check(og.cpu().numpy(), want, f"move {shape} x{nframes}")
    # a second call on the same context (the frame buffer reused, fewer frames)
    og2 = torch.zeros_like(xg)
    ctx.time_average_move(im, xg, og2, xf[:2], w[:2])
    ctx.synchronize()
    check(og2.cpu().numpy(), O.move_accumulate(x, oi, O.invert_transform(xf[:2]), w[:2],
                                              np.zeros_like(x)), "move x2")


def test_time_average_move_rejects_bad_arguments(gpu):
    im = yrtmi.ImageParams(8, 8, 8, 2.0, 2.0, 2.0)
    ctx = ctx_py()
    x = torch.zeros(im.shape, device=gpu)
    o = torch.zeros_like(x)
    xf = rigid(np.random.default_rng(0))[None]
    with pytest.raises(yrtmi.capi.YrtmiError, match="distinct"):
        ctx.time_average_move(im, x, x, xf, [1.0])
    with pytest.raises(yrtmi.capi.YrtmiError, match="transforms"):
        ctx.time_average_move(im, x, o, np.zeros((0, 12), np.float32), [])


def test_host_time_average_full_range_and_dynamic(gpu):
    rng = np.random.default_rng(9)
    ysc = Y.Scanner("m", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    ctx = Y.Context(ysc)
    p = Y.ImageParams.fromParams(30, 26, 18, 2.5, 3.0, 2.0)
    oi = O.ImageDef(30, 26, 18, 2.5, 3.0, 2.0)
    x = rng.uniform(0.5, 1.5, (1, 18, 26, 30)).astype(np.float32)
    img = Y.ImageDeviceOwned(ctx, p)
    img.copyFromHost(x)
    k = 25
    ts = np.cumsum(rng.integers(5, 100, k)).astype(np.int64) - 5
    xf = np.stack([rigid(rng, 0.3, 2.5) for _ in range(k)])
    assert Y.getFullTimeRange(ts) == O.full_time_range(ts)
    assert Y.getFullTimeRange(ts[:1]) == (int(ts[0]), int(ts[0]) + 1)
    full = Y.timeAverageMoveImage(ts, xf, img).transferToHostMemory()
    check(full.reshape(x.shape[1:]), O.time_average_move(x, oi, ts, xf), "full range")
    t0, t1 = int(ts[3]) + 1, int(ts[17])
    part = Y.timeAverageMoveImage(ts, xf, img, t0, t1).transferToHostMemory()
    check(part.reshape(x.shape[1:]), O.time_average_move(x, oi, ts, xf, t0, t1), "range")
    bounds = np.array([ts[0], ts[6], ts[12] + 3, ts[-1] + 40], np.uint32)
    dyn = Y.timeAverageMoveImageDynamic(ts, xf, img, bounds)
    assert dyn.getParams().nt == 3
    check(dyn.transferToHostMemory().reshape(3, 18, 26, 30),
          O.time_average_move_dynamic(x, oi, ts, xf, bounds), "dynamic")
    with pytest.raises(Exception, match="greater"):
        Y.timeAverageMoveImage(ts, xf, img, 50, 50)
    # a motion of one frame: its frame covers the whole range (weight 1),
    # over the full range and over a given one
    one = Y.timeAverageMoveImage(ts[:1], xf[:1], img).transferToHostMemory()
    want = O.move_accumulate(x[0], oi, O.invert_transform(xf[:1]), [1.0],
                             np.zeros(x.shape[1:], np.float32))
    check(one.reshape(x.shape[1:]), want, "one motion frame")
    one = Y.timeAverageMoveImage(ts[:1], xf[:1], img, int(ts[0]), int(ts[0]) + 1000)
    check(one.transferToHostMemory().reshape(x.shape[1:]), want, "one motion frame, range")
    check(one.transferToHostMemory().reshape(x.shape[1:]),
          O.time_average_move(x, oi, ts[:1], xf[:1], int(ts[0]), int(ts[0]) + 1000),
          "one motion frame vs oracle driver")


def test_cli_reconstruct_moves_the_sensitivity(gpu, tmp_path):
    """-m with a generated sensitivity image: the image is time-averaged
    over the list-mode's first to last timestamp before the OSEM runs
    (Reconstruct.cpp:556-660); --move_sens does the same for a given one."""
    from test_gpu_cli import files, listmode_file
    from yrtmi import cli
    sj, pj = files(tmp_path)
    sc_py = yrtmi.Scanner("cli", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
    lmf, ev, ts = listmode_file(tmp_path, sc_py, 20000, 11, None)
    mot = str(tmp_path / "motion.csv")
    mt = np.array([0, 15_000, 35_000], np.int64)
    xf = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0],
                   [1, 0, 0, 3.0, 0, 1, 0, -2.0, 0, 0, 1, 1.0],
                   [0.995, -0.0998, 0, 1.0, 0.0998, 0.995, 0, 0.5, 0, 0, 1, -1.0]], np.float32)
    Y.writeLORMotion(mot, mt, xf, np.zeros(3, np.float32))
    out, sens = str(tmp_path / "x.nii"), str(tmp_path / "sens.nii")
    assert cli.main(["reconstruct", "-s", sj, "-p", pj, "-i", lmf, "-f", "LM", "-m", mot,
                     "--num_iterations", "1", "--num_subsets", "2", "-o", out,
                     "--out_sens", sens]) == 0
    _, s_cli = Y.readImage(sens)
    _, x_cli = Y.readImage(out)
    # direct: the unmoved sensitivity, moved over [ts[0], ts[-1]]
    sc = Y.Scanner(sj)
    osem = Y.createOSEM(sc)
    ctx = osem.getContext()
    osem.setImageParams(Y.ImageParams(pj))
    osem.num_MLEM_iterations = 1
    osem.num_OSEM_subsets = 2
    t_rd, xf_rd, err = Y.readLORMotion(mot)
    motion = Y.motionFrames(ts, t_rd, xf_rd, err)
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               motion=motion)
    lm.setMotionTransforms(xf_rd)
    osem.setDataInput(lm)
    s0 = osem.generateSensitivityImages()[0]
    s0_host = s0.transferToHostMemory()
    moved = Y.timeAverageMoveImage(t_rd, xf_rd, s0, int(ts[0]), int(ts[-1]))
    s_dir = moved.transferToHostMemory()
    assert rel_l2(s_cli, s_dir) <= 1e-6
    assert rel_l2(s_dir, s0_host) > 1e-3  # the motion moved it
    p = Y.ImageParams(pj)
    check(s_dir.reshape(20, 40, 40),
          O.time_average_move(s0_host, O.ImageDef(40, 40, 20, 4.0, 4.0, 2.5), t_rd, xf_rd,
                              int(ts[0]), int(ts[-1])), "cli sensitivity")
    osem.setSensitivityImages([moved])
    x_dir = osem.reconstruct().transferToHostMemory()
    assert np.count_nonzero(x_dir) > 1000 and rel_l2(x_cli, x_dir) <= 1e-6
    assert p.nx == 40
