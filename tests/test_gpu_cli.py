"""The command-line front ends (yrtmi.cli) end to end on files: scanner and
image-parameter JSON, a list-mode file with TOF, NIfTI images and raw-data
histograms, against the same operations called directly through pyyrtmi
(whose numerics the other GPU tests pin against the oracle).  Bar: the
file-driven run equals the direct run to 1e-6 relative L2 (the same kernels
on the same inputs; only the order of FP32 atomic additions may differ)."""
import json

import numpy as np
import pytest

import pyyrtmi as Y
import yrtmi
from helpers import rel_l2
from yrtmi import cli, synth

pytestmark = pytest.mark.gpu

TOL = 1e-6


def files(tmp_path):
    sj = str(tmp_path / "scanner.json")
    with open(sj, "w") as f:
        json.dump({"VERSION": 3.2, "scannerName": "cli", "axialFOV": 48.0,
                   "crystalSize_trans": 4.0, "crystalSize_z": 4.0, "crystalDepth": 10.0,
                   "scannerRadius": 110.0, "dets_per_ring": 64, "num_rings": 12,
                   "num_doi": 1, "max_ring_diff": 11, "min_ang_diff": 1}, f)
    pj = str(tmp_path / "params.json")
    with open(pj, "w") as f:
        json.dump({"VERSION": 1.2, "nx": 40, "ny": 40, "nz": 20, "vx": 4.0, "vy": 4.0,
                   "vz": 2.5}, f)
    return sj, pj


def listmode_file(tmp_path, sc_py, n, seed, tof):
    im = yrtmi.ImageParams(40, 40, 20, 4.0, 4.0, 2.5)
    act = synth.cylinder_spheres(im, 60.0, 40.0, 1.0, [(10.0, 5.0, 0.0, 12.0, 4.0)],
                                 supersample=1)
    ev = synth.sample_events(sc_py, im, act, n, seed=seed, tof_fwhm_ps=tof)
    ts = np.sort(np.random.default_rng(seed).integers(0, 60_000, n)).astype(np.uint32)
    cols = [ts, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32)]
    if tof:
        cols.append(ev["tof"].astype(np.float32).view(np.uint32))
    f = str(tmp_path / "events.lmDat")
    np.stack(cols, axis=1).tofile(f)
    return f, ev, ts


def test_cli_reconstruct_listmode_tof(gpu, tmp_path):
    sj, pj = files(tmp_path)
    sc = Y.Scanner(sj)
    sc_py = yrtmi.Scanner("cli", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
    lmf, ev, ts = listmode_file(tmp_path, sc_py, 40000, 7, 300.0)
    out = str(tmp_path / "x.nii")
    sens = str(tmp_path / "sens.nii")
    rc = cli.main(["reconstruct", "-s", sj, "-p", pj, "-i", lmf, "-f", "LM", "--flag_tof",
                   "--num_rays", "2", "--tof_width_ps", "300", "--tof_n_std", "5",
                   "--num_iterations", "2", "--num_subsets", "3", "-o", out,
                   "--out_sens", sens])
    assert rc == 0
    p_out, x_cli = Y.readImage(out)
    _, s_cli = Y.readImage(sens)
    # the same reconstruction through the host layer directly
    osem = Y.createOSEM(sc)
    ctx = osem.getContext()
    osem.setImageParams(Y.ImageParams(pj))
    osem.setNumRays(2)
    osem.addTOF(300.0, 5)
    osem.num_MLEM_iterations = 2
    osem.num_OSEM_subsets = 3
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               tof=ev["tof"].astype(np.float32))
    osem.setDataInput(lm)
    s_dir = osem.generateSensitivityImages()[0].transferToHostMemory()
    x_dir = osem.reconstruct().transferToHostMemory()
    assert (p_out.nx, p_out.ny, p_out.nz) == (40, 40, 20)
    assert np.count_nonzero(x_dir) > 1000
    r1, r2 = rel_l2(s_cli, s_dir), rel_l2(x_cli, x_dir)
    print(f"cli vs direct: sensitivity {r1:.2e}, image {r2:.2e}", flush=True)
    assert r1 <= TOL and r2 <= TOL
    # the sensitivity image fed back (--sens) gives the same reconstruction
    out2 = str(tmp_path / "x2.nii.gz")
    assert cli.main(["reconstruct", "-s", sj, "--sens", sens, "-i", lmf, "-f", "LM",
                     "--flag_tof", "--num_rays", "2", "--tof_width_ps", "300",
                     "--num_iterations", "2", "--num_subsets", "3", "-o", out2]) == 0
    _, x2 = Y.readImage(out2)
    assert rel_l2(x2, x_dir) <= TOL


def test_cli_forward_and_backproject(gpu, tmp_path):
    sj, pj = files(tmp_path)
    sc = Y.Scanner(sj)
    ctx = Y.Context(sc)
    params = Y.ImageParams(pj)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, 1.0, (1, 20, 40, 40)).astype(np.float32)
    xin = str(tmp_path / "mu.nii")
    Y.writeImage(xin, params, x)
    # forward projection into a histogram, then the ACF form and one subset
    hf = str(tmp_path / "proj.rawd")
    assert cli.main(["forward_project", "-s", sj, "-i", xin, "-o", hf, "--projector", "D"]) == 0
    img = Y.ImageDeviceOwned(ctx, params)
    img.copyFromHost(x)
    pp = Y.ProjectorParams(sc)
    pp.setProjector("D")
    h = Y.Histogram3DDeviceOwned(ctx)
    Y.forwProject(ctx, pp, img, h)
    ctx.synchronize()
    v = h.transferToHostMemory().reshape(ctx.histogramShape())
    got = Y.readRawd(hf)
    assert got.shape == v.shape and rel_l2(got, v) <= TOL
    af = str(tmp_path / "acf.rawd")
    assert cli.main(["forward_project", "-s", sj, "-i", xin, "-o", af, "--projector", "D",
                     "--to_acf", "--num_subsets", "4", "--subset_id", "1"]) == 0
    acf = Y.readRawd(af)
    keep = (np.arange(v.shape[1]) % 4) == 1
    np.testing.assert_allclose(acf[:, keep, :], np.exp(-0.1 * v[:, keep, :]), rtol=2e-6)
    assert np.all(acf[:, ~keep, :] == 1.0)  # exp(-0) outside the subset
    # backprojection of the histogram file
    bf = str(tmp_path / "back.nii")
    assert cli.main(["backproject", "-s", sj, "-i", hf, "-f", "H", "-p", pj, "-o", bf,
                     "--projector", "D"]) == 0
    _, b_cli = Y.readImage(bf)
    b = Y.ImageDeviceOwned(ctx, params)
    b.setValue(0.0)
    Y.backProject(ctx, pp, b, h)
    ctx.synchronize()
    assert rel_l2(b_cli, b.transferToHostMemory()) <= TOL
    # list-mode backprojection (one per event), Siddon with TOF
    sc_py = yrtmi.Scanner("cli", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
    lmf, ev, ts = listmode_file(tmp_path, sc_py, 20000, 9, 300.0)
    lbf = str(tmp_path / "lmback.nii")
    assert cli.main(["backproject", "-s", sj, "-i", lmf, "-f", "LM", "--flag_tof", "-p", pj,
                     "-o", lbf, "--tof_width_ps", "300"]) == 0
    _, lb_cli = Y.readImage(lbf)
    pt = Y.ProjectorParams(sc)
    pt.addTOF(300.0, 5)
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               tof=ev["tof"].astype(np.float32))
    lm.setProjectionValues(np.ones(len(ev["d1"]), np.float32))
    lb = Y.ImageDeviceOwned(ctx, params)
    lb.setValue(0.0)
    Y.backProject(ctx, pt, lb, lm)
    ctx.synchronize()
    lbd = lb.transferToHostMemory()
    assert np.count_nonzero(lbd) > 1000 and rel_l2(lb_cli, lbd) <= TOL


def test_cli_dynamic_framing_and_motion(gpu, tmp_path):
    """--dyn splits the events into frames by timestamp (a 4D image);
    --lor_motion moves each event's LOR by its motion frame."""
    sj, pj = files(tmp_path)
    sc_py = yrtmi.Scanner("cli", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
    lmf, ev, ts = listmode_file(tmp_path, sc_py, 20000, 10, None)
    dyn = str(tmp_path / "frames.txt")
    Y.writeDynamicFraming(dyn, np.array([0, 20_000, 40_000, 60_000], np.uint32))
    mot = str(tmp_path / "motion.csv")
    xf = np.array([[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0],
                   [1, 0, 0, 3.0, 0, 1, 0, -2.0, 0, 0, 1, 1.0]], np.float32)
    Y.writeLORMotion(mot, np.array([0, 30_000], np.int64), xf, np.zeros(2, np.float32))
    out = str(tmp_path / "dyn.nii")
    assert cli.main(["backproject", "-s", sj, "-i", lmf, "-f", "LM", "-p", pj, "-o", out,
                     "--dyn", dyn, "-m", mot]) == 0
    p4, b4 = Y.readImage(out)
    assert p4.nt == 3 and b4.shape == (3, 20, 40, 40)
    # direct: frames and motion indices computed here
    sc = Y.Scanner(sj)
    ctx = Y.Context(sc)
    frame = np.searchsorted(np.array([20_000, 40_000, 60_000]), ts, side="right").astype(np.int32)
    motion = (ts >= 30_000).astype(np.uint8)
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               frame=frame, motion=motion)
    lm.numFrames = 3
    lm.setMotionTransforms(xf)
    lm.setProjectionValues(np.ones(len(ts), np.float32))
    params = Y.ImageParams(pj)
    p3 = Y.ImageParams(params.nx, params.ny, params.nz, params.length_x, params.length_y,
                       params.length_z, 0.0, 0.0, 0.0, 3)
    img = Y.ImageDeviceOwned(ctx, p3)
    img.setValue(0.0)
    Y.backProject(ctx, Y.ProjectorParams(sc), img, lm)
    ctx.synchronize()
    ref = img.transferToHostMemory()
    assert all(np.count_nonzero(ref[f]) > 100 for f in range(3))
    assert rel_l2(b4, ref) <= TOL
