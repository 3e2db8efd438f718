"""Command-line front ends with the reference's executable and flag names
(executables/recon/Reconstruct.cpp, ForwardProject.cpp, Backproject.cpp),
over the C++ host layer (pyyrtmi) on the GPU:

    python -m yrtmi.cli reconstruct      (yrtpet_reconstruct)
    python -m yrtmi.cli forward_project  (yrtpet_forward_project)
    python -m yrtmi.cli backproject      (yrtpet_backproject)

Files: scanner JSON (--scanner), image-parameters JSON (--params), NIfTI
images (.nii / .nii.gz), raw-data histograms (format H), list-mode files
(format LM with --flag_tof / --flag_randoms), motion CSV (--lor_motion) and
dynamic framing (--dyn).  Options of the reference outside this path
(variant image-space PSF, detector masks, sparse histograms, DOI list-mode,
in-vivo or hardware attenuation, ACF histograms, LR updaters) are refused with a
message rather than ignored; --gpu and --num_threads are accepted (the
projections always run on the GPU).
"""
import argparse
import sys

import numpy as np

UNSUPPORTED = ["varpsf", "detmask", "projector_updater", "sparse",
               "save_iter_step", "save_iter_ranges"]


def _Y():
    import pyyrtmi
    return pyyrtmi


def _common(p, need_params=False):
    p.add_argument("-s", "--scanner", required=True, help="Scanner parameters file")
    p.add_argument("--gpu", action="store_true", help="Use GPU acceleration (always on)")
    p.add_argument("--num_threads", type=int, default=-1, help="Number of threads (unused)")
    p.add_argument("--projector", default="S",
                   help="Projector to use, choices: Siddon (S), Distance-Driven (D)")
    p.add_argument("--num_rays", type=int, default=1,
                   help="Number of rays to use (for Siddon projector only)")
    p.add_argument("--proj_psf", default="",
                   help="Projection-space PSF kernel file (for DD projector only)")
    p.add_argument("--tof_width_ps", type=float, default=0.0, help="TOF width in picoseconds")
    p.add_argument("--tof_n_std", type=int, default=5,
                   help="Number of standard deviations for TOF's Gaussian. Default: 5")
    p.add_argument("--psf", default="", help="Image-space PSF kernel file")
    for name in UNSUPPORTED:
        p.add_argument("--" + name, default=None, help=argparse.SUPPRESS)


def _refuse_unsupported(a):
    for name in UNSUPPORTED:
        if getattr(a, name, None) not in (None, "", False):
            raise SystemExit(f"--{name}: not supported by this implementation")


def _input_args(p, required):
    p.add_argument("-i", "--input", required=required, default="", help="Input file")
    p.add_argument("-f", "--format", default="H",
                   help="Input file format. Possible values: H (Histogram3D), LM (list-mode)")
    p.add_argument("--flag_tof", action="store_true", help="List-mode: TOF column present")
    p.add_argument("--flag_randoms", action="store_true",
                   help="List-mode: randoms estimate column present")
    p.add_argument("-m", "--lor_motion", default="", help="Motion CSV file for motion correction")
    p.add_argument("--dyn", default="", help="Dynamic framing file")


def _projector_params(Y, sc, a):
    pp = Y.ProjectorParams(sc)
    pp.setProjector(a.projector)
    pp.numRays = a.num_rays
    if a.tof_width_ps > 0:
        pp.addTOF(a.tof_width_ps, a.tof_n_std)
    if a.proj_psf:
        pp.setProjPsf(a.proj_psf)
    return pp


def _load_input(Y, ctx, sc, a):
    """Projection data on the device: a Histogram3D (raw-data file) or a
    list-mode file with its motion and dynamic frames."""
    fmt = a.format.upper()
    if fmt == "H":
        h = Y.Histogram3DDeviceOwned(ctx)
        h.readFromFile(a.input)
        return h, None
    if fmt != "LM":
        raise SystemExit(f"--format {a.format}: only H and LM are supported")
    lm = Y.ListModeLUTOwned(sc, a.input, flag_tof=a.flag_tof, flag_randoms=a.flag_randoms)
    ts = lm.getTimestampArray()
    frame = motion = xf = None
    nframes = 1
    if a.dyn:
        bounds = Y.readDynamicFraming(a.dyn)
        frame = Y.dynamicFrames(ts, bounds)
        nframes = len(bounds) - 1
    if a.lor_motion:
        t, xf, err = Y.readLORMotion(a.lor_motion)
        motion = Y.motionFrames(ts, t, xf, err)
    dev = Y.ListModeDeviceOwned(ctx, lm, frame=frame, motion=motion)
    dev.numFrames = nframes
    if xf is not None:
        dev.setMotionTransforms(xf)
    return dev, lm


def _sens_name(out_sens, i, n):
    if n == 1:
        return out_sens
    for ext in (".nii.gz", ".nii"):
        if out_sens.endswith(ext):
            return out_sens[: -len(ext)] + f"_subset{i}" + ext
    return out_sens + f"_subset{i}"


def reconstruct(argv):
    p = argparse.ArgumentParser(prog="yrtpet_reconstruct",
                                description="OSEM reconstruction (MI355X)")
    _common(p)
    _input_args(p, required=False)
    p.add_argument("-p", "--params", default="", help="Image parameters file")
    p.add_argument("-o", "--out", default="", help="Output image filename")
    p.add_argument("--out_sens", default="", help="Sensitivity image output filename")
    p.add_argument("--sens", default="", help="Sensitivity image files (comma-separated)")
    p.add_argument("--sens_only", action="store_true", help="Only generate sensitivity image(s)")
    p.add_argument("--move_sens", action="store_true",
                   help="Move the provided sensitivity image based on motion")
    p.add_argument("--sensitivity", default="", help="Sensitivity histogram file")
    p.add_argument("--invert_sensitivity", action="store_true",
                   help="Invert the sensitivity histogram values")
    p.add_argument("--global_scale", type=float, default=1.0, help="Global scaling factor")
    p.add_argument("--num_iterations", type=int, default=10, help="Number of MLEM iterations")
    p.add_argument("--num_subsets", type=int, default=1, help="Number of OSEM subsets")
    p.add_argument("--initial_estimate", default="", help="Initial image estimate")
    p.add_argument("--randoms", default="", help="Randoms estimate histogram filename")
    p.add_argument("--scatter", default="", help="Scatter estimate histogram filename")
    p.add_argument("--att", default="", help="Total attenuation image filename")
    p.add_argument("--att_invivo", default="", help="In-vivo attenuation image filename")
    p.add_argument("--att_hardware", default="", help="Hardware attenuation image filename")
    p.add_argument("--acf", default="", help="Total ACF histogram filename")
    p.add_argument("--acf_invivo", default="", help="In-vivo ACF histogram filename")
    p.add_argument("--acf_hardware", default="", help="Hardware ACF histogram filename")
    p.add_argument("--hard_threshold", type=float, default=None, help="Hard threshold")
    p.add_argument("--denom_threshold", type=float, default=None,
                   help="Denominator threshold for the EM ratio")
    for fmt in ("sensitivity_format", "randoms_format", "scatter_format", "acf_format",
                "acf_invivo_format", "acf_hardware_format"):
        p.add_argument("--" + fmt, default="H", help=argparse.SUPPRESS)
    a = p.parse_args(argv)
    _refuse_unsupported(a)
    for fmt in ("sensitivity_format", "randoms_format", "scatter_format", "acf_format",
                "acf_invivo_format", "acf_hardware_format"):
        if getattr(a, fmt).upper() != "H":
            raise SystemExit(f"--{fmt}: only H is supported")
    if not a.sens_only and not a.out:
        raise SystemExit("--out is required unless --sens_only")
    Y = _Y()
    sc = Y.Scanner(a.scanner)
    osem = Y.createOSEM(sc)
    ctx = osem.getContext()
    osem.setProjector(a.projector)
    osem.setNumRays(a.num_rays)
    if a.tof_width_ps > 0:
        osem.addTOF(a.tof_width_ps, a.tof_n_std)
    if a.proj_psf:
        osem.addProjPSF(a.proj_psf)
    osem.num_MLEM_iterations = a.num_iterations
    osem.num_OSEM_subsets = a.num_subsets
    if a.hard_threshold is not None:
        osem.hardThreshold = a.hard_threshold
    if a.denom_threshold is not None:
        osem.denomThreshold = a.denom_threshold
    osem.setGlobalScalingFactor(a.global_scale)
    keep = []  # device objects the OSEM refers to
    sens_imgs = [Y.ImageDeviceFromFile(ctx, f) for f in a.sens.split(",") if f]
    if a.params:
        params = Y.ImageParams(a.params)
    elif sens_imgs:
        params = sens_imgs[0].getParams()
    else:
        raise SystemExit("--params is required when no sensitivity image is given")
    osem.setImageParams(params)
    if a.psf:
        osem.addImagePSF(a.psf)
    if a.input:
        data, lm = _load_input(Y, ctx, sc, a)
        keep += [data, lm]
        osem.setDataInput(data)
    elif not a.sens_only:
        raise SystemExit("--input is required unless --sens_only")
    if a.sensitivity:
        nh = Y.Histogram3DDeviceOwned(ctx)
        nh.readFromFile(a.sensitivity)
        keep.append(nh)
        osem.setSensitivityHistogram(nh)
        osem.setInvertSensitivity(a.invert_sensitivity)
    if a.att:
        att = Y.ImageDeviceFromFile(ctx, a.att)
        keep.append(att)
        osem.setAttenuationImage(att)
    for name, setter in (("att_invivo", osem.setInVivoAttenuationImage),
                         ("att_hardware", osem.setHardwareAttenuationImage)):
        if getattr(a, name):
            img = Y.ImageDeviceFromFile(ctx, getattr(a, name))
            keep.append(img)
            setter(img)
    for name, setter in (("randoms", osem.setRandomsHistogram),
                         ("scatter", osem.setScatterHistogram),
                         ("acf", osem.setACFHistogram),
                         ("acf_invivo", osem.setInVivoACFHistogram),
                         ("acf_hardware", osem.setHardwareACFHistogram)):
        if getattr(a, name):
            h = Y.Histogram3DDeviceOwned(ctx)
            h.readFromFile(getattr(a, name))
            keep.append(h)
            setter(h)
    if a.initial_estimate:
        x0 = Y.ImageDeviceFromFile(ctx, a.initial_estimate)
        keep.append(x0)
        osem.setInitialEstimate(x0)
    given = bool(sens_imgs)
    if given:
        move = a.move_sens
    else:
        sens_imgs = osem.generateSensitivityImages()
        move = True
        if a.out_sens:
            for i, img in enumerate(sens_imgs):
                img.writeToFile(_sens_name(a.out_sens, i, len(sens_imgs)))
    if a.lor_motion and move:
        # Reconstruct.cpp:556-660: the sensitivity image time-averaged over
        # the motion: per dynamic frame with --dyn, else over the list-mode's
        # first to last timestamp, else over the motion's full time range
        if len(sens_imgs) != 1:
            raise SystemExit("moving the sensitivity needs exactly one sensitivity image")
        t, xf, _ = Y.readLORMotion(a.lor_motion)
        if a.dyn:
            moved = Y.timeAverageMoveImageDynamic(t, xf, sens_imgs[0],
                                                  Y.readDynamicFraming(a.dyn))
        elif a.input and a.format.upper() == "LM":
            ts = keep[1].getTimestampArray()
            moved = Y.timeAverageMoveImage(t, xf, sens_imgs[0], int(ts[0]), int(ts[-1]))
        else:
            moved = Y.timeAverageMoveImage(t, xf, sens_imgs[0])
        sens_imgs = [moved]
        keep.append(moved)
        if a.out_sens:
            moved.writeToFile(a.out_sens)
        given = True
    if given:
        # setSensitivityImages copies; the generated images are the OSEM's own
        osem.setSensitivityImages(sens_imgs)
    if a.sens_only:
        return 0
    x = osem.reconstruct()
    x.writeToFile(a.out)
    return 0


def forward_project(argv):
    p = argparse.ArgumentParser(prog="yrtpet_forward_project",
                                description="Forward projection of an image into a histogram")
    _common(p)
    p.add_argument("-i", "--input", required=True, help="Input image file")
    p.add_argument("-o", "--out", required=True, help="Output histogram filename")
    p.add_argument("--num_subsets", type=int, default=1, help="Number of OSEM subsets")
    p.add_argument("--subset_id", type=int, default=0, help="Subset to project")
    p.add_argument("--to_acf", action="store_true", help="Generate ACF histogram")
    a = p.parse_args(argv)
    _refuse_unsupported(a)
    Y = _Y()
    sc = Y.Scanner(a.scanner)
    ctx = Y.Context(sc)
    pp = _projector_params(Y, sc, a)
    img = Y.ImageDeviceFromFile(ctx, a.input)
    if a.psf:
        # the image-space PSF applied to the image first (ForwardProject.cpp)
        blurred = Y.ImageDeviceOwned(ctx, img.getParams())
        Y.applyImagePSF(ctx, img, blurred, *Y.readImagePSF(a.psf))
        img = blurred
    h = Y.Histogram3DDeviceOwned(ctx)
    Y.forwProject(ctx, pp, img, h)
    ctx.synchronize()
    v = h.transferToHostMemory().reshape(ctx.histogramShape())
    if not (0 <= a.subset_id < a.num_subsets):
        raise SystemExit("--subset_id must be in [0, num_subsets)")
    if a.num_subsets > 1:
        # Histogram3D subsets take the phi planes phi = num_subsets * k + subset_id
        keep = (np.arange(v.shape[1]) % a.num_subsets) == a.subset_id
        v[:, ~keep, :] = 0.0
    if a.to_acf:
        # util::convertProjectionValuesToACF: acf = exp(-0.1 * integral of mu)
        v = np.exp(-0.1 * v, dtype=np.float32)
    Y.writeRawd(a.out, np.ascontiguousarray(v, np.float32))
    return 0


def backproject(argv):
    p = argparse.ArgumentParser(prog="yrtpet_backproject",
                                description="Backprojection of projection data into an image")
    _common(p)
    _input_args(p, required=True)
    p.add_argument("-o", "--out", required=True, help="Output image filename")
    p.add_argument("-p", "--params", required=True, help="Image parameters file")
    a = p.parse_args(argv)
    _refuse_unsupported(a)
    Y = _Y()
    sc = Y.Scanner(a.scanner)
    ctx = Y.Context(sc)
    pp = _projector_params(Y, sc, a)
    params = Y.ImageParams(a.params)
    data, lm = _load_input(Y, ctx, sc, a)
    if lm is not None:
        # list-mode events carry a value of one each
        data.setProjectionValues(np.ones(lm.count(), np.float32))
        if data.numFrames > 1:
            params = Y.ImageParams(params.nx, params.ny, params.nz, params.length_x,
                                   params.length_y, params.length_z, params.off_x,
                                   params.off_y, params.off_z, data.numFrames)
    img = Y.ImageDeviceOwned(ctx, params)
    img.setValue(0.0)
    Y.backProject(ctx, pp, img, data)
    ctx.synchronize()
    if a.psf:
        # the adjoint of the image-space PSF after the backprojection
        # (Backproject.cpp)
        blurred = Y.ImageDeviceOwned(ctx, img.getParams())
        Y.applyImagePSF(ctx, img, blurred, *Y.readImagePSF(a.psf), transpose=True)
        img = blurred
    img.writeToFile(a.out)
    return 0


COMMANDS = {"reconstruct": reconstruct, "forward_project": forward_project,
            "backproject": backproject}


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in COMMANDS:
        print("usage: python -m yrtmi.cli {" + ",".join(COMMANDS) + "} [options]",
              file=sys.stderr)
        return 2
    return COMMANDS[argv[0]](argv[1:])


if __name__ == "__main__":
    sys.exit(main())
