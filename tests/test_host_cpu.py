"""The C++ host layer's pybind11 module (pyyrtmi) without a GPU: the
reference's Python names for this path are there, the geometry classes
derive what the reference derives, and errors surface as RuntimeError as
the reference's ASSERTs do through pybind11."""
import os

import numpy as np
import pytest

import pyyrtmi as Y
import yrtmi

# names of pyyrtpet's OSEM binding (src/recon/OSEM.cpp:37-162) that this
# path implements
OSEM_METHODS = ["setImageParams", "getImageParams", "setDataInput", "isListModeEnabled",
                "setProjector", "setNumRays", "addTOF", "addProjPSF",
                "setGlobalScalingFactor", "setAttenuationImage", "setRandomsHistogram",
                "setScatterHistogram", "setSensitivityHistogram", "setInvertSensitivity",
                "setInitialEstimate", "setMaskImage", "getExpectedSensImagesAmount",
                "setSensitivityImages", "generateSensitivityImages", "reconstruct",
                "getProjectorParams"]
OSEM_MEMBERS = ["num_MLEM_iterations", "num_OSEM_subsets", "hardThreshold", "denomThreshold"]


def test_reference_names():
    for n in OSEM_METHODS + OSEM_MEMBERS:
        assert hasattr(Y.OSEM, n), n
    for n in ("setProjector", "addTOF", "removeTOF", "hasTOF", "getTOFWidth_ps",
              "getTOFNumStd", "setProjPsf", "numRays"):
        assert hasattr(Y.ProjectorParams, n), n
    for n in ("applyA", "applyAH"):
        assert hasattr(Y.OperatorProjectorMI, n), n
    for n in ("createOSEM", "forwProject", "backProject", "ImageDeviceOwned",
              "ImageDeviceAlias", "ListModeDeviceOwned", "ListModeDeviceAlias",
              "Histogram3DDeviceOwned", "Histogram3DDeviceAlias"):
        assert hasattr(Y, n), n
    assert Y.ProjectorParams.ProjectorType.SIDDON != Y.ProjectorParams.ProjectorType.DD


def test_image_params_match_python_layer():
    # ImageParams(nx, ny, nz, lengths...): v = length / n; FOV radius
    # max(lx, ly)/2 - max(vx, vy)/1000 (ImageParams.cpp:206-266)
    for args in [(20, 20, 12, 100.0, 100.0, 48.0), (61, 57, 35, 311.1, 302.1, 94.5)]:
        a = Y.ImageParams(*args)
        b = yrtmi.ImageParams.from_lengths(*args)
        assert (a.vx, a.vy, a.vz) == (np.float32(b.vx), np.float32(b.vy), np.float32(b.vz))
        assert np.float32(a.fovRadius) == np.float32(b.fovRadius)
    c = Y.ImageParams.fromParams(200, 200, 128, 2.0, 2.0, 2.0)
    assert (c.length_x, c.length_z, c.vx) == (400.0, 256.0, 2.0)
    with pytest.raises(RuntimeError):
        Y.ImageParams(0, 10, 10, 1.0, 1.0, 1.0)


def test_projector_params_and_psf_file(tmp_path):
    s = Y.Scanner("s", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16, 1, 15, 1, 1)
    pp = Y.ProjectorParams(s)
    assert not pp.hasTOF() and pp.numRays == 1
    pp.addTOF(250.0, 5)
    assert pp.hasTOF() and pp.getTOFWidth_ps() == 250.0 and pp.getTOFNumStd() == 5
    pp.removeTOF()
    assert not pp.hasTOF()
    pp.setProjector("DD")
    assert pp.projectorType == Y.ProjectorParams.ProjectorType.DD
    with pytest.raises(RuntimeError, match="Unknown projector"):
        pp.setProjector("bogus")
    # one-row CSV: sStep, kSpacing, K, K*numS taps (ProjectionPsfManager.cpp:27-49)
    f = tmp_path / "psf.csv"
    f.write_text("50.0,0.25,3,0.2,0.6,0.2,0.1,0.8,0.1\n")
    pp.setProjPsf(str(f))
    assert pp.projPsf_fname == str(f)
    f.write_text("50.0,0.25,4,0.2,0.6,0.2,0.1\n")
    with pytest.raises(RuntimeError, match="odd"):
        pp.setProjPsf(str(f))
    with pytest.raises(RuntimeError, match="open"):
        pp.setProjPsf(str(tmp_path / "missing.csv"))


def test_scanner_checks_and_cpu_osem_refused():
    with pytest.raises(RuntimeError, match="maxRingDiff"):
        Y.Scanner("s", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16, 1, 16, 1, 1)
    s = Y.Scanner("s", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16, 1, 15, 1, 1)
    assert s.getNumDets() == 96 * 16
    with pytest.raises(RuntimeError, match="GPU OSEM only"):
        Y.createOSEM(s, use_gpu=False)


def test_motion_frame_weights_match_oracle_and_one_frame():
    """The frames and weights of a time-averaged moved image
    (ReconstructionUtilsDevice.cu:146-190, LORMotion::getDuration): hand
    values, the oracle's selection, and a one-frame motion (full range,
    weight 1) that the reference's duration rule cannot serve."""
    import oracle as O
    ident = np.eye(3, 4, dtype=np.float32).reshape(1, -1)
    f0, w = Y.motionFrameWeights([0, 10, 30], 0, 50)
    assert f0 == 0 and np.allclose(w, [0.2, 0.4, 0.4])
    for t, t0, t1 in [([0, 10, 30], 0, 50), ([0, 10, 30], 10, 30), ([0, 10, 30], 5, 40),
                      ([3, 7, 8, 20], 3, 9), ([5], 5, 6), ([5], 0, 100)]:
        f0, w = Y.motionFrameWeights(t, t0, t1)
        frames, _, wo = O.time_average_frames(t, np.tile(ident[0], (len(t), 1)), t0, t1)
        assert f0 == frames[0] and len(w) == len(frames), (t, t0, t1)
        np.testing.assert_array_equal(np.float32(w), wo)
    # one frame: the whole range, weight 1, and the full time range of it
    assert Y.getFullTimeRange([5]) == (5, 6)
    assert Y.motionFrameWeights([5], 5, 6) == (0, [1.0])
    with pytest.raises(RuntimeError, match="greater"):
        Y.motionFrameWeights([0, 10], 20, 10)
    with pytest.raises(RuntimeError, match="no motion frame"):
        Y.motionFrameWeights([0, 10], 11, 12)
    with pytest.raises(RuntimeError, match="no motion frames"):
        Y.motionFrameWeights([], 0, 1)
