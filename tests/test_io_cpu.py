"""File formats of the C++ host layer (csrc/host/yrtmi_io.cpp through
pyyrtmi), no GPU.  Each reader is checked on bytes laid out here with numpy
from the reference's documented formats (list-mode_file.md, Array.hpp's
raw-data header, DetCoord.cpp's six floats per detector, motion_file.md,
NIfTI-1 fields as Image.cpp writes them), and the round trips follow the
reference's own unit tests (unit_tests/datastruct/test_ListModeLUT.cpp,
test_Image.cpp "image-readwrite", test_LORMotion.cpp "lor-motion",
test_DynamicFraming.cpp "dynamicframing-io", test_Array.cpp fileread)."""
import json
import os

import numpy as np
import pytest

import pyyrtmi as Y

RAWD_MAGIC = 732174000


def scanner():
    return Y.Scanner("t", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24, 1, 23, 1, 1)


# ---- raw-data arrays -------------------------------------------------------
def test_rawd_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.random((3, 4, 5), dtype=np.float32)
    f = str(tmp_path / "a.rawd")
    Y.writeRawd(f, a)
    raw = open(f, "rb").read()
    # int32 magic, int32 ndim, uint64 dims, float32 data (DataFileRawd.save)
    assert np.frombuffer(raw[:8], np.int32).tolist() == [RAWD_MAGIC, 3]
    assert np.frombuffer(raw[8:32], np.uint64).tolist() == [3, 4, 5]
    np.testing.assert_array_equal(np.frombuffer(raw[32:], np.float32).reshape(3, 4, 5), a)
    np.testing.assert_array_equal(Y.readRawd(f), a)
    # a file laid out by hand reads back
    g = str(tmp_path / "b.rawd")
    with open(g, "wb") as fh:
        np.array([RAWD_MAGIC, 2], np.int32).tofile(fh)
        np.array([2, 3], np.uint64).tofile(fh)
        np.arange(6, dtype=np.float32).tofile(fh)
    np.testing.assert_array_equal(Y.readRawd(g), np.arange(6, dtype=np.float32).reshape(2, 3))


def test_rawd_rejects_bad_files(tmp_path):
    f = str(tmp_path / "bad.rawd")
    with open(f, "wb") as fh:
        np.array([123, 1], np.int32).tofile(fh)
        np.array([4], np.uint64).tofile(fh)
        np.zeros(4, np.float32).tofile(fh)
    with pytest.raises(RuntimeError, match="MAGIC_NUMBER"):
        Y.readRawd(f)
    with open(f, "wb") as fh:
        np.array([RAWD_MAGIC, 1], np.int32).tofile(fh)
        np.array([10], np.uint64).tofile(fh)
        np.zeros(4, np.float32).tofile(fh)
    with pytest.raises(RuntimeError, match="shorter"):
        Y.readRawd(f)


# ---- list-mode files ---------------------------------------------------------
@pytest.mark.parametrize("tof,rnd", [(False, False), (True, False), (False, True),
                                     (True, True)])
def test_listmode_file_layout_and_round_trip(tmp_path, tof, rnd):
    sc = scanner()
    rng = np.random.default_rng(2)
    n = 1000
    ts = rng.integers(0, 1_000_000, n).astype(np.uint32)
    d1 = rng.integers(0, sc.getNumDets(), n).astype(np.uint32)
    d2 = rng.integers(0, sc.getNumDets(), n).astype(np.uint32)
    tv = rng.uniform(-5000, 5000, n).astype(np.float32)
    rv = rng.uniform(0, 1000, n).astype(np.float32)
    # the documented layout: uint32 t, d1, d2, [float32 tof], [float32 randoms]
    fields = [ts.view(np.uint32), d1, d2]
    if tof:
        fields.append(tv.view(np.uint32))
    if rnd:
        fields.append(rv.view(np.uint32))
    f = str(tmp_path / "lm.lmDat")
    np.stack(fields, axis=1).astype(np.uint32).tofile(f)
    lm = Y.ListModeLUTOwned(sc, f, flag_tof=tof, flag_randoms=rnd)
    assert lm.count() == n and lm.hasTOF() == tof and lm.hasRandomsEstimates() == rnd
    np.testing.assert_array_equal(lm.getTimestampArray(), ts)
    np.testing.assert_array_equal(lm.getDetector1Array(), d1)
    np.testing.assert_array_equal(lm.getDetector2Array(), d2)
    if tof:
        np.testing.assert_array_equal(lm.getTOFArray(), tv)
    else:
        with pytest.raises(RuntimeError, match="TOF"):
            lm.getTOFArray()
    if rnd:
        np.testing.assert_array_equal(lm.getRandomsEstimatesArray(), rv)
    # write and read back: identical bytes
    g = str(tmp_path / "lm2.lmDat")
    lm.writeToFile(g)
    assert open(g, "rb").read() == open(f, "rb").read()
    lm2 = Y.ListModeLUTOwned(ts, d1, d2, tof=tv if tof else None, randoms=rv if rnd else None)
    h = str(tmp_path / "lm3.lmDat")
    lm2.writeToFile(h)
    assert open(h, "rb").read() == open(f, "rb").read()


def test_listmode_file_errors(tmp_path):
    sc = scanner()
    f = str(tmp_path / "lm.lmDat")
    np.array([[0, 1, 2, 0]], np.uint32).tofile(f)
    with pytest.raises(RuntimeError, match="incorrect size"):
        Y.ListModeLUTOwned(sc, f, flag_tof=False)  # 16 bytes: not 12-byte events
    np.array([[0, 1, sc.getNumDets()]], np.uint32).tofile(f)
    with pytest.raises(RuntimeError, match="Detectors invalid in list-mode event 0"):
        Y.ListModeLUTOwned(sc, f)


# ---- NIfTI images -----------------------------------------------------------
def nifti_header(raw):
    h = {}
    h["sizeof_hdr"] = int(np.frombuffer(raw[0:4], np.int32)[0])
    h["dim"] = np.frombuffer(raw[40:56], np.int16).tolist()
    h["datatype"] = int(np.frombuffer(raw[70:72], np.int16)[0])
    h["pixdim"] = np.frombuffer(raw[76:108], np.float32).tolist()
    h["vox_offset"] = float(np.frombuffer(raw[108:112], np.float32)[0])
    h["xyzt_units"] = raw[123]
    h["qform_code"] = int(np.frombuffer(raw[252:254], np.int16)[0])
    h["sform_code"] = int(np.frombuffer(raw[254:256], np.int16)[0])
    h["srow"] = np.frombuffer(raw[280:328], np.float32).reshape(3, 4)
    h["magic"] = raw[344:348]
    return h


@pytest.mark.parametrize("nt,gz", [(1, False), (1, True), (3, False)])
def test_nifti_round_trip_and_fields(tmp_path, nt, gz):
    rng = np.random.default_rng(3)
    nx, ny, nz = 27, 31, 19
    p = Y.ImageParams(nx, ny, nz, 54.0, 46.5, 38.0, 3.25, -7.5, 1.0, nt)
    a = rng.random((nt, nz, ny, nx), dtype=np.float32)
    f = str(tmp_path / ("img.nii.gz" if gz else "img.nii"))
    Y.writeImage(f, p, a)
    q, b = Y.readImage(f)
    assert q.isSameAs(p) and q.nt == nt
    assert abs(q.off_x - p.off_x) < 1e-5 and abs(q.off_y - p.off_y) < 1e-5
    assert abs(q.off_z - p.off_z) < 1e-5
    np.testing.assert_array_equal(b, a)
    if gz:
        plain = str(tmp_path / "plain.nii")
        Y.writeImage(plain, p, a)
        assert os.path.getsize(f) < os.path.getsize(plain)  # as the reference checks
        return
    raw = open(f, "rb").read()
    h = nifti_header(raw)
    assert h["sizeof_hdr"] == 348 and h["magic"] == b"n+1\x00"
    assert h["dim"][:5] == [4 if nt > 1 else 3, nx, ny, nz, nt]
    assert h["datatype"] == 16 and h["vox_offset"] == 352.0
    np.testing.assert_allclose(h["pixdim"][1:5], [p.vx, p.vy, p.vz, 1.0])
    assert h["qform_code"] == 0 and h["sform_code"] == 1 and h["xyzt_units"] == (2 | 8)
    # sform as Image::writeToFile sets it: diag(-vx, -vy, vz), origin
    # (-o_x, -o_y, o_z), o = off - length / 2 + v / 2
    o = [p.off_x - 0.5 * p.length_x + 0.5 * p.vx, p.off_y - 0.5 * p.length_y + 0.5 * p.vy,
         p.off_z - 0.5 * p.length_z + 0.5 * p.vz]
    want = np.array([[-p.vx, 0, 0, -o[0]], [0, -p.vy, 0, -o[1]], [0, 0, p.vz, o[2]]],
                    np.float32)
    np.testing.assert_allclose(h["srow"], want, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(np.frombuffer(raw[352:], np.float32).reshape(a.shape), a)


def handmade_nifti(path, data, dtype_code, np_dtype, pixdim, slope=1.0, inter=0.0,
                   sform=None, qform=None, units=2):
    hdr = bytearray(352)
    hdr[0:4] = np.int32(348).tobytes()
    nz, ny, nx = data.shape
    hdr[40:56] = np.array([3, nx, ny, nz, 1, 1, 1, 1], np.int16).tobytes()
    hdr[70:72] = np.int16(dtype_code).tobytes()
    hdr[72:74] = np.int16(8 * np.dtype(np_dtype).itemsize).tobytes()
    hdr[76:108] = np.array(pixdim, np.float32).tobytes()
    hdr[108:112] = np.float32(352).tobytes()
    hdr[112:116] = np.float32(slope).tobytes()
    hdr[116:120] = np.float32(inter).tobytes()
    hdr[123] = units
    if sform is not None:
        hdr[254:256] = np.int16(1).tobytes()
        hdr[280:328] = np.asarray(sform, np.float32).tobytes()
    if qform is not None:
        hdr[252:254] = np.int16(1).tobytes()
        hdr[256:280] = np.asarray(qform, np.float32).tobytes()  # b c d qx qy qz
    hdr[344:348] = b"n+1\x00"
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(data.astype(np_dtype).tobytes())


def test_nifti_reads_other_types_scaling_and_qform(tmp_path):
    rng = np.random.default_rng(4)
    d = rng.integers(-300, 300, (4, 5, 6)).astype(np.int16)
    f = str(tmp_path / "i16.nii")
    sform = [[2.0, 0, 0, -5.0], [0, 3.0, 0, 6.0], [0, 0, 1.5, 7.0]]
    handmade_nifti(f, d, 4, np.int16, [0, 2.0, 3.0, 1.5, 1, 0, 0, 0], slope=0.5, inter=1.0,
                   sform=sform)
    p, img = Y.readImage(f)
    np.testing.assert_array_equal(img[0], d.astype(np.float32) * 0.5 + 1.0)
    assert (p.nx, p.ny, p.nz, p.nt) == (6, 5, 4, 1)
    assert (p.vx, p.vy, p.vz) == (2.0, 3.0, 1.5)
    # origin -> offset (Image::originToOffset): off = o + length / 2 - v / 2
    assert abs(p.off_x - (-5.0 + 6.0 - 1.0)) < 1e-6
    assert abs(p.off_y - (6.0 + 7.5 - 1.5)) < 1e-6
    assert abs(p.off_z - (7.0 + 3.0 - 0.75)) < 1e-6
    # a negative diagonal entry is flipped with its translation
    g = str(tmp_path / "flip.nii")
    handmade_nifti(g, d, 4, np.int16, [0, 2.0, 3.0, 1.5, 1, 0, 0, 0],
                   sform=[[-2.0, 0, 0, 5.0], [0, 3.0, 0, 6.0], [0, 0, 1.5, 7.0]])
    p2, _ = Y.readImage(g)
    assert abs(p2.off_x - p.off_x) < 1e-6
    # qform only (identity rotation b = c = d = 0): translation from qoffset
    q = str(tmp_path / "q.nii")
    handmade_nifti(q, d, 16, np.float32, [1, 2.0, 3.0, 1.5, 1, 0, 0, 0],
                   qform=[0, 0, 0, -5.0, 6.0, 7.0])
    p3, img3 = Y.readImage(q)
    assert abs(p3.off_x - p.off_x) < 1e-6 and abs(p3.off_z - p.off_z) < 1e-6
    np.testing.assert_array_equal(img3[0], d.astype(np.float32))
    # a rotation is rejected, as the reference does
    r = str(tmp_path / "rot.nii")
    handmade_nifti(r, d, 16, np.float32, [0, 2.0, 3.0, 1.5, 1, 0, 0, 0],
                   sform=[[2.0, 0.5, 0, 0], [0, 3.0, 0, 0], [0, 0, 1.5, 0]])
    with pytest.raises(RuntimeError, match="rotation"):
        Y.readImage(r)


# ---- scanner definitions -----------------------------------------------------
def scanner_json(path, **over):
    j = {"VERSION": 3.1, "scannerName": "json_scanner", "axialFOV": 96.0,
         "crystalSize_trans": 4.0, "crystalSize_z": 4.0, "crystalDepth": 10.0,
         "scannerRadius": 180.0, "dets_per_ring": 128, "num_rings": 24, "num_doi": 1,
         "max_ring_diff": 23, "min_ang_diff": 1, "dets_per_block": 1}
    j.update(over)
    for k, v in list(j.items()):
        if v is None:
            del j[k]
    with open(path, "w") as f:
        json.dump(j, f)
    return path


def test_scanner_json(tmp_path):
    f = scanner_json(str(tmp_path / "s.json"))
    s = Y.Scanner(f)
    ref = scanner()
    assert s.scannerName == "json_scanner" and s.getNumDets() == ref.getNumDets()
    for k in ("axialFOV", "crystalSize_z", "crystalSize_trans", "crystalDepth",
              "scannerRadius", "detsPerRing", "numRings", "numDOI", "maxRingDiff",
              "minAngDiff", "detsPerBlock"):
        assert getattr(s, k) == getattr(ref, k), k
    for a, b in zip(s.getDetectorLUT(), ref.getDetectorLUT()):
        np.testing.assert_array_equal(a, b)
    # camelCase aliases
    g = scanner_json(str(tmp_path / "s2.json"), dets_per_ring=None, detsPerRing=128)
    assert Y.Scanner(g).detsPerRing == 128
    with pytest.raises(RuntimeError, match="too recent"):
        Y.Scanner(scanner_json(str(tmp_path / "v.json"), VERSION=4.0))
    with pytest.raises(RuntimeError, match="Missing scanner radius"):
        Y.Scanner(scanner_json(str(tmp_path / "r.json"), scannerRadius=None))
    with pytest.raises(RuntimeError, match="Maximum Ring difference"):
        Y.Scanner(scanner_json(str(tmp_path / "m.json"), max_ring_diff=24))
    with pytest.raises(RuntimeError, match="detMask"):
        Y.Scanner(scanner_json(str(tmp_path / "k.json"), detMask="mask.raw"))


def test_scanner_json_with_detector_coordinates(tmp_path):
    ref = scanner()
    pos, ori = ref.getDetectorLUT()
    rng = np.random.default_rng(5)
    pos2 = (pos + rng.normal(0, 0.5, pos.shape)).astype(np.float32)
    ori2 = ori.astype(np.float32)
    lut = str(tmp_path / "det.lut")
    # six float32 per detector: x, y, z, ox, oy, oz
    np.concatenate([pos2, ori2], axis=1).astype(np.float32).tofile(lut)
    f = scanner_json(str(tmp_path / "s.json"), detCoord="det.lut")
    s = Y.Scanner(f)
    p, o = s.getDetectorLUT()
    np.testing.assert_array_equal(p, pos2)
    np.testing.assert_array_equal(o, ori2)
    Y.writeDetCoord(str(tmp_path / "det2.lut"), pos2, ori2)
    assert open(str(tmp_path / "det2.lut"), "rb").read() == open(lut, "rb").read()
    np.concatenate([pos2, ori2], axis=1)[:-1].astype(np.float32).tofile(lut)
    with pytest.raises(RuntimeError, match="number of detectors"):
        Y.Scanner(f)


# ---- motion and dynamic framing ------------------------------------------------
def test_lor_motion_round_trip(tmp_path):
    # the records of the reference's "lor-motion" test
    t = np.array([1000, 2000], np.int64)
    xf = np.array([[1.0, 0.0, 0.0, 10.0, 0.0, 1.0, 0.0, 20.0, 0.0, 0.0, 1.0, 30.0],
                   [0.707, 0.707, 0.0, 5.0, -0.707, 0.707, 0.0, 15.0, 0.0, 0.0, 1.0, 25.0]],
                  np.float32)
    e = np.array([1.0, 1.0], np.float32)
    f = str(tmp_path / "m.csv")
    Y.writeLORMotion(f, t, xf, e)
    first = open(f).read().splitlines()[0].split(",")
    assert first[0] == "1000" and len(first) == 14 and first[4] == "10.000000000"
    t2, xf2, e2 = Y.readLORMotion(f)
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_allclose(xf2, xf, atol=1.2e-7)
    np.testing.assert_allclose(e2, e)
    # lines that do not parse are skipped; spaces separate as commas do
    with open(f, "w") as fh:
        fh.write("1000 1 0 0 10 0 1 0 20 0 0 1 30 0.5\n")
        fh.write("x,1,0,0,10,0,1,0,20,0,0,1,30,0.5\n")
        fh.write("2000,1,0,0\n")
        fh.write("3000,1,0,0,10,0,1,0,20,0,0,1,30,0.5,9\n")
        fh.write("\n4000,1,0,0,1,0,1,0,2,0,0,1,3,0.25\n")
    t3, xf3, e3 = Y.readLORMotion(f)
    np.testing.assert_array_equal(t3, [1000, 4000])
    np.testing.assert_allclose(e3, [0.5, 0.25])


def ref_motion_frames(ts, starts):
    """ListMode::addLORMotion's event walk (ListMode.cpp:95-125), with 255
    for the events before the first frame."""
    out = np.full(len(ts), 255, np.uint8)
    e, n, nf = 0, len(ts), len(starts)
    while e < n and ts[e] < starts[0]:
        e += 1
    f = 0
    while f < nf - 1:
        while e < n and ts[e] < starts[f + 1]:
            out[e] = f
            e += 1
        f += 1
    out[e:] = f
    return out


def ref_dynamic_frames(ts, bounds):
    """ListMode::addDynamicFraming's event walk (ListMode.cpp:128-180)."""
    out = np.full(len(ts), -1, np.int32)
    e, n = 0, len(ts)
    while e < n and ts[e] < bounds[0]:
        e += 1
    for f in range(len(bounds) - 1):
        while e < n and ts[e] < bounds[f + 1]:
            out[e] = f
            e += 1
    return out


def test_event_frames(tmp_path):
    rng = np.random.default_rng(6)
    ts = np.sort(rng.integers(0, 10_000, 5000)).astype(np.uint32)
    starts = np.array([500, 2000, 2000, 7000], np.int64)
    xf = np.tile(np.eye(3, 4, dtype=np.float32).reshape(1, 12), (4, 1))
    m = Y.motionFrames(ts, starts, xf, np.zeros(4, np.float32))
    np.testing.assert_array_equal(m, ref_motion_frames(ts, starts))
    assert m[0] == 255 and m[-1] == 3 and not np.any(m == 1)  # an empty frame
    bounds = np.array([1000, 3000, 6500, 9000], np.uint32)
    d = Y.dynamicFrames(ts, bounds)
    np.testing.assert_array_equal(d, ref_dynamic_frames(ts, bounds))
    assert d[0] == -1 and d[-1] == -1 and set(np.unique(d)) == {-1, 0, 1, 2}
    # events out of order are walked as the reference walks them
    tsu = rng.integers(0, 10_000, 500).astype(np.uint32)
    np.testing.assert_array_equal(Y.dynamicFrames(tsu, bounds), ref_dynamic_frames(tsu, bounds))
    np.testing.assert_array_equal(Y.motionFrames(tsu, starts, xf, np.zeros(4, np.float32)),
                                  ref_motion_frames(tsu, starts))
    # framing file: one timestamp per line (DynamicFraming::writeToFile)
    f = str(tmp_path / "framing.txt")
    Y.writeDynamicFraming(f, bounds)
    assert open(f).read() == "1000\n3000\n6500\n9000\n"
    np.testing.assert_array_equal(Y.readDynamicFraming(f), bounds)


# ---- command line --------------------------------------------------------------
def test_cli_refuses_options_outside_the_path(tmp_path):
    from yrtmi import cli
    f = scanner_json(str(tmp_path / "s.json"))
    for argv in (["reconstruct", "-s", f, "-o", "x.nii", "--varpsf", "k.csv"],
                 ["forward_project", "-s", f, "-i", "x.nii", "-o", "h.rawd", "--sparse", "1"],
                 ["backproject", "-s", f, "-i", "lm", "-o", "x.nii", "-p", "p.json",
                  "--detmask", "m"]):
        with pytest.raises(SystemExit, match="not supported"):
            cli.main(argv)
    with pytest.raises(SystemExit, match="only H"):
        cli.main(["reconstruct", "-s", f, "-o", "x.nii", "--randoms_format", "SH"])
    assert cli.main([]) == 2
    with pytest.raises(SystemExit):
        cli.main(["backproject", "-s", f])  # --input, --out and --params are required


def test_image_params_json(tmp_path):
    f = str(tmp_path / "p.json")
    with open(f, "w") as fh:
        json.dump({"VERSION": 1.2, "nx": 10, "ny": 12, "nz": 8, "vx": 2.5, "length_y": 30.0,
                   "vz": 3.0, "offset_x": 1.5, "off_z": -2.0, "num_frames": 2}, fh)
    p = Y.ImageParams(f)
    assert (p.nx, p.ny, p.nz, p.nt) == (10, 12, 8, 2)
    assert (p.length_x, p.length_y, p.length_z) == (25.0, 30.0, 24.0)
    assert (p.vx, p.vy, p.vz) == (2.5, 2.5, 3.0)
    assert (p.off_x, p.off_y, p.off_z) == (1.5, 0.0, -2.0)
    g = str(tmp_path / "q.json")
    p.serialize(g)
    q = Y.ImageParams(g)
    assert q.isSameAs(p) and q.nt == 2 and q.off_x == p.off_x and q.off_z == p.off_z
    with open(f, "w") as fh:
        json.dump({"VERSION": 2.0, "nx": 1, "ny": 1, "nz": 1, "vx": 1, "vy": 1, "vz": 1}, fh)
    with pytest.raises(RuntimeError, match="Wrong version"):
        Y.ImageParams(f)
    with open(f, "w") as fh:
        json.dump({"VERSION": 1.2, "nx": 1, "ny": 1, "nz": 1, "vx": 1, "vy": 1}, fh)
    with pytest.raises(RuntimeError, match="voxel size"):
        Y.ImageParams(f)
