"""Pins of the oracle's time-averaged moved image (oracle.move_accumulate,
time_average_move, time_average_move_dynamic, restating
timeAverageMoveImage_kernel, ImageSpaceKernels.cu:191-264, and its drivers,
ReconstructionUtilsDevice.cu:78-300) and of its helpers invert_transform
(TransformUtils.hpp:16-40), full_time_range (ReconstructionUtils.cpp:
468-495) and motion_duration (LORMotion.cpp:73-88).  The checks are fixed
by geometry, not by retyping the formula: whole-voxel shifts and a quarter
turn are permutations of the voxels, a random rigid motion equals scipy's
order-1 spline interpolation with zero padding, the average is linear in
its frame weights.  CPU only."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from helpers import rel_l2


def rigid(rng, max_angle=np.pi, max_shift=5.0):
    """A random rotation (z-y-x Euler angles) and translation as transform_t rows."""
    a, b, c = rng.uniform(-max_angle, max_angle, 3)
    cz, sz, cy, sy, cx, sx = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    R = Rz @ Ry @ Rx
    t = rng.uniform(-max_shift, max_shift, 3)
    return np.concatenate([np.c_[R, t].reshape(-1)]).astype(np.float32)


def xf_of(R, t):
    return np.c_[np.asarray(R, np.float64), np.asarray(t, np.float64)].reshape(-1).astype(np.float32)


IDENT = xf_of(np.eye(3), [0, 0, 0])


def test_invert_transform_composes_to_identity():
    rng = np.random.default_rng(3)
    xf = np.stack([rigid(rng) for _ in range(20)])
    inv = O.invert_transform(xf)
    p = rng.uniform(-50, 50, (20, 3)).astype(np.float32)
    q = O.apply_motion(O.apply_motion(p, xf), inv)
    np.testing.assert_allclose(q, p, atol=2e-4)
    # the rotation part is the transpose, exactly
    R = xf[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3, 3)
    Ri = inv[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3, 3)
    np.testing.assert_array_equal(Ri, np.transpose(R, (0, 2, 1)))


def test_full_time_range_and_durations():
    assert O.full_time_range([7]) == (7, 8)
    assert O.full_time_range([0, 10, 30]) == (0, 50)
    assert [float(O.motion_duration([0, 10, 30], f)) for f in range(3)] == [10.0, 20.0, 20.0]
    frames, inv, w = O.time_average_frames([0, 10, 30], np.tile(IDENT, (3, 1)), 0, 50)
    assert list(frames) == [0, 1, 2]
    assert abs(float(w.sum()) - 1.0) < 1e-6
    # both ends included: the frame starting at t1 counts, with its full duration
    frames, _, w = O.time_average_frames([0, 10, 30], np.tile(IDENT, (3, 1)), 10, 30)
    assert list(frames) == [1, 2] and list(w) == [1.0, 1.0]
    # a frame starting before t0 is left out
    frames, _, _ = O.time_average_frames([0, 10, 30], np.tile(IDENT, (3, 1)), 5, 40)
    assert list(frames) == [1, 2]
    with pytest.raises(ValueError):
        O.time_average_frames([0, 10], np.tile(IDENT, (2, 1)), 20, 10)


def image(rng, nz, ny, nx):
    return rng.uniform(0.5, 1.5, (nz, ny, nx)).astype(np.float32)


def test_identity_and_whole_voxel_shifts():
    rng = np.random.default_rng(4)
    im = O.ImageDef(12, 10, 8, 2.0, 2.5, 3.0)
    x = image(rng, 8, 10, 12)
    got = O.move_accumulate(x, im, IDENT[None], [1.0], np.zeros_like(x))
    np.testing.assert_allclose(got, x, rtol=1e-5)
    # moving the object by +2 voxels in x and -1 in z: out(p) = in(p - t)
    fwd = xf_of(np.eye(3), [2 * 2.0, 0.0, -1 * 3.0])
    got = O.move_accumulate(x, im, O.invert_transform(fwd)[None], [1.0], np.zeros_like(x))
    want = np.zeros_like(x)
    want[:-1, :, 2:] = x[1:, :, :-2]
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


def test_quarter_turn_is_a_voxel_permutation():
    rng = np.random.default_rng(5)
    n = 10
    im = O.ImageDef(n, n, 6, 2.0, 2.0, 2.0)
    x = image(rng, 6, n, n)
    R = [[0, -1, 0], [1, 0, 0], [0, 0, 1]]  # +90 degrees about z
    got = O.move_accumulate(x, im, O.invert_transform(xf_of(R, [0, 0, 0]))[None], [1.0],
                            np.zeros_like(x))
    # out at (x_i, y_j) = in at R^T (x_i, y_j) = (y_j, -x_i): x index j, y index n-1-i
    want = np.empty_like(x)
    for j in range(n):
        for i in range(n):
            want[:, j, i] = x[:, n - 1 - i, j]
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed", [6, 7])
def test_random_rigid_motion_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    im = O.ImageDef(14, 11, 9, 2.0, 1.5, 2.5, ox=1.0, oy=-2.0, oz=0.5)
    x = image(rng, 9, 11, 14)
    fwd = rigid(rng, max_angle=0.6, max_shift=3.0)
    got = O.move_accumulate(x, im, O.invert_transform(fwd)[None], [1.0], np.zeros_like(x))
    # the textbook form in float64: q = R^T (p - t), index = (q - c0) / v
    R = fwd.reshape(3, 4)[:, :3].astype(np.float64)
    t = fwd.reshape(3, 4)[:, 3].astype(np.float64)
    v = np.array([2.0, 1.5, 2.5])
    n = np.array([14, 11, 9])
    off = np.array([1.0, -2.0, 0.5])
    c0 = off - n * v / 2 + v / 2
    Z, Yy, X = np.meshgrid(*[c0[a] + v[a] * np.arange(n[a]) for a in (2, 1, 0)], indexing="ij")
    P = np.stack([X, Yy, Z], axis=-1)
    Q = (P - t) @ R  # rows: R^T (p - t)
    idx = (Q - c0) / v
    want = ndimage.map_coordinates(x.astype(np.float64), [idx[..., 2], idx[..., 1], idx[..., 0]],
                                   order=1, mode="grid-constant", cval=0.0, prefilter=False)
    assert np.count_nonzero(want) > x.size // 2
    assert rel_l2(got, want) < 1e-5


def test_average_is_linear_in_frame_weights():
    rng = np.random.default_rng(8)
    im = O.ImageDef(9, 8, 7, 2.0, 2.0, 2.0)
    x = image(rng, 7, 8, 9)
    xf = np.stack([rigid(rng, 0.3, 2.0) for _ in range(3)])
    inv = O.invert_transform(xf)
    w = np.array([0.2, 0.5, 0.3], np.float32)
    all3 = O.move_accumulate(x, im, inv, w, np.zeros_like(x))
    parts = sum(w[f] * O.move_accumulate(x, im, inv[f:f + 1], [1.0], np.zeros_like(x))
                for f in range(3))
    assert rel_l2(all3, parts) < 1e-6
    # the drivers: full range weights sum to one; the dynamic form per frame
    ts = [0, 100, 250]
    full = O.time_average_move(x, im, ts, xf)
    assert rel_l2(full, O.move_accumulate(x, im, inv, [100 / 400, 150 / 400, 150 / 400],
                                          np.zeros_like(x))) < 1e-6
    dyn = O.time_average_move_dynamic(x, im, ts, xf, [0, 100, 400])
    assert dyn.shape == (2,) + x.shape
    assert rel_l2(dyn[1], O.time_average_move(x, im, ts, xf, 100, 400)) == 0.0
