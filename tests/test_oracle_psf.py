"""Pins of the oracle's image-space PSF (oracle.psf_convolve, restating
OperatorPsf::convolve, OperatorPsf.cpp:290-375) and Gaussian kernel
(oracle.gaussian_kernel_1d, OperatorPsf::createGaussianKernel1D,
OperatorPsf.cpp:158-185), and of the host layer's PSF file reader
(OperatorPsf::readFromFileInternal's layout).  CPU only."""
import numpy as np
import pytest

import oracle as O
import pyyrtmi as Y


def brute(img, kx, ky, kz):
    """The reference's triple loop, per axis with its line buffer, in FP32."""
    x = np.array(img, np.float32)
    nt, nz, ny, nx = x.shape
    for k, ax, n in ((kx, 3, nx), (ky, 2, ny), (kz, 1, nz)):
        k = np.asarray(k, np.float32)
        h = len(k) // 2
        y = np.empty_like(x)
        for idx in np.ndindex(*[s for a, s in enumerate(x.shape) if a != ax]):
            line = x[idx[:ax] + (slice(None),) + idx[ax:]].copy()
            out = np.empty(n, np.float32)
            for i in range(n):
                s = np.float32(0)
                for kk in range(-h, h + 1):
                    r = i - kk
                    r = r + n if r < 0 else (r - n if r >= n else r)
                    s = np.float32(s + np.float32(k[kk + h] * line[r]))
                out[i] = s
            y[idx[:ax] + (slice(None),) + idx[ax:]] = out
        x = y
    return x


def test_psf_convolve_matches_the_reference_loop():
    rng = np.random.default_rng(1)
    img = rng.random((2, 5, 6, 7), dtype=np.float32)
    kx = rng.random(5).astype(np.float32)
    ky = rng.random(3).astype(np.float32)
    kz = rng.random(7).astype(np.float32)  # half length 3 < nz = 5
    np.testing.assert_array_equal(O.psf_convolve(img, kx, ky, kz), brute(img, kx, ky, kz))
    np.testing.assert_array_equal(O.psf_convolve(img, kx, ky, kz, transpose=True),
                                  brute(img, kx[::-1], ky[::-1], kz[::-1]))


def test_psf_identity_shift_and_adjoint():
    rng = np.random.default_rng(2)
    img = rng.random((1, 6, 8, 9), dtype=np.float32)
    one = np.array([0, 1, 0], np.float32)
    np.testing.assert_array_equal(O.psf_convolve(img, one, one, one), img)
    # a one-hot tap shifts circularly: tap 0 pairs with kk = -h, so [1, 0, 0]
    # gives out[i] = in[i + 1] along its axis (here z)
    shift = np.array([1, 0, 0], np.float32)
    np.testing.assert_array_equal(O.psf_convolve(img, one, one, shift),
                                  np.roll(img, -1, axis=1))
    np.testing.assert_array_equal(O.psf_convolve(img, shift, one, one),
                                  np.roll(img, -1, axis=3))
    # <P x, y> = <x, P^T y>
    k = [rng.random(n).astype(np.float32) for n in (5, 3, 5)]
    y = rng.random(img.shape, dtype=np.float32)
    a = np.vdot(O.psf_convolve(img, *k).astype(np.float64), y.astype(np.float64))
    b = np.vdot(img.astype(np.float64), O.psf_convolve(y, *k, transpose=True).astype(np.float64))
    assert abs(a - b) <= 1e-5 * abs(a)
    with pytest.raises(ValueError):
        O.psf_convolve(img, np.ones(4, np.float32), one, one)


def test_gaussian_kernel_closed_form():
    for sigma, v, size in ((2.0, 1.5, 0), (3.0, 2.0, 9), (1.2, 2.5, 5)):
        k = O.gaussian_kernel_1d(sigma, v, size)
        n = size or int(np.rint(5 * sigma / v))
        assert len(k) == n and abs(float(k.sum()) - 1.0) < 1e-6
        x = np.arange(n) - n // 2
        want = np.exp(-0.5 * (x * v / sigma) ** 2)
        np.testing.assert_allclose(k, want / want.sum(), rtol=2e-6)
        np.testing.assert_allclose(k, k[::-1], rtol=1e-7)


def test_image_psf_file_reader(tmp_path):
    f = str(tmp_path / "psf.csv")
    with open(f, "w") as fh:
        fh.write("0.25,0.5,0.25,9\n")
        fh.write("0.1 0.2 0.4 0.2 0.1\n")
        fh.write("1\n")
        fh.write("3,5,1\n")
    kx, ky, kz = Y.readImagePSF(f)
    np.testing.assert_array_equal(kx, np.float32([0.25, 0.5, 0.25]))
    np.testing.assert_array_equal(ky, np.float32([0.1, 0.2, 0.4, 0.2, 0.1]))
    np.testing.assert_array_equal(kz, np.float32([1.0]))
    with open(f, "w") as fh:
        fh.write("1,2\n1,2\n1,2\n2,2,2\n")
    with pytest.raises(RuntimeError, match="odd"):
        Y.readImagePSF(f)
