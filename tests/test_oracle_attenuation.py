"""Pin of the oracle's in-vivo attenuation precorrection (em_subset's
precorr, restating OSEM_GPU.cu:1303-1308: the EM denominator, randoms and
scatter included, multiplied by the factor): m / ((A x) a + r) p equals
(m / p) / ((A x) a + r), so precorr p must equal measurements m / p.  CPU
only."""
import numpy as np

import oracle as O
from helpers import rel_l2


def test_precorrection_equals_scaled_measurements():
    sc = O.ScannerDef(40.0, 4.0, 4.0, 10.0, 100.0, 48, 10)
    im = O.ImageDef(24, 24, 10, 4.0, 4.0, 4.0)
    pos, ori = O.lut(sc)
    rng = np.random.default_rng(2)
    n = 3000
    d1 = rng.integers(0, sc.num_dets, n)
    d2 = (d1 + sc.dets_per_ring // 2 + rng.integers(-8, 8, n)) % sc.num_dets
    L = O.lors_from_pairs(pos, ori, d1, d2)
    proj = O.proj_for_scanner(sc)
    x = rng.uniform(0.5, 1.5, im.shape).astype(np.float32)
    m = rng.integers(0, 5, n).astype(np.float32)
    a = rng.uniform(0.3, 1.0, n).astype(np.float32)
    r = rng.uniform(0.0, 0.2, n).astype(np.float32)
    p = rng.uniform(0.2, 1.0, n).astype(np.float32)
    u1 = np.zeros(im.shape, np.float32)
    O.em_subset(im, x, u1, proj, L, meas=m, att=a, rnd=r, precorr=p)
    u2 = np.zeros(im.shape, np.float32)
    O.em_subset(im, x, u2, proj, L, meas=(m / p).astype(np.float32), att=a, rnd=r)
    assert np.count_nonzero(u1) > 1000
    assert rel_l2(u1, u2) < 1e-6
    u3 = np.zeros(im.shape, np.float32)
    O.em_subset(im, x, u3, proj, L, meas=m, att=a, rnd=r)
    assert rel_l2(u1, u3) > 1e-2  # the factor is applied
