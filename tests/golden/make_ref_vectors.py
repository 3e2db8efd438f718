"""Write tests/golden/ref_vectors.npz from the reference's own code.

Runs the C-ABI harness of oracle/_ref (the reference's TimeOfFlightHelper,
util::get_alpha, util::calculateIntegral, MultiRayGenerator and
ProjectionPsfManager compiled from the reference tree by `make -C oracle
ref`) on seeded inputs and stores inputs + outputs.  The fixture lets the
oracle be pinned where the reference tree is absent.

    python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
import oracle as O  # noqa: E402

P = lambda a: a.ctypes.data_as(C.c_void_p)


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libyrtref.so not built (make -C oracle ref)")
    rng = np.random.default_rng(20240601)
    out = {}
    # TOF helper constants
    widths = np.array([50.0, 95.0, 250.0, 350.0, 500.0, 1000.0], np.float32)
    nstds = np.array([-1, 0, 1, 3, 5], np.int32)
    hv = np.zeros((widths.size, nstds.size, 3), np.float32)
    for i, w in enumerate(widths):
        for j, k in enumerate(nstds):
            a, b, c = C.c_float(), C.c_float(), C.c_float()
            R.yr_tof_helper(float(w), int(k), C.byref(a), C.byref(b), C.byref(c))
            hv[i, j] = (a.value, b.value, c.value)
    out.update(tof_widths=widths, tof_nstds=nstds, tof_helper=hv)
    # alpha range and weight on random inputs
    n = 2000
    tw = rng.choice(widths, n).astype(np.float32)
    tn = rng.choice(nstds, n).astype(np.int32)
    ln = rng.uniform(50, 900, n).astype(np.float32)
    tp = rng.uniform(-2000, 2000, n).astype(np.float32)
    lo = (rng.uniform(0, 1, n) * ln).astype(np.float32)
    hi = (lo + rng.uniform(0, 5, n)).astype(np.float32)
    ar = np.zeros((n, 2), np.float32)
    wt = np.zeros(n, np.float32)
    for i in range(n):
        a, b = C.c_float(), C.c_float()
        R.yr_tof_alpha_range(float(tw[i]), int(tn[i]), float(ln[i]), float(tp[i]), C.byref(a), C.byref(b))
        ar[i] = (a.value, b.value)
        wt[i] = R.yr_tof_weight(float(tw[i]), int(tn[i]), float(ln[i]), float(tp[i]), float(lo[i]), float(hi[i]))
    out.update(tof_in_width=tw, tof_in_nstd=tn, tof_in_norm=ln, tof_in_ps=tp,
               tof_in_lo=lo, tof_in_hi=hi, tof_alpha_range=ar, tof_weight=wt)
    # get_alpha
    ga_in = rng.uniform(-10, 10, (n, 5)).astype(np.float32)
    ga_in[: n // 4, 3] = ga_in[: n // 4, 2]  # flat rays
    ga_in[:, 1] = np.abs(ga_in[:, 1]) + 0.1
    ga_in[:, 0] = -ga_in[:, 1]
    ga_out = np.zeros((n, 2), np.float32)
    for i in range(n):
        r0, r1, p1, p2, _ = ga_in[i]
        inv = np.float32(0.0) if p1 == p2 else np.float32(1.0) / (p2 - p1)
        ga_in[i, 4] = inv
        a, b = C.c_float(), C.c_float()
        R.yr_get_alpha(float(r0), float(r1), float(p1), float(p2), float(inv), C.byref(a), C.byref(b))
        ga_out[i] = (a.value, b.value)
    out.update(get_alpha_in=ga_in, get_alpha_out=ga_out)
    # calculateIntegral on random kernels / intervals
    m = 1000
    ks = np.zeros((m, 31), np.float32)
    kk = rng.choice([1, 3, 5, 11, 21, 31], m).astype(np.int32)
    kst = rng.uniform(0.1, 1.0, m).astype(np.float32)
    x0 = rng.uniform(-12, 12, m).astype(np.float32)
    x1 = (x0 + rng.uniform(-1, 8, m)).astype(np.float32)
    iv = np.zeros(m, np.float32)
    for i in range(m):
        ks[i, :kk[i]] = rng.uniform(0, 1, kk[i])
        iv[i] = R.yr_calc_integral(P(np.ascontiguousarray(ks[i])), int(kk[i]), float(kst[i]),
                                   float(x0[i]), float(x1[i]))
    out.update(psf_kernels=ks, psf_k=kk, psf_step=kst, psf_x0=x0, psf_x1=x1, psf_integral=iv)
    # multi-ray lines (default generator: not parallel, no depth)
    q = 200
    mr_p1 = rng.uniform(-300, 300, (q, 3)).astype(np.float32)
    mr_p2 = rng.uniform(-300, 300, (q, 3)).astype(np.float32)
    ang1 = rng.uniform(0, 2 * np.pi, q)
    ang2 = rng.uniform(0, 2 * np.pi, q)
    mr_n1 = np.stack([np.cos(ang1), np.sin(ang1), np.zeros(q)], 1).astype(np.float32)
    mr_n2 = np.stack([np.cos(ang2), np.sin(ang2), np.zeros(q)], 1).astype(np.float32)
    mr_lines = np.zeros((q, 15, 6), np.float32)
    for i in range(q):
        buf = np.zeros(15 * 6, np.float32)
        R.yr_multiray_lines(P(mr_p1[i]), P(mr_p2[i]), P(mr_n1[i]), P(mr_n2[i]), 5.0, 3.2, 20.0, 15, P(buf))
        mr_lines[i] = buf.reshape(15, 6)
    out.update(mr_p1=mr_p1, mr_p2=mr_p2, mr_n1=mr_n1, mr_n2=mr_n2, mr_thick=np.array([5.0, 3.2, 20.0], np.float32),
               mr_lines=mr_lines)
    # projection-PSF manager on a CSV file: 3 rows of 13 taps, sStep 40
    K, S = 13, 3
    taps = rng.uniform(0, 1, (S, K)).astype(np.float32)
    row = np.concatenate([[40.0, 0.3, K], taps.ravel()]).astype(np.float32)
    with tempfile.NamedTemporaryFile("w", suffix=".csv", delete=False) as f:
        f.write(",".join(repr(float(v)) for v in row) + "\n")
        csv = f.name
    pq = 300
    pp1 = rng.uniform(-200, 200, (pq, 3)).astype(np.float32)
    pp2 = rng.uniform(-200, 200, (pq, 3)).astype(np.float32)
    pfl = rng.integers(0, 2, pq).astype(np.int32)
    px0 = rng.uniform(-5, 5, pq).astype(np.float32)
    px1 = (px0 + rng.uniform(-1, 6, pq)).astype(np.float32)
    pres = np.zeros((pq, 3), np.float32)
    for i in range(pq):
        hw, wv, ro = C.c_float(), C.c_float(), C.c_int()
        R.yr_psf_query(csv.encode(), P(pp1[i]), P(pp2[i]), int(pfl[i]), float(px0[i]), float(px1[i]),
                       C.byref(hw), C.byref(wv), C.byref(ro))
        pres[i] = (hw.value, wv.value, ro.value)
    os.unlink(csv)
    out.update(pm_row=row, pm_p1=pp1, pm_p2=pp2, pm_flip=pfl, pm_x0=px0, pm_x1=px1, pm_out=pres)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
