// yrtmi device functions for gfx950: LOR fetch from SoA events and the
// resident detector LUT, TOF Gaussian, projection-space PSF integral, and the
// Siddon / distance-driven ray kernels.
//
// Arithmetic: FP32 storage and accumulation; the ray-plane intersection
// setup (FOV-cylinder clip, slab alphas, TOF window, first crossings, first
// voxel index) in FP32 in the reference's operation order, or in FP64 on
// request (near-grazing stability).
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace yrtmi
{

// FP32 operations each rounded once to nearest and never fused with a
// neighbouring one (contraction is off in these bodies, so the compiler
// cannot turn a product and a sum into an FMA even after inlining): the
// reference's separately rounded float arithmetic where a result must match
// it bit for bit.  The HIP headers' __fadd_rn / __fmul_rn are plain,
// contractible operators and __fsqrt_rn a native (not correctly rounded)
// square root, so they cannot serve here.  sqrtf and '/' are correctly
// rounded in this build (hipcc's default for FP32 division and sqrt).
__device__ __forceinline__ float fmulRn(float a, float b)
{
#pragma clang fp contract(off)
	return a * b;
}
__device__ __forceinline__ float faddRn(float a, float b)
{
#pragma clang fp contract(off)
	return a + b;
}
__device__ __forceinline__ float fsubRn(float a, float b)
{
#pragma clang fp contract(off)
	return a - b;
}
__device__ __forceinline__ float fdivRn(float a, float b)
{
#pragma clang fp contract(off)
	return a / b;
}
__device__ __forceinline__ float fsqrtRn(float a)
{
	return sqrtf(a);
}

constexpr int kMaxRays = 16;

enum Mode : int
{
	MODE_FWD = 0,
	MODE_BACK = 1,
	MODE_EM = 2,
	MODE_SENS = 3,
	MODE_COUNT = 4
};

// Image geometry (ImageParams).
struct DImage
{
	int32_t nx, ny, nz, nt;
	float vx, vy, vz, lx, ly, lz, ox, oy, oz, fovR;
	int32_t nxy;
	int64_t nvox;
};

// TimeOfFlightHelper constants (TimeOfFlight.hpp:19-34): sigma, truncation
// width (<= 0: none) and Gaussian norm, all mm; rsigma = RN(1/sigma), with
// which the distance-driven slab weight divides by sigma exactly
// (divBySigma).
struct DTof
{
	float sigma, trunc, norm, rsigma;
};

// The float nearest 1/s (host side): of the double quotient rounded to float
// and its two neighbours, the one with the smallest |c*s - 1|, a residual
// that double evaluates exactly (a product of two floats has 48 bits).
inline float rcpNearest(float s)
{
	const float c = (float)(1.0 / (double)s);
	float best = c;
	double be = std::fabs((double)c * (double)s - 1.0);
	const float cand[2] = {std::nextafter(c, 0.f), std::nextafter(c, 2.f * c)};
	for (float k : cand)
	{
		const double e = std::fabs((double)k * (double)s - 1.0);
		if (e < be)
		{
			be = e;
			best = k;
		}
	}
	return best;
}

// Prefix sums of the trapezoids of each row of `taps` (numS rows of K):
// out[r * (K + 2) + m + 1] = sum over xi in [-1, m) of kSpacing * (tap(xi) +
// tap(xi + 1)) / 2 in FP64, taps outside [0, K) being 0, out[r * (K + 2)] =
// 0.  The integral of a row's linear interpolant over whole kernel segments
// [a, b) is then out[b + 1] - out[a + 1] (DPsf::cum).
inline void psfPrefixSums(const float* taps, int64_t numS, int64_t K, double kSpacing,
                          double* out)
{
	const int64_t KC = K + 2;
	for (int64_t i = 0; i < numS; i++)
	{
		const float* row = taps + i * K;
		double* c = out + i * KC;
		c[0] = 0.0;
		for (int64_t m = 0; m <= K; m++)
		{
			const double a = (m - 1 >= 0) ? (double)row[m - 1] : 0.0;
			const double b = (m < K) ? (double)row[m] : 0.0;
			c[m + 1] = c[m] + kSpacing * 0.5 * (a + b);
		}
	}
}

// The same constants in FP64 (the FP64 projector).
struct DTof64
{
	double sigma, trunc, norm;
};

// Projection-space PSF tables (ProjectionPsfManager.cpp:27-103).
struct DPsf
{
	// numS rows of K + 4 floats: the row's taps with two zeros on either
	// side (entry i + 2 = tap i), so that psfWeight reads the taps of any
	// clamped index without a test
	const float* k;
	const float* kflip;  // the same of the reversed taps (backprojection)
	// numS x (K + 2) prefix sums of the kernel's trapezoids (FP64): row r,
	// entry m + 1 = sum over xi in [-1, m) of kSpacing * (tap(xi) +
	// tap(xi + 1)) / 2, taps outside [0, K) being 0; cumFlip for kflip
	const double* cum;
	const double* cumFlip;
	int32_t numS, K;
	float sStep, kSpacing, halfWidth;
	float rks;  // RN(1 / kSpacing)
	// bytes of one direction's tables (the prefix sums, then the padded
	// taps) when the DD kernels stage them in LDS, else 0 (too large: they
	// read them from global memory)
	int32_t ldsBytes;
};

// Row stride of DPsf::k and DPsf::kflip.
__host__ __device__ constexpr int psfRowTaps(int K) { return K + 4; }
// LDS tables of at most this many bytes are staged.
constexpr int kPsfLdsMax = 16384;

struct DProj
{
	int32_t type;
	int32_t numRays;
	int32_t useTof;
	int32_t setup64;  // FP64 ray-plane setup instead of the reference FP32 order
	DTof tof;
	DTof64 tof64;
	float tz, tt;  // crystal size z / transaxial
	float rayOff[4 * (kMaxRays - 1)];
	DPsf psf;
};

// Device-resident scanner tables.
struct DScanner
{
	const float* lut;  // SoA: x[n], y[n], z[n], ox[n], oy[n], oz[n]
	uint32_t ndet;
	// Histogram3D geometry
	uint32_t numR, numPhi, numZBin, rCut, numDOI, numDOIPoss, dpr, nrings;
	const uint32_t* zbin;  // z1 | (z2 << 16) per z bin
	const float* motion;   // k x 12 (transform_t)
	int32_t nmotion;
};

struct DLors
{
	int32_t kind;  // 0 pairs, 1 bin list, 2 histogram subset iterator
	int64_t n;
	const uint32_t* d1;
	const uint32_t* d2;
	const uint32_t* bins;
	uint32_t S, s, nps;  // subset iterator: S subsets, subset s, phis in it
	const float* tof;
	const int32_t* frame;
	const uint8_t* motion;
};

struct DCorr
{
	const float* meas;
	const float* sens;
	const float* att;
	const float* randoms;
	const float* scatter;
	float g;
	float denomThr;
	const float* precorr;  // in-vivo attenuation precorrection (nullable)
};

struct KArgs
{
	DImage im;
	DProj pr;
	DScanner sc;
	DLors L;
	DCorr c;
	const float* x;   // image read by forward projection
	float* y;         // image accumulated by backprojection
	const float* vals;  // MODE_BACK values / MODE_SENS per-LOR factor
	float* out;       // MODE_FWD output
	DImage attIm;     // MODE_SENS attenuation image geometry
	const float* att;  // MODE_SENS attenuation image (nullable)
	unsigned long long* counter;  // MODE_COUNT
	// Backprojection target: when `part` is set, each XCD accumulates into
	// its own copy part + xcc * partStride (folded into `y` afterwards).
	float* part;
	int64_t partStride;
	int32_t nxcc;
	// Ray table of a tile plan for these LORs (two float4 per LOR and ray:
	// end points after offsets, TOF value, frame), or null: the forward of
	// the EM ratio then reads it instead of resolving each LOR.
	const float4* rays;
	// The Siddon EM ratio's forward reads the bricked copy of x (xb, bvox
	// floats per frame, brickIndex), made by launchToBricks.
	const float* xb;
	int64_t bvox;
};

// Hardware id of the XCD (accelerator complex die) running this wave.
__device__ __forceinline__ uint32_t xccId()
{
	uint32_t x;
	asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
	return x & 0xfu;
}

struct Lor
{
	float3 p1, p2, n1, n2;
	float tof;
	int32_t frame;
};

// ---------------------------------------------------------------------------
// Histogram3D closed-form bin -> detector pair
// (Histogram3D::getDetPairInSameRing, Histogram3D.cpp:428-450, and
//  getDetPairFromCoords 299-363 with the z loop tabulated).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void histoPair(const DScanner& S, uint32_t r,
                                          uint32_t phi, uint32_t zb,
                                          uint32_t& d1, uint32_t& d2)
{
	const int32_t ntot = (int32_t)S.dpr;
	const int32_t rr = (int32_t)(r / S.numDOIPoss);
	const int32_t d02 = ntot / 2 + (int32_t)(phi & 1u);
	const int32_t sh = rr - ntot / 4 + (int32_t)S.rCut;
	int32_t dr1 = sh;
	int32_t dr2 = d02 - sh;
	if (dr1 < 0)
		dr1 += ntot;
	if (dr2 < 0)
		dr2 += ntot;
	int32_t a = (dr1 + (int32_t)(phi >> 1)) % ntot;
	int32_t b = (dr2 + (int32_t)(phi >> 1)) % ntot;
	if (a > b)
	{
		const int32_t t = a;
		a = b;
		b = t;
	}
	if (a == b)
	{
		d1 = d2 = 0;
		return;
	}
	const uint32_t doiCase = r % S.numDOIPoss;
	const uint32_t doi1 = doiCase % S.numDOI;
	const uint32_t doi2 = doiCase / S.numDOI;
	const uint32_t z = S.zbin[zb];
	const uint32_t z1 = z & 0xffffu, z2 = z >> 16;
	d1 = (uint32_t)a + z1 * S.dpr + doi1 * (S.dpr * S.nrings);
	d2 = (uint32_t)b + z2 * S.dpr + doi2 * (S.dpr * S.nrings);
}

// Returns false for an event that cannot be resolved (detector id outside
// the LUT, bin outside the histogram); such events are skipped.
__device__ __forceinline__ bool eventPair(const DScanner& S, const DLors& L,
                                          int64_t e, uint32_t& d1,
                                          uint32_t& d2)
{
	if (L.kind == 0)
	{
		d1 = L.d1[e];
		d2 = L.d2[e];
		return d1 < S.ndet && d2 < S.ndet;
	}
	uint32_t r, phi, zb;
	const uint32_t npr = S.numPhi * S.numR;
	if (L.kind == 1)
	{
		const uint32_t bin = L.bins[e];
		zb = bin / npr;
		if (zb >= S.numZBin)
		{
			d1 = d2 = 0;
			return false;
		}
		const uint32_t rem = bin - zb * npr;
		phi = rem / S.numR;
		r = rem - phi * S.numR;
	}
	else
	{
		// BinIteratorRangeHistogram3D::getSafe (BinIterator.cpp:191-198)
		const uint32_t idx = (uint32_t)e;
		const uint32_t per = L.nps * S.numR;
		zb = idx / per;
		const uint32_t rem = idx - zb * per;
		const uint32_t ps = rem / S.numR;
		r = rem - ps * S.numR;
		phi = L.S * ps + L.s;
	}
	histoPair(S, r, phi, zb, d1, d2);
	return true;
}

__device__ __forceinline__ float3 lutPos(const DScanner& S, uint32_t d)
{
	return make_float3(S.lut[d], S.lut[S.ndet + d], S.lut[2 * S.ndet + d]);
}
__device__ __forceinline__ float3 lutOri(const DScanner& S, uint32_t d)
{
	return make_float3(S.lut[3 * S.ndet + d], S.lut[4 * S.ndet + d],
	                   S.lut[5 * S.ndet + d]);
}

// Matrix * Vector3D + translation (Matrix.cpp:177-186,
// ProjectionData.cpp:307-318).
__device__ __forceinline__ float3 moveP(const float* m, float3 p)
{
	float3 o;
	o.x = faddRn(faddRn(faddRn(fmulRn(m[0], p.x), fmulRn(m[1], p.y)),
	                          fmulRn(m[2], p.z)),
	                m[3]);
	o.y = faddRn(faddRn(faddRn(fmulRn(m[4], p.x), fmulRn(m[5], p.y)),
	                          fmulRn(m[6], p.z)),
	                m[7]);
	o.z = faddRn(faddRn(faddRn(fmulRn(m[8], p.x), fmulRn(m[9], p.y)),
	                          fmulRn(m[10], p.z)),
	                m[11]);
	return o;
}

// ProjectionData::collectProjectionProperties (ProjectionData.cpp:198-273):
// endpoints (with motion), detector orientations, TOF, dynamic frame.
// Unresolvable events (bad detector id / bin / frame / motion index) come
// back with frame = -1 so that every mode skips them.
__device__ __forceinline__ Lor fetchLor(const DScanner& S, const DLors& L,
                                        int64_t e, int32_t nt)
{
	uint32_t d1, d2;
	bool ok = eventPair(S, L, e, d1, d2);
	if (!ok)
		d1 = d2 = 0;
	Lor l;
	l.p1 = lutPos(S, d1);
	l.p2 = lutPos(S, d2);
	l.n1 = lutOri(S, d1);
	l.n2 = lutOri(S, d2);
	if (L.motion != nullptr)
	{
		const int mi = (int)L.motion[e];
		if (mi < S.nmotion)
		{
			const float* m = S.motion + 12 * mi;
			l.p1 = moveP(m, l.p1);
			l.p2 = moveP(m, l.p2);
		}
		else
			ok = false;
	}
	l.tof = (L.tof != nullptr) ? L.tof[e] : 0.f;
	l.frame = (L.frame != nullptr) ? L.frame[e] : 0;
	if (!ok || l.frame >= nt)
		l.frame = -1;
	return l;
}

// ---------------------------------------------------------------------------
// multi-ray line (MultiRayGenerator::getRandomLine, isParallel = hasDepth =
// false, MultiRayGenerator.cpp:54-114)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float3 transDir(float3 n)
{
	// normalize(n x z) with Vector3D::normalize's 1e-24 guard
	const float cx = fsubRn(fmulRn(n.y, 1.f), fmulRn(n.z, 0.f));
	const float cy = fsubRn(fmulRn(n.z, 0.f), fmulRn(n.x, 1.f));
	const float cz = fsubRn(fmulRn(n.x, 0.f), fmulRn(n.y, 0.f));
	const float nn = fsqrtRn(
	    faddRn(faddRn(fmulRn(cx, cx), fmulRn(cy, cy)), fmulRn(cz, cz)));
	if ((double)nn > 10e-25)
		return make_float3(fdivRn(cx, nn), fdivRn(cy, nn), fdivRn(cz, nn));
	return make_float3(0.f, 0.f, 0.f);
}

__device__ __forceinline__ void rayLine(const DProj& pr, const Lor& l, int r,
                                        float3& q1, float3& q2)
{
	if (r == 0)
	{
		q1 = l.p1;
		q2 = l.p2;
		return;
	}
	const float* o = pr.rayOff + 4 * (r - 1);
	const float3 t1 = transDir(l.n1);
	const float3 t2 = transDir(l.n2);
	const float s1 = fmulRn(o[0], pr.tz), s2 = fmulRn(o[1], pr.tt);
	const float s3 = fmulRn(o[2], pr.tz), s4 = fmulRn(o[3], pr.tt);
	q1.x = faddRn(faddRn(l.p1.x, fmulRn(0.f, s1)), fmulRn(t1.x, s2));
	q1.y = faddRn(faddRn(l.p1.y, fmulRn(0.f, s1)), fmulRn(t1.y, s2));
	q1.z = faddRn(faddRn(l.p1.z, fmulRn(1.f, s1)), fmulRn(t1.z, s2));
	q2.x = fsubRn(faddRn(l.p2.x, fmulRn(0.f, s3)), fmulRn(t2.x, s4));
	q2.y = fsubRn(faddRn(l.p2.y, fmulRn(0.f, s3)), fmulRn(t2.y, s4));
	q2.z = fsubRn(faddRn(l.p2.z, fmulRn(1.f, s3)), fmulRn(t2.z, s4));
}

// ---------------------------------------------------------------------------
// TOF (TimeOfFlight.hpp:153-185).  Segment weight = Gaussian at the segment
// midpoint: with a = alpha, x_c = (a_mid*d - pc)/sigma, pc = d/2 + tof_mm.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tofMm(float tof_ps)
{
	return (float)((double)tof_ps * (double)0.299792458f * 0.5);
}

__device__ __forceinline__ float gaussW(float xc, float norm)
{
	return __expf(-0.5f * xc * xc) * norm;
}

// ---------------------------------------------------------------------------
// Projection-PSF integral (ProjectorUtils.hpp:43-98 via
// ProjectionPsfManager::getWeight, ProjectionPsfManager.cpp:93-103).
// ---------------------------------------------------------------------------
// x / y from ry = RN(1/y): q = RN(x*ry) corrected by one FMA residual step
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990): RN(x/y) for finite x, y away
// from the overflow and underflow thresholds.
__device__ __forceinline__ float divByRecip(float x, float y, float ry)
{
#pragma clang fp contract(off)
	const float q = x * ry;
	return fmaf(fmaf(-q, y, x), ry, q);
}

// The integral of the taps' linear interpolant over [x0, x1]: the partial
// trapezoids at both ends as the reference computes them, the whole
// trapezoids between them from the row's FP64 prefix table `cum` (the
// reference adds them one by one in FP32; the two sums differ by rounding).
// kp: the row's padded taps (kp[i + 2] = tap i, zero for i outside [0, K)).
// Every index is clamped into the padded row, so the six table reads have
// no tests and are issued together: a tap index below -2 or above K reads
// zeros, as the unpadded taps' bounds tests give, and the prefix entries
// read for an empty middle range are not used.
__device__ __forceinline__ float psfWeight(const DPsf& P, const float* kp, const double* cum,
                                           float x0, float x1)
{
#pragma clang fp contract(off)
	const int hw = (P.K + 1) / 2;
	const float ks = P.kSpacing;
	// (K odd: hw = half + 1, the reference's second range test is the same)
	if (x0 > hw * ks || x1 < -hw * ks || x0 >= x1)
		return 0.f;
	const int K = P.K;
	const int half = (K - 1) / 2;
	const float x0s = divByRecip(x0, ks, P.rks);
	const int x0i = (int)floorf(x0s + (float)half);
	const float x1s = divByRecip(x1, ks, P.rks);
	const int x1i = (int)floorf(x1s + (float)half);
	const int j0 = min(max(x0i, -2), K) + 2, j1 = min(max(x1i, -2), K) + 2;
	const int ia = max(x0i + 1, -1), ib = min(x1i, K);
	const float v0lo = kp[j0], v0hi = kp[j0 + 1];
	const float v1lo = kp[j1], v1hi = kp[j1 + 1];
	const double ca = cum[min(ia, K) + 1], cb = cum[max(ib, -1) + 1];
	const float tau0 = x0s + (float)(-(-half + x0i));
	const float v0 = v0lo * (1.f - tau0) + v0hi * tau0;
	const float tau1 = x1s - (float)(-half + x1i);
	const float v1 = v1lo * (1.f - tau1) + v1hi * tau1;
	float out = 0.f;
	if (x0i == x1i)
	{
		out = 0.5f * (v0 + v1) * (x1 - x0);
	}
	else
	{
		out += 0.5f * (v0 + v0hi) * ks * (1.f - tau0);
		out += 0.5f * (v1lo + v1) * ks * tau1;
		// whole trapezoids xi in [x0i + 1, x1i); those below xi = -1 and
		// from xi = K on are zero (x0 may lie far below the kernel)
		if (ib > ia)
			out += (float)(cb - ca);
	}
	return out;
}

}  // namespace yrtmi
