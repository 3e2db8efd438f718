// Siddon and distance-driven ray kernels (device functions).
//
// Two setup precisions, chosen per projector (template S64):
//  - S64 = false (default): the ray-plane setup in FP32 with the reference's
//    operation order and no FMA contraction, so that projections agree with
//    the reference CPU path to rounding level even on rays that graze voxel
//    planes, where FP32 setup itself is inaccurate.
//  - S64 = true: the setup (FOV-cylinder clip, slab alphas, TOF window, first
//    crossings, first voxel) in FP64; it matches the FP64 oracle on such
//    grazing rays.
// Traversal, weights and accumulation are FP32 in both.
#pragma once

#include "yrtmi_device.hpp"

namespace yrtmi
{

// x / sigma, correctly rounded, from the correctly rounded reciprocal
// rs = RN(1/sigma): q = RN(x*rs) is within an ulp and one FMA residual step
// rounds it to RN(x/sigma) (P. Markstein, IBM J. Res. Dev. 34(1), 1990;
// finite x away from the overflow and underflow thresholds).
__device__ __forceinline__ float divBySigma(float x, float sigma, float rs)
{
#pragma clang fp contract(off)
	const float q = x * rs;
	return fmaf(fmaf(-q, sigma, x), rs, q);
}

// util::get_alpha (ProjectorUtils.hpp:15-41): slab [r0, r1] entry / exit
// alphas; a ray flat in this axis and outside the slab gives the empty range
// amin = 1, amax = 0 (callers decide what an empty range means).
template <typename T>
__device__ __forceinline__ void getAlpha(T r0, T r1, T p1, T p2, T inv, T& amin,
                                         T& amax)
{
#pragma clang fp contract(off)
	amin = T(0);
	amax = T(1);
	if (p1 != p2)
	{
		const T a0 = (r0 - p1) * inv;
		const T a1 = (r1 - p1) * inv;
		if (a0 < a1)
		{
			amin = a0;
			amax = a1;
		}
		else
		{
			amin = a1;
			amax = a0;
		}
	}
	else if (p1 < r0 || p1 > r1)
	{
		amax = T(0);
		amin = T(1);
	}
}

// sqrt(log2(e) / 2): exp(-x^2 / 2) = 2^(-(x kTofC)^2)
constexpr double kTofC = 0.84932180028801904272;

// Siddon ray after its setup: the clipped alpha range, the next plane
// crossing of each axis and what the walk needs to advance them
// (ProjectorSiddon::project_helper<IS_FWD, FLAG_INCR=true, FLAG_TOF>,
// ProjectorSiddon.cpp:478-767).
struct SidRay
{
	float3 p1, p2;
	float invx, invy, invz;
	float a_cur, amax;
	float axn, ayn, azn;  // alpha of the next plane per axis (FMAX: flat)
	float xc, yc, zc;     // coordinate of that plane
	float dnorm;
	// A segment [a0, a1] weighs (a1 - a0) * g * ws: ws = d (no TOF) or
	// d * norm (TOF), a constant of the ray that the walk's callers apply once
	// per ray; g = 1 (no TOF) or the TOF Gaussian exp(-x_c^2 / 2) at the
	// segment's midpoint, x_c = (a_mid d - pc) / sigma, pc = d/2 + tof_mm.
	// The walk forms u = (a0 + a1) tA - tB = x_c sqrt(log2(e) / 2) with one
	// FMA (tA = kTofC d / (2 sigma), tB = kTofC pc / sigma, rounded once
	// from FP64) and g = 2^(-u^2): the reference's value to a few FP32 ulps
	// (it forms x_c in FP32).
	float tA, tB, ws;
	int dirx, diry, dirz;
	// index of the first plane per axis, counted from the image face the
	// ray enters through (x_start + k * dir * v)
	int kx, ky, kz;
	// first voxel (S64 only)
	double x1d, y1d, z1d, ddx, ddy, ddz;
};

// Ray-plane setup: FOV-cylinder clip, image slabs, TOF window, first plane
// crossing of each axis.  Returns false when the ray misses the image.
template <bool TOF, bool S64>
__device__ __forceinline__ bool sidSetup(const DImage& P, float3 p1, float3 p2,
                                         const DTof& tof, float tofv, SidRay& r)
{
#pragma clang fp contract(off)
	const float FMAX = 3.402823466e+38f;
	const bool flat_x = (p1.x == p2.x);
	const bool flat_y = (p1.y == p2.y);
	const bool flat_z = (p1.z == p2.z);
	r.p1 = p1;
	r.p2 = p2;
	r.axn = FMAX;
	r.ayn = FMAX;
	r.azn = FMAX;
	r.tA = 0.f;
	r.tB = 0.f;
	r.ws = 0.f;
	r.kx = r.ky = r.kz = 0;
	r.x1d = r.y1d = r.z1d = r.ddx = r.ddy = r.ddz = 0.0;
	if constexpr (S64)
	{
		const double x1d = p1.x, y1d = p1.y, z1d = p1.z;
		const double ddx = (double)p2.x - x1d;
		const double ddy = (double)p2.y - y1d;
		const double ddz = (double)p2.z - z1d;
		r.x1d = x1d;
		r.y1d = y1d;
		r.z1d = z1d;
		r.ddx = ddx;
		r.ddy = ddy;
		r.ddz = ddz;
		double amin = 0.0, amax = 1.0;
		const double A = ddx * ddx + ddy * ddy;
		if (A != 0.0)
		{
			const double R = P.fovR;
			const double B = 2.0 * (ddx * x1d + ddy * y1d);
			const double Cq = x1d * x1d + y1d * y1d - R * R;
			const double Delta = B * B - 4.0 * A * Cq;
			if (Delta <= 0.0)
				return false;
			const double sq = sqrt(Delta);
			amin = fmax(amin, (-B - sq) / (2.0 * A));
			amax = fmin(amax, (-B + sq) / (2.0 * A));
		}
		const double dnormd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
		const double invx = flat_x ? 0.0 : 1.0 / ddx;
		const double invy = flat_y ? 0.0 : 1.0 / ddy;
		const double invz = flat_z ? 0.0 : 1.0 / ddz;
		r.dirx = (invx >= 0.0) ? 1 : -1;
		r.diry = (invy >= 0.0) ? 1 : -1;
		r.dirz = (invz >= 0.0) ? 1 : -1;
		const double hx = 0.5 * (double)P.lx, hy = 0.5 * (double)P.ly,
		             hz = 0.5 * (double)P.lz;
		// util::get_alpha (ProjectorUtils.hpp:15-41)
		if (!flat_x)
		{
			const double a0 = (-hx - x1d) * invx, a1 = (hx - x1d) * invx;
			amin = fmax(amin, fmin(a0, a1));
			amax = fmin(amax, fmax(a0, a1));
		}
		else if (x1d < -hx || x1d > hx)
			return false;
		if (!flat_y)
		{
			const double a0 = (-hy - y1d) * invy, a1 = (hy - y1d) * invy;
			amin = fmax(amin, fmin(a0, a1));
			amax = fmin(amax, fmax(a0, a1));
		}
		else if (y1d < -hy || y1d > hy)
			return false;
		if (!flat_z)
		{
			const double a0 = (-hz - z1d) * invz, a1 = (hz - z1d) * invz;
			amin = fmax(amin, fmin(a0, a1));
			amax = fmin(amax, fmax(a0, a1));
		}
		else if (z1d < -hz || z1d > hz)
			return false;
		if (TOF)
		{
			const double tmm = (double)tofMm(tofv);
			if (tof.trunc > 0.f)
			{
				amin = fmax(amin, 0.5 + (tmm - (double)tof.trunc) / dnormd);
				amax = fmin(amax, 0.5 + (tmm + (double)tof.trunc) / dnormd);
			}
			r.tB = (float)((0.5 * dnormd + tmm) / (double)tof.sigma * kTofC);
			r.tA = (float)(0.5 * dnormd / (double)tof.sigma * kTofC);
		}
		if ((invx >= 0.0 && x1d > hx) || (invx < 0.0 && x1d < -hx) ||
		    (invy >= 0.0 && y1d > hy) || (invy < 0.0 && y1d < -hy) ||
		    (invz >= 0.0 && z1d > hz) || (invz < 0.0 && z1d < -hz))
			return false;
		if (!(amin < amax))
			return false;
		double xc = (invx > 0.0) ? -hx : hx;
		double yc = (invy > 0.0) ? -hy : hy;
		double zc = (invz > 0.0) ? -hz : hz;
		if (!flat_x)
		{
			const double k = ceil((double)r.dirx * (amin * ddx - xc + x1d) / (double)P.vx);
			xc += k * r.dirx * (double)P.vx;
			r.axn = (float)((xc - x1d) * invx);
			r.kx = (int)k;
		}
		if (!flat_y)
		{
			const double k = ceil((double)r.diry * (amin * ddy - yc + y1d) / (double)P.vy);
			yc += k * r.diry * (double)P.vy;
			r.ayn = (float)((yc - y1d) * invy);
			r.ky = (int)k;
		}
		if (!flat_z)
		{
			const double k = ceil((double)r.dirz * (amin * ddz - zc + z1d) / (double)P.vz);
			zc += k * r.dirz * (double)P.vz;
			r.azn = (float)((zc - z1d) * invz);
			r.kz = (int)k;
		}
		r.xc = (float)xc;
		r.yc = (float)yc;
		r.zc = (float)zc;
		r.invx = (float)invx;
		r.invy = (float)invy;
		r.invz = (float)invz;
		r.dnorm = (float)dnormd;
		r.ws = TOF ? (float)(dnormd * (double)tof.norm) : r.dnorm;
		r.amax = (float)amax;
		r.a_cur = (float)amin;
	}
	else
	{
		// reference FP32 order (ProjectorSiddon.cpp:499-594)
		const float A = (p2.x - p1.x) * (p2.x - p1.x) + (p2.y - p1.y) * (p2.y - p1.y);
		const float B = 2.0f * ((p2.x - p1.x) * p1.x + (p2.y - p1.y) * p1.y);
		const float Cq = p1.x * p1.x + p1.y * p1.y - P.fovR * P.fovR;
		const float Delta = B * B - 4.0f * A * Cq;
		float t0 = 0.f, t1 = 1.f;
		if (A != 0.f)
		{
			if (Delta <= 0.f)
				return false;
			t0 = (-B - sqrtf(Delta)) / (2.0f * A);
			t1 = (-B + sqrtf(Delta)) / (2.0f * A);
		}
		const float ex = p1.x - p2.x, ey = p1.y - p2.y, ez = p1.z - p2.z;
		r.dnorm = sqrtf(ex * ex + ey * ey + ez * ez);
		r.invx = flat_x ? 0.0f : 1.0f / (p2.x - p1.x);
		r.invy = flat_y ? 0.0f : 1.0f / (p2.y - p1.y);
		r.invz = flat_z ? 0.0f : 1.0f / (p2.z - p1.z);
		r.dirx = (r.invx >= 0.0f) ? 1 : -1;
		r.diry = (r.invy >= 0.0f) ? 1 : -1;
		r.dirz = (r.invz >= 0.0f) ? 1 : -1;
		const float x0 = -P.lx * 0.5f, x1 = P.lx * 0.5f;
		const float y0 = -P.ly * 0.5f, y1 = P.ly * 0.5f;
		const float z0 = -P.lz * 0.5f, z1 = P.lz * 0.5f;
		float amin = fmaxf(0.0f, t0), amax = fminf(1.0f, t1);
		if (!flat_x)
		{
			const float a0 = (-0.5f * P.lx - p1.x) * r.invx, a1 = (0.5f * P.lx - p1.x) * r.invx;
			amin = fmaxf(amin, (a0 < a1) ? a0 : a1);
			amax = fminf(amax, (a0 < a1) ? a1 : a0);
		}
		else if (p1.x < -0.5f * P.lx || p1.x > 0.5f * P.lx)
			return false;
		if (!flat_y)
		{
			const float a0 = (-0.5f * P.ly - p1.y) * r.invy, a1 = (0.5f * P.ly - p1.y) * r.invy;
			amin = fmaxf(amin, (a0 < a1) ? a0 : a1);
			amax = fminf(amax, (a0 < a1) ? a1 : a0);
		}
		else if (p1.y < -0.5f * P.ly || p1.y > 0.5f * P.ly)
			return false;
		if (!flat_z)
		{
			const float a0 = (-0.5f * P.lz - p1.z) * r.invz, a1 = (0.5f * P.lz - p1.z) * r.invz;
			amin = fmaxf(amin, (a0 < a1) ? a0 : a1);
			amax = fminf(amax, (a0 < a1) ? a1 : a0);
		}
		else if (p1.z < -0.5f * P.lz || p1.z > 0.5f * P.lz)
			return false;
		if (TOF)
		{
			// TimeOfFlightHelper::getAlphaRange (TimeOfFlight.hpp:153-173)
			const float tmm = tofMm(tofv);
			if (tof.trunc > 0.f)
			{
				const float lo = (float)fmax(0.0, 0.5 + (double)((tmm - tof.trunc) / r.dnorm));
				const float hi = (float)fmin(1.0, 0.5 + (double)((tmm + tof.trunc) / r.dnorm));
				amin = fmaxf(amin, lo);
				amax = fminf(amax, hi);
			}
			// the Gaussian's constants from FP64 (the S64 form)
			r.tB = (float)((0.5 * (double)r.dnorm + (double)tmm) / (double)tof.sigma * kTofC);
			r.tA = (float)(0.5 * (double)r.dnorm / (double)tof.sigma * kTofC);
		}
		r.ws = TOF ? (float)((double)r.dnorm * (double)tof.norm) : r.dnorm;
		r.xc = (r.invx > 0.0f) ? x0 : x1;
		r.yc = (r.invy > 0.0f) ? y0 : y1;
		r.zc = (r.invz > 0.0f) ? z0 : z1;
		if ((r.invx >= 0.0f && p1.x > x1) || (r.invx < 0.0f && p1.x < x0) ||
		    (r.invy >= 0.0f && p1.y > y1) || (r.invy < 0.0f && p1.y < y0) ||
		    (r.invz >= 0.0f && p1.z > z1) || (r.invz < 0.0f && p1.z < z0))
			return false;
		if (!(amin < amax))
			return false;
		r.a_cur = amin;
		r.amax = amax;
		if (!flat_x)
		{
			const float kx = ceilf((float)r.dirx * (amin * (p2.x - p1.x) - r.xc + p1.x) / P.vx);
			r.xc += (float)((int)kx * r.dirx) * P.vx;
			r.axn = (r.xc - p1.x) * r.invx;
			r.kx = (int)kx;
		}
		if (!flat_y)
		{
			const float ky = ceilf((float)r.diry * (amin * (p2.y - p1.y) - r.yc + p1.y) / P.vy);
			r.yc += (float)((int)ky * r.diry) * P.vy;
			r.ayn = (r.yc - p1.y) * r.invy;
			r.ky = (int)ky;
		}
		if (!flat_z)
		{
			const float kz = ceilf((float)r.dirz * (amin * (p2.z - p1.z) - r.zc + p1.z) / P.vz);
			r.zc += (float)((int)kz * r.dirz) * P.vz;
			r.azn = (r.zc - p1.z) * r.invz;
			r.kz = (int)kz;
		}
	}
	return true;
}

// The part of an FP32 set-up (sidSetup<TOF, false>) that a walk re-entered
// at a tile record needs beyond the ray's end points, in two float4 (a
// plan's ray table): the per-axis reciprocals, the exit alpha, the length
// and the TOF constants.  sidLoad rebuilds the SidRay from the end points
// and them: every field sidRestart and sidWalk read equals the set-up's
// (the planes are recomputed by sidRestart, with the set-up's own
// expression in the FP32 set-up, so kx/ky/kz are set to never match and
// every non-flat axis is recomputed).
__device__ __forceinline__ void sidStore(float4* e, const SidRay& r)
{
	e[0] = make_float4(r.invx, r.invy, r.invz, r.amax);
	e[1] = make_float4(r.dnorm, r.tA, r.tB, r.ws);
}

__device__ __forceinline__ void sidLoad(float3 p1, float3 p2, const float4& s0,
                                        const float4& s1, SidRay& r)
{
	const float FMAX = 3.402823466e+38f;
	r.p1 = p1;
	r.p2 = p2;
	r.invx = s0.x;
	r.invy = s0.y;
	r.invz = s0.z;
	r.amax = s0.w;
	r.dnorm = s1.x;
	r.tA = s1.y;
	r.tB = s1.z;
	r.ws = s1.w;
	r.dirx = (r.invx >= 0.0f) ? 1 : -1;
	r.diry = (r.invy >= 0.0f) ? 1 : -1;
	r.dirz = (r.invz >= 0.0f) ? 1 : -1;
	r.axn = FMAX;
	r.ayn = FMAX;
	r.azn = FMAX;
	r.xc = r.yc = r.zc = 0.f;
	r.kx = r.ky = r.kz = -0x40000000;
	r.a_cur = 0.f;
	r.x1d = r.y1d = r.z1d = r.ddx = r.ddy = r.ddz = 0.0;
}

// The voxel holding the midpoint of segment [a0, a1] (the walk's rule for
// its first voxel, ProjectorSiddon.cpp:669-683).
template <bool S64>
__device__ __forceinline__ void sidMidVoxel(const DImage& P, const SidRay& r, float a0,
                                            float a1, int& vx, int& vy, int& vz)
{
#pragma clang fp contract(off)
	if constexpr (S64)
	{
		const double am = 0.5 * ((double)a0 + (double)a1);
		vx = (int)((r.x1d + am * r.ddx + 0.5 * (double)P.lx) / (double)P.vx);
		vy = (int)((r.y1d + am * r.ddy + 0.5 * (double)P.ly) / (double)P.vy);
		vz = (int)((r.z1d + am * r.ddz + 0.5 * (double)P.lz) / (double)P.vz);
	}
	else
	{
		const float am = 0.5f * (a0 + a1);
		vx = (int)((r.p1.x + am * (r.p2.x - r.p1.x) + P.lx / 2.0f) * (1.0f / P.vx));
		vy = (int)((r.p1.y + am * (r.p2.y - r.p1.y) + P.ly / 2.0f) * (1.0f / P.vy));
		vz = (int)((r.p1.z + am * (r.p2.z - r.p1.z) + P.lz / 2.0f) * (1.0f / P.vz));
	}
}

// Plane-index shift of a ray: the walk numbers planes from its first plane
// and moves its voxel index by one per plane crossed, so along the whole ray
// (walk's next plane index) - (index of the voxel's far face) is a constant
// per axis.  It is 0 except where the first voxel, taken at the midpoint of
// a first segment an ulp long at a plane corner, lands one voxel over; the
// walk then carries that offset to the end of the ray.  Computed by running
// the walk's selection to its first nonempty segment.  Returns false when
// the ray has no segment.
template <bool S64>
__device__ __forceinline__ bool sidPlaneShift(const DImage& P, const SidRay& r, int& dx,
                                              int& dy, int& dz)
{
#pragma clang fp contract(off)
	const float sdx = (float)r.dirx * P.vx, sdy = (float)r.diry * P.vy,
	            sdz = (float)r.dirz * P.vz;
	float axn = r.axn, ayn = r.ayn, azn = r.azn, xc = r.xc, yc = r.yc, zc = r.zc;
	float axp = axn, ayp = ayn, azp = azn;
	int sx = 0, sy = 0, sz = 0;  // planes consumed before the first segment
	const float a_cur = r.a_cur;
	while (a_cur < r.amax)
	{
		float a_next = -1.f;
		if (axp <= ayp && axp <= azp)
		{
			a_next = axn;
			xc += sdx;
			axn = (xc - r.p1.x) * r.invx;
			sx++;
		}
		if (ayp <= axp && ayp <= azp)
		{
			a_next = ayn;
			yc += sdy;
			ayn = (yc - r.p1.y) * r.invy;
			sy++;
		}
		if (azp <= axp && azp <= ayp)
		{
			a_next = azn;
			zc += sdz;
			azn = (zc - r.p1.z) * r.invz;
			sz++;
		}
		if (a_next > r.amax)
			a_next = r.amax;
		if (a_cur >= a_next)
		{
			axp = axn;
			ayp = ayn;
			azp = azn;
			continue;
		}
		// planes advanced in this (first nonempty) iteration belong to its
		// end, not to its voxel
		if (axp <= ayp && axp <= azp)
			sx--;
		if (ayp <= axp && ayp <= azp)
			sy--;
		if (azp <= axp && azp <= ayp)
			sz--;
		int vx, vy, vz;
		sidMidVoxel<S64>(P, r, a_cur, a_next, vx, vy, vz);
		// far face of voxel v: k = v + 1 (dir > 0) or n - v (dir < 0);
		// a flat axis has no planes
		dx = (r.p1.x == r.p2.x) ? 0 : r.kx + sx - ((r.dirx > 0) ? vx + 1 : P.nx - vx);
		dy = (r.p1.y == r.p2.y) ? 0 : r.ky + sy - ((r.diry > 0) ? vy + 1 : P.ny - vy);
		dz = (r.p1.z == r.p2.z) ? 0 : r.kz + sz - ((r.dirz > 0) ? vz + 1 : P.nz - vz);
		return true;
	}
	return false;
}

// Re-enters the walk of `r` (fresh from sidSetup) at alpha `a0` inside voxel
// (vx, vy, vz), whose segment starts there.  The next plane of each axis is
// that voxel's far face in the direction of travel plus the ray's plane
// shift (sidPlaneShift), counted from the entry face of the image as the
// walk counts it (x_start + k * dir * v).  On voxel sizes that are sums of
// few powers of two this reproduces the walk's accumulated plane positions
// exactly; otherwise to an ulp, and a plane recomputed an ulp before a0 is
// taken at a0: the walk's first step there is an empty segment, then the
// move across it, as where the planes coincide.
template <bool S64>
__device__ __forceinline__ void sidRestart(const DImage& P, SidRay& r, float a0,
                                           int vx, int vy, int vz, int dx, int dy,
                                           int dz)
{
#pragma clang fp contract(off)
	r.a_cur = a0;
	if (r.p1.x != r.p2.x)
	{
		const int k = ((r.dirx > 0) ? vx + 1 : P.nx - vx) + dx;
		float c;
		if constexpr (S64)
			c = (float)(((r.dirx > 0) ? -0.5 : 0.5) * (double)P.lx +
			            (double)(k * r.dirx) * (double)P.vx);
		else
			c = ((r.dirx > 0) ? -P.lx * 0.5f : P.lx * 0.5f) +
			    (float)(k * r.dirx) * P.vx;
		if (k != r.kx)  // else: the setup's first plane, kept as set up
		{
			r.xc = c;
			r.axn = fmaxf((c - r.p1.x) * r.invx, a0);
		}
	}
	if (r.p1.y != r.p2.y)
	{
		const int k = ((r.diry > 0) ? vy + 1 : P.ny - vy) + dy;
		float c;
		if constexpr (S64)
			c = (float)(((r.diry > 0) ? -0.5 : 0.5) * (double)P.ly +
			            (double)(k * r.diry) * (double)P.vy);
		else
			c = ((r.diry > 0) ? -P.ly * 0.5f : P.ly * 0.5f) +
			    (float)(k * r.diry) * P.vy;
		if (k != r.ky)
		{
			r.yc = c;
			r.ayn = fmaxf((c - r.p1.y) * r.invy, a0);
		}
	}
	if (r.p1.z != r.p2.z)
	{
		const int k = ((r.dirz > 0) ? vz + 1 : P.nz - vz) + dz;
		float c;
		if constexpr (S64)
			c = (float)(((r.dirz > 0) ? -0.5 : 0.5) * (double)P.lz +
			            (double)(k * r.dirz) * (double)P.vz);
		else
			c = ((r.dirz > 0) ? -P.lz * 0.5f : P.lz * 0.5f) +
			    (float)(k * r.dirz) * P.vz;
		if (k != r.kz)
		{
			r.zc = c;
			r.azn = fmaxf((c - r.p1.z) * r.invz, a0);
		}
	}
}

// Voxel box a walk stays in: [lo, hi) per axis, and the index
// base + vz*sz + vy*sy + vx*sx of a voxel in the caller's units, modulo 2^32
// (the image: voxels from its origin; an LDS tile: bytes from the tile's).
struct WalkBox
{
	int x0, x1, y0, y1, z0, z1;
	int sx, sy, sz, base;
};

__device__ __forceinline__ WalkBox imageBox(const DImage& P)
{
	return WalkBox{0, P.nx, 0, P.ny, 0, P.nz, 1, P.nx, P.nxy, 0};
}

// Bricked copy of an image frame for the forward's gathers: bricks of
// 4 x 4 x 2 voxels (128 bytes, one cache line), x-major within a brick and
// brick-x-major across bricks, the frame padded to whole bricks.  Rays a
// voxel or two apart (the rays of one LOR) then mostly read the same line
// in one load instruction, where rows of 32 voxels along x split them.
__host__ __device__ constexpr int brickNx(int nx) { return (nx + 3) >> 2; }
__host__ __device__ constexpr int brickNy(int ny) { return (ny + 3) >> 2; }
__host__ __device__ constexpr int brickNz(int nz) { return (nz + 1) >> 1; }
__host__ __device__ inline int64_t brickFrameVoxels(int nx, int ny, int nz)
{
	return (int64_t)brickNx(nx) * brickNy(ny) * brickNz(nz) * 32;
}
__device__ __forceinline__ uint32_t brickIndex(int nx, int ny, int vx, int vy, int vz)
{
	const uint32_t b = ((uint32_t)(vz >> 1) * (uint32_t)brickNy(ny) + (uint32_t)(vy >> 2)) *
	                       (uint32_t)brickNx(nx) +
	                   (uint32_t)(vx >> 2);
	return b * 32u + (uint32_t)(((vz & 1) << 4) | ((vy & 3) << 2) | (vx & 3));
}

// FP32 walk (ProjectorSiddon.cpp:622-766) from r.a_cur to r.amax.  With
// `first` the walk starts at the ray's entry and its first voxel is the one
// holding the midpoint of the first nonempty segment; otherwise it starts in
// (vx, vy, vz) (after sidRestart).  It ends when the voxel leaves `B`.
// fn(vx, vy, vz, idx, a_start, w, nidx) is called for every segment, w
// without the ray's constant factor r.ws, nidx the index of the walk's next
// voxel when that is inside `B` (else idx again, whose line is cached): a
// forward may fetch it one segment ahead.  onEnter(idx) is called with the
// first voxel's index before its segment.
//
// Each step takes the nearest next plane m = min(axn, ayn, azn); every axis
// whose plane is at m advances (the reference's three `<=` tests: ties
// advance together), and the segment ends at min(m, amax).  The steady
// loop is branch-free per axis (selects), so a wave runs one instruction
// stream whatever the axes its lanes cross.  The voxel is kept as its index
// and, per axis, the crossings left before it leaves the box (the walk ends
// when one goes negative); its coordinates are derived from those where fn
// uses them.
//
// BRICK: the index is the voxel's in the bricked frame (brickIndex; `B`
// must be the image): a crossing moves it by the in-brick stride, or across
// to the next brick when the voxel was the brick's last in the direction of
// travel, which the crossings left tell (per axis, (off - c) mod the brick
// size is 0 there, off = n along +, 0 along -).
template <bool TOF, bool S64, bool BRICK = false, class Fn, class Enter>
__device__ __forceinline__ void sidWalk(const DImage& P, const SidRay& r,
                                        const DTof& tof, const WalkBox& B,
                                        bool first, int vx, int vy, int vz,
                                        Fn&& fn, Enter&& onEnter)
{
#pragma clang fp contract(off)
	const float sdx = (float)r.dirx * P.vx, sdy = (float)r.diry * P.vy,
	            sdz = (float)r.dirz * P.vz;
	const float amaxf = r.amax;
	float axn = r.axn, ayn = r.ayn, azn = r.azn;
	float xc = r.xc, yc = r.yc, zc = r.zc;
	float a_cur = r.a_cur, a_next = 0.f;
	// planes crossed by the last step (ties advance together)
	bool mx = false, my = false, mz = false;
	// FP32 set-up: an axis's next crossing is (plane - origin) * factor of
	// its plane alone (the set-up computes it so, sidRestart too where it
	// does not clamp a plane to the restart's alpha, and a clamped plane's
	// axis moves in the first step, its crossing being the least), so every
	// axis's is recomputed each step, the planes that did not move giving
	// the same value.  A flat axis (next crossing FMAX, factor 0) takes the
	// origin -FMAX and the factor 1: plane + FMAX rounds to FMAX.
	const float FMAX = 3.402823466e+38f;
	const float oxr = (r.p1.x == r.p2.x) ? -FMAX : r.p1.x,
	            oyr = (r.p1.y == r.p2.y) ? -FMAX : r.p1.y,
	            ozr = (r.p1.z == r.p2.z) ? -FMAX : r.p1.z;
	const float fxr = (r.p1.x == r.p2.x) ? 1.f : r.invx, fyr = (r.p1.y == r.p2.y) ? 1.f : r.invy,
	            fzr = (r.p1.z == r.p2.z) ? 1.f : r.invz;
	auto step = [&]() {
		const float m = fminf(axn, fminf(ayn, azn));
		mx = axn == m;
		my = ayn == m;
		mz = azn == m;
		if constexpr (!S64)
		{
			xc = mx ? xc + sdx : xc;
			yc = my ? yc + sdy : yc;
			zc = mz ? zc + sdz : zc;
			axn = (xc - oxr) * fxr;
			ayn = (yc - oyr) * fyr;
			azn = (zc - ozr) * fzr;
		}
		else
		{
			const float xn = xc + sdx, yn = yc + sdy, zn = zc + sdz;
			xc = mx ? xn : xc;
			yc = my ? yn : yc;
			zc = mz ? zn : zc;
			axn = mx ? (xn - r.p1.x) * r.invx : axn;
			ayn = my ? (yn - r.p1.y) * r.invy : ayn;
			azn = mz ? (zn - r.p1.z) * r.invz : azn;
		}
		// min(m, amax) as an integer min of the bit patterns: equal to the
		// float min here (amax > 0, m no NaN: a negative m has the smaller
		// pattern as a signed integer, and non-negative floats order as
		// their patterns), and one instruction where fminf would first
		// quiet a possible signalling NaN
		a_next = __int_as_float(min(__float_as_int(m), __float_as_int(amaxf)));
	};
	// the voxel: crossings left per axis in its direction of travel, and
	// its index
	int cx = 0, cy = 0, cz = 0;
	uint32_t idx = 0u;
	const uint32_t xs = (uint32_t)(r.dirx * (BRICK ? 1 : B.sx)),
	               ys = (uint32_t)(r.diry * (BRICK ? 4 : B.sy)),
	               zs = (uint32_t)(r.dirz * (BRICK ? 16 : B.sz));
	// BRICK: the strides across a brick boundary and the offsets of the
	// boundary test
	const uint32_t bxs = (uint32_t)(r.dirx * 29),
	               bys = (uint32_t)(r.diry * (brickNx(P.nx) * 32 - 12)),
	               bzs = (uint32_t)(r.dirz * (brickNx(P.nx) * brickNy(P.ny) * 32 - 16));
	// the crossing counter's phase at a brick's last voxel: (off - c) mod
	// the brick size is 0 there, i.e. c mod the size equals off mod the size
	const int ox = ((r.dirx > 0) ? P.nx : 0) & 3, oy = ((r.diry > 0) ? P.ny : 0) & 3,
	          oz = ((r.dirz > 0) ? P.nz : 0) & 1;
	auto enter = [&]() {
		cx = (r.dirx > 0) ? B.x1 - 1 - vx : vx - B.x0;
		cy = (r.diry > 0) ? B.y1 - 1 - vy : vy - B.y0;
		cz = (r.dirz > 0) ? B.z1 - 1 - vz : vz - B.z0;
		if constexpr (BRICK)
			idx = brickIndex(P.nx, P.ny, vx, vy, vz);
		else
			idx = (uint32_t)B.base + (uint32_t)vz * (uint32_t)B.sz +
			      (uint32_t)vy * (uint32_t)B.sy + (uint32_t)vx * (uint32_t)B.sx;
	};
	auto emit = [&](int nx, int ny, int nz, uint32_t ni) {
		float w = a_next - a_cur;
		if (TOF)
		{
			// TimeOfFlightHelper::getWeight (TimeOfFlight.hpp:175-185): the
			// Gaussian at the segment's midpoint, u = (a0 + a1) tA - tB
			const float u = fmaf(a_cur + a_next, r.tA, -r.tB);
			w = w * __builtin_amdgcn_exp2f(-(u * u));
		}
		const int gx = (r.dirx > 0) ? B.x1 - 1 - cx : B.x0 + cx;
		const int gy = (r.diry > 0) ? B.y1 - 1 - cy : B.y0 + cy;
		const int gz = (r.dirz > 0) ? B.z1 - 1 - cz : B.z0 + cz;
		fn(gx, gy, gz, idx, a_cur, w, ((nx | ny | nz) >= 0) ? ni : idx);
		a_cur = a_next;
	};
	// the voxel across the planes of the last step: crossings left and
	// index after the move (computed beside the step's compares)
	auto moved = [&](int& nx, int& ny, int& nz, uint32_t& ni) {
		if constexpr (BRICK)
		{
			const uint32_t sx = ((cx & 3) == ox) ? bxs : xs;
			const uint32_t sy = ((cy & 3) == oy) ? bys : ys;
			const uint32_t sz = ((cz & 1) == oz) ? bzs : zs;
			ni = idx + (mx ? sx : 0u) + (my ? sy : 0u) + (mz ? sz : 0u);
		}
		else
			ni = idx + (mx ? xs : 0u) + (my ? ys : 0u) + (mz ? zs : 0u);
		nx = cx - (mx ? 1 : 0);
		ny = cy - (my ? 1 : 0);
		nz = cz - (mz ? 1 : 0);
	};
	if (first)
	{
		// planes at or before the ray's start are consumed without a voxel
		for (;;)
		{
			if (!(a_cur < amaxf))
				return;
			step();
			if (a_cur < a_next)
				break;
		}
		sidMidVoxel<S64>(P, r, a_cur, a_next, vx, vy, vz);
		if (vx < B.x0 || vx >= B.x1 || vy < B.y0 || vy >= B.y1 || vz < B.z0 || vz >= B.z1)
			return;
		enter();
		onEnter(idx);
		int nx, ny, nz;
		uint32_t ni;
		moved(nx, ny, nz, ni);
		emit(nx, ny, nz, ni);
		cx = nx;
		cy = ny;
		cz = nz;
		idx = ni;
	}
	else
	{
		enter();
		onEnter(idx);
	}
	// one exit test per segment; unrolled twice, so the loop-carried index
	// and alpha need no register copies
#pragma unroll 2
	while ((a_cur < amaxf) & ((cx | cy | cz) >= 0))
	{
		step();
		int nx, ny, nz;
		uint32_t ni;
		moved(nx, ny, nz, ni);
		emit(nx, ny, nz, ni);
		cx = nx;
		cy = ny;
		cz = nz;
		idx = ni;
	}
}

template <bool TOF, bool S64, class Fn>
__device__ __forceinline__ void sidWalk(const DImage& P, const SidRay& r,
                                        const DTof& tof, const WalkBox& B,
                                        bool first, int vx, int vy, int vz,
                                        Fn&& fn)
{
	sidWalk<TOF, S64, false>(P, r, tof, B, first, vx, vy, vz, fn, [](uint32_t) {});
}

// Siddon ray over the whole image.  `img` points at the frame's voxels.
// MODE_FWD returns the line integral; MODE_BACK adds val*w atomically;
// MODE_COUNT counts applied segments.
// BRICK (forward only): `img` is the frame's bricked copy (brickIndex).
template <int MODE, bool TOF, bool S64, bool BRICK = false>
__device__ __forceinline__ float siddonRay(const DImage& P,
                                           float* __restrict__ img, float3 p1,
                                           float3 p2, float val,
                                           const DTof& tof, float tofv,
                                           uint32_t& cnt)
{
	SidRay r;
	if (!sidSetup<TOF, S64>(P, p1, p2, tof, tofv, r))
		return 0.f;
	if constexpr (MODE == MODE_FWD)
	{
		// one FMA per segment in the walk's order, the voxel's value
		// fetched one segment ahead (its gather overlaps a whole step)
		float acc = 0.f, xv = 0.f;
		sidWalk<TOF, S64, BRICK>(
		    P, r, tof, imageBox(P), true, 0, 0, 0,
		    [&](int, int, int, uint32_t, float, float w, uint32_t nidx) {
			    acc = fmaf(w, xv, acc);
			    xv = img[nidx];
		    },
		    [&](uint32_t idx) { xv = img[idx]; });
		return acc * r.ws;
	}
	const float vs = val * r.ws;
	sidWalk<TOF, S64>(P, r, tof, imageBox(P), true, 0, 0, 0,
	                  [&](int, int, int, uint32_t idx, float, float w, uint32_t) {
		                  if (MODE == MODE_BACK)
			                  unsafeAtomicAdd(img + idx, vs * w);
		                  else
			                  cnt++;
	                  });
	return 0.f;
}

// Distance-driven ray (ProjectorDD::dd_project_ref, ProjectorDD.cpp:260-578),
// split into its per-LOR set-up (ddSetup) and the sweep over the slices of
// the major axis (ddVisit), so that the sweep can be restricted to a box of
// voxels (a tile) with the same arithmetic.  d1/d2: endpoints with the image
// offset subtracted.  Footprint arithmetic keeps the reference's FP32 order
// without contraction: the voxel range of the PSF-extended footprint is a
// rounding decision (rintf) whose outcome must not change with FMA fusion.
struct DDRay
{
	bool usePsf;
	int psfRow;         // row of the PSF table for this LOR
	const float* pk;    // PSF taps, padded (flipped for backprojection)
	const double* pcum; // the row's trapezoid prefix sums (flipped likewise)
	float ext, d_norm, d1_i, d2_i;
	// reciprocals of the slice-invariant denominators of ddSlice: the
	// footprint edges' (a_xy_lo, a_xy_hi, a_z_lo, a_z_hi) and the TOF slab's
	// (d2_i - d1_i); rOk: all five usable (finite, normal), else ddSlice
	// divides
	float r_xy_lo, r_xy_hi, r_z_lo, r_z_hi, r_i;
	bool rOk;
	int xy_i_0, xy_i_1, nyx, strideXY, strideYX;
	bool flag_y;  // major axis y (else x)
	float lxy, lyx, dxy, dyx;
	float d1_xy_lo, d1_xy_hi, d2_xy_lo, d2_xy_hi;
	float d1_yx_lo, d1_yx_hi, d2_yx_lo, d2_yx_hi;
	float d1zloI, d1zloZ, d1zhiI, d1zhiZ, d2zloI, d2zloZ, d2zhiI, d2zhiZ;
	float dxy_cos_theta, yx_off, inv_dyx, z_off, inv_dz, half_dyx, hvz, pc;
};

// The reciprocals of ddSlice's slice-invariant denominators.
__device__ __forceinline__ bool recipOk(float d, float r)
{
	return fabsf(d) >= 1.17549435e-38f && fabsf(d) < 3.0e38f && fabsf(r) >= 1.17549435e-38f &&
	       fabsf(r) < 3.0e38f;
}

__device__ __forceinline__ void ddRecips(DDRay& R)
{
#pragma clang fp contract(off)
	const float dxl = R.d2_xy_hi - R.d1_xy_lo, dxh = R.d2_xy_lo - R.d1_xy_hi;
	const float dzl = R.d2zloI - R.d1zloI, dzh = R.d2zhiI - R.d1zhiI;
	const float di = R.d2_i - R.d1_i;
	R.r_xy_lo = 1.0f / dxl;
	R.r_xy_hi = 1.0f / dxh;
	R.r_z_lo = 1.0f / dzl;
	R.r_z_hi = 1.0f / dzh;
	R.r_i = 1.0f / di;
	R.rOk = recipOk(dxl, R.r_xy_lo) && recipOk(dxh, R.r_xy_hi) && recipOk(dzl, R.r_z_lo) &&
	        recipOk(dzh, R.r_z_hi) && recipOk(di, R.r_i);
}

template <bool TOF, bool S64>
__device__ __forceinline__ void ddSetup(const DImage& P, float3 d1, float3 d2, float3 n1,
                                        float3 n2, const DProj& pr, float tofv,
                                        bool flipPsf, DDRay& R)
{
#pragma clang fp contract(off)
	const float d12x = d1.x - d2.x, d12y = d1.y - d2.y, d12z = d1.z - d2.z;
	const bool flag_y = fabsf(d12y) > fabsf(d12x);
	const float d_norm = sqrtf(d12x * d12x + d12y * d12y + d12z * d12z);
	const bool usePsf = pr.psf.k != nullptr;
	const float* pk = nullptr;
	const double* pcum = nullptr;
	float ext = 0.f;
	int psfRow = 0;
	if (usePsf)
	{
		// ProjectionPsfManager::getKernel (ProjectionPsfManager.cpp:63-91)
		float npx = d2.y - d1.y;
		float npy = d1.x - d2.x;
		const float nn = sqrtf(npx * npx + npy * npy);
		int s = 0;
		if (nn != 0.f)
		{
			npx /= nn;
			npy /= nn;
			const float sv = fabsf(d1.x * npx + d1.y * npy);
			s = (int)floorf(sv / pr.psf.sStep);
			if (s > pr.psf.numS - 1)
				s = pr.psf.numS - 1;
		}
		psfRow = s;
		pk = (flipPsf ? pr.psf.kflip : pr.psf.k) + s * psfRowTaps(pr.psf.K);
		pcum = (flipPsf ? pr.psf.cumFlip : pr.psf.cum) + s * (pr.psf.K + 2);
		ext = pr.psf.halfWidth;
	}
	const float dx = P.vx, dy = P.vy, dz = P.vz;
	int x_i_0, y_i_0, x_i_1, y_i_1;
	if constexpr (S64)
	{
		const double x1d = d1.x, y1d = d1.y, z1d = d1.z;
		const double ddx = (double)d2.x - x1d, ddy = (double)d2.y - y1d,
		             ddz = (double)d2.z - z1d;
		const double rx = 0.5 * ((double)P.lx - dx), ry = 0.5 * ((double)P.ly - dy),
		             rz = 0.5 * ((double)P.lz - dz);
		double axmin, axmax, aymin, aymax, azmin, azmax;
		getAlpha<double>(-rx, rx, x1d, d2.x, (d1.x == d2.x) ? 0.0 : 1.0 / ddx, axmin, axmax);
		getAlpha<double>(-ry, ry, y1d, d2.y, (d1.y == d2.y) ? 0.0 : 1.0 / ddy, aymin, aymax);
		getAlpha<double>(-rz, rz, z1d, d2.z, (d1.z == d2.z) ? 0.0 : 1.0 / ddz, azmin, azmax);
		// an empty range is not an early exit: the reference still sweeps the
		// clamped slice range (ProjectorDD.cpp:319-336,447)
		double amin = fmax(fmax(0.0, axmin), fmax(aymin, azmin));
		double amax = fmin(fmin(1.0, axmax), fmin(aymax, azmax));
		if (TOF && pr.tof.trunc > 0.f)
		{
			const double tmm = (double)tofMm(tofv);
			const double dn = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
			amin = fmax(amin, 0.5 + (tmm - (double)pr.tof.trunc) / dn);
			amax = fmin(amax, 0.5 + (tmm + (double)pr.tof.trunc) / dn);
		}
		x_i_0 = (int)floor((x1d + amin * ddx) / dx + 0.5 * (double)(P.nx - 1) + 0.5);
		y_i_0 = (int)floor((y1d + amin * ddy) / dy + 0.5 * (double)(P.ny - 1) + 0.5);
		x_i_1 = (int)floor((x1d + amax * ddx) / dx + 0.5 * (double)(P.nx - 1) + 0.5);
		y_i_1 = (int)floor((y1d + amax * ddy) / dy + 0.5 * (double)(P.ny - 1) + 0.5);
	}
	else
	{
		// reference FP32 setup (ProjectorDD.cpp:305-336)
		const float ix = (d1.x == d2.x) ? 0.0f : 1.0f / (d2.x - d1.x);
		const float iy = (d1.y == d2.y) ? 0.0f : 1.0f / (d2.y - d1.y);
		const float iz = (d1.z == d2.z) ? 0.0f : 1.0f / (d2.z - d1.z);
		float axmin, axmax, aymin, aymax, azmin, azmax;
		getAlpha<float>(-0.5f * (P.lx - dx), 0.5f * (P.lx - dx), d1.x, d2.x, ix, axmin, axmax);
		getAlpha<float>(-0.5f * (P.ly - dy), 0.5f * (P.ly - dy), d1.y, d2.y, iy, aymin, aymax);
		getAlpha<float>(-0.5f * (P.lz - dz), 0.5f * (P.lz - dz), d1.z, d2.z, iz, azmin, azmax);
		// an empty range is not an early exit (see above)
		float amin = fmaxf(fmaxf(0.0f, axmin), fmaxf(aymin, azmin));
		float amax = fminf(fminf(1.0f, axmax), fminf(aymax, azmax));
		if (TOF && pr.tof.trunc > 0.f)
		{
			const float tmm = tofMm(tofv);
			const float lo = (float)fmax(0.0, 0.5 + (double)((tmm - pr.tof.trunc) / d_norm));
			const float hi = (float)fmin(1.0, 0.5 + (double)((tmm + pr.tof.trunc) / d_norm));
			amin = fmaxf(amin, lo);
			amax = fminf(amax, hi);
		}
		const float x_0 = d1.x + amin * (d2.x - d1.x);
		const float y_0 = d1.y + amin * (d2.y - d1.y);
		const float x_1 = d1.x + amax * (d2.x - d1.x);
		const float y_1 = d1.y + amax * (d2.y - d1.y);
		x_i_0 = (int)floorf(x_0 / dx + 0.5f * (float)(P.nx - 1) + 0.5f);
		y_i_0 = (int)floorf(y_0 / dy + 0.5f * (float)(P.ny - 1) + 0.5f);
		x_i_1 = (int)floorf(x_1 / dx + 0.5f * (float)(P.nx - 1) + 0.5f);
		y_i_1 = (int)floorf(y_1 / dy + 0.5f * (float)(P.ny - 1) + 0.5f);
	}
	// ---- FP32 footprint (338-577) ----
	float d1_i, d2_i, n1_i, n2_i;
	if (flag_y)
	{
		d1_i = d1.y;
		d2_i = d2.y;
		n1_i = n1.y;
		n2_i = n2.y;
	}
	else
	{
		d1_i = d1.x;
		d2_i = d2.x;
		n1_i = n1.x;
		n2_i = n2.x;
	}
	const float n1xy2 = n1.x * n1.x + n1.y * n1.y;
	const float n1xy = sqrtf(n1xy2);
	const float n1px = n1.y / n1xy, n1py = -n1.x / n1xy;
	const float n1zn = sqrtf((n1.x * n1.z) * (n1.x * n1.z) +
	                         (n1.y * n1.z) * (n1.y * n1.z) + n1xy2);
	const float n1pi = (n1_i * n1.z) / n1zn, n1pz = -n1xy2 / n1zn;
	const float n2xy2 = n2.x * n2.x + n2.y * n2.y;
	const float n2xy = sqrtf(n2xy2);
	const float n2px = n2.y / n2xy, n2py = -n2.x / n2xy;
	const float n2zn = sqrtf((n2.x * n2.z) * (n2.x * n2.z) +
	                         (n2.y * n2.z) * (n2.y * n2.z) + n2xy2);
	const float n2pi = (n2_i * n2.z) / n2zn, n2pz = -n2xy2 / n2zn;
	const float htt = pr.tt * 0.5f;
	const float d1xyloX = d1.x - htt * n1px, d1xyloY = d1.y - htt * n1py;
	const float d1xyhiX = d1.x + htt * n1px, d1xyhiY = d1.y + htt * n1py;
	const float d2xyloX = d2.x - htt * n2px, d2xyloY = d2.y - htt * n2py;
	const float d2xyhiX = d2.x + htt * n2px, d2xyhiY = d2.y + htt * n2py;
	const float htz = pr.tz * 0.5f;
	const float d1zloI = d1_i - htz * n1pi, d1zloZ = d1.z - htz * n1pz;
	const float d1zhiI = d1_i + htz * n1pi, d1zhiZ = d1.z + htz * n1pz;
	const float d2zloI = d2_i - htz * n2pi, d2zloZ = d2.z - htz * n2pz;
	const float d2zhiI = d2_i + htz * n2pi, d2zhiZ = d2.z + htz * n2pz;
	int xy_i_0, xy_i_1, nyx;
	float lxy, lyx, dxy, dyx;
	float d1_xy_lo, d1_xy_hi, d2_xy_lo, d2_xy_hi;
	float d1_yx_lo, d1_yx_hi, d2_yx_lo, d2_yx_hi;
	int strideXY, strideYX;
	if (flag_y)
	{
		xy_i_0 = max(0, min(y_i_0, y_i_1));
		xy_i_1 = min(P.ny - 1, max(y_i_0, y_i_1));
		lxy = P.ly;
		dxy = P.vy;
		lyx = P.lx;
		dyx = P.vx;
		nyx = P.nx;
		d1_xy_lo = d1xyloY;
		d1_xy_hi = d1xyhiY;
		d2_xy_lo = d2xyloY;
		d2_xy_hi = d2xyhiY;
		d1_yx_lo = d1xyloX;
		d1_yx_hi = d1xyhiX;
		d2_yx_lo = d2xyloX;
		d2_yx_hi = d2xyhiX;
		strideXY = P.nx;
		strideYX = 1;
	}
	else
	{
		xy_i_0 = max(0, min(x_i_0, x_i_1));
		xy_i_1 = min(P.nx - 1, max(x_i_0, x_i_1));
		lxy = P.lx;
		dxy = P.vx;
		lyx = P.ly;
		dyx = P.vy;
		nyx = P.ny;
		d1_xy_lo = d1xyloX;
		d1_xy_hi = d1xyhiX;
		d2_xy_lo = d2xyloX;
		d2_xy_hi = d2xyhiX;
		d1_yx_lo = d1xyloY;
		d1_yx_hi = d1xyhiY;
		d2_yx_lo = d2xyloY;
		d2_yx_hi = d2xyhiY;
		strideXY = 1;
		strideYX = P.nx;
	}
	const float dxy_cos_theta =
	    (d1_i != d2_i) ? dxy / (fabsf(d1_i - d2_i) / d_norm) : dxy;
	const float yx_off = (float)(nyx - 1) / 2.f;
	const float inv_dyx = 1.0f / dyx;
	const float z_off = (float)(P.nz - 1) * 0.5f;
	const float inv_dz = 1.0f / dz;
	const float half_dyx = dyx * 0.5f;
	const float hvz = P.vz * 0.5f;
	float pc = 0.f;
	if (TOF)
		pc = (float)(0.5 * (double)d_norm + (double)tofMm(tofv));
	R.usePsf = usePsf;
	R.psfRow = psfRow;
	R.pk = pk;
	R.pcum = pcum;
	R.ext = ext;
	R.d_norm = d_norm;
	R.d1_i = d1_i;
	R.d2_i = d2_i;
	R.xy_i_0 = xy_i_0;
	R.xy_i_1 = xy_i_1;
	R.nyx = nyx;
	R.strideXY = strideXY;
	R.strideYX = strideYX;
	R.flag_y = flag_y;
	R.lxy = lxy;
	R.lyx = lyx;
	R.dxy = dxy;
	R.dyx = dyx;
	R.d1_xy_lo = d1_xy_lo;
	R.d1_xy_hi = d1_xy_hi;
	R.d2_xy_lo = d2_xy_lo;
	R.d2_xy_hi = d2_xy_hi;
	R.d1_yx_lo = d1_yx_lo;
	R.d1_yx_hi = d1_yx_hi;
	R.d2_yx_lo = d2_yx_lo;
	R.d2_yx_hi = d2_yx_hi;
	R.d1zloI = d1zloI;
	R.d1zloZ = d1zloZ;
	R.d1zhiI = d1zhiI;
	R.d1zhiZ = d1zhiZ;
	R.d2zloI = d2zloI;
	R.d2zloZ = d2zloZ;
	R.d2zhiI = d2zhiI;
	R.d2zhiZ = d2zhiZ;
	R.dxy_cos_theta = dxy_cos_theta;
	R.yx_off = yx_off;
	R.inv_dyx = inv_dyx;
	R.z_off = z_off;
	R.inv_dz = inv_dz;
	R.half_dyx = half_dyx;
	R.hvz = hvz;
	R.pc = pc;
	ddRecips(R);
}

// The per-LOR part of a DDRay in kDDVec float4 (a plan's DD ray table, 128
// bytes: one cache line), and back: the LOR's set-up is computed once
// (ddSetup) and the sweeps that use it read it, its five reciprocals
// included.  The image- and PSF-derived fields are recomputed from the
// image and projector exactly as ddSetup computes them, so a loaded DDRay
// equals the one ddSetup makes for the same LOR.  frame -1: skipped LOR.
constexpr int kDDVec = 8;

__device__ __forceinline__ void ddStore(float4* __restrict__ tab, uint64_t t, const DDRay& R,
                                        int32_t frame)
{
	float4* o = tab + kDDVec * t;
	o[0] = make_float4(R.d1_xy_lo, R.d1_xy_hi, R.d2_xy_lo, R.d2_xy_hi);
	o[1] = make_float4(R.d1_yx_lo, R.d1_yx_hi, R.d2_yx_lo, R.d2_yx_hi);
	o[2] = make_float4(R.d1zloI, R.d1zloZ, R.d1zhiI, R.d1zhiZ);
	o[3] = make_float4(R.d2zloI, R.d2zloZ, R.d2zhiI, R.d2zhiZ);
	o[4] = make_float4(R.dxy_cos_theta, R.d_norm, R.d1_i, R.d2_i);
	// slice range and major axis in one word: a non-empty range lies within
	// the image (below 2048 slices); an empty one is stored as [1, 0]
	const bool empty = R.xy_i_0 > R.xy_i_1;
	const uint32_t lo = empty ? 1u : (uint32_t)R.xy_i_0, hi = empty ? 0u : (uint32_t)R.xy_i_1;
	const uint32_t rg = lo | (hi << 11) | (R.flag_y ? 1u << 22 : 0u);
	o[5] = make_float4(R.pc, __uint_as_float(rg), __int_as_float(R.psfRow), __int_as_float(frame));
	o[6] = make_float4(R.r_xy_lo, R.r_xy_hi, R.r_z_lo, R.r_z_hi);
	o[7] = make_float4(R.r_i, R.inv_dyx, __int_as_float(R.rOk ? 1 : 0), 0.f);
}

// An empty entry: frame -1 and an empty slice range.
__device__ __forceinline__ void ddStoreSkip(float4* __restrict__ tab, uint64_t t)
{
	float4* o = tab + kDDVec * t;
	const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
	o[0] = z;
	o[1] = z;
	o[2] = z;
	o[3] = z;
	o[4] = z;
	o[5] = make_float4(0.f, __uint_as_float(1u), 0.f, __int_as_float(-1));
	o[6] = z;
	o[7] = z;
}

__device__ __forceinline__ int32_t ddFrame(const float4& e5) { return __float_as_int(e5.w); }

__device__ __forceinline__ void ddLoad(const DImage& P, const DProj& pr,
                                       const float4 (&e)[kDDVec], bool flipPsf, DDRay& R)
{
#pragma clang fp contract(off)
	R.d1_xy_lo = e[0].x;
	R.d1_xy_hi = e[0].y;
	R.d2_xy_lo = e[0].z;
	R.d2_xy_hi = e[0].w;
	R.d1_yx_lo = e[1].x;
	R.d1_yx_hi = e[1].y;
	R.d2_yx_lo = e[1].z;
	R.d2_yx_hi = e[1].w;
	R.d1zloI = e[2].x;
	R.d1zloZ = e[2].y;
	R.d1zhiI = e[2].z;
	R.d1zhiZ = e[2].w;
	R.d2zloI = e[3].x;
	R.d2zloZ = e[3].y;
	R.d2zhiI = e[3].z;
	R.d2zhiZ = e[3].w;
	R.dxy_cos_theta = e[4].x;
	R.d_norm = e[4].y;
	R.d1_i = e[4].z;
	R.d2_i = e[4].w;
	R.pc = e[5].x;
	const uint32_t rg = __float_as_uint(e[5].y);
	R.xy_i_0 = (int)(rg & 2047u);
	R.xy_i_1 = (int)((rg >> 11) & 2047u);
	R.flag_y = ((rg >> 22) & 1u) != 0;
	R.psfRow = __float_as_int(e[5].z);
	R.usePsf = pr.psf.k != nullptr;
	R.pk = R.usePsf ? (flipPsf ? pr.psf.kflip : pr.psf.k) + R.psfRow * psfRowTaps(pr.psf.K)
	                : nullptr;
	R.pcum = R.usePsf ? (flipPsf ? pr.psf.cumFlip : pr.psf.cum) + R.psfRow * (pr.psf.K + 2)
	                  : nullptr;
	R.ext = R.usePsf ? pr.psf.halfWidth : 0.f;
	const bool fy = R.flag_y;
	R.lxy = fy ? P.ly : P.lx;
	R.dxy = fy ? P.vy : P.vx;
	R.lyx = fy ? P.lx : P.ly;
	R.dyx = fy ? P.vx : P.vy;
	R.nyx = fy ? P.nx : P.ny;
	R.strideXY = fy ? P.nx : 1;
	R.strideYX = fy ? 1 : P.nx;
	R.yx_off = (float)(R.nyx - 1) / 2.f;
	R.inv_dyx = e[7].y;
	R.z_off = (float)(P.nz - 1) * 0.5f;
	R.inv_dz = 1.0f / P.vz;
	R.half_dyx = R.dyx * 0.5f;
	R.hvz = P.vz * 0.5f;
	R.r_xy_lo = e[6].x;
	R.r_xy_hi = e[6].y;
	R.r_z_lo = e[6].z;
	R.r_z_hi = e[6].w;
	R.r_i = e[7].x;
	R.rOk = __float_as_int(e[7].z) != 0;
}

// Footprint range of slice xyi: minor-axis voxels [yx0, yx1], z voxels
// [zi0, zi1] (clamped to the image), and the quantities the voxel weights
// need.  The quotients by slice-invariant denominators use the LOR's
// reciprocals (divByRecip: the same correctly rounded quotients).
struct DDSlice
{
	float r0, r1, r0ov, r1ov, widthFrac_yx, z0r, z1r, widthFrac_z, tw;
	int yx0, yx1, zi0, zi1;
};

template <bool TOF>
__device__ __forceinline__ void ddSlice(const DImage& P, const DDRay& R, const DProj& pr,
                                        int xyi, DDSlice& q)
{
#pragma clang fp contract(off)
	const float pix_xy = -0.5f * R.lxy + ((float)xyi + 0.5f) * R.dxy;
	float a_xy_lo, a_xy_hi, a_z_lo, a_z_hi;
	if (R.rOk)
	{
		a_xy_lo = divByRecip(pix_xy - R.d1_xy_lo, R.d2_xy_hi - R.d1_xy_lo, R.r_xy_lo);
		a_xy_hi = divByRecip(pix_xy - R.d1_xy_hi, R.d2_xy_lo - R.d1_xy_hi, R.r_xy_hi);
		a_z_lo = divByRecip(pix_xy - R.d1zloI, R.d2zloI - R.d1zloI, R.r_z_lo);
		a_z_hi = divByRecip(pix_xy - R.d1zhiI, R.d2zhiI - R.d1zhiI, R.r_z_hi);
	}
	else
	{
		a_xy_lo = (pix_xy - R.d1_xy_lo) / (R.d2_xy_hi - R.d1_xy_lo);
		a_xy_hi = (pix_xy - R.d1_xy_hi) / (R.d2_xy_lo - R.d1_xy_hi);
		a_z_lo = (pix_xy - R.d1zloI) / (R.d2zloI - R.d1zloI);
		a_z_hi = (pix_xy - R.d1zhiI) / (R.d2zhiI - R.d1zhiI);
	}
	float r0 = R.d1_yx_lo + a_xy_lo * (R.d2_yx_hi - R.d1_yx_lo);
	float r1 = R.d1_yx_hi + a_xy_hi * (R.d2_yx_lo - R.d1_yx_hi);
	if (r0 > r1)
	{
		const float t = r0;
		r0 = r1;
		r1 = t;
	}
	q.widthFrac_yx = r1 - r0;
	q.r0ov = r0;
	q.r1ov = r1;
	r0 -= R.ext;
	r1 += R.ext;
	q.r0 = r0;
	q.r1 = r1;
	q.yx0 = max(0, (int)rintf(r0 * R.inv_dyx + R.yx_off));
	q.yx1 = min(R.nyx - 1, (int)rintf(r1 * R.inv_dyx + R.yx_off));
	float z0r = R.d1zloZ + a_z_lo * (R.d2zloZ - R.d1zloZ);
	float z1r = R.d1zhiZ + a_z_hi * (R.d2zhiZ - R.d1zhiZ);
	if (z0r > z1r)
	{
		const float t = z0r;
		z0r = z1r;
		z1r = t;
	}
	q.z0r = z0r;
	q.z1r = z1r;
	q.widthFrac_z = z1r - z0r;
	q.zi0 = max(0, (int)rintf(z0r * R.inv_dz + R.z_off));
	q.zi1 = min(P.nz - 1, (int)rintf(z1r * R.inv_dz + R.z_off));
	q.tw = 1.f;
	if (TOF)
	{
		// getWeight at the slab's [a_lo, a_hi] (ProjectorDD.cpp:507-515)
		const float di = R.d2_i - R.d1_i;
		const float nlo = pix_xy - R.d1_i - 0.5f * R.dxy;
		const float nhi = pix_xy - R.d1_i + 0.5f * R.dxy;
		const float alo = R.rOk ? divByRecip(nlo, di, R.r_i) : nlo / di;
		const float ahi = R.rOk ? divByRecip(nhi, di, R.r_i) : nhi / di;
		const float xcn =
		    divBySigma(0.5f * (alo * R.d_norm + ahi * R.d_norm) - R.pc, pr.tof.sigma,
		               pr.tof.rsigma);
		q.tw = gaussW(xcn, pr.tof.norm);
	}
}

// A slice's footprint weight x / width from the slice's reciprocal of the
// width (ok: usable), else by division.
__device__ __forceinline__ float ddDivW(float x, float width, float rw, bool ok)
{
	return ok ? divByRecip(x, width, rw) : x / width;
}

// z weights of a slice computed once for all its minor-axis voxels when the
// slice spans at most this many z planes (else per voxel)
constexpr int DD_ZW = 4;

// The footprint voxels of slices [xlo, xhi] of the major axis, minor-axis
// index in [ylo, yhi] and z in [zlo, zhi] (each also clamped to the ray's own
// range): fn(io, xyi, yxi, zi, w, xv), io the voxel's index in the frame.  The
// whole image (ranges covering it) is ProjectorDD's sweep.  The weight of
// voxel (xyi, yxi, zi) is ProjectorDD's: (w_yx / width_yx) * (w_z / width_z)
// * dxy / |cos|, times the TOF weight of the slice; the z factor depends on
// the slice and zi only, so it is computed once per slice and z plane.
// PRE: xv = img[io], the loads of a minor-axis voxel's z planes issued
// before its weight is computed (the PSF integral then covers their
// latency); else xv = 0.
template <bool TOF, bool PRE, class Fn>
__device__ __forceinline__ void ddVisitX(const DImage& P, const DDRay& R, const DProj& pr,
                                         int xlo, int xhi, int ylo, int yhi, int zlo, int zhi,
                                         const float* __restrict__ img, Fn&& fn)
{
#pragma clang fp contract(off)
	const int x0 = max(R.xy_i_0, xlo), x1 = min(R.xy_i_1, xhi);
	for (int xyi = x0; xyi <= x1; xyi++)
	{
		DDSlice q;
		ddSlice<TOF>(P, R, pr, xyi, q);
		const int yx0 = max(q.yx0, ylo), yx1 = min(q.yx1, yhi);
		const int zi0 = max(q.zi0, zlo), zi1 = min(q.zi1, zhi);
		if (yx0 > yx1 || zi0 > zi1)
			continue;
		const float ry = 1.0f / q.widthFrac_yx, rz = 1.0f / q.widthFrac_z;
		const bool oky = recipOk(q.widthFrac_yx, ry), okz = recipOk(q.widthFrac_z, rz);
		const int nzv = zi1 - zi0 + 1;
		// z weights of the slice (at most DD_ZW planes) and which planes the
		// footprint overlaps
		float wz[DD_ZW];
		uint32_t zmask = 0u;
		if (nzv <= DD_ZW)
		{
#pragma unroll
			for (int j = 0; j < DD_ZW; j++)
			{
				const int zi = zi0 + j;
				const float pix_z = -0.5f * P.lz + ((float)zi + 0.5f) * P.vz;
				const float zp0 = pix_z - R.hvz;
				const float zp1 = pix_z + R.hvz;
				wz[j] = ddDivW(fminf(zp1, q.z1r) - fmaxf(zp0, q.z0r), q.widthFrac_z, rz, okz);
				if (j < nzv && q.z1r >= zp0 && q.z0r < zp1)
					zmask |= 1u << j;
			}
		}
		for (int yxi = yx0; yxi <= yx1; yxi++)
		{
			const float pix_yx = -0.5f * R.lyx + ((float)yxi + 0.5f) * R.dyx;
			const float p0 = pix_yx - R.half_dyx;
			const float p1 = pix_yx + R.half_dyx;
			if (!(q.r1 >= p0 && q.r0 < p1))
				continue;
			const int ib = xyi * R.strideXY + yxi * R.strideYX;
			float xv[DD_ZW];
#pragma unroll
			for (int j = 0; j < DD_ZW; j++)
				xv[j] = (PRE && nzv <= DD_ZW && ((zmask >> j) & 1u)) ? img[(zi0 + j) * P.nxy + ib]
				                                                      : 0.f;
			const float wxy =
			    R.usePsf ? psfWeight(pr.psf, R.pk, R.pcum, p0 - q.r1ov, p1 - q.r0ov)
			             : (fminf(p1, q.r1ov) - fmaxf(p0, q.r0ov));
			const float wxy_s = ddDivW(wxy, q.widthFrac_yx, ry, oky);
			if (nzv <= DD_ZW)
			{
#pragma unroll
				for (int j = 0; j < DD_ZW; j++)
				{
					if (!((zmask >> j) & 1u))
						continue;
					const int zi = zi0 + j;
					float w = wxy_s * wz[j] * R.dxy_cos_theta;
					if (TOF)
						w *= q.tw;
					fn(zi * P.nxy + ib, xyi, yxi, zi, w, xv[j]);
				}
			}
			else
			{
				for (int zi = zi0; zi <= zi1; zi++)
				{
					const float pix_z = -0.5f * P.lz + ((float)zi + 0.5f) * P.vz;
					const float zp0 = pix_z - R.hvz;
					const float zp1 = pix_z + R.hvz;
					if (!(q.z1r >= zp0 && q.z0r < zp1))
						continue;
					const float wz_s = ddDivW(fminf(zp1, q.z1r) - fmaxf(zp0, q.z0r),
					                          q.widthFrac_z, rz, okz);
					float w = wxy_s * wz_s * R.dxy_cos_theta;
					if (TOF)
						w *= q.tw;
					const int io = zi * P.nxy + ib;
					fn(io, xyi, yxi, zi, w, PRE ? img[io] : 0.f);
				}
			}
		}
	}
}

// ddVisitX without the image loads: fn(io, xyi, yxi, zi, w).
template <bool TOF, class Fn>
__device__ __forceinline__ void ddVisit(const DImage& P, const DDRay& R, const DProj& pr,
                                        int xlo, int xhi, int ylo, int yhi, int zlo, int zhi,
                                        Fn&& fn)
{
	ddVisitX<TOF, false>(P, R, pr, xlo, xhi, ylo, yhi, zlo, zhi, nullptr,
	                     [&](int io, int xyi, int yxi, int zi, float w, float) {
		                     fn(io, xyi, yxi, zi, w);
	                     });
}

// Forward sweep of a set-up DD LOR over the whole image: the sum of w * x
// in the sweep's voxel order.
template <bool TOF>
__device__ __forceinline__ float ddSweepFwd(const DImage& P, const float* __restrict__ img,
                                            const DDRay& R, const DProj& pr)
{
	float acc = 0.f;
	ddVisitX<TOF, true>(P, R, pr, 0, 1 << 30, 0, 1 << 30, 0, 1 << 30, img,
	                    [&](int, int, int, int, float w, float xv) { acc = fmaf(w, xv, acc); });
	return acc;
}

template <int MODE, bool TOF, bool S64>
__device__ float ddRay(const DImage& P, float* __restrict__ img, float3 d1,
                       float3 d2, float3 n1, float3 n2, float val,
                       const DProj& pr, float tofv, uint32_t& cnt)
{
	DDRay R;
	ddSetup<TOF, S64>(P, d1, d2, n1, n2, pr, tofv, !(MODE == MODE_FWD || MODE == MODE_COUNT), R);
	if constexpr (MODE == MODE_FWD)
		return ddSweepFwd<TOF>(P, img, R, pr);
	ddVisit<TOF>(P, R, pr, 0, 1 << 30, 0, 1 << 30, 0, 1 << 30,
	             [&](int io, int, int, int, float w) {
		             if (MODE == MODE_BACK)
			             unsafeAtomicAdd(img + io, val * w);
		             else
			             cnt++;
	             });
	return 0.f;
}

}  // namespace yrtmi
