/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle for the OSEM projection path.
 *
 * A restatement (not a copy) of the reference CPU algorithms of YRT-PET 2.0.7
 * for the projection hot path.  Each function cites the reference file:line
 * whose behaviour it follows.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path never
 * does.
 *
 * FP32 instantiations keep the reference's operation order (float storage,
 * the double-promoted TOF constants).  FP64 instantiations run the same
 * algorithm in double; they are the "FP64 oracle".
 *
 * Pins (tests/test_oracle_*.py):
 *   get_alpha, TOF, PSF integral     - vs the reference headers compiled in
 *                                      oracle/_ref + TOF known answer
 *   multi-ray offsets / line gen     - vs reference MultiRayGenerator in _ref
 *                                      + bounding-box invariants
 *   Siddon                           - closed forms (isocentre, chord, z rays,
 *                                      outside rays), FP64 brute-force voxel
 *                                      intersection, regression coordinates,
 *                                      adjoint identity
 *   DD (+TOF, +PSF)                  - getOverlap known answers, uniform-image
 *                                      footprint identity, adjoint, per-voxel
 *                                      closed forms for LORs whose crystals
 *                                      face each other along an axis (unequal
 *                                      crystal sizes, anisotropic voxels,
 *                                      footprint-end rounding, flat and
 *                                      asymmetric PSF rows with the flip on
 *                                      backprojection, TOF slab weights)
 *   LUT / Histogram3D / subsets      - closed-form geometry, bin<->pair
 *                                      round trip, subset partition
 *   EM subset / update / threshold   - elementwise closed forms, MLEM
 *                                      invariants (count preservation, fixed
 *                                      point), one-LOR hand computation
 */
#include "yrt_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <thread>
#include <type_traits>
#include <vector>

namespace
{

constexpr double kPi = 3.14159265358979;          // Constants.hpp:14
constexpr double kC_mm_ps = 0.299792458;          // Constants.hpp:18
constexpr double kC_mm_ps_flt = (double)0.299792458f;  // Constants.hpp:19
constexpr double kDoublePrecision = 10e-25;       // Constants.hpp:20

template <class T>
struct V3
{
	T x, y, z;
};
template <class T>
inline V3<T> operator-(const V3<T>& a, const V3<T>& b)
{
	return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
inline V3<T> operator+(const V3<T>& a, const V3<T>& b)
{
	return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
inline V3<T> operator*(const V3<T>& a, T s)
{
	return {a.x * s, a.y * s, a.z * s};
}
// Vector3D::getNorm (src/geometry/Vector3D.cpp:127-137)
template <class T>
inline T norm3(const V3<T>& a)
{
	return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
// Vector3D::normalize (src/geometry/Vector3D.cpp:156-177)
template <class T>
inline V3<T> normalized(V3<T> a)
{
	const T n = norm3(a);
	if (n > kDoublePrecision)
	{
		return {a.x / n, a.y / n, a.z / n};
	}
	return {T(0), T(0), T(0)};
}
// Vector3D::crossProduct (src/geometry/Vector3D.cpp:237-243)
template <class T>
inline V3<T> cross(const V3<T>& a, const V3<T>& b)
{
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class T>
struct ImgP
{
	int64_t nx, ny, nz, nt;
	T vx, vy, vz, lx, ly, lz, ox, oy, oz, fovR;
};

template <class T>
ImgP<T> toImg(const yo_image* im)
{
	return {im->nx, im->ny, im->nz, im->nt,     T(im->vx), T(im->vy),
	        T(im->vz), T(im->lx), T(im->ly),    T(im->lz), T(im->ox),
	        T(im->oy), T(im->oz), T(im->fov_radius)};
}

// util::get_alpha (include/yrt-pet/operators/ProjectorUtils.hpp:15-41)
template <class T>
inline void get_alpha(T r0, T r1, T p1, T p2, T inv_p12, T& amin, T& amax)
{
	amin = T(0);
	amax = T(1);
	if (p1 != p2)
	{
		const T a0 = (r0 - p1) * inv_p12;
		const T a1 = (r1 - p1) * inv_p12;
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

// TimeOfFlightHelper (include/yrt-pet/operators/TimeOfFlight.hpp:19-84)
template <class T>
struct Tof
{
	T sigma, trunc, norm;
	Tof(float width_ps, int n_std)
	{
		if constexpr (std::is_same_v<T, float>)
		{
			const double width_mm = width_ps * kC_mm_ps * 0.5;
			sigma = static_cast<float>(width_mm /
			                           (2 * sqrtf(2 * logf(2.0f))));
			trunc = (n_std <= 0) ? -1.f : n_std * sigma;
			norm = static_cast<float>(1 / (std::sqrt(2 * kPi) * sigma));
		}
		else
		{
			const double width_mm = double(width_ps) * kC_mm_ps * 0.5;
			sigma = width_mm / (2 * std::sqrt(2 * std::log(2.0)));
			trunc = (n_std <= 0) ? -1.0 : n_std * sigma;
			norm = 1 / (std::sqrt(2 * M_PI) * sigma);
		}
	}
	static T tof_mm(T tof_ps)
	{
		if constexpr (std::is_same_v<T, float>)
			return static_cast<float>(tof_ps * kC_mm_ps_flt * 0.5);
		else
			return tof_ps * kC_mm_ps * 0.5;
	}
	void alphaRange(T& amin, T& amax, T lorNorm, T tof_ps) const
	{
		if (trunc <= T(0))
		{
			amin = T(0);
			amax = T(1);
		}
		else
		{
			const T tmm = tof_mm(tof_ps);
			amin = static_cast<T>(
			    std::max(0.0, 0.5 + static_cast<double>((tmm - trunc) / lorNorm)));
			amax = static_cast<T>(
			    std::min(1.0, 0.5 + static_cast<double>((tmm + trunc) / lorNorm)));
		}
	}
	T weight(T lorNorm, T tof_ps, T lo, T hi) const
	{
		const T tmm = tof_mm(tof_ps);
		const T pc = static_cast<T>(0.5 * static_cast<double>(lorNorm) + tmm);
		const T xc = (T(0.5) * (lo + hi) - pc) / sigma;
		return std::exp(T(-0.5) * xc * xc) * norm;
	}
};

// util::calculateIntegral (include/yrt-pet/operators/ProjectorUtils.hpp:43-98)
template <class T>
T psfIntegral(const float* kernel, int K, T kstep, T x0, T x1)
{
	const int half = (K - 1) / 2;
	if (x0 > (half + 1) * kstep || x1 < -(half + 1) * kstep || x0 >= x1)
	{
		return T(0);
	}
	auto tap = [&](int i) -> T
	{ return (i < 0 || i >= K) ? T(0) : T(kernel[i]); };
	const T x0s = x0 / kstep;
	const int x0i = static_cast<int>(std::floor(x0s + half));
	const T tau0 = x0s + -(-half + x0i);
	const T v0lo = tap(x0i);
	const T v0hi = (x0i < -1 || x0i >= K - 1) ? T(0) : T(kernel[x0i + 1]);
	const T v0 = v0lo * (1 - tau0) + v0hi * tau0;
	const T x1s = x1 / kstep;
	const int x1i = static_cast<int>(std::floor(x1s + half));
	const T tau1 = x1s - (-half + x1i);
	const T v1lo = tap(x1i);
	const T v1hi = (x1i < -1 || x1i >= K - 1) ? T(0) : T(kernel[x1i + 1]);
	const T v1 = v1lo * (1 - tau1) + v1hi * tau1;
	T out = T(0);
	if (x0i == x1i)
	{
		out = T(0.5) * (v0 + v1) * (x1 - x0);
	}
	else
	{
		out += T(0.5) * (v0 + v0hi) * kstep * (1 - tau0);
		out += T(0.5) * (v1lo + v1) * kstep * tau1;
		for (int xi = x0i + 1; xi < x1i; xi++)
		{
			out += kstep * T(0.5) * (tap(xi) + tap(xi + 1));
		}
	}
	return out;
}

// ProjectionPsfManager (src/operators/ProjectionPsfManager.cpp:52-103)
template <class T>
struct Psf
{
	const float* kernels = nullptr;
	int numS = 0, K = 0;
	T sStep = 0, kSpacing = 0;
	std::vector<float> flipped;
	bool active() const { return kernels != nullptr; }
	void init(const yo_projector* pr)
	{
		if (pr->psf_kernels == nullptr || pr->psf_ksize <= 0)
			return;
		kernels = pr->psf_kernels;
		numS = pr->psf_num_s;
		K = pr->psf_ksize;
		sStep = pr->psf_s_step;
		kSpacing = pr->psf_k_spacing;
		flipped.resize(size_t(numS) * K);
		for (int i = 0; i < numS; i++)
			for (int j = 0; j < K; j++)
				flipped[size_t(i) * K + (K - 1 - j)] = kernels[size_t(i) * K + j];
	}
	T halfWidth() const { return T((K + 1) / 2) * kSpacing; }
	const float* kernel(const V3<T>& p1, const V3<T>& p2, bool flip) const
	{
		T npx = p2.y - p1.y;
		T npy = p1.x - p2.x;
		const T nn = std::sqrt(npx * npx + npy * npy);
		size_t s = 0;
		if (nn != T(0))
		{
			npx /= nn;
			npy /= nn;
			const T sv = std::abs(p1.x * npx + p1.y * npy);
			s = std::min(static_cast<size_t>(std::floor(sv / sStep)),
			             static_cast<size_t>(numS - 1));
		}
		return flip ? flipped.data() + s * K : kernels + s * K;
	}
	T weight(const float* k, T x0, T x1) const
	{
		const int hw = (K + 1) / 2;
		if (x0 > hw * kSpacing || x1 < -hw * kSpacing || x0 >= x1)
			return T(0);
		return psfIntegral<T>(k, K, kSpacing, x0, x1);
	}
};

template <class T>
inline void accum(T* p, T v, bool atomic)
{
	if (atomic)
	{
		std::atomic_ref<T> a(*p);
		a.fetch_add(v, std::memory_order_relaxed);
	}
	else
	{
		*p += v;
	}
}

// ProjectorSiddon::project_helper<IS_FWD, FLAG_INCR=true, FLAG_TOF>
// (src/operators/ProjectorSiddon.cpp:478-767).  Counts applied segments in
// *nseg when non-null.
template <class T, bool FWD, bool TOF>
void siddonRay(const ImgP<T>& P, T* img, const V3<T>& p1, const V3<T>& p2,
               T& val, int64_t frame, const Tof<T>* tof, T tofv, bool atomic,
               int64_t* nseg)
{
	if (FWD)
		val = T(0);
	// 1. FOV cylinder (499-518)
	const T A = (p2.x - p1.x) * (p2.x - p1.x) + (p2.y - p1.y) * (p2.y - p1.y);
	const T B = T(2) * ((p2.x - p1.x) * p1.x + (p2.y - p1.y) * p1.y);
	const T C = p1.x * p1.x + p1.y * p1.y - P.fovR * P.fovR;
	const T Delta = B * B - T(4) * A * C;
	T t0, t1;
	if (A != T(0))
	{
		if (Delta <= T(0))
			return;
		t0 = (-B - std::sqrt(Delta)) / (T(2) * A);
		t1 = (-B + std::sqrt(Delta)) / (T(2) * A);
	}
	else
	{
		t0 = T(0);
		t1 = T(1);
	}
	const T d_norm = norm3(p1 - p2);
	const bool flat_x = (p1.x == p2.x);
	const bool flat_y = (p1.y == p2.y);
	const bool flat_z = (p1.z == p2.z);
	const T inv_x = flat_x ? T(0) : T(1) / (p2.x - p1.x);
	const T inv_y = flat_y ? T(0) : T(1) / (p2.y - p1.y);
	const T inv_z = flat_z ? T(0) : T(1) / (p2.z - p1.z);
	const int64_t dir_x = (inv_x >= T(0)) ? 1 : -1;
	const int64_t dir_y = (inv_y >= T(0)) ? 1 : -1;
	const int64_t dir_z = (inv_z >= T(0)) ? 1 : -1;
	// 2. Box (531-553)
	const T dx = P.vx, dy = P.vy, dz = P.vz;
	const T inv_dx = T(1) / dx, inv_dy = T(1) / dy, inv_dz = T(1) / dz;
	const T x0 = -P.lx * T(0.5), x1 = P.lx * T(0.5);
	const T y0 = -P.ly * T(0.5), y1 = P.ly * T(0.5);
	const T z0 = -P.lz * T(0.5), z1 = P.lz * T(0.5);
	T axmin, axmax, aymin, aymax, azmin, azmax;
	get_alpha<T>(T(-0.5) * P.lx, T(0.5) * P.lx, p1.x, p2.x, inv_x, axmin, axmax);
	get_alpha<T>(T(-0.5) * P.ly, T(0.5) * P.ly, p1.y, p2.y, inv_y, aymin, aymax);
	get_alpha<T>(T(-0.5) * P.lz, T(0.5) * P.lz, p1.z, p2.z, inv_z, azmin, azmax);
	T amin = std::max({T(0), t0, axmin, aymin, azmin});
	T amax = std::min({T(1), t1, axmax, aymax, azmax});
	if (TOF)
	{
		T a0t, a1t;
		tof->alphaRange(a0t, a1t, d_norm, tofv);
		amin = std::max(amin, a0t);
		amax = std::min(amax, a1t);
	}
	T a_cur = amin;
	T a_next = T(-1);
	T x_cur = (inv_x > T(0)) ? x0 : x1;
	T y_cur = (inv_y > T(0)) ? y0 : y1;
	T z_cur = (inv_z > T(0)) ? z0 : z1;
	// early out (567-572)
	if ((inv_x >= T(0) && p1.x > x1) || (inv_x < T(0) && p1.x < x0) ||
	    (inv_y >= T(0) && p1.y > y1) || (inv_y < T(0) && p1.y < y0) ||
	    (inv_z >= T(0) && p1.z > z1) || (inv_z < T(0) && p1.z < z0))
	{
		return;
	}
	// first crossings (574-594)
	T ax_next = flat_x ? std::numeric_limits<T>::max() : axmin;
	if (!flat_x)
	{
		const int64_t kx = static_cast<int64_t>(
		    std::ceil(T(dir_x) * (a_cur * (p2.x - p1.x) - x_cur + p1.x) / dx));
		x_cur += T(kx * dir_x) * dx;
		ax_next = (x_cur - p1.x) * inv_x;
	}
	T ay_next = flat_y ? std::numeric_limits<T>::max() : aymin;
	if (!flat_y)
	{
		const int64_t ky = static_cast<int64_t>(
		    std::ceil(T(dir_y) * (a_cur * (p2.y - p1.y) - y_cur + p1.y) / dy));
		y_cur += T(ky * dir_y) * dy;
		ay_next = (y_cur - p1.y) * inv_y;
	}
	T az_next = flat_z ? std::numeric_limits<T>::max() : azmin;
	if (!flat_z)
	{
		const int64_t kz = static_cast<int64_t>(
		    std::ceil(T(dir_z) * (a_cur * (p2.z - p1.z) - z_cur + p1.z) / dz));
		z_cur += T(kz * dir_z) * dz;
		az_next = (z_cur - p1.z) * inv_z;
	}
	bool flag_first = true;
	int64_t vx = -1, vy = -1, vz = -1;
	int dir_prev = -1;
	int dir_next = -1;
	int64_t off = 0;
	const int64_t nx = P.nx, ny = P.ny, nz = P.nz, nxy = nx * ny;
	const int64_t nvox = nx * ny * nz;
	T axp = ax_next, ayp = ay_next, azp = az_next;
	bool done = false;
	// 3. traversal (622-766)
	while (a_cur < amax && !done)
	{
		dir_next = 0;
		if (axp <= ayp && axp <= azp)
		{
			a_next = ax_next;
			x_cur += T(dir_x) * dx;
			ax_next = (x_cur - p1.x) * inv_x;
			dir_next |= 1;
		}
		if (ayp <= axp && ayp <= azp)
		{
			a_next = ay_next;
			y_cur += T(dir_y) * dy;
			ay_next = (y_cur - p1.y) * inv_y;
			dir_next |= 2;
		}
		if (azp <= axp && azp <= ayp)
		{
			a_next = az_next;
			z_cur += T(dir_z) * dz;
			az_next = (z_cur - p1.z) * inv_z;
			dir_next |= 4;
		}
		if (a_next > amax)
			a_next = amax;
		if (a_cur >= a_next)
		{
			axp = ax_next;
			ayp = ay_next;
			azp = az_next;
			continue;
		}
		T tw = T(1);
		const T a_mid = T(0.5) * (a_cur + a_next);
		if (TOF)
			tw = tof->weight(d_norm, tofv, a_cur * d_norm, a_next * d_norm);
		if (flag_first)
		{
			vx = static_cast<int64_t>(
			    (p1.x + a_mid * (p2.x - p1.x) + P.lx / T(2)) * inv_dx);
			vy = static_cast<int64_t>(
			    (p1.y + a_mid * (p2.y - p1.y) + P.ly / T(2)) * inv_dy);
			vz = static_cast<int64_t>(
			    (p1.z + a_mid * (p2.z - p1.z) + P.lz / T(2)) * inv_dz);
			off = vz * nxy + vy * nx;
			flag_first = false;
			if (vx < 0 || vx >= nx || vy < 0 || vy >= ny || vz < 0 || vz >= nz)
				done = true;
		}
		else
		{
			if (dir_prev & 1)
			{
				vx += dir_x;
				if (vx < 0 || vx >= nx)
					done = true;
			}
			if (dir_prev & 2)
			{
				vy += dir_y;
				if (vy < 0 || vy >= ny)
					done = true;
				else
					off += dir_y * nx;
			}
			if (dir_prev & 4)
			{
				vz += dir_z;
				if (vz < 0 || vz >= nz)
					done = true;
				else
					off += dir_z * nxy;
			}
		}
		if (done)
			continue;
		dir_prev = dir_next;
		T w = (a_next - a_cur) * d_norm;
		const int64_t idx = frame * nvox + vx + off;
		if (TOF)
			w *= tw;
		if (nseg)
			(*nseg)++;
		if (img)
		{
			if (FWD)
				val += w * img[idx];
			else
				accum(&img[idx], val * w, atomic);
		}
		a_cur = a_next;
		axp = ax_next;
		ayp = ay_next;
		azp = az_next;
	}
}

// ProjectorDD::getOverlap_safe (src/operators/ProjectorDD.cpp:234-249)
template <class T>
inline T ddOverlap(T p0, T p1, T d0, T d1)
{
	return std::min(p1, d1) - std::max(p0, d0);
}

template <class T>
inline T rintT(T v)
{
	if constexpr (std::is_same_v<T, float>)
		return std::rintf(v);
	else
		return std::rint(v);
}

// ProjectorDD::dd_project_ref<IS_FWD, FLAG_TOF>
// (src/operators/ProjectorDD.cpp:260-578).  d1/d2 are the LOR endpoints with
// the image offset already subtracted (275-277).
template <class T, bool FWD, bool TOF>
void ddRay(const ImgP<T>& P, T* img, const V3<T>& d1, const V3<T>& d2,
           const V3<T>& n1, const V3<T>& n2, T& val, int64_t frame,
           const Tof<T>* tof, T tofv, const Psf<T>* psf, T thick_z,
           T thick_trans, bool atomic, int64_t* nseg)
{
	if (FWD)
		val = T(0);
	const V3<T> d12 = d1 - d2;
	const bool flag_y = std::abs(d12.y) > std::abs(d12.x);
	const T d_norm = norm3(d12);
	const float* pk = nullptr;
	T ext = T(0);
	const bool usePsf = psf != nullptr && psf->active();
	if (usePsf)
	{
		pk = psf->kernel(d1, d2, !FWD);
		ext = psf->halfWidth();
	}
	const int64_t nxy = P.nx * P.ny;
	const int64_t nvox = nxy * P.nz;
	const T dx = P.vx, dy = P.vy, dz = P.vz;
	const T inv_x = (d1.x == d2.x) ? T(0) : T(1) / (d2.x - d1.x);
	const T inv_y = (d1.y == d2.y) ? T(0) : T(1) / (d2.y - d1.y);
	const T inv_z = (d1.z == d2.z) ? T(0) : T(1) / (d2.z - d1.z);
	T axmin, axmax, aymin, aymax, azmin, azmax;
	get_alpha<T>(T(-0.5) * (P.lx - dx), T(0.5) * (P.lx - dx), d1.x, d2.x,
	             inv_x, axmin, axmax);
	get_alpha<T>(T(-0.5) * (P.ly - dy), T(0.5) * (P.ly - dy), d1.y, d2.y,
	             inv_y, aymin, aymax);
	get_alpha<T>(T(-0.5) * (P.lz - dz), T(0.5) * (P.lz - dz), d1.z, d2.z,
	             inv_z, azmin, azmax);
	T amin = std::max({T(0), axmin, aymin, azmin});
	T amax = std::min({T(1), axmax, aymax, azmax});
	if (TOF)
	{
		T a0t, a1t;
		tof->alphaRange(a0t, a1t, d_norm, tofv);
		amin = std::max(amin, a0t);
		amax = std::min(amax, a1t);
	}
	const T x_0 = d1.x + amin * (d2.x - d1.x);
	const T y_0 = d1.y + amin * (d2.y - d1.y);
	const T x_1 = d1.x + amax * (d2.x - d1.x);
	const T y_1 = d1.y + amax * (d2.y - d1.y);
	const int64_t x_i_0 = static_cast<int64_t>(
	    std::floor(x_0 / dx + T(0.5) * T(P.nx - 1) + T(0.5)));
	const int64_t y_i_0 = static_cast<int64_t>(
	    std::floor(y_0 / dy + T(0.5) * T(P.ny - 1) + T(0.5)));
	const int64_t x_i_1 = static_cast<int64_t>(
	    std::floor(x_1 / dx + T(0.5) * T(P.nx - 1) + T(0.5)));
	const int64_t y_i_1 = static_cast<int64_t>(
	    std::floor(y_1 / dy + T(0.5) * T(P.ny - 1) + T(0.5)));
	T d1_i, d2_i, n1_i, n2_i;
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
	// detector-edge normals (354-372)
	const T n1xy2 = n1.x * n1.x + n1.y * n1.y;
	const T n1xy = std::sqrt(n1xy2);
	const T n1px = n1.y / n1xy;
	const T n1py = -n1.x / n1xy;
	const T n1zn = std::sqrt((n1.x * n1.z) * (n1.x * n1.z) +
	                         (n1.y * n1.z) * (n1.y * n1.z) + n1xy2);
	const T n1pi = (n1_i * n1.z) / n1zn;
	const T n1pz = -n1xy2 / n1zn;
	const T n2xy2 = n2.x * n2.x + n2.y * n2.y;
	const T n2xy = std::sqrt(n2xy2);
	const T n2px = n2.y / n2xy;
	const T n2py = -n2.x / n2xy;
	const T n2zn = std::sqrt((n2.x * n2.z) * (n2.x * n2.z) +
	                         (n2.y * n2.z) * (n2.y * n2.z) + n2xy2);
	const T n2pi = (n2_i * n2.z) / n2zn;
	const T n2pz = -n2xy2 / n2zn;
	// in-plane / through-plane detector edges (374-394)
	const T htt = thick_trans * T(0.5);
	const T d1xyloX = d1.x - htt * n1px, d1xyloY = d1.y - htt * n1py;
	const T d1xyhiX = d1.x + htt * n1px, d1xyhiY = d1.y + htt * n1py;
	const T d2xyloX = d2.x - htt * n2px, d2xyloY = d2.y - htt * n2py;
	const T d2xyhiX = d2.x + htt * n2px, d2xyhiY = d2.y + htt * n2py;
	const T htz = thick_z * T(0.5);
	const T d1zloI = d1_i - htz * n1pi, d1zloZ = d1.z - htz * n1pz;
	const T d1zhiI = d1_i + htz * n1pi, d1zhiZ = d1.z + htz * n1pz;
	const T d2zloI = d2_i - htz * n2pi, d2zloZ = d2.z - htz * n2pz;
	const T d2zhiI = d2_i + htz * n2pi, d2zhiZ = d2.z + htz * n2pz;
	int64_t xy_i_0, xy_i_1, nyx;
	T lxy, lyx, dxy, dyx;
	T d1_xy_lo, d1_xy_hi, d2_xy_lo, d2_xy_hi;
	T d1_yx_lo, d1_yx_hi, d2_yx_lo, d2_yx_hi;
	if (flag_y)
	{
		xy_i_0 = std::max<int64_t>(0, std::min(y_i_0, y_i_1));
		xy_i_1 = std::min<int64_t>(P.ny - 1, std::max(y_i_0, y_i_1));
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
	}
	else
	{
		xy_i_0 = std::max<int64_t>(0, std::min(x_i_0, x_i_1));
		xy_i_1 = std::min<int64_t>(P.nx - 1, std::max(x_i_0, x_i_1));
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
	}
	T dxy_cos_theta;
	if (d1_i != d2_i)
		dxy_cos_theta = dxy / (std::abs(d1_i - d2_i) / d_norm);
	else
		dxy_cos_theta = dxy;

	for (int64_t xyi = xy_i_0; xyi <= xy_i_1; xyi++)
	{
		const T pix_xy = T(-0.5) * lxy + (T(xyi) + T(0.5)) * dxy;
		const T a_xy_lo = (pix_xy - d1_xy_lo) / (d2_xy_hi - d1_xy_lo);
		const T a_xy_hi = (pix_xy - d1_xy_hi) / (d2_xy_lo - d1_xy_hi);
		const T a_z_lo = (pix_xy - d1zloI) / (d2zloI - d1zloI);
		const T a_z_hi = (pix_xy - d1zhiI) / (d2zhiI - d1zhiI);
		T r0 = d1_yx_lo + a_xy_lo * (d2_yx_hi - d1_yx_lo);
		T r1 = d1_yx_hi + a_xy_hi * (d2_yx_lo - d1_yx_hi);
		if (r0 > r1)
			std::swap(r0, r1);
		const T widthFrac_yx = r1 - r0;
		const T r0ov = r0, r1ov = r1;
		r0 -= ext;
		r1 += ext;
		const T yx_off = T(nyx - 1) / T(2);
		const T inv_dyx = T(1) / dyx;
		const int64_t yx0 = std::max<int64_t>(
		    0, static_cast<int64_t>(rintT<T>(r0 * inv_dyx + yx_off)));
		const int64_t yx1 = std::min<int64_t>(
		    nyx - 1, static_cast<int64_t>(rintT<T>(r1 * inv_dyx + yx_off)));
		for (int64_t yxi = yx0; yxi <= yx1; yxi++)
		{
			const T pix_yx = T(-0.5) * lyx + (T(yxi) + T(0.5)) * dyx;
			T z0r = d1zloZ + a_z_lo * (d2zloZ - d1zloZ);
			T z1r = d1zhiZ + a_z_hi * (d2zhiZ - d1zhiZ);
			if (z0r > z1r)
				std::swap(z0r, z1r);
			const T widthFrac_z = z1r - z0r;
			const T half_dyx = dyx * T(0.5);
			const T p0 = pix_yx - half_dyx;
			const T p1 = pix_yx + half_dyx;
			if (r1 >= p0 && r0 < p1)
			{
				const T wxy = usePsf ? psf->weight(pk, p0 - r1ov, p1 - r0ov)
				                     : ddOverlap<T>(p0, p1, r0ov, r1ov);
				const T wxy_s = wxy / widthFrac_yx;
				const T z_off = T(P.nz - 1) * T(0.5);
				const T inv_dz = T(1) / dz;
				const int64_t zi0 = std::max<int64_t>(
				    0, static_cast<int64_t>(rintT<T>(z0r * inv_dz + z_off)));
				const int64_t zi1 = std::min<int64_t>(
				    P.nz - 1,
				    static_cast<int64_t>(rintT<T>(z1r * inv_dz + z_off)));
				for (int64_t zi = zi0; zi <= zi1; zi++)
				{
					const T pix_z = T(-0.5) * P.lz + (T(zi) + T(0.5)) * P.vz;
					T tw = T(1);
					if (TOF)
					{
						const T alo = (pix_xy - d1_i - T(0.5) * dxy) / (d2_i - d1_i);
						const T ahi = (pix_xy - d1_i + T(0.5) * dxy) / (d2_i - d1_i);
						tw = tof->weight(d_norm, tofv, alo * d_norm, ahi * d_norm);
					}
					const T zp0 = pix_z - P.vz * T(0.5);
					const T zp1 = pix_z + P.vz * T(0.5);
					if (z1r >= zp0 && z0r < zp1)
					{
						const T wz = ddOverlap<T>(zp0, zp1, z0r, z1r);
						const T wz_s = wz / widthFrac_z;
						int64_t io = zi * nxy;
						if (flag_y)
							io += P.nx * xyi + yxi;
						else
							io += P.nx * yxi + xyi;
						T w = wxy_s * wz_s * dxy_cos_theta;
						if (TOF)
							w *= tw;
						if (nseg)
							(*nseg)++;
						if (img)
						{
							if (FWD)
								val += w * img[frame * nvox + io];
							else
								accum(&img[frame * nvox + io], val * w, atomic);
						}
					}
				}
			}
		}
	}
}

template <class T>
struct Lor
{
	V3<T> p1, p2, n1, n2;
	T tof;
	int64_t frame;
};

template <class T>
Lor<T> getLor(const yo_lors* L, int64_t i)
{
	Lor<T> l;
	l.p1 = {T(L->p1[3 * i]), T(L->p1[3 * i + 1]), T(L->p1[3 * i + 2])};
	l.p2 = {T(L->p2[3 * i]), T(L->p2[3 * i + 1]), T(L->p2[3 * i + 2])};
	if (L->n1)
		l.n1 = {T(L->n1[3 * i]), T(L->n1[3 * i + 1]), T(L->n1[3 * i + 2])};
	else
		l.n1 = {T(0), T(0), T(0)};
	if (L->n2)
		l.n2 = {T(L->n2[3 * i]), T(L->n2[3 * i + 1]), T(L->n2[3 * i + 2])};
	else
		l.n2 = {T(0), T(0), T(0)};
	l.tof = L->tof ? T(L->tof[i]) : T(0);
	l.frame = L->frame ? L->frame[i] : 0;
	return l;
}

template <class T>
struct Projector
{
	ImgP<T> P;
	const yo_projector* pr;
	Tof<T> tof;
	Psf<T> psf;
	bool useTof;
	Projector(const yo_image* im, const yo_projector* p)
	    : P(toImg<T>(im)),
	      pr(p),
	      tof(p->use_tof ? p->tof_width_ps : 0.f, p->tof_n_std),
	      useTof(p->use_tof != 0)
	{
		psf.init(p);
	}

	// MultiRayGenerator::getRandomLine (src/geometry/MultiRayGenerator.cpp:
	// 72-114) with isParallel = hasDepth = false, the offsets being the
	// rand_r(seed=13) sequence restarted for every LOR
	// (ProjectorSiddon.cpp:258,330).
	void rayLine(const Lor<T>& l, int r, V3<T>& q1, V3<T>& q2) const
	{
		if (r == 0)
		{
			q1 = l.p1;
			q2 = l.p2;
			return;
		}
		const T tz = pr->crystal_size_z, tt = pr->crystal_size_trans;
		const float* o = pr->ray_offsets + 4 * (r - 1);
		const V3<T> vz{T(0), T(0), T(1)};
		const V3<T> t1 = normalized(cross(l.n1, vz));
		const V3<T> t2 = normalized(cross(l.n2, vz));
		q1 = l.p1 + vz * (T(o[0]) * tz) + t1 * (T(o[1]) * tt);
		q2 = (l.p2 + vz * (T(o[2]) * tz)) - t2 * (T(o[3]) * tt);
	}

	// ProjectorSiddon::forwardProjection / backProjection (lor overloads,
	// src/operators/ProjectorSiddon.cpp:239-369) and ProjectorDD
	// (src/operators/ProjectorDD.cpp:154-232).
	template <bool FWD>
	T project(T* img, const Lor<T>& l, T value, bool atomic,
	          int64_t* nseg) const
	{
		if (l.frame < 0)
			return T(0);
		const V3<T> off{P.ox, P.oy, P.oz};
		if (pr->type == 1)
		{
			const V3<T> d1 = l.p1 - off, d2 = l.p2 - off;
			T v = value;
			if (useTof)
				ddRay<T, FWD, true>(P, img, d1, d2, l.n1, l.n2, v, l.frame, &tof,
				                    l.tof, &psf, T(pr->crystal_size_z),
				                    T(pr->crystal_size_trans), atomic, nseg);
			else
				ddRay<T, FWD, false>(P, img, d1, d2, l.n1, l.n2, v, l.frame,
				                     &tof, l.tof, &psf, T(pr->crystal_size_z),
				                     T(pr->crystal_size_trans), atomic, nseg);
			return FWD ? v : T(0);
		}
		const int nr = pr->num_rays > 0 ? pr->num_rays : 1;
		T acc = T(0);
		const T perRay = (nr > 1) ? value / T(nr) : value;
		for (int r = 0; r < nr; r++)
		{
			V3<T> q1, q2;
			rayLine(l, r, q1, q2);
			q1 = q1 - off;
			q2 = q2 - off;
			T v = FWD ? T(0) : perRay;
			if (useTof)
				siddonRay<T, FWD, true>(P, img, q1, q2, v, l.frame, &tof, l.tof,
				                        atomic, nseg);
			else
				siddonRay<T, FWD, false>(P, img, q1, q2, v, l.frame, &tof,
				                         l.tof, atomic, nseg);
			if (FWD)
				acc += v;
		}
		if (FWD && nr > 1)
			acc = acc / T(nr);
		return acc;
	}
};

// util::parallelForChunked (include/yrt-pet/utils/Concurrency.hpp:42-67):
// contiguous chunks, one per thread.
template <class F>
void parallelFor(int64_t n, int nthreads, F&& fn)
{
	if (nthreads <= 1 || n < 2 * nthreads)
	{
		for (int64_t i = 0; i < n; i++)
			fn(i);
		return;
	}
	std::vector<std::thread> th;
	const int64_t chunk = (n + nthreads - 1) / nthreads;
	for (int t = 0; t < nthreads; t++)
	{
		const int64_t b = t * chunk;
		const int64_t e = std::min(n, b + chunk);
		if (b >= e)
			break;
		th.emplace_back(
		    [b, e, &fn]()
		    {
			    for (int64_t i = b; i < e; i++)
				    fn(i);
		    });
	}
	for (auto& t : th)
		t.join();
}

template <class T>
int fwdImpl(const yo_image* im, const T* img, const yo_projector* pr,
            const yo_lors* L, T* out, int nthreads)
{
	Projector<T> P(im, pr);
	parallelFor(L->n, nthreads,
	            [&](int64_t i)
	            {
		            const Lor<T> l = getLor<T>(L, i);
		            out[i] = P.template project<true>(const_cast<T*>(img), l,
		                                              T(0), false, nullptr);
	            });
	return 0;
}

template <class T>
int backImpl(const yo_image* im, T* img, const yo_projector* pr,
             const yo_lors* L, const T* vals, int nthreads)
{
	Projector<T> P(im, pr);
	const bool atomic = nthreads > 1;
	parallelFor(L->n, nthreads,
	            [&](int64_t i)
	            {
		            const Lor<T> l = getLor<T>(L, i);
		            P.template project<false>(img, l, vals[i], atomic, nullptr);
	            });
	return 0;
}

template <class T>
int emSubsetImpl(const yo_image* im, const T* x, T* u, const yo_projector* pr,
                 const yo_lors* L, const T* meas, const T* sens, const T* att,
                 const T* rnd, const T* scat, const T* pre, T g, T thr, int nthreads)
{
	Projector<T> P(im, pr);
	const bool atomic = nthreads > 1;
	// OSEM_CPU::computeEMUpdateImage lambda (src/recon/OSEM_CPU.cpp:326-374)
	parallelFor(L->n, nthreads,
	            [&](int64_t i)
	            {
		            const Lor<T> l = getLor<T>(L, i);
		            T update = P.template project<true>(const_cast<T*>(x), l,
		                                                T(0), false, nullptr);
		            if (sens)
			            update *= sens[i];
		            if (att)
			            update *= att[i];
		            update *= g;
		            if (rnd)
			            update += rnd[i];
		            if (scat)
			            update += scat[i];
		            // in-vivo attenuation precorrection (OSEM_GPU.cu:1303-1308)
		            if (pre)
			            update *= pre[i];
		            if (std::abs(update) > thr)
		            {
			            const T m = meas ? meas[i] : T(1);
			            update = m / update;
			            if (sens)
				            update *= sens[i];
			            if (att)
				            update *= att[i];
			            update *= g;
			            P.template project<false>(u, l, update, atomic, nullptr);
		            }
	            });
	return 0;
}

// Histogram3D geometry (src/datastruct/projection/Histogram3D.cpp:165-203)
struct H3
{
	int64_t numR, numPhi, numZBin, numZBinDiff, rCut, numDOIPoss;
	const yo_scanner* sc;
	explicit H3(const yo_scanner* s) : sc(s)
	{
		rCut = s->min_ang_diff / 2;
		if (s->min_ang_diff % 2 != 0 && s->dets_per_ring % 4 == 0)
			rCut++;
		numDOIPoss = s->num_doi * s->num_doi;
		numR = numDOIPoss * (s->dets_per_ring / 2 + 1 - s->min_ang_diff);
		numPhi = s->dets_per_ring;
		const int64_t dzmax = s->max_ring_diff;
		numZBin = (dzmax + 1) * s->num_rings - (dzmax * (dzmax + 1)) / 2;
		numZBinDiff = numZBin - s->num_rings;
		numZBin = numZBin + numZBinDiff;
	}
	// getDetPairInSameRing (Histogram3D.cpp:428-450)
	void sameRing(int64_t r_ring, int64_t phi, int64_t& d1r, int64_t& d2r) const
	{
		const int ntot = (int)sc->dets_per_ring;
		const int r = (int)r_ring;
		const int d01 = 0;
		int d02 = ntot / 2;
		if (phi % 2 != 0)
			d02 = ntot / 2 + 1;
		int dr1 = d01 + (r - ntot / 4 + (int)rCut);
		int dr2 = d02 - (r - ntot / 4 + (int)rCut);
		if (dr1 < 0)
			dr1 += ntot;
		if (dr2 < 0)
			dr2 += ntot;
		int64_t a = (dr1 + phi / 2) % ntot;
		int64_t b = (dr2 + phi / 2) % ntot;
		if (a > b)
			std::swap(a, b);
		d1r = a;
		d2r = b;
	}
	// getDetPairFromCoords (Histogram3D.cpp:299-363)
	void pair(int64_t bin, uint32_t& d1, uint32_t& d2) const
	{
		const int64_t zb = bin / (numPhi * numR);
		const int64_t phi = (bin % (numPhi * numR)) / numR;
		const int64_t r = (bin % (numPhi * numR)) % numR;
		const int64_t r_ring = r / numDOIPoss;
		int64_t d1r, d2r;
		sameRing(r_ring, phi, d1r, d2r);
		if (d1r == d2r)
		{
			d1 = d2 = 0;
			return;
		}
		const int64_t doi_case = r % numDOIPoss;
		const int64_t doi1 = doi_case % sc->num_doi;
		const int64_t doi2 = doi_case / sc->num_doi;
		int64_t z1, z2;
		const int64_t nr = sc->num_rings;
		if (zb < nr)
		{
			z1 = z2 = zb;
		}
		else
		{
			int64_t cur = (zb - nr) % numZBinDiff + nr;
			int64_t planes = nr;
			int64_t dz = 0;
			while (cur - planes >= 0)
			{
				cur -= planes;
				planes -= 1;
				dz++;
			}
			z1 = cur;
			z2 = z1 + dz;
			if (zb - nr >= numZBinDiff)
				std::swap(z1, z2);
		}
		const int64_t dpr = sc->dets_per_ring;
		d1 = (uint32_t)(d1r + z1 * dpr + doi1 * (dpr * nr));
		d2 = (uint32_t)(d2r + z2 * dpr + doi2 * (dpr * nr));
	}
};

// Image::updateEMThresholdStaticInternal (src/datastruct/image/Image.cpp:
// 808-833)
template <class T>
int emUpdate(T* x, const T* u, const T* s, int64_t n, T thr)
{
	for (int64_t j = 0; j < n; j++)
	{
		const T sv = s[j];
		if (sv > thr)
			x[j] *= u[j] / sv;
	}
	return 0;
}

}  // namespace

extern "C" {

// DetRegular::generateLUT (src/datastruct/scanner/DetRegular.cpp:113-182)
int yo_lut_regular(const yo_scanner* sc, float* pos, float* orient)
{
	const int64_t numBlocks = sc->dets_per_ring / sc->dets_per_block;
	const float blockLength = sc->crystal_size_trans * sc->dets_per_block;
	for (int64_t doi = 0; doi < sc->num_doi; doi++)
	{
		const float blockDistance = sc->radius + sc->crystal_depth * doi;
		for (int64_t ring = 0; ring < sc->num_rings; ring++)
		{
			const float zGap = sc->axial_fov / ((float)sc->num_rings);
			const float zPos = static_cast<float>(-sc->axial_fov / 2.0 +
			                                      (ring + 0.5f) * zGap);
			for (int64_t block = 0; block < numBlocks; block++)
			{
				// TWOPI is the literal 6.28318530717959 (Constants.hpp:16)
				const float ang = static_cast<float>(
				    double(block) * 6.28318530717959 / double(numBlocks));
				const float xB =
				    (blockDistance + 0.5f * sc->crystal_depth) * std::cos(ang);
				const float yB =
				    (blockDistance + 0.5f * sc->crystal_depth) * std::sin(ang);
				for (int64_t det = 0; det < sc->dets_per_block; det++)
				{
					float rel = 0.0f;
					if (sc->dets_per_block != 1)
						rel = static_cast<float>(
						    -0.5 + (det + 0.5) / (float)sc->dets_per_block);
					const float relX =
					    static_cast<float>(rel * std::cos(ang + kPi / 2.0));
					const float relY =
					    static_cast<float>(rel * std::sin(ang + kPi / 2.0));
					const float xD = xB + blockLength * relX;
					const float yD = yB + blockLength * relY;
					const int64_t idx = det + block * sc->dets_per_block +
					                    ring * sc->dets_per_ring +
					                    doi * (sc->num_rings * sc->dets_per_ring);
					pos[3 * idx] = xD;
					pos[3 * idx + 1] = yD;
					pos[3 * idx + 2] = zPos;
					orient[3 * idx] = std::cos(ang);
					orient[3 * idx + 1] = std::sin(ang);
					orient[3 * idx + 2] = 0.0f;
				}
			}
		}
	}
	return 0;
}

int yo_histo3d_dims(const yo_scanner* sc, int64_t* num_r, int64_t* num_phi,
                    int64_t* num_z_bin)
{
	H3 h(sc);
	*num_r = h.numR;
	*num_phi = h.numPhi;
	*num_z_bin = h.numZBin;
	return 0;
}

int yo_histo3d_pairs(const yo_scanner* sc, const int64_t* bins, int64_t n,
                     uint32_t* d1, uint32_t* d2)
{
	H3 h(sc);
	for (int64_t i = 0; i < n; i++)
		h.pair(bins[i], d1[i], d2[i]);
	return 0;
}

// Histogram3D::getCoordsFromDetPair / getBinIdFromCoords
// (src/datastruct/projection/Histogram3D.cpp:267-271,366-405); the ring map
// (setupHistogram, 286-297) is rebuilt as a dense (d1_ring, d2_ring) table.
int yo_histo3d_bins_from_pairs(const yo_scanner* sc, const uint32_t* d1in,
                               const uint32_t* d2in, int64_t n, int64_t* bins)
{
	H3 h(sc);
	const int64_t dpr = sc->dets_per_ring;
	std::vector<int64_t> ringR(dpr * dpr, -1), ringPhi(dpr * dpr, -1);
	for (int64_t phi = 0; phi < h.numPhi; phi++)
		for (int64_t rr = 0; rr < h.numR / h.numDOIPoss; rr++)
		{
			int64_t a, b;
			h.sameRing(rr, phi, a, b);
			ringR[a * dpr + b] = rr;
			ringPhi[a * dpr + b] = phi;
		}
	const int64_t nr = sc->num_rings;
	for (int64_t i = 0; i < n; i++)
	{
		uint32_t d1 = d1in[i], d2 = d2in[i];
		if (d1 > d2)
			std::swap(d1, d2);
		uint32_t r1 = d1 % dpr, r2 = d2 % dpr;
		if (r1 > r2)
		{
			std::swap(d1, d2);
			std::swap(r1, r2);
		}
		const int64_t rr = ringR[r1 * dpr + r2];
		const int64_t phi = ringPhi[r1 * dpr + r2];
		if (rr < 0)
		{
			bins[i] = -1;
			continue;
		}
		int64_t doi1 = d1 / (nr * dpr), doi2 = d2 / (nr * dpr);
		const int64_t r = rr * h.numDOIPoss + (doi1 + doi2 * sc->num_doi);
		const int64_t z1 = (d1 / dpr) % nr, z2 = (d2 / dpr) % nr;
		const int64_t dz = std::abs(z2 - z1);
		if (dz > sc->max_ring_diff)
		{
			bins[i] = -1;
			continue;
		}
		int64_t zb = dz * nr + std::min(z1, z2) - dz * (dz - 1) / 2;
		if (dz > 0 && z1 > z2)
			zb += h.numZBinDiff;
		bins[i] = zb * h.numPhi * h.numR + phi * h.numR + r;
	}
	return 0;
}

// BinIteratorRangeHistogram3D (src/datastruct/projection/BinIterator.cpp:
// 147-198)
int64_t yo_histo3d_subset_size(const yo_scanner* sc, int S, int s)
{
	H3 h(sc);
	int64_t nps = h.numPhi / S;
	if (s + nps * S < h.numPhi)
		nps += 1;
	return h.numR * nps * h.numZBin;
}

int yo_histo3d_subset_bins(const yo_scanner* sc, int S, int s, int64_t* bins)
{
	H3 h(sc);
	int64_t nps = h.numPhi / S;
	if (s + nps * S < h.numPhi)
		nps += 1;
	const int64_t size = h.numR * nps * h.numZBin;
	for (int64_t idx = 0; idx < size; idx++)
	{
		const int64_t zb = idx / (nps * h.numR);
		int64_t phi = (idx % (nps * h.numR)) / h.numR;
		const int64_t r = (idx % (nps * h.numR)) % h.numR;
		phi = S * phi + s;
		bins[idx] = zb * h.numPhi * h.numR + phi * h.numR + r;
	}
	return 0;
}

// The 4 uniform numbers per extra ray drawn by MultiRayGenerator::
// getRandomLine (src/geometry/MultiRayGenerator.cpp:78-92) from the platform
// rand_r with the per-LOR seed 13 (ProjectorSiddon.cpp:258).
int yo_multiray_offsets(int num_rays, float* out)
{
	unsigned int seed = 13;
	for (int i = 0; i < 4 * (num_rays - 1); i++)
		out[i] = static_cast<float>(rand_r(&seed)) /
		             static_cast<float>(RAND_MAX) -
		         0.5f;
	return 0;
}

// Lines 1..num_lines of the multi-ray generator for one LOR (the same code
// path as the projectors' rays 1..numRays-1), out: num_lines x 6.
int yo_multiray_lines(const float* p1, const float* p2, const float* n1,
                      const float* n2, float tz, float tt, int num_lines,
                      float* out)
{
	std::vector<float> offs(4 * num_lines);
	yo_multiray_offsets(num_lines + 1, offs.data());
	yo_projector pr{};
	pr.num_rays = num_lines + 1;
	pr.ray_offsets = offs.data();
	pr.crystal_size_z = tz;
	pr.crystal_size_trans = tt;
	yo_image im{1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 1};
	Projector<float> P(&im, &pr);
	Lor<float> l{};
	l.p1 = {p1[0], p1[1], p1[2]};
	l.p2 = {p2[0], p2[1], p2[2]};
	l.n1 = {n1[0], n1[1], n1[2]};
	l.n2 = {n2[0], n2[1], n2[2]};
	for (int r = 1; r <= num_lines; r++)
	{
		V3<float> q1, q2;
		P.rayLine(l, r, q1, q2);
		float* o = out + 6 * (r - 1);
		o[0] = q1.x;
		o[1] = q1.y;
		o[2] = q1.z;
		o[3] = q2.x;
		o[4] = q2.y;
		o[5] = q2.z;
	}
	return 0;
}

int yo_tof_helper(float width_ps, int n_std, float* sigma, float* trunc,
                  float* norm)
{
	Tof<float> t(width_ps, n_std);
	*sigma = t.sigma;
	*trunc = t.trunc;
	*norm = t.norm;
	return 0;
}

int yo_tof_alpha_range(float width_ps, int n_std, float lor_norm, float tof_ps,
                       float* amin, float* amax)
{
	Tof<float> t(width_ps, n_std);
	t.alphaRange(*amin, *amax, lor_norm, tof_ps);
	return 0;
}

float yo_tof_weight(float width_ps, int n_std, float lor_norm, float tof_ps,
                    float lo_mm, float hi_mm)
{
	Tof<float> t(width_ps, n_std);
	return t.weight(lor_norm, tof_ps, lo_mm, hi_mm);
}

float yo_psf_integral(const float* kernel, int ksize, float kstep, float x0,
                      float x1)
{
	return psfIntegral<float>(kernel, ksize, kstep, x0, x1);
}

// ProjectorDD::getOverlap (src/operators/ProjectorDD.cpp:251-258)
float yo_dd_overlap(float p0, float p1, float d0, float d1)
{
	return std::max(0.f, ddOverlap<float>(p0, p1, d0, d1));
}

int yo_get_alpha(float r0, float r1, float p1, float p2, float inv_p12,
                 float* amin, float* amax)
{
	get_alpha<float>(r0, r1, p1, p2, inv_p12, *amin, *amax);
	return 0;
}

int yo_fwd_f32(const yo_image* im, const float* img, const yo_projector* pr,
               const yo_lors* lors, float* out, int nthreads)
{
	return fwdImpl<float>(im, img, pr, lors, out, nthreads);
}
int yo_back_f32(const yo_image* im, float* img, const yo_projector* pr,
                const yo_lors* lors, const float* vals, int nthreads)
{
	return backImpl<float>(im, img, pr, lors, vals, nthreads);
}
int yo_fwd_f64(const yo_image* im, const double* img, const yo_projector* pr,
               const yo_lors* lors, double* out, int nthreads)
{
	return fwdImpl<double>(im, img, pr, lors, out, nthreads);
}
int yo_back_f64(const yo_image* im, double* img, const yo_projector* pr,
                const yo_lors* lors, const double* vals, int nthreads)
{
	return backImpl<double>(im, img, pr, lors, vals, nthreads);
}

int yo_count_segments(const yo_image* im, const yo_projector* pr,
                      const yo_lors* L, int64_t* counts, int nthreads)
{
	Projector<float> P(im, pr);
	parallelFor(L->n, nthreads,
	            [&](int64_t i)
	            {
		            const Lor<float> l = getLor<float>(L, i);
		            int64_t c = 0;
		            P.template project<true>(nullptr, l, 0.f, false, &c);
		            counts[i] = c;
	            });
	return 0;
}

int yo_em_subset_f32(const yo_image* im, const float* x, float* u,
                     const yo_projector* pr, const yo_lors* lors,
                     const float* meas, const float* sens, const float* att,
                     const float* rnd, const float* scat, const float* pre, float g,
                     float thr, int nthreads)
{
	return emSubsetImpl<float>(im, x, u, pr, lors, meas, sens, att, rnd, scat, pre,
	                           g, thr, nthreads);
}
int yo_em_subset_f64(const yo_image* im, const double* x, double* u,
                     const yo_projector* pr, const yo_lors* lors,
                     const double* meas, const double* sens, const double* att,
                     const double* rnd, const double* scat, const double* pre,
                     double g, double thr, int nthreads)
{
	return emSubsetImpl<double>(im, x, u, pr, lors, meas, sens, att, rnd, scat, pre,
	                            g, thr, nthreads);
}

int yo_em_update_f32(float* x, const float* u, const float* s, int64_t n,
                     float thr)
{
	return emUpdate<float>(x, u, s, n, thr);
}
int yo_em_update_f64(double* x, const double* u, const double* s, int64_t n,
                     double thr)
{
	return emUpdate<double>(x, u, s, n, thr);
}

int yo_apply_threshold_f32(float* x, const float* mask, int64_t n, float thr,
                           float le_scale, float le_off, float gt_scale,
                           float gt_off)
{
	for (int64_t k = 0; k < n; k++)
	{
		if (mask[k] <= thr)
			x[k] = x[k] * le_scale + le_off;
		else
			x[k] = x[k] * gt_scale + gt_off;
	}
	return 0;
}

}  // extern "C"
