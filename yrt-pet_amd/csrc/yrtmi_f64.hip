// FP64 Siddon projector (single and multi-ray, TOF, dynamic frames, LOR
// motion): the whole ray in double precision, for double-precision gradient
// checks of the differentiable projector and as an accuracy reference.  It
// follows ProjectorSiddon::project_helper (ProjectorSiddon.cpp:478-767)
// with every quantity in FP64, as the oracle's FP64 instantiation does;
// detector positions are the FP32 LUT values widened.
#include "yrtmi_f64.hpp"

namespace yrtmi
{

namespace
{
struct D3
{
	double x, y, z;
};

__device__ __forceinline__ D3 widen(float3 a)
{
	return D3{(double)a.x, (double)a.y, (double)a.z};
}

// normalized(cross(n, z)) in FP64 (MultiRayGenerator's transaxial direction)
__device__ __forceinline__ D3 transDir64(float3 nf)
{
#pragma clang fp contract(off)
	const D3 n = widen(nf);
	const D3 c{n.y * 1.0 - n.z * 0.0, n.z * 0.0 - n.x * 1.0, n.x * 0.0 - n.y * 0.0};
	const double nn = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
	if (nn > 10e-25)
		return D3{c.x / nn, c.y / nn, c.z / nn};
	return D3{0.0, 0.0, 0.0};
}

__device__ __forceinline__ void rayLine64(const DProj& pr, const Lor& l, int r, D3& q1,
                                          D3& q2)
{
#pragma clang fp contract(off)
	q1 = widen(l.p1);
	q2 = widen(l.p2);
	if (r == 0)
		return;
	const float* o = pr.rayOff + 4 * (r - 1);
	const D3 t1 = transDir64(l.n1), t2 = transDir64(l.n2);
	const double tz = pr.tz, tt = pr.tt;
	const double s1 = (double)o[0] * tz, s2 = (double)o[1] * tt;
	const double s3 = (double)o[2] * tz, s4 = (double)o[3] * tt;
	q1 = D3{(q1.x + 0.0 * s1) + t1.x * s2, (q1.y + 0.0 * s1) + t1.y * s2,
	        (q1.z + 1.0 * s1) + t1.z * s2};
	q2 = D3{(q2.x + 0.0 * s3) - t2.x * s4, (q2.y + 0.0 * s3) - t2.y * s4,
	        (q2.z + 1.0 * s3) - t2.z * s4};
}

__device__ __forceinline__ void getAlpha64(double r0, double r1, double p1, double p2,
                                           double inv, double& amin, double& amax)
{
#pragma clang fp contract(off)
	amin = 0.0;
	amax = 1.0;
	if (p1 != p2)
	{
		const double a0 = (r0 - p1) * inv, a1 = (r1 - p1) * inv;
		amin = a0 < a1 ? a0 : a1;
		amax = a0 < a1 ? a1 : a0;
	}
	else if (p1 < r0 || p1 > r1)
	{
		amax = 0.0;
		amin = 1.0;
	}
}

// One ray.  FWD: returns the line integral of img; else adds val * w.
template <bool FWD, bool TOF>
__device__ double siddon64(const DImage& P, const double* img, double* out, D3 p1, D3 p2,
                           double val, const DTof64& tof, float tofps)
{
#pragma clang fp contract(off)
	const double A = (p2.x - p1.x) * (p2.x - p1.x) + (p2.y - p1.y) * (p2.y - p1.y);
	const double B = 2.0 * ((p2.x - p1.x) * p1.x + (p2.y - p1.y) * p1.y);
	const double fovR = P.fovR;
	const double C = p1.x * p1.x + p1.y * p1.y - fovR * fovR;
	const double Delta = B * B - 4.0 * A * C;
	double t0 = 0.0, t1 = 1.0;
	if (A != 0.0)
	{
		if (Delta <= 0.0)
			return 0.0;
		t0 = (-B - sqrt(Delta)) / (2.0 * A);
		t1 = (-B + sqrt(Delta)) / (2.0 * A);
	}
	const double ex = p1.x - p2.x, ey = p1.y - p2.y, ez = p1.z - p2.z;
	const double dnorm = sqrt(ex * ex + ey * ey + ez * ez);
	const bool flat_x = p1.x == p2.x, flat_y = p1.y == p2.y, flat_z = p1.z == p2.z;
	const double inv_x = flat_x ? 0.0 : 1.0 / (p2.x - p1.x);
	const double inv_y = flat_y ? 0.0 : 1.0 / (p2.y - p1.y);
	const double inv_z = flat_z ? 0.0 : 1.0 / (p2.z - p1.z);
	const int dir_x = inv_x >= 0.0 ? 1 : -1, dir_y = inv_y >= 0.0 ? 1 : -1,
	          dir_z = inv_z >= 0.0 ? 1 : -1;
	const double dx = P.vx, dy = P.vy, dz = P.vz;
	const double inv_dx = 1.0 / dx, inv_dy = 1.0 / dy, inv_dz = 1.0 / dz;
	const double lx = P.lx, ly = P.ly, lz = P.lz;
	const double x0 = -lx * 0.5, x1 = lx * 0.5, y0 = -ly * 0.5, y1 = ly * 0.5,
	             z0 = -lz * 0.5, z1 = lz * 0.5;
	double axmin, axmax, aymin, aymax, azmin, azmax;
	getAlpha64(-0.5 * lx, 0.5 * lx, p1.x, p2.x, inv_x, axmin, axmax);
	getAlpha64(-0.5 * ly, 0.5 * ly, p1.y, p2.y, inv_y, aymin, aymax);
	getAlpha64(-0.5 * lz, 0.5 * lz, p1.z, p2.z, inv_z, azmin, azmax);
	double amin = fmax(fmax(fmax(fmax(0.0, t0), axmin), aymin), azmin);
	double amax = fmin(fmin(fmin(fmin(1.0, t1), axmax), aymax), azmax);
	double tmm = 0.0;
	if (TOF)
	{
		tmm = (double)tofps * 0.299792458 * 0.5;
		if (tof.trunc > 0.0)
		{
			amin = fmax(amin, fmax(0.0, 0.5 + (tmm - tof.trunc) / dnorm));
			amax = fmin(amax, fmin(1.0, 0.5 + (tmm + tof.trunc) / dnorm));
		}
	}
	double a_cur = amin;
	double x_cur = inv_x > 0.0 ? x0 : x1, y_cur = inv_y > 0.0 ? y0 : y1,
	       z_cur = inv_z > 0.0 ? z0 : z1;
	if ((inv_x >= 0.0 && p1.x > x1) || (inv_x < 0.0 && p1.x < x0) ||
	    (inv_y >= 0.0 && p1.y > y1) || (inv_y < 0.0 && p1.y < y0) ||
	    (inv_z >= 0.0 && p1.z > z1) || (inv_z < 0.0 && p1.z < z0))
		return 0.0;
	const double DMAX = 1.7976931348623157e308;
	double axn = flat_x ? DMAX : axmin, ayn = flat_y ? DMAX : aymin,
	       azn = flat_z ? DMAX : azmin;
	if (!flat_x)
	{
		const long kx = (long)ceil((double)dir_x * (a_cur * (p2.x - p1.x) - x_cur + p1.x) / dx);
		x_cur += (double)(kx * dir_x) * dx;
		axn = (x_cur - p1.x) * inv_x;
	}
	if (!flat_y)
	{
		const long ky = (long)ceil((double)dir_y * (a_cur * (p2.y - p1.y) - y_cur + p1.y) / dy);
		y_cur += (double)(ky * dir_y) * dy;
		ayn = (y_cur - p1.y) * inv_y;
	}
	if (!flat_z)
	{
		const long kz = (long)ceil((double)dir_z * (a_cur * (p2.z - p1.z) - z_cur + p1.z) / dz);
		z_cur += (double)(kz * dir_z) * dz;
		azn = (z_cur - p1.z) * inv_z;
	}
	bool first = true;
	int vx = -1, vy = -1, vz = -1, dprev = -1;
	long off = 0;
	double axp = axn, ayp = ayn, azp = azn, acc = 0.0;
	const double pc = 0.5 * dnorm + tmm;
	while (a_cur < amax)
	{
		int dn = 0;
		double a_next = -1.0;
		if (axp <= ayp && axp <= azp)
		{
			a_next = axn;
			x_cur += (double)dir_x * dx;
			axn = (x_cur - p1.x) * inv_x;
			dn |= 1;
		}
		if (ayp <= axp && ayp <= azp)
		{
			a_next = ayn;
			y_cur += (double)dir_y * dy;
			ayn = (y_cur - p1.y) * inv_y;
			dn |= 2;
		}
		if (azp <= axp && azp <= ayp)
		{
			a_next = azn;
			z_cur += (double)dir_z * dz;
			azn = (z_cur - p1.z) * inv_z;
			dn |= 4;
		}
		if (a_next > amax)
			a_next = amax;
		if (a_cur >= a_next)
		{
			axp = axn;
			ayp = ayn;
			azp = azn;
			continue;
		}
		if (first)
		{
			const double am = 0.5 * (a_cur + a_next);
			vx = (int)((p1.x + am * (p2.x - p1.x) + lx / 2.0) * inv_dx);
			vy = (int)((p1.y + am * (p2.y - p1.y) + ly / 2.0) * inv_dy);
			vz = (int)((p1.z + am * (p2.z - p1.z) + lz / 2.0) * inv_dz);
			off = (long)vz * P.nxy + (long)vy * P.nx;
			first = false;
			if (vx < 0 || vx >= P.nx || vy < 0 || vy >= P.ny || vz < 0 || vz >= P.nz)
				break;
		}
		else
		{
			if (dprev & 1)
			{
				vx += dir_x;
				if (vx < 0 || vx >= P.nx)
					break;
			}
			if (dprev & 2)
			{
				vy += dir_y;
				if (vy < 0 || vy >= P.ny)
					break;
				off += dir_y * P.nx;
			}
			if (dprev & 4)
			{
				vz += dir_z;
				if (vz < 0 || vz >= P.nz)
					break;
				off += (long)dir_z * P.nxy;
			}
		}
		dprev = dn;
		double w = (a_next - a_cur) * dnorm;
		if (TOF)
		{
			const double xc = (0.5 * (a_cur * dnorm + a_next * dnorm) - pc) / tof.sigma;
			w *= exp(-0.5 * xc * xc) * tof.norm;
		}
		const long idx = off + vx;
		if (FWD)
			acc += w * img[idx];
		else
			unsafeAtomicAdd(out + idx, val * w);
		a_cur = a_next;
		axp = axn;
		ayp = ayn;
		azp = azn;
	}
	return acc;
}

template <bool FWD, bool TOF>
__global__ void __launch_bounds__(256) kProj64(KArgs a, const double* __restrict__ x,
                                               double* __restrict__ y,
                                               const double* __restrict__ vals,
                                               double* __restrict__ out)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	const int R = a.pr.numRays;
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.L.n; e += stride)
	{
		const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
		if (l.frame < 0)
		{
			if (FWD)
				out[e] = 0.0;
			continue;
		}
		const D3 off{(double)a.im.ox, (double)a.im.oy, (double)a.im.oz};
		const int64_t fo = (int64_t)l.frame * a.im.nvox;
		const double v = FWD ? 0.0 : ((R > 1) ? vals[e] / (double)R : vals[e]);
		double s = 0.0;
		for (int r = 0; r < R; r++)
		{
			D3 q1, q2;
			rayLine64(a.pr, l, r, q1, q2);
			q1 = D3{q1.x - off.x, q1.y - off.y, q1.z - off.z};
			q2 = D3{q2.x - off.x, q2.y - off.y, q2.z - off.z};
			s += siddon64<FWD, TOF>(a.im, FWD ? x + fo : nullptr, FWD ? nullptr : y + fo, q1,
			                        q2, v, a.pr.tof64, l.tof);
		}
		if (FWD)
			out[e] = (R > 1) ? s / (double)R : s;
	}
}
}  // namespace

hipError_t launchProj64(bool fwd, const KArgs& a, const double* x, double* y,
                        const double* vals, double* out, hipStream_t st)
{
	int64_t nb = (a.L.n + 255) / 256;
	if (nb > 65536)
		nb = 65536;
	if (nb < 1)
		nb = 1;
	const bool tof = a.pr.useTof != 0;
	if (fwd)
	{
		if (tof)
			kProj64<true, true><<<(unsigned)nb, 256, 0, st>>>(a, x, y, vals, out);
		else
			kProj64<true, false><<<(unsigned)nb, 256, 0, st>>>(a, x, y, vals, out);
	}
	else
	{
		if (tof)
			kProj64<false, true><<<(unsigned)nb, 256, 0, st>>>(a, x, y, vals, out);
		else
			kProj64<false, false><<<(unsigned)nb, 256, 0, st>>>(a, x, y, vals, out);
	}
	return hipGetLastError();
}

}  // namespace yrtmi
