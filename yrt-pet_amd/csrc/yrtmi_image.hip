// Image-space kernels for gfx950: the separable image-space PSF
// (OperatorPsf::convolve, OperatorPsf.cpp:290-375; its device form
// convolve3DSeparable_kernel, ImageSpaceKernels.cu:296-341).
//
// One launch per axis over every voxel of every frame, x fastest in the
// thread index, so that the loads of each tap are coalesced whatever the
// axis; the taps are kernel arguments (uniform: scalar loads).  The sum runs
// over kk = -h..h ascending with a separate multiply and add, the order of
// the reference's CPU loop, so the result is bitwise the same.  HBM-bound:
// 8 B per voxel per axis (the taps' neighbours hit L2).
#include "yrtmi_kernels.hpp"

namespace yrtmi
{

template <int AXIS>
__global__ void __launch_bounds__(256) kPsfAxis(const float* __restrict__ in,
                                                float* __restrict__ out, PsfTaps t,
                                                int nx, int ny, int nz, int total)
{
#pragma clang fp contract(off)
	const int h = t.n / 2;
	const int stride = (int)(gridDim.x * blockDim.x);
	for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < total; i += stride)
	{
		const int x = i % nx;
		const int r = i / nx;
		const int y = r % ny;
		const int z = (r / ny) % nz;
		// the voxel's line along AXIS: its first element and stride, the
		// coordinate along it and its length
		const int c = AXIS == 0 ? x : (AXIS == 1 ? y : z);
		const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
		const int st = AXIS == 0 ? 1 : (AXIS == 1 ? nx : nx * ny);
		const int base = i - c * st;
		float sum = 0.f;
		for (int kk = -h; kk <= h; kk++)
		{
			// util::circular (Tools.cpp:160-167): one wrap (h < n)
			int q = c - kk;
			q = q < 0 ? q + n : (q >= n ? q - n : q);
			sum = sum + t.k[kk + h] * in[base + q * st];
		}
		out[i] = sum;
	}
}

hipError_t launchPsfAxis(int axis, const float* in, float* out, const PsfTaps& t, int nx,
                         int ny, int nz, int total, hipStream_t st)
{
	if (total <= 0)
		return hipSuccess;
	const unsigned g = (unsigned)std::min(std::max((total + 255) / 256, 1), 65536);
	if (axis == 0)
		kPsfAxis<0><<<g, 256, 0, st>>>(in, out, t, nx, ny, nz, total);
	else if (axis == 1)
		kPsfAxis<1><<<g, 256, 0, st>>>(in, out, t, nx, ny, nz, total);
	else
		kPsfAxis<2><<<g, 256, 0, st>>>(in, out, t, nx, ny, nz, total);
	return hipGetLastError();
}

// Time-averaged moved image (timeAverageMoveImage_kernel<true>,
// ImageSpaceKernels.cu:191-264, with util::trilinearInterpolateCore,
// ImageUtils.cuh:15-70).  One thread per output voxel, x fastest: the voxel
// centre ((i v - L/2) + off) + v/2 (util::indexToPosition, ImageUtils.cuh:
// 96-101), then per frame in order q = inv * p as nested fmaf (the
// reference kernel's form), c = (q - origin) * (1/v), i0 = floor(c),
// d = c - i0, and the 8 corners in the reference's order, each adding
// fmaf(in[corner] * w_corner, frame_weight, acc) (corners outside the image
// weigh 0 and read voxel 0, as the reference).  The frames are uniform
// across the grid (scalar loads); the gathers of one frame from 64 adjacent
// x voxels touch two or three cache lines per corner row, so the kernel is
// bound by L2 gathers, 32 B of input per frame per voxel from cache and
// 8 B per voxel from HBM.
__global__ void __launch_bounds__(256) kTimeAverageMove(const float* __restrict__ in,
                                                        float* __restrict__ out,
                                                        const float4* __restrict__ frames,
                                                        int n, MoveGeom g)
{
#pragma clang fp contract(off)
	const int total = g.nx * g.ny * g.nz;
	const int stride = (int)(gridDim.x * blockDim.x);
	// the grid coordinate of the interpolation is clamped to [-4, n + 4]:
	// beyond it every corner is outside (weight 0) either way, and the float
	// to int conversion of floor() stays in range
	const float cmax[3] = {(float)g.nx + 4.f, (float)g.ny + 4.f, (float)g.nz + 4.f};
	for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < total; i += stride)
	{
		const int ix = i % g.nx;
		const int r = i / g.nx;
		const int iy = r % g.ny;
		const int iz = r / g.ny;
		const int id[3] = {ix, iy, iz};
		float p[3];
		for (int a = 0; a < 3; a++)
			p[a] = (((float)id[a] * g.v[a] - 0.5f * g.len[a]) + g.off[a]) + 0.5f * g.v[a];
		float acc = out[i];
		for (int f = 0; f < n; f++)
		{
			const float4 r0 = frames[4 * f], r1 = frames[4 * f + 1], r2 = frames[4 * f + 2];
			const float fw = frames[4 * f + 3].x;
			const float q[3] = {fmaf(r0.x, p[0], fmaf(r0.y, p[1], fmaf(r0.z, p[2], r0.w))),
			                    fmaf(r1.x, p[0], fmaf(r1.y, p[1], fmaf(r1.z, p[2], r1.w))),
			                    fmaf(r2.x, p[0], fmaf(r2.y, p[1], fmaf(r2.z, p[2], r2.w)))};
			int i0[3];
			float d[3];
			for (int a = 0; a < 3; a++)
			{
				const float c = fminf(fmaxf((q[a] - g.origin[a]) * g.inv[a], -4.f), cmax[a]);
				i0[a] = (int)floorf(c);
				d[a] = c - (float)i0[a];
			}
			const bool inx0 = i0[0] >= 0 && i0[0] < g.nx, inx1 = i0[0] + 1 >= 0 && i0[0] + 1 < g.nx;
			const bool iny0 = i0[1] >= 0 && i0[1] < g.ny, iny1 = i0[1] + 1 >= 0 && i0[1] + 1 < g.ny;
			const bool inz0 = i0[2] >= 0 && i0[2] < g.nz, inz1 = i0[2] + 1 >= 0 && i0[2] + 1 < g.nz;
			const int base = i0[0] + i0[1] * g.nx + i0[2] * g.nx * g.ny;
			for (int k = 0; k < 8; k++)
			{
				const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
				const float wx = ox ? d[0] : 1.f - d[0];
				const float wy = oy ? d[1] : 1.f - d[1];
				const float wz = oz ? d[2] : 1.f - d[2];
				const bool ok = (ox ? inx1 : inx0) && (oy ? iny1 : iny0) && (oz ? inz1 : inz0);
				const int idx = ok ? base + ox + oy * g.nx + oz * g.nx * g.ny : 0;
				const float w = ok ? (wx * wy) * wz : 0.f;
				acc = fmaf(in[idx] * w, fw, acc);
			}
		}
		out[i] = acc;
	}
}

hipError_t launchTimeAverageMove(const float* in, float* out, const float4* frames, int n,
                                 const MoveGeom& g, hipStream_t st)
{
	const int64_t total = (int64_t)g.nx * g.ny * g.nz;
	if (total <= 0 || n <= 0)
		return hipSuccess;
	const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
	kTimeAverageMove<<<grid, 256, 0, st>>>(in, out, frames, n, g);
	return hipGetLastError();
}

}  // namespace yrtmi
