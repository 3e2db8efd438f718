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

}  // namespace yrtmi
