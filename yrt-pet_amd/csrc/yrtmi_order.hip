// Locality order of a LOR set, and per-LOR array permutation.
//
// The EM update of a subset does not depend on the order in which its LORs
// are projected (OSEM_CPU.cpp:329-374 sums their backprojections; the order
// changes FP32 rounding only).  A list-mode subset holds its events in
// acquisition order, so the lines that concurrently running waves walk are
// unrelated and every voxel gather misses L2.  Sorted by the key below,
// consecutive LORs lie in the same thin slab of the image (same axial
// midpoint and tilt) at neighbouring transaxial angles and offsets, and the
// voxels a window of concurrent waves reads fit in L2.
#include "yrtmi_order.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

namespace yrtmi
{

// Key of one LOR, most significant first: dynamic frame (16 bits), axial
// midpoint as a voxel slice of the image (10), axial tilt (10), transaxial
// angle in [0, pi) (12), signed transaxial offset from the image axis (12).
// Skipped LORs (frame < 0) get the largest key.
__device__ __forceinline__ uint64_t lorKey(const DImage& im, const Lor& l)
{
	if (l.frame < 0)
		return ~0ull;
	const float kPi = 3.14159265f;
	float dx = l.p2.x - l.p1.x, dy = l.p2.y - l.p1.y, dz = l.p2.z - l.p1.z;
	if (dy < 0.f || (dy == 0.f && dx < 0.f))
	{
		dx = -dx;
		dy = -dy;
		dz = -dz;
	}
	const float txy = sqrtf(dx * dx + dy * dy);
	const float mx = 0.5f * (l.p1.x + l.p2.x) - im.ox;
	const float my = 0.5f * (l.p1.y + l.p2.y) - im.oy;
	const float mz = 0.5f * (l.p1.z + l.p2.z) - im.oz;
	const float phi = atan2f(dy, dx);                       // [0, pi]
	const float tilt = atan2f(dz, txy);                     // [-pi/2, pi/2]
	const float s = (txy > 0.f) ? (mx * dy - my * dx) / txy : 0.f;
	const float rmax = 0.5f * sqrtf(im.lx * im.lx + im.ly * im.ly);
	auto q = [](float u, int levels) {  // u in [0, 1) -> [0, levels)
		const int i = (int)floorf(u * (float)levels);
		return (uint64_t)min(max(i, 0), levels - 1);
	};
	const uint64_t fq = (uint64_t)min(l.frame, 65535);
	const uint64_t zq = q((mz + 0.5f * im.lz) / im.lz * (float)im.nz / 1024.f, 1024);
	const uint64_t tq = q(tilt / kPi + 0.5f, 1024);
	const uint64_t pq = q(phi / kPi, 4096);
	const uint64_t sq = q(0.5f + 0.5f * s / rmax, 4096);
	return (fq << 44) | (zq << 34) | (tq << 24) | (pq << 12) | sq;
}

__global__ void __launch_bounds__(256)
    kLorKey(DScanner S, DLors L, DImage im, uint64_t* __restrict__ key,
            uint32_t* __restrict__ idx)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < L.n; e += stride)
	{
		key[e] = lorKey(im, fetchLor(S, L, e, im.nt));
		idx[e] = (uint32_t)e;
	}
}

template <class T>
__global__ void __launch_bounds__(256)
    kPermute(const T* __restrict__ src, const uint32_t* __restrict__ perm, int64_t n,
             T* __restrict__ dst)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		dst[i] = src[perm[i]];
}

static unsigned gridFor(int64_t n)
{
	int64_t nb = (n + 255) / 256;
	return (unsigned)std::min<int64_t>(std::max<int64_t>(nb, 1), 65536);
}

hipError_t lorOrder(const DScanner& S, const DLors& L, const DImage& im, uint32_t* perm,
                    hipStream_t st)
{
	const int64_t n = L.n;
	if (n == 0)
		return hipSuccess;
	uint64_t *kin = nullptr, *kout = nullptr;
	uint32_t* iin = nullptr;
	void* tmp = nullptr;
	size_t tmpBytes = 0;
	hipError_t e = hipMalloc(&kin, sizeof(uint64_t) * n);
	if (e == hipSuccess)
		e = hipMalloc(&kout, sizeof(uint64_t) * n);
	if (e == hipSuccess)
		e = hipMalloc(&iin, sizeof(uint32_t) * n);
	if (e == hipSuccess)
	{
		kLorKey<<<gridFor(n), 256, 0, st>>>(S, L, im, kin, iin);
		e = hipGetLastError();
	}
	if (e == hipSuccess)
		e = rocprim::radix_sort_pairs(nullptr, tmpBytes, kin, kout, iin, perm, (size_t)n, 0,
		                              60, st);
	if (e == hipSuccess)
		e = hipMalloc(&tmp, std::max<size_t>(tmpBytes, 1));
	if (e == hipSuccess)
		e = rocprim::radix_sort_pairs(tmp, tmpBytes, kin, kout, iin, perm, (size_t)n, 0, 60,
		                              st);
	if (e == hipSuccess)
		e = hipStreamSynchronize(st);
	// hipFree waits for the device; the buffers are released on every path
	(void)hipFree(kin);
	(void)hipFree(kout);
	(void)hipFree(iin);
	(void)hipFree(tmp);
	return e;
}

hipError_t launchPermute(const void* src, int elemBytes, const uint32_t* perm, int64_t n,
                         void* dst, hipStream_t st)
{
	if (n == 0)
		return hipSuccess;
	const unsigned g = gridFor(n);
	switch (elemBytes)
	{
	case 1:
		kPermute<uint8_t><<<g, 256, 0, st>>>((const uint8_t*)src, perm, n, (uint8_t*)dst);
		break;
	case 2:
		kPermute<uint16_t><<<g, 256, 0, st>>>((const uint16_t*)src, perm, n, (uint16_t*)dst);
		break;
	case 4:
		kPermute<uint32_t><<<g, 256, 0, st>>>((const uint32_t*)src, perm, n, (uint32_t*)dst);
		break;
	default:
		kPermute<uint64_t><<<g, 256, 0, st>>>((const uint64_t*)src, perm, n, (uint64_t*)dst);
		break;
	}
	return hipGetLastError();
}

}  // namespace yrtmi
