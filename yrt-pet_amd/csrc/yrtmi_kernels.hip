// yrtmi kernels for gfx950: one event (or histogram bin) per group of
// `LPE` lanes, the lanes splitting the rays of multi-ray Siddon and
// reducing the forward sum with cross-lane shuffles.  Events stream from HBM
// as SoA (d1, d2, tof, frame); detector positions come from the resident
// LUT; images are FP32 float[nt][nz][ny][nx].
#include "yrtmi_kernels.hpp"

namespace yrtmi
{

template <int LPE>
__device__ __forceinline__ float groupSum(float v)
{
#pragma unroll
	for (int o = 1; o < LPE; o <<= 1)
		v += __shfl_xor(v, o, LPE);
	return v;
}

// Forward projection (MODE_FWD) or segment count (MODE_COUNT, `img`
// unused) of the rays of one event handled by this lane (rays lane,
// lane+LPE, ...), Siddon or DD.  Returns this lane's partial sum.
template <int TYPE, bool TOF, int LPE, bool S64, int FMODE = MODE_FWD>
__device__ __forceinline__ float fwdLane(const KArgs& a, const Lor& l,
                                         int lane, const float* img,
                                         uint32_t& cnt)
{
	static_assert(FMODE == MODE_FWD || FMODE == MODE_COUNT, "forward modes only");
	const float3 off = make_float3(a.im.ox, a.im.oy, a.im.oz);
	if (TYPE == 1)
	{
		if (lane != 0)
			return 0.f;
		const float3 d1 = make_float3(l.p1.x - off.x, l.p1.y - off.y, l.p1.z - off.z);
		const float3 d2 = make_float3(l.p2.x - off.x, l.p2.y - off.y, l.p2.z - off.z);
		return ddRay<FMODE, TOF, S64>(a.im, const_cast<float*>(img), d1, d2, l.n1,
		                         l.n2, 0.f, a.pr, l.tof, cnt);
	}
	float s = 0.f;
	for (int r = lane; r < a.pr.numRays; r += LPE)
	{
		float3 q1, q2;
		rayLine(a.pr, l, r, q1, q2);
		q1 = make_float3(q1.x - off.x, q1.y - off.y, q1.z - off.z);
		q2 = make_float3(q2.x - off.x, q2.y - off.y, q2.z - off.z);
		s += siddonRay<FMODE, TOF, S64>(a.im, const_cast<float*>(img), q1, q2,
		                           0.f, a.pr.tof, l.tof, cnt);
	}
	return s;
}

template <int TYPE, bool TOF, int LPE, bool S64, int MODE>
__device__ __forceinline__ void backLane(const KArgs& a, const Lor& l,
                                         int lane, float* img, float v,
                                         uint32_t& cnt)
{
	const float3 off = make_float3(a.im.ox, a.im.oy, a.im.oz);
	if (TYPE == 1)
	{
		if (lane != 0)
			return;
		const float3 d1 = make_float3(l.p1.x - off.x, l.p1.y - off.y, l.p1.z - off.z);
		const float3 d2 = make_float3(l.p2.x - off.x, l.p2.y - off.y, l.p2.z - off.z);
		ddRay<MODE, TOF, S64>(a.im, img, d1, d2, l.n1, l.n2, v, a.pr, l.tof, cnt);
		return;
	}
	// ProjectorSiddon::backProjection splits the value over the rays
	// (ProjectorSiddon.cpp:321-328).
	const float vr = (a.pr.numRays > 1) ? v / (float)a.pr.numRays : v;
	for (int r = lane; r < a.pr.numRays; r += LPE)
	{
		float3 q1, q2;
		rayLine(a.pr, l, r, q1, q2);
		q1 = make_float3(q1.x - off.x, q1.y - off.y, q1.z - off.z);
		q2 = make_float3(q2.x - off.x, q2.y - off.y, q2.z - off.z);
		siddonRay<MODE, TOF, S64>(a.im, img, q1, q2, vr, a.pr.tof, l.tof, cnt);
	}
}

template <int MODE, int TYPE, bool TOF, int LPE, bool S64>
__global__ void __launch_bounds__(256) kEvents(KArgs a)
{
	const int64_t total = a.L.n * LPE;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	uint32_t cnt = 0;
	// backprojection target: this XCD's private partial image when enabled
	float* ydst = a.y;
	if (MODE == MODE_BACK || MODE == MODE_EM || MODE == MODE_SENS)
	{
		if (a.part != nullptr)
		{
			const int k = min((int)xccId(), a.nxcc - 1);
			ydst = a.part + (int64_t)k * a.partStride;
		}
	}
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
	     t += stride)
	{
		const int64_t e = t / LPE;
		const int lane = (int)(t & (LPE - 1));
		const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
		const int64_t fo = (int64_t)l.frame * a.im.nvox;
		if (MODE == MODE_FWD)
		{
			float s = 0.f;
			if (l.frame >= 0)
				s = fwdLane<TYPE, TOF, LPE, S64>(a, l, lane, a.x + fo, cnt);
			s = groupSum<LPE>(s);
			if (lane == 0)
				a.out[e] = (TYPE == 0 && a.pr.numRays > 1)
				               ? s / (float)a.pr.numRays
				               : s;
		}
		else if (MODE == MODE_BACK)
		{
			if (l.frame >= 0)
				backLane<TYPE, TOF, LPE, S64, MODE_BACK>(a, l, lane, ydst + fo,
				                                    a.vals[e], cnt);
		}
		else if (MODE == MODE_COUNT)
		{
			if (l.frame >= 0)
				fwdLane<TYPE, TOF, LPE, S64, MODE_COUNT>(a, l, lane, nullptr, cnt);
		}
		else if (MODE == MODE_EM)
		{
			// computeEMUpdateImage (OSEM_CPU.cpp:329-374); skipping zero
			// measurement / zero factor / negative frame as the GPU
			// reference does (OSEM_GPU.cu:1222-1225) contributes exactly
			// what the CPU path backprojects (a zero value).
			const float m = a.c.meas ? a.c.meas[e] : 1.0f;
			const float sf = a.c.sens ? a.c.sens[e] : 1.0f;
			const float af = a.c.att ? a.c.att[e] : 1.0f;
			if (m == 0.f || sf == 0.f || af == 0.f || l.frame < 0)
				continue;
			float s = fwdLane<TYPE, TOF, LPE, S64>(a, l, lane, a.x + fo, cnt);
			s = groupSum<LPE>(s);
			float upd = (TYPE == 0 && a.pr.numRays > 1) ? s / (float)a.pr.numRays : s;
			if (a.c.sens)
				upd *= sf;
			if (a.c.att)
				upd *= af;
			upd *= a.c.g;
			if (a.c.randoms)
				upd += a.c.randoms[e];
			if (a.c.scatter)
				upd += a.c.scatter[e];
			if (a.c.precorr)
				upd *= a.c.precorr[e];
			if (fabsf(upd) > a.c.denomThr)
			{
				float v = m / upd;
				if (a.c.sens)
					v *= sf;
				if (a.c.att)
					v *= af;
				v *= a.c.g;
				backLane<TYPE, TOF, LPE, S64, MODE_BACK>(a, l, lane, ydst + fo, v, cnt);
			}
		}
		else if (MODE == MODE_SENS)
		{
			// generateSensImage (OSEM_CPU.cpp:146-156, Corrector_CPU.cpp:
			// 17-45): value = acf * sens; acf from a single-ray Siddon of the
			// mu image without the image offset (Corrector.cpp:601-611).
			float v = a.vals ? a.vals[e] : 1.0f;
			if (a.att)
			{
				uint32_t c2 = 0;
				const float mu = siddonRay<MODE_FWD, false, S64>(
				    a.attIm, const_cast<float*>(a.att), l.p1, l.p2, 0.f,
				    a.pr.tof, 0.f, c2);
				v *= expf(-mu * 0.1f);
			}
			if (l.frame >= 0)
				backLane<TYPE, false, LPE, S64, MODE_BACK>(a, l, lane, ydst, v, cnt);
		}
	}
	if (MODE == MODE_COUNT)
	{
		unsigned long long c = cnt;
#pragma unroll
		for (int o = 32; o >= 1; o >>= 1)
			c += __shfl_xor(c, o);
		if ((threadIdx.x & 63) == 0 && c != 0)
			atomicAdd(a.counter, c);
	}
}

// Per-LOR attenuation factor exp(-0.1 * Siddon(mu)) (Corrector.cpp:601-611).
__global__ void __launch_bounds__(256) kAttFactors(KArgs a)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.L.n;
	     e += stride)
	{
		const Lor l = fetchLor(a.sc, a.L, e, 1);
		uint32_t c2 = 0;
		float mu = 0.f;
		if (l.frame >= 0)
			mu = siddonRay<MODE_FWD, false, false>(a.attIm, const_cast<float*>(a.att),
			                                l.p1, l.p2, 0.f, a.pr.tof, 0.f, c2);
		a.out[e] = expf(-mu * 0.1f);
	}
}

// updateEM(Dynamic)_kernel (ImageSpaceKernels.cu:13-66): x *= u / s where
// s > thr, s broadcast over frames when sFrames == 1.
__global__ void __launch_bounds__(256) kEmUpdate(float* __restrict__ x,
                                                 const float* __restrict__ u,
                                                 const float* __restrict__ s,
                                                 int64_t nvox, int64_t nt,
                                                 int64_t sFrames, float thr)
{
	const int64_t n = nvox * nt;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
	{
		const float sv = s[sFrames == 1 ? i % nvox : i];
		if (sv > thr)
			x[i] *= u[i] / sv;
	}
}

// applyThreshold(Broadcast)_kernel (ImageSpaceKernels.cu:79-130).
__global__ void __launch_bounds__(256) kThreshold(float* __restrict__ x,
                                                  const float* __restrict__ m,
                                                  int64_t nvox, int64_t nt,
                                                  int64_t mFrames, float thr,
                                                  float les, float leo,
                                                  float gts, float gto)
{
	const int64_t n = nvox * nt;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
	{
		const float mv = m[mFrames == 1 ? i % nvox : i];
		x[i] = (mv <= thr) ? x[i] * les + leo : x[i] * gts + gto;
	}
}

__global__ void __launch_bounds__(256) kFill(float* __restrict__ x, int64_t n,
                                             float v)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
		x[i] = v;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
template <int MODE, int TYPE, bool TOF, bool S64>
static hipError_t launchLpe(const KArgs& a, int lpe, dim3 g, dim3 b,
                            hipStream_t st)
{
	if (TYPE == 1)
	{
		// DD: one lane per event
		kEvents<MODE, TYPE, TOF, 1, S64><<<g, b, 0, st>>>(a);
		return hipGetLastError();
	}
	switch (lpe)
	{
	case 1:
		kEvents<MODE, TYPE, TOF, 1, S64><<<g, b, 0, st>>>(a);
		break;
	case 2:
		kEvents<MODE, TYPE, TOF, 2, S64><<<g, b, 0, st>>>(a);
		break;
	case 4:
		kEvents<MODE, TYPE, TOF, 4, S64><<<g, b, 0, st>>>(a);
		break;
	case 8:
		kEvents<MODE, TYPE, TOF, 8, S64><<<g, b, 0, st>>>(a);
		break;
	default:
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

template <int MODE, int TYPE, bool TOF>
static hipError_t launchSetup(const KArgs& a, int lpe, dim3 g, dim3 b,
                              hipStream_t st)
{
	return a.pr.setup64 ? launchLpe<MODE, TYPE, TOF, true>(a, lpe, g, b, st)
	                    : launchLpe<MODE, TYPE, TOF, false>(a, lpe, g, b, st);
}

template <int MODE>
static hipError_t launchMode(const KArgs& a, int lpe, dim3 g, dim3 b,
                             hipStream_t st)
{
	const bool tof = a.pr.useTof != 0 && MODE != MODE_SENS;
	if (a.pr.type == 1)
		return tof ? launchSetup<MODE, 1, true>(a, 1, g, b, st)
		           : launchSetup<MODE, 1, false>(a, 1, g, b, st);
	return tof ? launchSetup<MODE, 0, true>(a, lpe, g, b, st)
	           : launchSetup<MODE, 0, false>(a, lpe, g, b, st);
}

hipError_t launchEvents(int mode, const KArgs& a, int lpe, int block,
                        int64_t grid, hipStream_t st)
{
	if (a.pr.type == 1)
		lpe = 1;
	const int64_t threads = a.L.n * lpe;
	int64_t nb = (threads + block - 1) / block;
	if (grid > 0 && grid < nb)
		nb = grid;
	if (nb < 1)
		nb = 1;
	if (nb > 0x7fffffff)
		nb = 0x7fffffff;
	const dim3 g((unsigned)nb), b((unsigned)block);
	switch (mode)
	{
	case MODE_FWD:
		return launchMode<MODE_FWD>(a, lpe, g, b, st);
	case MODE_BACK:
		return launchMode<MODE_BACK>(a, lpe, g, b, st);
	case MODE_EM:
		return launchMode<MODE_EM>(a, lpe, g, b, st);
	case MODE_SENS:
		return launchMode<MODE_SENS>(a, lpe, g, b, st);
	case MODE_COUNT:
		return launchMode<MODE_COUNT>(a, lpe, g, b, st);
	default:
		return hipErrorInvalidValue;
	}
}

static dim3 gridFor(int64_t n)
{
	int64_t nb = (n + 255) / 256;
	if (nb > 65536)
		nb = 65536;
	if (nb < 1)
		nb = 1;
	return dim3((unsigned)nb);
}

hipError_t launchAttFactors(const KArgs& a, hipStream_t st)
{
	kAttFactors<<<gridFor(a.L.n), 256, 0, st>>>(a);
	return hipGetLastError();
}

hipError_t launchEmUpdate(float* x, const float* u, const float* s,
                          int64_t nvox, int64_t nt, int64_t sFrames, float thr,
                          hipStream_t st)
{
	kEmUpdate<<<gridFor(nvox * nt), 256, 0, st>>>(x, u, s, nvox, nt, sFrames,
	                                              thr);
	return hipGetLastError();
}

hipError_t launchThreshold(float* x, const float* m, int64_t nvox, int64_t nt,
                           int64_t mFrames, float thr, float les, float leo,
                           float gts, float gto, hipStream_t st)
{
	kThreshold<<<gridFor(nvox * nt), 256, 0, st>>>(x, m, nvox, nt, mFrames,
	                                               thr, les, leo, gts, gto);
	return hipGetLastError();
}

// y[j] += sum_k part[k][j]; part[k][j] = 0 (ready for the next
// backprojection).  float4 over n4 = n/4 quads, scalar tail.
__global__ void __launch_bounds__(256) kFoldPartials(float* __restrict__ y,
                                                     float* __restrict__ part,
                                                     int64_t stride, int nk,
                                                     int64_t n, int vec)
{
	const int64_t n4 = vec ? (n >> 2) : 0;
	const int64_t gs = (int64_t)gridDim.x * blockDim.x;
	const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
	for (int64_t i = t0; i < n4; i += gs)
	{
		float4 s = z;
		for (int k = 0; k < nk; k++)
		{
			float4* p = reinterpret_cast<float4*>(part + k * stride) + i;
			const float4 v = *p;
			s.x += v.x;
			s.y += v.y;
			s.z += v.z;
			s.w += v.w;
			*p = z;
		}
		float4* q = reinterpret_cast<float4*>(y) + i;
		float4 o = *q;
		o.x += s.x;
		o.y += s.y;
		o.z += s.z;
		o.w += s.w;
		*q = o;
	}
	for (int64_t i = 4 * n4 + t0; i < n; i += gs)
	{
		float s = 0.f;
		for (int k = 0; k < nk; k++)
		{
			s += part[k * stride + i];
			part[k * stride + i] = 0.f;
		}
		y[i] += s;
	}
}

hipError_t launchFoldPartials(float* y, float* part, int64_t stride, int nk,
                              int64_t n, hipStream_t st)
{
	// float4 path needs 16-byte aligned rows (stride is a multiple of 64)
	const int vec = ((uintptr_t)y % 16 == 0 && (uintptr_t)part % 16 == 0 &&
	                 stride % 4 == 0) ? 1 : 0;
	kFoldPartials<<<gridFor(vec ? (n + 3) / 4 : n), 256, 0, st>>>(y, part, stride,
	                                                               nk, n, vec);
	return hipGetLastError();
}

// Element i of histogram subset s of S (BinIteratorRangeHistogram3D::get,
// BinIterator.cpp:191-197): r fastest, then the subset's phis, then z bins.
__device__ __forceinline__ uint32_t subsetBin(uint32_t i, const SubsetGeom& g)
{
	const uint32_t per = g.nps * g.numR;
	const uint32_t zb = i / per;
	const uint32_t rem = i - zb * per;
	const uint32_t ps = rem / g.numR;
	const uint32_t r = rem - ps * g.numR;
	return (zb * g.numPhi + g.S * ps + g.s) * g.numR + r;
}

__global__ void __launch_bounds__(256) kCountNonzero(const float* __restrict__ v,
                                                     SubsetGeom g, int64_t n,
                                                     int* __restrict__ blockCount)
{
	__shared__ int wc[4];
	int c = 0;
	for (int j = 0; j < kCompactRounds; j++)
	{
		const int64_t i = (int64_t)blockIdx.x * kCompactPerBlock + j * 256 + threadIdx.x;
		if (i < n && v[subsetBin((uint32_t)i, g)] != 0.f)
			c++;
	}
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1)
		c += __shfl_xor(c, o);
	if ((threadIdx.x & 63) == 0)
		wc[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0)
		blockCount[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ void __launch_bounds__(256) kWriteNonzero(const float* __restrict__ v,
                                                     SubsetGeom g, int64_t n,
                                                     const int64_t* __restrict__ blockOff,
                                                     uint32_t* __restrict__ bins,
                                                     float* __restrict__ meas)
{
	__shared__ int wc[4];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	int64_t run = blockOff[blockIdx.x];
	for (int j = 0; j < kCompactRounds; j++)
	{
		const int64_t i = (int64_t)blockIdx.x * kCompactPerBlock + j * 256 + threadIdx.x;
		uint32_t bin = 0;
		float val = 0.f;
		if (i < n)
		{
			bin = subsetBin((uint32_t)i, g);
			val = v[bin];
		}
		const bool keep = val != 0.f;
		const unsigned long long m = __ballot(keep);
		if (lane == 0)
			wc[w] = __popcll(m);
		__syncthreads();
		int before = 0, total = 0;
#pragma unroll
		for (int k = 0; k < 4; k++)
		{
			before += (k < w) ? wc[k] : 0;
			total += wc[k];
		}
		if (keep)
		{
			const int64_t o = run + before + __popcll(m & ((1ull << lane) - 1ull));
			bins[o] = bin;
			meas[o] = val;
		}
		run += total;
		__syncthreads();
	}
}

hipError_t launchCountNonzero(const float* v, const SubsetGeom& g, int64_t n,
                              int* blockCount, int64_t nBlocks, hipStream_t st)
{
	kCountNonzero<<<(unsigned)nBlocks, 256, 0, st>>>(v, g, n, blockCount);
	return hipGetLastError();
}

hipError_t launchWriteNonzero(const float* v, const SubsetGeom& g, int64_t n,
                              const int64_t* blockOff, int64_t nBlocks,
                              uint32_t* bins, float* meas, hipStream_t st)
{
	kWriteNonzero<<<(unsigned)nBlocks, 256, 0, st>>>(v, g, n, blockOff, bins, meas);
	return hipGetLastError();
}

__device__ __forceinline__ float maybeInvert(float v, int invert)
{
	return (invert && v != 0.0f) ? 1.0f / v : v;
}

__global__ void __launch_bounds__(256) kGather(const float* __restrict__ src,
                                               int64_t srcN,
                                               const uint32_t* __restrict__ idx,
                                               int64_t n, int invert,
                                               float* __restrict__ out)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
	{
		const uint32_t k = idx[i];
		out[i] = maybeInvert(((int64_t)k < srcN) ? src[k] : 0.f, invert);
	}
}

hipError_t launchGather(const float* src, int64_t srcN, const uint32_t* idx,
                        int64_t n, int invert, float* out, hipStream_t st)
{
	kGather<<<gridFor(n), 256, 0, st>>>(src, srcN, idx, n, invert, out);
	return hipGetLastError();
}

__global__ void __launch_bounds__(256) kSubsetGather(const float* __restrict__ h,
                                                     SubsetGeom g, int64_t n,
                                                     int invert,
                                                     float* __restrict__ out)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
		out[i] = maybeInvert(h[subsetBin((uint32_t)i, g)], invert);
}

hipError_t launchSubsetGather(const float* h, const SubsetGeom& g, int64_t n,
                              int invert, float* out, hipStream_t st)
{
	kSubsetGather<<<gridFor(n), 256, 0, st>>>(h, g, n, invert, out);
	return hipGetLastError();
}

// x and y may alias.
__global__ void __launch_bounds__(256) kAxpby(int64_t n, float a, const float* x,
                                              float b, float* y)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
		y[i] = (x != nullptr) ? a * x[i] + b * y[i] : b * y[i];
}

hipError_t launchAxpby(int64_t n, float a, const float* x, float b, float* y,
                       hipStream_t st)
{
	kAxpby<<<gridFor(n), 256, 0, st>>>(n, a, x, b, y);
	return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) kInterleaved(const T* __restrict__ src,
                                                    int64_t n, int64_t S, int64_t s,
                                                    T* __restrict__ dst)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
	     i += stride)
		dst[i] = src[s + i * S];
}

hipError_t launchInterleaved(const void* src, int elemBytes, int64_t n, int64_t S,
                             int64_t s, void* dst, hipStream_t st)
{
	switch (elemBytes)
	{
	case 1:
		kInterleaved<<<gridFor(n), 256, 0, st>>>((const uint8_t*)src, n, S, s, (uint8_t*)dst);
		break;
	case 2:
		kInterleaved<<<gridFor(n), 256, 0, st>>>((const uint16_t*)src, n, S, s, (uint16_t*)dst);
		break;
	case 4:
		kInterleaved<<<gridFor(n), 256, 0, st>>>((const uint32_t*)src, n, S, s, (uint32_t*)dst);
		break;
	case 8:
		kInterleaved<<<gridFor(n), 256, 0, st>>>((const uint64_t*)src, n, S, s, (uint64_t*)dst);
		break;
	default:
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t launchFill(float* x, int64_t n, float v, hipStream_t st)
{
	kFill<<<gridFor(n), 256, 0, st>>>(x, n, v);
	return hipGetLastError();
}

// Histogram3D::getBinIdFromDetPair (Histogram3D.cpp:416-421, getCoordsFromDetPair
// 366-405, getBinIdFromCoords 267-271): the pair ordered by id then by
// position in its ring, (r, phi) from the in-ring table (ringMap[a * dpr + b]
// = rr | phi << 16, -1 where no bin), the z bin from the two rings.  False
// for a pair outside the histogram (a detector beyond the scanner, no
// in-ring bin, a ring difference above the limit), where the reference
// throws.
__device__ __forceinline__ bool pairBin(const PairGeom& g, const int32_t* __restrict__ ringMap,
                                        uint32_t d1, uint32_t d2, int64_t& bin)
{
	bool ok = d1 < g.ndet && d2 < g.ndet;
	if (d1 > d2)
	{
		const uint32_t t = d1;
		d1 = d2;
		d2 = t;
	}
	uint32_t r1 = d1 % g.dpr, r2 = d2 % g.dpr;
	if (r1 > r2)
	{
		uint32_t t = d1;
		d1 = d2;
		d2 = t;
		t = r1;
		r1 = r2;
		r2 = t;
	}
	const int32_t m = ok ? ringMap[r1 * g.dpr + r2] : -1;
	const int32_t z1 = (int32_t)((d1 / g.dpr) % g.nrings);
	const int32_t z2 = (int32_t)((d2 / g.dpr) % g.nrings);
	const int32_t dz = z2 > z1 ? z2 - z1 : z1 - z2;
	ok = ok && m >= 0 && dz <= g.maxRingDiff;
	if (!ok)
		return false;
	const uint32_t doi1 = d1 / (g.nrings * g.dpr), doi2 = d2 / (g.nrings * g.dpr);
	const uint32_t r = (uint32_t)(m & 0xffff) * g.numDOIPoss + doi1 + doi2 * g.numDOI;
	const uint32_t phi = (uint32_t)m >> 16;
	int64_t zb = (int64_t)dz * g.nrings + (z1 < z2 ? z1 : z2) - (int64_t)dz * (dz - 1) / 2;
	if (dz > 0 && z1 > z2)
		zb += g.numZBinDiff;
	bin = (zb * g.numPhi + phi) * g.numR + r;
	return true;
}

// Histogram3D::getProjectionValueFromHistogramBin with a detector pair
// (Histogram3D.cpp:588-606): the value of the pair's bin.  The pair (0, 0)
// reads 0; a pair outside the histogram sets *err and reads 0.
__global__ void __launch_bounds__(256) kGatherPairs(const float* __restrict__ histo,
                                                    const uint32_t* __restrict__ d1s,
                                                    const uint32_t* __restrict__ d2s, int64_t n,
                                                    const int32_t* __restrict__ ringMap,
                                                    PairGeom g, float* __restrict__ out,
                                                    int* __restrict__ err)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const uint32_t d1 = d1s[i], d2 = d2s[i];
		float v = 0.f;
		if (d1 != 0 || d2 != 0)
		{
			int64_t bin = 0;
			if (pairBin(g, ringMap, d1, d2, bin))
				v = histo[bin];
			else
				atomicOr(err, 1);
		}
		out[i] = v;
	}
}

// The bin of each pair; a pair outside the histogram (or the disabled pair
// (0, 0), ProjectionData's null LOR) gets 0xffffffff and sets *err.
__global__ void __launch_bounds__(256) kPairBins(const uint32_t* __restrict__ d1s,
                                                 const uint32_t* __restrict__ d2s, int64_t n,
                                                 const int32_t* __restrict__ ringMap,
                                                 PairGeom g, uint32_t* __restrict__ out,
                                                 int* __restrict__ err)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const uint32_t d1 = d1s[i], d2 = d2s[i];
		int64_t bin = 0;
		const bool ok = (d1 != 0 || d2 != 0) && pairBin(g, ringMap, d1, d2, bin);
		if (!ok)
			atomicOr(err, 1);
		out[i] = ok ? (uint32_t)bin : 0xffffffffu;
	}
}

hipError_t launchGatherPairs(const float* histo, const uint32_t* d1, const uint32_t* d2,
                             int64_t n, const int32_t* ringMap, const PairGeom& g, float* out,
                             int* err, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	kGatherPairs<<<gridFor(n), 256, 0, st>>>(histo, d1, d2, n, ringMap, g, out, err);
	return hipGetLastError();
}

hipError_t launchPairBins(const uint32_t* d1, const uint32_t* d2, int64_t n,
                          const int32_t* ringMap, const PairGeom& g, uint32_t* out, int* err,
                          hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	kPairBins<<<gridFor(n), 256, 0, st>>>(d1, d2, n, ringMap, g, out, err);
	return hipGetLastError();
}

__global__ void __launch_bounds__(256) kFillI32(int32_t* __restrict__ x, int64_t n, int32_t v)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		x[i] = v;
}

hipError_t launchFillI32(int32_t* x, int64_t n, int32_t v, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	kFillI32<<<gridFor(n), 256, 0, st>>>(x, n, v);
	return hipGetLastError();
}

}  // namespace yrtmi
