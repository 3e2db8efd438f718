// Tile-binned Siddon backprojection for gfx950.
//
// Random FP32 atomics into a 20 MB image run at a few tens of G/s on the
// chip, an order of magnitude below the rate at which rays are walked, so
// the direct backprojection is bound by them.  Here the walk of every
// (LOR, ray) is cut at the boundaries of TX x TY x TZ image tiles:
//
//  plan (once per LOR set; geometry only):
//    kBin<COUNT>  walk every ray, count tile entries per tile (LDS counters,
//                 one global atomic per tile and workgroup)
//    host         prefix sum -> tile ranges, chunk table
//    kBin<FILL>   walk again, write one record per tile entry into its
//                 tile's range: key = lor * R + ray, the alpha where the
//                 walk enters the tile, the voxel it enters and the ray's
//                 plane shift (sidPlaneShift)
//  per backprojection:
//    kTileBack    one workgroup per chunk of a tile's records: re-enter each
//                 ray's walk at its record (sidRestart), accumulate in the
//                 tile held in LDS (ds_add_f32), then add the tile to the
//                 image with coalesced atomics (one per touched voxel).
#include "yrtmi_tiles.hpp"

namespace yrtmi
{

__device__ __forceinline__ void ldsAdd(float* p, float v)
{
	__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int tileOf(const DTiles& T, int f, int vx, int vy,
                                      int vz)
{
	return ((f * T.ntz + vz / kTileZ) * T.nty + vy / kTileY) * T.ntx + vx / kTileX;
}

// Walks ray r of LOR e (global ray index t = e * R + r) and calls
// onEnter(tile, alpha, packed entry) each time the walk enters a tile.  A
// plane shift beyond the record's range sets *err (the ray is not binned).
template <bool TOF, bool S64, class Fn>
__device__ __forceinline__ void binRay(const KArgs& a, const DTiles& T,
                                       int64_t e, int r, int* err, Fn&& onEnter)
{
	const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
	if (l.frame < 0)
		return;
	float3 q1, q2;
	rayLine(a.pr, l, r, q1, q2);
	q1 = make_float3(q1.x - a.im.ox, q1.y - a.im.oy, q1.z - a.im.oz);
	q2 = make_float3(q2.x - a.im.ox, q2.y - a.im.oy, q2.z - a.im.oz);
	SidRay s;
	if (!sidSetup<TOF, S64>(a.im, q1, q2, a.pr.tof, l.tof, s))
		return;
	int dx, dy, dz;
	if (!sidPlaneShift<S64>(a.im, s, dx, dy, dz))
		return;
	if (abs(dx) > kMaxShift || abs(dy) > kMaxShift || abs(dz) > kMaxShift)
	{
		atomicOr(err, 2);
		return;
	}
	int last = -1;
	sidWalk<TOF, S64>(a.im, s, a.pr.tof, imageBox(a.im), true, 0, 0, 0,
	                  [&](int vx, int vy, int vz, int, float a0, float) {
		                  const int t = tileOf(T, l.frame, vx, vy, vz);
		                  if (t != last)
		                  {
			                  onEnter(t, a0,
			                          packEntry(vx % kTileX, vy % kTileY, vz % kTileZ, dx,
			                                    dy, dz, last < 0));
			                  last = t;
		                  }
	                  });
}

// PASS 0: count[t] += entries of tile t.  PASS 1: write the records; each
// workgroup first counts its own entries, reserves a range per tile from
// cursor[t], then walks again to fill it.  LDS: cnt[nTiles], base[nTiles].
template <bool TOF, bool S64, int PASS>
__global__ void __launch_bounds__(256) kBin(KArgs a, DTiles T, DPlan pl)
{
	extern __shared__ int sm[];
	int* cnt = sm;
	int* base = sm + T.nTiles;
	const int64_t total = a.L.n * (int64_t)pl.R;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	for (int i = threadIdx.x; i < T.nTiles; i += blockDim.x)
		cnt[i] = 0;
	__syncthreads();
	for (int64_t t = t0; t < total; t += stride)
	{
		const int64_t e = t / pl.R;
		const int r = (int)(t - e * pl.R);
		binRay<TOF, S64>(a, T, e, r, pl.err, [&](int tile, float, uint32_t) {
			atomicAdd(&cnt[tile], 1);
		});
	}
	__syncthreads();
	if (PASS == 0)
	{
		for (int i = threadIdx.x; i < T.nTiles; i += blockDim.x)
			if (cnt[i] != 0)
				atomicAdd(&pl.count[i], (unsigned long long)cnt[i]);
		return;
	}
	for (int i = threadIdx.x; i < T.nTiles; i += blockDim.x)
	{
		const int c = cnt[i];
		base[i] = (c != 0) ? (int)atomicAdd(&pl.cursor[i], (unsigned int)c) : 0;
		cnt[i] = 0;
	}
	__syncthreads();
	for (int64_t t = t0; t < total; t += stride)
	{
		const int64_t e = t / pl.R;
		const int r = (int)(t - e * pl.R);
		binRay<TOF, S64>(a, T, e, r, pl.err, [&](int tile, float a0, uint32_t pv) {
			const int slot = base[tile] + atomicAdd(&cnt[tile], 1);
			if ((unsigned long long)slot < pl.count[tile])
			{
				const int64_t j = pl.tileOff[tile] + slot;
				pl.key[j] = (uint32_t)t;
				pl.alpha[j] = a0;
				pl.vox[j] = pv;
			}
			else
				atomicOr(pl.err, 1);
		});
	}
}

// One workgroup per chunk: accumulate the chunk's ray pieces in the LDS
// tile, then add the tile into y (frame f of the chunk's tile).  v[e] is
// the value backprojected along LOR e (split over the rays as
// ProjectorSiddon::backProjection does, ProjectorSiddon.cpp:321-328);
// LORs with v == 0 are skipped.
template <bool TOF, bool S64>
__global__ void __launch_bounds__(kTileThreads) kTileBack(KArgs a, DTiles T,
                                                          DPlanRead pl,
                                                          const float* __restrict__ v,
                                                          float* __restrict__ y)
{
	__shared__ float acc[kTileVox];
	const TileChunk ch = pl.chunks[blockIdx.x];
	int t = ch.tile;
	const int txi = t % T.ntx;
	t /= T.ntx;
	const int tyi = t % T.nty;
	t /= T.nty;
	const int tzi = t % T.ntz;
	const int f = t / T.ntz;
	WalkBox B;
	B.x0 = txi * kTileX;
	B.y0 = tyi * kTileY;
	B.z0 = tzi * kTileZ;
	B.x1 = min(B.x0 + kTileX, a.im.nx);
	B.y1 = min(B.y0 + kTileY, a.im.ny);
	B.z1 = min(B.z0 + kTileZ, a.im.nz);
	B.sy = kTileX;
	B.sz = kTileX * kTileY;
	B.base = -(B.z0 * B.sz + B.y0 * B.sy + B.x0);
	for (int i = threadIdx.x; i < kTileVox; i += kTileThreads)
		acc[i] = 0.f;
	__syncthreads();
	const int R = pl.R;
	for (int i = threadIdx.x; i < ch.count; i += kTileThreads)
	{
		const int64_t j = ch.begin + i;
		const uint32_t key = pl.key[j];
		const int64_t e = key / (uint32_t)R;
		const int r = (int)(key - (uint32_t)e * (uint32_t)R);
		if (e >= a.L.n)
			continue;
		float val = v[e];
		if (val == 0.f)
			continue;
		if (R > 1)
			val = val / (float)R;
		const uint32_t pv = pl.vox[j];
		const int vx = B.x0 + (int)(pv & 31u), vy = B.y0 + (int)((pv >> 5) & 31u),
		          vz = B.z0 + (int)((pv >> 10) & 15u);
		const int dx = (int)((pv >> 14) & 15u) - 8, dy = (int)((pv >> 18) & 15u) - 8,
		          dz = (int)((pv >> 22) & 15u) - 8;
		const bool first = (pv >> 31) != 0;
		if (!first && (vx >= B.x1 || vy >= B.y1 || vz >= B.z1))
			continue;
		const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
		if (l.frame != f)
			continue;
		float3 q1, q2;
		rayLine(a.pr, l, r, q1, q2);
		q1 = make_float3(q1.x - a.im.ox, q1.y - a.im.oy, q1.z - a.im.oz);
		q2 = make_float3(q2.x - a.im.ox, q2.y - a.im.oy, q2.z - a.im.oz);
		SidRay s;
		if (!sidSetup<TOF, S64>(a.im, q1, q2, a.pr.tof, l.tof, s))
			continue;
		if (!first)
			sidRestart<S64>(a.im, s, pl.alpha[j], vx, vy, vz, dx, dy, dz);
		sidWalk<TOF, S64>(a.im, s, a.pr.tof, B, first, vx, vy, vz,
		                  [&](int, int, int, int idx, float, float w) {
			                  ldsAdd(&acc[idx], val * w);
		                  });
	}
	__syncthreads();
	float* yf = y + (int64_t)f * a.im.nvox;
	for (int i = threadIdx.x; i < kTileVox; i += kTileThreads)
	{
		const int lx = i % kTileX, ly = (i / kTileX) % kTileY, lz = i / (kTileX * kTileY);
		const int gx = B.x0 + lx, gy = B.y0 + ly, gz = B.z0 + lz;
		const float s = acc[i];
		if (s != 0.f && gx < B.x1 && gy < B.y1 && gz < B.z1)
			unsafeAtomicAdd(yf + (int64_t)gz * a.im.nxy + (int64_t)gy * a.im.nx + gx, s);
	}
}

// EM ratio per LOR (computeEMUpdateImage, OSEM_CPU.cpp:329-374): the
// forward projection of x, corrected, against the measurement; v[e] = 0
// where the LOR is skipped (zero measurement / factor, invalid LOR, or a
// denominator at or below the threshold).
template <bool TOF, int LPE, bool S64>
__global__ void __launch_bounds__(256) kEmRatio(KArgs a, float* __restrict__ v)
{
	const int64_t total = a.L.n * LPE;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
	     t += stride)
	{
		const int64_t e = t / LPE;
		const int lane = (int)(t & (LPE - 1));
		const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
		const float m = a.c.meas ? a.c.meas[e] : 1.0f;
		const float sf = a.c.sens ? a.c.sens[e] : 1.0f;
		const float af = a.c.att ? a.c.att[e] : 1.0f;
		const bool skip = (m == 0.f || sf == 0.f || af == 0.f || l.frame < 0);
		float s = 0.f;
		if (!skip)
		{
			const float* img = a.x + (int64_t)l.frame * a.im.nvox;
			for (int r = lane; r < a.pr.numRays; r += LPE)
			{
				float3 q1, q2;
				rayLine(a.pr, l, r, q1, q2);
				q1 = make_float3(q1.x - a.im.ox, q1.y - a.im.oy, q1.z - a.im.oz);
				q2 = make_float3(q2.x - a.im.ox, q2.y - a.im.oy, q2.z - a.im.oz);
				uint32_t c = 0;
				s += siddonRay<MODE_FWD, TOF, S64>(a.im, const_cast<float*>(img), q1,
				                                   q2, 0.f, a.pr.tof, l.tof, c);
			}
		}
#pragma unroll
		for (int o = 1; o < LPE; o <<= 1)
			s += __shfl_xor(s, o, LPE);
		if (lane != 0)
			continue;
		float out = 0.f;
		if (!skip)
		{
			float upd = (a.pr.numRays > 1) ? s / (float)a.pr.numRays : s;
			if (a.c.sens)
				upd *= sf;
			if (a.c.att)
				upd *= af;
			upd *= a.c.g;
			if (a.c.randoms)
				upd += a.c.randoms[e];
			if (a.c.scatter)
				upd += a.c.scatter[e];
			if (fabsf(upd) > a.c.denomThr)
			{
				out = m / upd;
				if (a.c.sens)
					out *= sf;
				if (a.c.att)
					out *= af;
				out *= a.c.g;
			}
		}
		v[e] = out;
	}
}

// Sensitivity value per LOR (generateSensImage, OSEM_CPU.cpp:146-156):
// v[e] = sens factor * exp(-0.1 * Siddon(mu)) (Corrector.cpp:601-611).
template <bool S64>
__global__ void __launch_bounds__(256) kSensValue(KArgs a, float* __restrict__ v)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.L.n;
	     e += stride)
	{
		const Lor l = fetchLor(a.sc, a.L, e, 1);
		float val = a.vals ? a.vals[e] : 1.0f;
		if (a.att)
		{
			uint32_t c2 = 0;
			const float mu = siddonRay<MODE_FWD, false, S64>(
			    a.attIm, const_cast<float*>(a.att), l.p1, l.p2, 0.f, a.pr.tof, 0.f, c2);
			val *= expf(-mu * 0.1f);
		}
		v[e] = (l.frame >= 0) ? val : 0.f;
	}
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <bool TOF, bool S64>
static hipError_t binT(int pass, const KArgs& a, const DTiles& T,
                       const DPlan& pl, int grid, hipStream_t st)
{
	const size_t lds = (size_t)T.nTiles * 2 * sizeof(int);
	if (pass == 0)
		kBin<TOF, S64, 0><<<grid, 256, lds, st>>>(a, T, pl);
	else
		kBin<TOF, S64, 1><<<grid, 256, lds, st>>>(a, T, pl);
	return hipGetLastError();
}

hipError_t launchBin(int pass, const KArgs& a, const DTiles& T, const DPlan& pl,
                     int grid, hipStream_t st)
{
	const bool tof = a.pr.useTof != 0;
	if (a.pr.setup64)
		return tof ? binT<true, true>(pass, a, T, pl, grid, st)
		           : binT<false, true>(pass, a, T, pl, grid, st);
	return tof ? binT<true, false>(pass, a, T, pl, grid, st)
	           : binT<false, false>(pass, a, T, pl, grid, st);
}

hipError_t launchTileBack(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                          int64_t nChunks, const float* v, float* y,
                          hipStream_t st)
{
	if (nChunks <= 0)
		return hipSuccess;
	const dim3 g((unsigned)nChunks);
	const bool tof = a.pr.useTof != 0;
	if (a.pr.setup64)
	{
		if (tof)
			kTileBack<true, true><<<g, kTileThreads, 0, st>>>(a, T, pl, v, y);
		else
			kTileBack<false, true><<<g, kTileThreads, 0, st>>>(a, T, pl, v, y);
	}
	else
	{
		if (tof)
			kTileBack<true, false><<<g, kTileThreads, 0, st>>>(a, T, pl, v, y);
		else
			kTileBack<false, false><<<g, kTileThreads, 0, st>>>(a, T, pl, v, y);
	}
	return hipGetLastError();
}

template <bool TOF, bool S64>
static hipError_t ratioT(const KArgs& a, int lpe, dim3 g, float* v, hipStream_t st)
{
	switch (lpe)
	{
	case 1:
		kEmRatio<TOF, 1, S64><<<g, 256, 0, st>>>(a, v);
		break;
	case 2:
		kEmRatio<TOF, 2, S64><<<g, 256, 0, st>>>(a, v);
		break;
	case 4:
		kEmRatio<TOF, 4, S64><<<g, 256, 0, st>>>(a, v);
		break;
	case 8:
		kEmRatio<TOF, 8, S64><<<g, 256, 0, st>>>(a, v);
		break;
	default:
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t launchEmRatio(const KArgs& a, int lpe, float* v, hipStream_t st)
{
	const int64_t threads = a.L.n * lpe;
	int64_t nb = (threads + 255) / 256;
	if (nb < 1)
		nb = 1;
	if (nb > 0x7fffffff)
		nb = 0x7fffffff;
	const dim3 g((unsigned)nb);
	const bool tof = a.pr.useTof != 0;
	if (a.pr.setup64)
		return tof ? ratioT<true, true>(a, lpe, g, v, st)
		           : ratioT<false, true>(a, lpe, g, v, st);
	return tof ? ratioT<true, false>(a, lpe, g, v, st)
	           : ratioT<false, false>(a, lpe, g, v, st);
}

hipError_t launchSensValue(const KArgs& a, float* v, hipStream_t st)
{
	int64_t nb = (a.L.n + 255) / 256;
	if (nb > 65536)
		nb = 65536;
	if (nb < 1)
		nb = 1;
	if (a.pr.setup64)
		kSensValue<true><<<(unsigned)nb, 256, 0, st>>>(a, v);
	else
		kSensValue<false><<<(unsigned)nb, 256, 0, st>>>(a, v);
	return hipGetLastError();
}

}  // namespace yrtmi
