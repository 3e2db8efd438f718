// Tile-binned backprojection (Siddon and distance-driven) for gfx950.
//
// Random FP32 atomics into a 20 MB image run at a few tens of G/s on the
// chip, an order of magnitude below the rate at which rays are walked, so
// the direct backprojection is bound by them.  Here the walk of every
// (LOR, ray) is cut at the boundaries of TX x TY x TZ image tiles:
//
//  plan (once per LOR set; geometry only):
//    kRayPrep     resolve every (LOR, ray) once into the ray table
//    kBin<COUNT>  walk every ray, count tile entries per tile (LDS counters,
//                 one global atomic per tile and workgroup)
//    host         prefix sum -> tile ranges, chunk table
//    kBin<FILL>   walk again, write one record per tile entry into its
//                 tile's range: key = lor * R + ray, the alpha where the
//                 walk enters the tile, the voxel it enters and the ray's
//                 plane shift (sidPlaneShift); a DD record is the LOR alone
//  per backprojection:
//    kTileBack    one workgroup per chunk of a tile's records: re-enter each
//                 ray's walk at its record (sidRestart) or sweep the DD
//                 footprint over the tile, accumulate in the tile held in
//                 LDS, then add the tile to the image with coalesced atomics
//                 (one per touched voxel).
#include "yrtmi_tiles.hpp"

#include <algorithm>
#include <type_traits>

namespace yrtmi
{

__device__ __forceinline__ void ldsAdd(float* p, float v)
{
	__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void ldsAdd64(unsigned long long* p, long long v)
{
	__hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED,
	                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int tileOf(const DTiles& T, int f, int vx, int vy,
                                      int vz)
{
	return ((f * T.ntz + (vz >> T.lz)) * T.nty + (vy >> T.ly)) * T.ntx + (vx >> T.lx);
}

// Ray table of a LOR set (rayEntry for every (LOR, ray) t = e * R + r):
// the LOR resolved once (detector pair or histogram bin, LUT, motion,
// frame), the multi-ray line and the image offset.  Skipped LORs get frame -1.
// Distance-driven (DD): the LOR's footprint set-up (ddSetup, ddStore).
template <bool DD, bool TOF, bool S64>
__global__ void __launch_bounds__(256) kRayPrep(KArgs a, float4* __restrict__ tab)
{
	const int R = a.pr.numRays;
	const int64_t total = a.L.n * (int64_t)R;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride)
	{
		const int64_t e = t / R;
		const int r = (int)(t - e * R);
		const Lor l = fetchLor(a.sc, a.L, e, a.im.nt);
		if (DD)
		{
			if (l.frame < 0)
			{
				ddStoreSkip(tab, (uint64_t)t);
				continue;
			}
			const float3 d1 = make_float3(l.p1.x - a.im.ox, l.p1.y - a.im.oy, l.p1.z - a.im.oz);
			const float3 d2 = make_float3(l.p2.x - a.im.ox, l.p2.y - a.im.oy, l.p2.z - a.im.oz);
			DDRay D;
			ddSetup<TOF, S64>(a.im, d1, d2, l.n1, l.n2, a.pr, l.tof, false, D);
			ddStore(tab, (uint64_t)t, D, l.frame);
			continue;
		}
		constexpr int V = rayTabVec(S64);
		const float3 z = make_float3(0.f, 0.f, 0.f);
		if (l.frame < 0)
		{
			rayEntry(tab, V, (uint32_t)t, z, z, 0.f, -1);
			if (!S64)
				tab[V * (uint64_t)t + 2] = tab[V * (uint64_t)t + 3] = make_float4(0.f, 0.f, 0.f, 0.f);
			continue;
		}
		float3 q1, q2;
		rayLine(a.pr, l, r, q1, q2);
		q1 = make_float3(q1.x - a.im.ox, q1.y - a.im.oy, q1.z - a.im.oz);
		q2 = make_float3(q2.x - a.im.ox, q2.y - a.im.oy, q2.z - a.im.oz);
		rayEntry(tab, V, (uint32_t)t, q1, q2, l.tof, l.frame);
		if (!S64)
		{
			// the ray's FP32 set-up, for walks re-entered at a tile record
			SidRay sr;
			if (sidSetup<TOF, false>(a.im, q1, q2, a.pr.tof, l.tof, sr))
				sidStore(tab + V * (uint64_t)t + 2, sr);
			else
				tab[V * (uint64_t)t + 2] = tab[V * (uint64_t)t + 3] = make_float4(0.f, 0.f, 0.f, 0.f);
		}
	}
}

// Walks ray t of the ray table and calls onRecord(tile, alpha, packed
// entry, segments) for each tile the walk passes through, once the walk has
// left it (so the piece's number of segments is known).  A plane shift
// beyond the record's range sets *err (the ray is not binned).
template <bool TOF, bool S64, class Fn>
__device__ __forceinline__ void binRay(const DImage& im, const DTof& tof, const DTiles& T,
                                       const float4* __restrict__ tab, uint32_t t, int* err,
                                       Fn&& onRecord)
{
	constexpr int V = rayTabVec(S64);
	const float4 A = tab[V * (uint64_t)t], Bq = tab[V * (uint64_t)t + 1];
	const int frame = __float_as_int(Bq.w);
	if (frame < 0)
		return;
	const float3 q1 = make_float3(A.x, A.y, A.z), q2 = make_float3(A.w, Bq.x, Bq.y);
	SidRay s;
	if (!sidSetup<TOF, S64>(im, q1, q2, tof, Bq.z, s))
		return;
	int dx, dy, dz;
	if (!sidPlaneShift<S64>(im, s, dx, dy, dz))
		return;
	if (abs(dx) > kMaxShift || abs(dy) > kMaxShift || abs(dz) > kMaxShift)
	{
		atomicOr(err, 2);
		return;
	}
	int last = -1, start = 0, nseg = 0;
	float pa = 0.f;
	uint32_t pv = 0;
	sidWalk<TOF, S64>(im, s, tof, imageBox(im), true, 0, 0, 0,
	                  [&](int vx, int vy, int vz, uint32_t, float a0, float, uint32_t) {
		                  const int tl = tileOf(T, frame, vx, vy, vz);
		                  if (tl != last)
		                  {
			                  if (last >= 0)
				                  onRecord(last, pa, pv, nseg - start);
			                  pa = a0;
			                  pv = packEntry(vx & (T.tx - 1), vy & (T.ty - 1), vz & (T.tz - 1),
			                                 dx, dy, dz, last < 0);
			                  start = nseg;
			                  last = tl;
		                  }
		                  nseg++;
	                  });
	if (last >= 0)
		onRecord(last, pa, pv, nseg - start);
}

// DD: the tiles of LOR t's footprint.  The slices of the major axis are
// taken one tile slab at a time, with the footprint ranges of each slice
// (ddSlice, the ranges the sweep itself visits).  Each tile of the bounding
// box of a slab's ranges that some slice reaches gives a record:
// onRecord(tile, 0, first | last << 6, voxels), first and last the slices
// (less the slab's first) that reach it — the sweep of the record goes over
// these slices only — and voxels its number of footprint voxels in the tile.
template <class Fn>
__device__ __forceinline__ void binRayDD(const DImage& im, const DProj& pr, const DTiles& T,
                                         const float4* __restrict__ tab, uint32_t t,
                                         Fn&& onRecord)
{
	const float4* te = tab + kDDVec * (uint64_t)t;
	float4 e[kDDVec];
#pragma unroll
	for (int i = 0; i < kDDVec; i++)
		e[i] = te[i];
	const int frame = ddFrame(e[5]);
	if (frame < 0)
		return;
	DDRay R;
	ddLoad(im, pr, e, false, R);
	const int lm = R.flag_y ? T.ly : T.lx;  // major axis
	const int ln = R.flag_y ? T.lx : T.ly;  // minor axis
	for (int s0 = R.xy_i_0; s0 <= R.xy_i_1;)
	{
		const int m = s0 >> lm, mb = m << lm;
		const int s1 = min(R.xy_i_1, mb + (1 << lm) - 1);
		int nlo = 1 << 30, nhi = -1, zlo = 1 << 30, zhi = -1;
		for (int xyi = s0; xyi <= s1; xyi++)
		{
			DDSlice q;
			ddSlice<false>(im, R, pr, xyi, q);
			if (q.yx0 <= q.yx1 && q.zi0 <= q.zi1)
			{
				nlo = min(nlo, q.yx0);
				nhi = max(nhi, q.yx1);
				zlo = min(zlo, q.zi0);
				zhi = max(zhi, q.zi1);
			}
		}
		if (nhi >= 0)
		{
			for (int zt = zlo >> T.lz; zt <= zhi >> T.lz; zt++)
				for (int nt = nlo >> ln; nt <= nhi >> ln; nt++)
				{
					const int y0 = nt << ln, y1 = y0 + (1 << ln) - 1;
					const int z0 = zt << T.lz, z1 = z0 + (1 << T.lz) - 1;
					int first = -1, last = -1, vox = 0;
					for (int xyi = s0; xyi <= s1; xyi++)
					{
						DDSlice q;
						ddSlice<false>(im, R, pr, xyi, q);
						const int a0 = max(q.yx0, y0), a1 = min(q.yx1, y1);
						const int b0 = max(q.zi0, z0), b1 = min(q.zi1, z1);
						if (a0 <= a1 && b0 <= b1)
						{
							if (first < 0)
								first = xyi;
							last = xyi;
							vox += (a1 - a0 + 1) * (b1 - b0 + 1);
						}
					}
					if (vox > 0)
					{
						const int tx = R.flag_y ? nt : m, ty = R.flag_y ? m : nt;
						onRecord(((frame * T.ntz + zt) * T.nty + ty) * T.ntx + tx, 0.f,
						         (uint32_t)(first - mb) | ((uint32_t)(last - mb) << 6), vox);
					}
				}
		}
		s0 = s1 + 1;
	}
}

__device__ __forceinline__ int lengthClass(const DPlan& pl, int segs)
{
	int c = 0;
	for (int k = 0; k < kMaxClasses - 1; k++)
		c += (k < pl.C - 1 && segs >= pl.cut[k]) ? 1 : 0;
	return c;
}

// Each workgroup bins one contiguous range of rays (they are in locality
// order), so a bin's records from one workgroup come from neighbouring rays
// and the tile backprojection's reads of their ray-table entries coalesce.
// PASS 0: each workgroup counts the records of its rays per bin (LDS),
// keeps its counts in blockCnt[block][bin] and adds them to count[bin].
// PASS 1 (same grid, same rays per workgroup): each workgroup reserves its
// counts in every bin from cursor[bin] and walks its rays again, writing
// each record into its range (an LDS countdown per bin, so a record beyond
// the reservation is caught without a global read).  DD records have no
// alpha.
// LDS: cnt[nB] ints, then (PASS 1) base[nB] 32-bit record indices.
template <bool DD, bool TOF, bool S64, int PASS>
__global__ void __launch_bounds__(256) kBin(DImage im, DProj pr, int64_t nRays, DTiles T,
                                            DPlan pl)
{
	extern __shared__ int sm[];
	const int nB = T.nTiles * pl.C;
	int* cnt = sm;
	uint32_t* base = reinterpret_cast<uint32_t*>(sm + nB);
	const int64_t per = (nRays + gridDim.x - 1) / gridDim.x;
	const int64_t tEnd = min(nRays, (int64_t)(blockIdx.x + 1) * per);
	const int64_t t0 = (int64_t)blockIdx.x * per + threadIdx.x;
	const int64_t stride = blockDim.x;
	unsigned int* myCnt = pl.blockCnt + (int64_t)blockIdx.x * nB;
	auto binOne = [&](uint32_t t, auto&& onRecord) {
		if constexpr (DD)
			binRayDD(im, pr, T, pl.ray, t, onRecord);
		else
			binRay<TOF, S64>(im, pr.tof, T, pl.ray, t, pl.err, onRecord);
	};
	if (PASS == 0)
	{
		for (int i = threadIdx.x; i < nB; i += blockDim.x)
			cnt[i] = 0;
		__syncthreads();
		for (int64_t t = t0; t < tEnd; t += stride)
			binOne((uint32_t)t, [&](int tile, float, uint32_t, int segs) {
				atomicAdd(&cnt[tile * pl.C + lengthClass(pl, segs)], 1);
			});
		__syncthreads();
		for (int i = threadIdx.x; i < nB; i += blockDim.x)
		{
			myCnt[i] = (unsigned int)cnt[i];
			if (cnt[i] != 0)
				atomicAdd(&pl.count[i], (unsigned long long)cnt[i]);
		}
		return;
	}
	for (int i = threadIdx.x; i < nB; i += blockDim.x)
	{
		const int c = (int)myCnt[i];
		base[i] = (c != 0) ? (uint32_t)(pl.tileOff[i] +
		                                (int64_t)atomicAdd(&pl.cursor[i], (unsigned int)c))
		                   : 0u;
		cnt[i] = c;
	}
	__syncthreads();
	for (int64_t t = t0; t < tEnd; t += stride)
		binOne((uint32_t)t, [&](int tile, float a0, uint32_t pv, int segs) {
			const int b = tile * pl.C + lengthClass(pl, segs);
			const int left = atomicSub(&cnt[b], 1);
			if (left > 0)
			{
				const int64_t j = (int64_t)base[b] + left - 1;
				// one record of consecutive words (one vector store)
				uint32_t* o = pl.rec + recWords(DD) * j;
				o[0] = (uint32_t)t;
				if (!DD)
					o[1] = __float_as_uint(a0);
				o[recWords(DD) - 1] = withFineClass(DD, pv, segs, T.tx + T.ty + T.tz);
			}
			else
				atomicOr(pl.err, 1);
		});
}

// Records of every chunk ordered by their fine length class (withFineClass),
// once when the plan is built: the tile backprojection's waves then take
// records of about the same length, where the plan's bins can hold only a
// few classes per tile.  Each window of kSortWin records of a chunk is
// sorted in place, stably by (class, position): one workgroup per chunk,
// each lane holding its records in registers between the reads and the
// writes; a record's rank in its class within a group of 64 is counted
// with wave ballots.  The tile sums do not depend on the records' order
// (fixed point), so the backprojection is unchanged.
constexpr int kSortWin = 2048;

template <bool DD>
__global__ void __launch_bounds__(256) kChunkSort(DPlan pl, const TileChunk* __restrict__ chunks)
{
	constexpr int G = kSortWin / 64, GPW = G / 4;  // groups of 64, per wave
	__shared__ int cntG[G][16];
	__shared__ int offG[G][16];
	__shared__ int base[16];
	const TileChunk ch = chunks[blockIdx.x];
	const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
	const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	for (int w0 = 0; w0 < ch.count; w0 += kSortWin)
	{
		const int n = min(kSortWin, ch.count - w0);
		const int64_t j0 = ch.begin + w0;
		uint32_t k[GPW], v[GPW];
		uint32_t al[GPW];
		int cls[GPW];
#pragma unroll
		for (int i = 0; i < GPW; i++)
		{
			const int g = wv + 4 * i, r = g * 64 + lane;
			k[i] = 0u;
			v[i] = 0u;
			al[i] = 0u;
			cls[i] = 16;  // no record
			if (r < n)
			{
				const uint32_t* o = pl.rec + recWords(DD) * (j0 + r);
				k[i] = o[0];
				v[i] = o[recWords(DD) - 1];
				if (!DD)
					al[i] = o[1];
				cls[i] = fineClassOf(DD, v[i]);
			}
			for (int c = 0; c < 16; c++)
			{
				const int m = __popcll(__ballot(cls[i] == c));
				if (lane == 0)
					cntG[g][c] = m;
			}
		}
		__syncthreads();
		if (threadIdx.x < 16)
		{
			const int c = (int)threadIdx.x;
			int sum = 0;
			for (int g = 0; g < G; g++)
			{
				offG[g][c] = sum;
				sum += cntG[g][c];
			}
			base[c] = sum;
		}
		__syncthreads();
		if (threadIdx.x == 0)
		{
			int sum = 0;
			for (int c = 0; c < 16; c++)
			{
				const int t = base[c];
				base[c] = sum;
				sum += t;
			}
		}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < GPW; i++)
		{
			const int g = wv + 4 * i;
			uint64_t mine = 0ull;
			for (int c = 0; c < 16; c++)
			{
				const uint64_t m = __ballot(cls[i] == c);
				if (cls[i] == c)
					mine = m;
			}
			if (cls[i] < 16)
			{
				const int64_t j = j0 + base[cls[i]] + offG[g][cls[i]] + __popcll(mine & below);
				uint32_t* o = pl.rec + recWords(DD) * j;
				o[0] = k[i];
				if (!DD)
					o[1] = al[i];
				o[recWords(DD) - 1] = v[i];
			}
		}
		__syncthreads();
	}
}

hipError_t launchChunkSort(bool dd, const DPlan& pl, const TileChunk* chunks, int64_t nChunks,
                           hipStream_t st)
{
	for (int64_t c0 = 0; c0 < nChunks; c0 += 65535)
	{
		const unsigned g = (unsigned)std::min<int64_t>(65535, nChunks - c0);
		if (dd)
			kChunkSort<true><<<g, 256, 0, st>>>(pl, chunks + c0);
		else
			kChunkSort<false><<<g, 256, 0, st>>>(pl, chunks + c0);
	}
	return hipGetLastError();
}

// One chunk: accumulate its ray pieces in the LDS tile `acc`, then add the
// tile into y (frame f of the chunk's tile).  v[e] is the value
// backprojected along LOR e (split over the rays as
// ProjectorSiddon::backProjection does, ProjectorSiddon.cpp:321-328); LORs
// with v == 0 are skipped.  Each lane runs a three-stage pipeline over its
// records: the fields of the record two ahead are read, the ray-table entry
// and value of the next one are in flight, and the current one is set up
// from its table entry and walked (the walk itself issues no global loads).
//
// Tile accumulation.  FP32 atomic adds into LDS run at about half a lane per
// clock per CU and bounded the kernel; 64-bit integer LDS atomics do not.
// So the tile is accumulated in fixed point, each contribution c = v/R * w
// rounded once to a multiple of 2^-k and added as an integer (exact and
// order-free, so the tile sum is deterministic); the flush converts back
// (int64 -> float, times 2^-k: exact scaling).  A voxel takes at most one
// contribution per record, so with every term below 2^lim in magnitude,
// lim = min(62 - e2, 51) (records < 2^e2; 2^51 for the conversion), no
// partial sum can overflow; a term that does not fit — and any non-finite
// one — is added to the image directly with a global FP32 atomic, as the
// direct backprojection does.  k is chosen from the launch's values: E, the
// binade of the median nonzero |v| (kVExpHist, kPickExp), 2^10 of headroom,
// and wmax, the largest segment weight (voxel diagonal, times the TOF norm):
// k = lim - (E + 10 + ew), wmax < 2^ew.
// Typical contributions then keep ~35 bits below their magnitude; values
// more than ~1000 times the median (the heavy tail of EM ratios m / Ax on
// LORs with tiny projections) cost a global atomic instead of precision.
// The scale only trades precision against that fallback: no choice of it
// can overflow the tile.  The 32^3 shape (whose 64-bit tile would need
// 256 KB of LDS) accumulates in FP32.
//
// DD records (DD = true) hold the LOR alone: its table entry (ddEntry) is set
// up (ddSetup) and its footprint swept over the slices, minor-axis voxels and
// z planes of the tile (ddVisit), the voxel weights exactly those of the
// direct sweep.  A voxel again takes one contribution per record.
// The projector's PSF tables of one direction (flip: the backprojection's),
// the prefix sums then the padded taps, copied into the workgroup's dynamic
// LDS (DPsf::ldsBytes, given by the DD kernels' launches) by all its
// threads; the caller waits at a barrier before reading them.  Returns the
// LDS prefix sums, or null when the tables stay in global memory.
__device__ __forceinline__ const double* psfStage(const DPsf& P, bool flip, double* l)
{
	if (P.ldsBytes == 0)
		return nullptr;
	const int nc = P.numS * (P.K + 2), nk = P.numS * psfRowTaps(P.K);
	const double* c = flip ? P.cumFlip : P.cum;
	const float* k = flip ? P.kflip : P.k;
	for (int i = threadIdx.x; i < nc; i += blockDim.x)
		l[i] = c[i];
	float* lk = reinterpret_cast<float*>(l + nc);
	for (int i = threadIdx.x; i < nk; i += blockDim.x)
		lk[i] = k[i];
	return l;
}

// A DD ray's PSF rows pointed at the staged tables.
__device__ __forceinline__ void psfRowsLds(const DPsf& P, const double* lc, DDRay& D)
{
	const float* lk = reinterpret_cast<const float*>(lc + P.numS * (P.K + 2));
	D.pcum = lc + D.psfRow * (P.K + 2);
	D.pk = lk + D.psfRow * psfRowTaps(P.K);
}

template <bool DD, bool TOF, bool S64, int SHAPE>
__device__ __forceinline__ void tileChunk(const KArgs& a, const DTiles& T,
                                          const DPlanRead& pl, const TileChunk ch,
                                          const float* __restrict__ v,
                                          float* __restrict__ y, unsigned long long* acc64,
                                          const double* psfL)
{
	float* acc = reinterpret_cast<float*>(acc64);
	constexpr int TX = tileDim(SHAPE, 0), TY = tileDim(SHAPE, 1), TZ = tileDim(SHAPE, 2);
	constexpr int NV = TX * TY * TZ, NT = tileThreads(SHAPE);
	int t = ch.tile;
	const int txi = t % T.ntx;
	t /= T.ntx;
	const int tyi = t % T.nty;
	t /= T.nty;
	const int tzi = t % T.ntz;
	const int f = t / T.ntz;
	WalkBox B;
	B.x0 = txi * TX;
	B.y0 = tyi * TY;
	B.z0 = tzi * TZ;
	B.x1 = min(B.x0 + TX, a.im.nx);
	B.y1 = min(B.y0 + TY, a.im.ny);
	B.z1 = min(B.z0 + TZ, a.im.nz);
	// voxel index in bytes of the LDS tile's words (8: fixed point, 4: FP32)
	auto wordBytes = [&](int esz) {
		B.sx = esz;
		B.sy = TX * esz;
		B.sz = TX * TY * esz;
		B.base = -(B.z0 * B.sz + B.y0 * B.sy + B.x0 * B.sx);
	};
	constexpr int NW = (SHAPE == 2) ? NV / 2 : NV;  // 64-bit words of the LDS tile
	for (int i = threadIdx.x; i < NW; i += NT)
		acc64[i] = 0ull;
	// fixed-point scale of this chunk (uniform)
	int fixed = 0, k = 0;
	float climit = 0.f, wbound = 0.f;
	if constexpr (SHAPE != 2)
	{
		if (pl.vexp != nullptr)
		{
			// largest voxel weight: a Siddon segment is at most the voxel
			// diagonal; a DD weight is at most the slice width over the cosine
			// of the LOR to the major axis (taken as up to 1/4 here: a larger
			// weight only takes the global-atomic path)
			float wmax = DD ? 4.f * fmaxf(a.im.vx, a.im.vy)
			                : 1.01f * sqrtf(a.im.vx * a.im.vx + a.im.vy * a.im.vy +
			                                a.im.vz * a.im.vz);
			if (TOF)
				wmax *= 1.01f * a.pr.tof.norm;
			int ew = 0;
			(void)frexpf(wmax, &ew);
			// terms below 2^lim, lim = min(62 - e2, 51): no partial sum of
			// the chunk's at most 2^e2 terms overflows, and every term
			// converts exactly through the 2^52 bias below
			const int e2 = 32 - __clz(ch.count);
			const int lim = min(62 - e2, 51);
			k = lim - (*pl.vexp + 10 + ew);
			// |c| < 2^(lim - k) <=> |c * 2^k| < 2^lim (inf when above the
			// float range: then every finite c qualifies)
			climit = ldexpf(1.f, lim - k);
			wbound = wmax;
			fixed = 1;
		}
	}
	fixed = __builtin_amdgcn_readfirstlane(fixed);
	k = __builtin_amdgcn_readfirstlane(k);
	float* yf = y + (int64_t)f * a.im.nvox;
	__syncthreads();
	const uint32_t R = (uint32_t)pl.R;
	const bool rPow2 = (R & (R - 1u)) == 0u;
	const int rShift = __ffs((int)R) - 1;
	const float rInvPow2 = ldexpf(1.f, -rShift);
	const uint64_t nRays = (uint64_t)a.L.n * R;
	// no record (keys are < n * R < 2^32); a key is checked against n * R
	// where it is used, not where it is read, so that its load stays in
	// flight during the walk
	constexpr uint32_t kNone = 0xffffffffu;
	auto lorOf = [&](uint32_t key) { return rPow2 ? key >> rShift : key / R; };
	struct Rec
	{
		uint32_t key, pv;
		float al;
	};
	// Siddon: the ray's end points, TOF value and frame (two float4), with
	// the FP32 set-up two more (sidStore); DD: kDDVec (ddStore)
	constexpr int NE = DD ? kDDVec : rayTabVec(S64);
	struct Ray
	{
		float4 e[NE];
		float val;
	};
	auto readRecord = [&](int k, Rec& r) {
		r.key = kNone;
		r.al = 0.f;
		r.pv = 0u;
		if (k < ch.count)
		{
			const int64_t j = ch.begin + k;
			const uint32_t* o = pl.rec + recWords(DD) * j;
			r.key = o[0];
			if (!DD)
				r.al = __uint_as_float(o[1]);
			r.pv = o[recWords(DD) - 1];
		}
	};
	auto readRay = [&](uint32_t key, Ray& q) {
#pragma unroll
		for (int i = 0; i < NE; i++)
			q.e[i] = make_float4(0.f, 0.f, 0.f, 0.f);
		q.val = 0.f;
		if ((uint64_t)key < nRays)
		{
#pragma unroll
			for (int i = 0; i < NE; i++)
				q.e[i] = pl.ray[NE * (uint64_t)key + i];
			q.val = v[DD ? key : lorOf(key)];
		}
	};
	// Siddon: a record whose value times the largest term factor (wmax:
	// every term c = v * w of its walk is at most |v| * wmax) is within
	// the fixed-point range walks with `addIn`, which skips the per-term
	// range test; other records (and the FP32 tile) walk with `add`
	auto walk = [&](const Rec& rc, const Ray& q, auto&& add, auto&& addIn, auto inRange) {
		float val = q.val;
		if ((uint64_t)rc.key >= nRays || val == 0.f)
			return;
		{
			const float4 qa = q.e[0], qb = q.e[1];
			if (__float_as_int(qb.w) != f)
				return;
			const float3 q1 = make_float3(qa.x, qa.y, qa.z), q2 = make_float3(qa.w, qb.x, qb.y);
			// v / R: for R a power of two the product with 1/R is the same
			// correctly rounded value (1/R exact), one multiply
			if (R > 1u)
				val = rPow2 ? val * rInvPow2 : val / (float)R;
			const uint32_t pv = rc.pv;
			const int vx = B.x0 + (int)(pv & 31u), vy = B.y0 + (int)((pv >> 5) & 31u),
			          vz = B.z0 + (int)((pv >> 10) & 31u);
			const int dx = (int)((pv >> 15) & 15u) - 8, dy = (int)((pv >> 19) & 15u) - 8,
			          dz = (int)((pv >> 23) & 15u) - 8;
			// FP32 set-up: every record, a ray's first included (its voxel is
			// the walk's first, its alpha the ray's start), is re-entered from
			// the ray's stored set-up, one instruction stream for the wave;
			// FP64 set-up: the first record walks from the ray's entry
			const bool entry = S64 && (pv >> 31) != 0;
			if (!entry && (vx >= B.x1 || vy >= B.y1 || vz >= B.z1))
				return;
			SidRay s;
			if (!S64)
				sidLoad(q1, q2, q.e[2 % NE], q.e[3 % NE], s);
			else if (!sidSetup<TOF, S64>(a.im, q1, q2, a.pr.tof, qb.z, s))
				return;
			if (!entry)
				sidRestart<S64>(a.im, s, rc.al, vx, vy, vz, dx, dy, dz);
			const float vs = val * s.ws;
			auto go = [&](auto&& fn) {
				sidWalk<TOF, S64>(a.im, s, a.pr.tof, B, entry, vx, vy, vz,
				                  [&](int gx, int gy, int gz, uint32_t idx, float, float w,
				                      uint32_t) { fn(idx, gx, gy, gz, vs * w); });
			};
			if constexpr (decltype(inRange)::value)
			{
				if (fabsf(val) * wbound < climit)
					go(addIn);
				else
					go(add);
			}
			else
				go(add);
		}
	};
	// DD: a wave's 64 records are swept one (record, slice) pair per lane.
	// The pairs are numbered by the wave's prefix sum of the records' slice
	// counts (last - first + 1 of the record's word), and the lane of pair s
	// finds its record, the last whose first pair is at most s, by a binary
	// search over the prefix sums; so the lanes differ in the footprint of
	// their slice only, not in their record's number of slices.  Each pair
	// reads its LOR's set-up (the pairs of one record read the same lines in
	// one load).  The sums and searches run in the wave's uniform control
	// flow (the trip counts are wave-uniform), with no workgroup barrier.
	auto pipelineDD = [&](auto&& add) {
		const int lane = (int)(threadIdx.x & 63u);
		for (int i0 = (int)threadIdx.x - lane; i0 < ch.count; i0 += NT)
		{
			Rec rc;
			readRecord(i0 + lane, rc);
			const int ns = (uint64_t)rc.key < nRays
			                   ? (int)((rc.pv >> 6) & 63u) - (int)(rc.pv & 63u) + 1
			                   : 0;
			int incl = ns;
#pragma unroll
			for (int d = 1; d < 64; d <<= 1)
			{
				const int t = __shfl_up(incl, (unsigned int)d, 64);
				if (lane >= d)
					incl += t;
			}
			const int excl = incl - ns;
			const int total = __shfl(incl, 63, 64);
			for (int s0 = 0; s0 < total; s0 += 64)
			{
				const int s = s0 + lane;
				int r = 0;
#pragma unroll
				for (int b = 32; b > 0; b >>= 1)
					r += (__shfl(excl, r + b, 64) <= s) ? b : 0;
				const uint32_t key = (uint32_t)__shfl((int)rc.key, r, 64);
				const uint32_t pv = (uint32_t)__shfl((int)rc.pv, r, 64);
				const int slot = s - __shfl(excl, r, 64);
				if (s >= total)
					continue;
				const float val = v[key];
				float4 e[kDDVec];
#pragma unroll
				for (int j = 0; j < kDDVec; j++)
					e[j] = pl.ray[kDDVec * (uint64_t)key + j];
				if (val == 0.f || ddFrame(e[5]) != f)
					continue;
				DDRay D;
				ddLoad(a.im, a.pr, e, true, D);
				// the pair's slice: first + slot of the tile's slab
				const int xyi = (D.flag_y ? B.y0 : B.x0) + (int)(pv & 63u) + slot;
				if (xyi >= (D.flag_y ? B.y1 : B.x1))
					continue;
				const int ylo = D.flag_y ? B.x0 : B.y0, yhi = (D.flag_y ? B.x1 : B.y1) - 1;
				auto sweep = [&](const DDRay& Q) {
					ddVisit<TOF>(a.im, Q, a.pr, xyi, xyi, ylo, yhi, B.z0, B.z1 - 1,
					             [&](int, int xs, int yxi, int zi, float w) {
						             const int gx = Q.flag_y ? yxi : xs,
						                       gy = Q.flag_y ? xs : yxi;
						             add((uint32_t)(B.base + zi * B.sz + gy * B.sy + gx * B.sx),
						                 gx, gy, zi, val * w);
					             });
				};
				// (two copies of the sweep: one reads the PSF rows from LDS)
				if (psfL != nullptr)
				{
					psfRowsLds(a.pr.psf, psfL, D);
					sweep(D);
				}
				else
					sweep(D);
			}
		}
	};
	auto pipeline = [&](auto&& add, auto&& addIn, auto inRange) {
		if constexpr (DD)
			pipelineDD(add);
		else
		{
			int i = threadIdx.x;
			Rec r0, r1;
			Ray q0;
			readRecord(i, r0);
			readRay(r0.key, q0);
			readRecord(i + NT, r1);
			for (; i < ch.count; i += NT)
			{
				Ray q1;
				readRay(r1.key, q1);
				Rec r2;
				readRecord(i + 2 * NT, r2);
				walk(r0, q0, add, addIn, inRange);
				r0 = r1;
				q0 = q1;
				r1 = r2;
			}
		}
	};
	// c * 2^k rounded to an integer: one FP64 fma with the bias 1.5 * 2^52
	// (|c * 2^k| < 2^51, so the sum stays in [2^52, 2^53), whose ulp is 1:
	// the fma's rounding is the round-to-nearest-even of c * 2^k, which is
	// exact in FP64), the integer being the sum's bits less the bias's.
	// The bias's low word is zero, so that is one 32-bit subtract.
	const double scale = ldexp(1.0, k);
	constexpr double kBias = 6755399441055744.0;  // 1.5 * 2^52
	constexpr unsigned long long kBiasBits = 0x4338000000000000ull;
	if (fixed)
	{
		wordBytes(8);
		auto addFixed = [&](uint32_t idx, float c) {
			const double d = fma((double)c, scale, kBias);
			ldsAdd64(reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(acc64) + idx),
			         (long long)((unsigned long long)__double_as_longlong(d) - kBiasBits));
		};
		pipeline(
		    [&](uint32_t idx, int gx, int gy, int gz, float c) {
			    if (fabsf(c) < climit)
				    addFixed(idx, c);
			    else  // out of the fixed-point range, or not finite
				    unsafeAtomicAdd(yf + (int64_t)gz * a.im.nxy + (int64_t)gy * a.im.nx + gx, c);
		    },
		    [&](uint32_t idx, int, int, int, float c) { addFixed(idx, c); },
		    std::integral_constant<bool, !DD>{});
	}
	else
	{
		wordBytes(4);
		auto addFloat = [&](uint32_t idx, int, int, int, float c) {
			ldsAdd(reinterpret_cast<float*>(reinterpret_cast<char*>(acc) + idx), c);
		};
		pipeline(addFloat, addFloat, std::integral_constant<bool, false>{});
	}
	__syncthreads();
	for (int i = threadIdx.x; i < NV; i += NT)
	{
		const int lx = i % TX, ly = (i / TX) % TY, lz = i / (TX * TY);
		const int gx = B.x0 + lx, gy = B.y0 + ly, gz = B.z0 + lz;
		const float s = fixed ? ldexpf((float)(long long)acc64[i], -k) : acc[i];
		if (s != 0.f && gx < B.x1 && gy < B.y1 && gz < B.z1)
			unsafeAtomicAdd(yf + (int64_t)gz * a.im.nxy + (int64_t)gy * a.im.nx + gx, s);
	}
	__syncthreads();  // acc is reused by the next chunk
}

// One workgroup per chunk (pl.queue == null), or a persistent grid taking
// chunks from the plan's queue until it is empty.  Every workgroup leaves
// once the queue index passes the chunk count, so the grid always drains.
template <bool DD, bool TOF, bool S64, int SHAPE>
__global__ void __launch_bounds__(tileThreads(SHAPE))
    kTileBack(KArgs a, DTiles T, DPlanRead pl, const float* __restrict__ v,
              float* __restrict__ y)
{
	// NV 64-bit words (fixed point), or NV floats for the 32^3 shape
	constexpr int NV = tileDim(SHAPE, 0) * tileDim(SHAPE, 1) * tileDim(SHAPE, 2);
	__shared__ unsigned long long acc[SHAPE == 2 ? NV / 2 : NV];
	extern __shared__ double psfLds[];
	const double* psfL = nullptr;
	if constexpr (DD)
	{
		psfL = psfStage(a.pr.psf, true, psfLds);
		__syncthreads();
	}
	if (pl.queue == nullptr)
	{
		tileChunk<DD, TOF, S64, SHAPE>(a, T, pl, pl.chunks[blockIdx.x], v, y, acc, psfL);
		return;
	}
	// the next chunk's index is passed through the tile's first word (read
	// between two barriers, before the chunk clears the tile), so the
	// workgroup's LDS is the tile alone: 32 KB for 16^3, five per CU
	for (;;)
	{
		if (threadIdx.x == 0)
			acc[0] = (unsigned long long)atomicAdd(&pl.queue[0], 1u);
		__syncthreads();
		const int64_t c = (int64_t)acc[0];
		__syncthreads();
		if (c >= pl.nChunks)
			break;
		tileChunk<DD, TOF, S64, SHAPE>(a, T, pl, pl.chunks[c], v, y, acc, psfL);
	}
	if (threadIdx.x == 0)
	{
		__threadfence();
		if (atomicAdd(&pl.queue[1], 1u) == gridDim.x - 1)
		{
			atomicExch(&pl.queue[0], 0u);
			atomicExch(&pl.queue[1], 0u);
		}
	}
}

// EM ratio per LOR (computeEMUpdateImage, OSEM_CPU.cpp:329-374): the
// forward projection of x, corrected, against the measurement; v[e] = 0
// where the LOR is skipped (zero measurement / factor, invalid LOR, or a
// denominator at or below the threshold).
// The hardware deals workgroups to the XCDs round-robin (workgroup b runs on
// XCD b % kXcds).  With the grid a multiple of kXcds, workgroup b takes the
// range (b % kXcds) * (grid / kXcds) + b / kXcds instead, so each XCD works
// through one contiguous stretch of the LORs: in locality order their voxels
// then share that XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcdRange(int64_t b, int64_t grid)
{
	return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

// The EM ratio of LOR e from its forward projection s (summed over its rays).
__device__ __forceinline__ float emRatioOut(const KArgs& a, int64_t e, float s, float m,
                                           float sf, float af)
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
	// in-vivo attenuation precorrection (OSEM_GPU.cu:1303-1308)
	if (a.c.precorr)
		upd *= a.c.precorr[e];
	float out = 0.f;
	if (fabsf(upd) > a.c.denomThr)
	{
		out = m / upd;
		if (a.c.sens)
			out *= sf;
		if (a.c.att)
			out *= af;
		out *= a.c.g;
	}
	return out;
}

// Siddon: LPE lanes per LOR splitting its rays; the rays read from the plan's
// table (TAB) or made from the resolved LOR.
template <bool TOF, int LPE, bool S64, bool TAB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S64 ? 4 : 5, 8)))
kEmRatio(KArgs a, float* __restrict__ v)
{
	const int64_t total = a.L.n * LPE;
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t t = xcdRange(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; t < total;
	     t += stride)
	{
		const int64_t e = t / LPE;
		const int lane = (int)(t & (LPE - 1));
		const int R = a.pr.numRays;
		// the LOR from the plan's ray table (consecutive LORs: coalesced), or
		// resolved from the detector pair
		Lor l;
		float4 A, B;
		if (TAB)
		{
			const int64_t t0 = e * R + min(lane, R - 1);
			A = a.rays[rayTabVec(S64) * t0];
			B = a.rays[rayTabVec(S64) * t0 + 1];
			l.frame = __float_as_int(B.w);
			l.tof = B.z;
		}
		else
			l = fetchLor(a.sc, a.L, e, a.im.nt);
		const float m = a.c.meas ? a.c.meas[e] : 1.0f;
		const float sf = a.c.sens ? a.c.sens[e] : 1.0f;
		const float af = a.c.att ? a.c.att[e] : 1.0f;
		const bool skip = (m == 0.f || sf == 0.f || af == 0.f || l.frame < 0);
		float s = 0.f;
		if (!skip)
		{
			const float* img = a.xb + (int64_t)l.frame * a.bvox;
			for (int r = lane; r < R; r += LPE)
			{
				float3 q1, q2;
				if (TAB)
				{
					if (r != lane)
					{
						A = a.rays[rayTabVec(S64) * (e * R + r)];
						B = a.rays[rayTabVec(S64) * (e * R + r) + 1];
					}
					q1 = make_float3(A.x, A.y, A.z);
					q2 = make_float3(A.w, B.x, B.y);
				}
				else
				{
					rayLine(a.pr, l, r, q1, q2);
					q1 = make_float3(q1.x - a.im.ox, q1.y - a.im.oy, q1.z - a.im.oz);
					q2 = make_float3(q2.x - a.im.ox, q2.y - a.im.oy, q2.z - a.im.oz);
				}
				uint32_t c = 0;
				s += siddonRay<MODE_FWD, TOF, S64, true>(a.im, const_cast<float*>(img), q1, q2,
				                                         0.f, a.pr.tof, l.tof, c);
			}
		}
#pragma unroll
		for (int o = 1; o < LPE; o <<= 1)
			s += __shfl_xor(s, o, LPE);
		if (lane == 0)
			v[e] = skip ? 0.f : emRatioOut(a, e, s, m, sf, af);
	}
}

// DD: one lane per LOR, its set-up read from the plan's table (TAB) or made
// from the resolved LOR.
template <bool TOF, bool S64, bool TAB>
__global__ void __launch_bounds__(256) kEmRatioDD(KArgs a, float* __restrict__ v)
{
	extern __shared__ double psfLds[];
	const double* psfL = psfStage(a.pr.psf, false, psfLds);
	__syncthreads();
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t e = xcdRange(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; e < a.L.n;
	     e += stride)
	{
		const float m = a.c.meas ? a.c.meas[e] : 1.0f;
		const float sf = a.c.sens ? a.c.sens[e] : 1.0f;
		const float af = a.c.att ? a.c.att[e] : 1.0f;
		int frame;
		Lor l;
		float4 de[kDDVec];
		if (TAB)
		{
			de[5] = a.rays[kDDVec * e + 5];
			frame = ddFrame(de[5]);
		}
		else
		{
			l = fetchLor(a.sc, a.L, e, a.im.nt);
			frame = l.frame;
		}
		const bool skip = (m == 0.f || sf == 0.f || af == 0.f || frame < 0);
		float out = 0.f;
		if (!skip)
		{
			DDRay D;
			if (TAB)
			{
#pragma unroll
				for (int i = 0; i < kDDVec; i++)
					if (i != 5)
						de[i] = a.rays[kDDVec * e + i];
				ddLoad(a.im, a.pr, de, false, D);
			}
			else
			{
				const float3 d1 = make_float3(l.p1.x - a.im.ox, l.p1.y - a.im.oy, l.p1.z - a.im.oz);
				const float3 d2 = make_float3(l.p2.x - a.im.ox, l.p2.y - a.im.oy, l.p2.z - a.im.oz);
				ddSetup<TOF, S64>(a.im, d1, d2, l.n1, l.n2, a.pr, l.tof, false, D);
			}
			const float* img = a.x + (int64_t)frame * a.im.nvox;
			// (two copies of the sweep: one reads the PSF rows from LDS)
			float s;
			if (psfL != nullptr)
			{
				psfRowsLds(a.pr.psf, psfL, D);
				s = ddSweepFwd<TOF>(a.im, img, D, a.pr);
			}
			else
				s = ddSweepFwd<TOF>(a.im, img, D, a.pr);
			out = emRatioOut(a, e, s, m, sf, af);
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
template <bool DD, bool TOF, bool S64>
static hipError_t binT(int pass, const KArgs& a, const DTiles& T,
                       const DPlan& pl, int grid, hipStream_t st)
{
	const int nB = T.nTiles * pl.C;
	const int64_t nRays = a.L.n * (int64_t)pl.R;
	if (pass == 0)
		kBin<DD, TOF, S64, 0><<<grid, 256, (size_t)nB * sizeof(int), st>>>(a.im, a.pr, nRays,
		                                                                    T, pl);
	else
		kBin<DD, TOF, S64, 1><<<grid, 256, binLdsBytes(nB), st>>>(a.im, a.pr, nRays, T, pl);
	return hipGetLastError();
}

hipError_t launchRayPrep(const KArgs& a, float4* tab, hipStream_t st)
{
	const int64_t total = a.L.n * (int64_t)a.pr.numRays;
	if (total == 0)
		return hipSuccess;
	const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1), 65536);
	// Siddon: the FP64 set-up stores the end points only (two float4 per
	// ray); the FP32 set-up stores its set-up too (four), with the TOF
	// constants when the projector has TOF
	if (a.pr.type != 1)
	{
		if (a.pr.setup64)
			kRayPrep<false, false, true><<<g, 256, 0, st>>>(a, tab);
		else if (a.pr.useTof)
			kRayPrep<false, true, false><<<g, 256, 0, st>>>(a, tab);
		else
			kRayPrep<false, false, false><<<g, 256, 0, st>>>(a, tab);
	}
	else if (a.pr.useTof)
	{
		if (a.pr.setup64)
			kRayPrep<true, true, true><<<g, 256, 0, st>>>(a, tab);
		else
			kRayPrep<true, true, false><<<g, 256, 0, st>>>(a, tab);
	}
	else
	{
		if (a.pr.setup64)
			kRayPrep<true, false, true><<<g, 256, 0, st>>>(a, tab);
		else
			kRayPrep<true, false, false><<<g, 256, 0, st>>>(a, tab);
	}
	return hipGetLastError();
}

template <bool DD>
static hipError_t binKind(int pass, const KArgs& a, const DTiles& T, const DPlan& pl,
                          int grid, hipStream_t st)
{
	const bool tof = a.pr.useTof != 0;
	if (a.pr.setup64)
		return tof ? binT<DD, true, true>(pass, a, T, pl, grid, st)
		           : binT<DD, false, true>(pass, a, T, pl, grid, st);
	return tof ? binT<DD, true, false>(pass, a, T, pl, grid, st)
	           : binT<DD, false, false>(pass, a, T, pl, grid, st);
}

hipError_t launchBin(int pass, const KArgs& a, const DTiles& T, const DPlan& pl,
                     int grid, hipStream_t st)
{
	return a.pr.type == 1 ? binKind<true>(pass, a, T, pl, grid, st)
	                      : binKind<false>(pass, a, T, pl, grid, st);
}

template <bool DD, bool TOF, bool S64, int SHAPE>
static void tileBackT(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                      const float* v, float* y, unsigned grid, hipStream_t st)
{
	kTileBack<DD, TOF, S64, SHAPE>
	    <<<grid, tileThreads(SHAPE), DD ? a.pr.psf.ldsBytes : 0, st>>>(a, T, pl, v, y);
}

template <bool DD, bool TOF, bool S64>
static void tileBackShape(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                          const float* v, float* y, unsigned grid, hipStream_t st)
{
	switch (T.shape)
	{
	case 1:
		tileBackT<DD, TOF, S64, 1>(a, T, pl, v, y, grid, st);
		break;
	case 2:
		tileBackT<DD, TOF, S64, 2>(a, T, pl, v, y, grid, st);
		break;
	case 3:
		tileBackT<DD, TOF, S64, 3>(a, T, pl, v, y, grid, st);
		break;
	default:
		tileBackT<DD, TOF, S64, 0>(a, T, pl, v, y, grid, st);
		break;
	}
}

template <bool DD>
static void tileBackKind(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                         const float* v, float* y, unsigned grid, hipStream_t st)
{
	const bool tof = a.pr.useTof != 0;
	if (a.pr.setup64)
	{
		if (tof)
			tileBackShape<DD, true, true>(a, T, pl, v, y, grid, st);
		else
			tileBackShape<DD, false, true>(a, T, pl, v, y, grid, st);
	}
	else
	{
		if (tof)
			tileBackShape<DD, true, false>(a, T, pl, v, y, grid, st);
		else
			tileBackShape<DD, false, false>(a, T, pl, v, y, grid, st);
	}
}

__global__ void __launch_bounds__(256) kZeroWords(unsigned int* p, int n)
{
	for (int i = threadIdx.x; i < n; i += blockDim.x)
		p[i] = 0u;
}

// hist[b] += number of i with biased exponent field b of |v[i]| (255: not
// finite).
__global__ void __launch_bounds__(256) kVExpHist(const float* __restrict__ v, int64_t n,
                                                 unsigned int* hist)
{
	__shared__ unsigned int h[256];
	h[threadIdx.x] = 0u;
	__syncthreads();
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		atomicAdd(&h[(__float_as_uint(v[i]) >> 23) & 0xffu], 1u);
	__syncthreads();
	if (h[threadIdx.x] != 0u)
		atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// *e = E, the binade of the median of the nonzero finite |v|: the smallest
// exponent with |v| < 2^E for at least half of them (|v| < 2^(b - 126) for
// biased exponent b); -126 when every value is zero or subnormal.
__global__ void kPickExp(const unsigned int* __restrict__ hist, int64_t n, int* e)
{
	if (threadIdx.x != 0)
		return;
	uint64_t total = 0;
	for (int b = 1; b <= 254; b++)
		total += hist[b];
	uint64_t above = 0;
	int E = -126;
	for (int b = 254; b >= 1 && total > 0; b--)
	{
		above += hist[b];
		if (2 * above >= total)
		{
			E = b - 126;
			break;
		}
	}
	*e = E;
}

hipError_t launchTileBack(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                          const float* v, float* y, int persistent, hipStream_t st)
{
	if (pl.nChunks <= 0)
		return hipSuccess;
	if (pl.vexp != nullptr)
	{
		// the scale word is followed by the 256-bin exponent histogram
		int* ve = const_cast<int*>(pl.vexp);
		unsigned int* hist = reinterpret_cast<unsigned int*>(ve) + 2;
		const int64_t n = a.L.n;
		kZeroWords<<<1, 256, 0, st>>>(hist, 256);
		const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>((n + 2047) / 2048, 1), 2048);
		kVExpHist<<<g, 256, 0, st>>>(v, n, hist);
		kPickExp<<<1, 64, 0, st>>>(hist, n, ve);
		const hipError_t e = hipGetLastError();
		if (e != hipSuccess)
			return e;
	}
	DPlanRead q = pl;
	unsigned grid;
	if (persistent > 0 && pl.queue != nullptr)
		grid = (unsigned)std::min<int64_t>(persistent, pl.nChunks);
	else
	{
		q.queue = nullptr;
		grid = (unsigned)pl.nChunks;
	}
	if (a.pr.type == 1)
		tileBackKind<true>(a, T, q, v, y, grid, st);
	else
		tileBackKind<false>(a, T, q, v, y, grid, st);
	return hipGetLastError();
}

template <bool TOF, bool S64, bool TAB>
static hipError_t ratioLpe(const KArgs& a, int lpe, dim3 g, float* v, hipStream_t st)
{
	switch (lpe)
	{
	case 1:
		kEmRatio<TOF, 1, S64, TAB><<<g, 256, 0, st>>>(a, v);
		break;
	case 2:
		kEmRatio<TOF, 2, S64, TAB><<<g, 256, 0, st>>>(a, v);
		break;
	case 4:
		kEmRatio<TOF, 4, S64, TAB><<<g, 256, 0, st>>>(a, v);
		break;
	case 8:
		kEmRatio<TOF, 8, S64, TAB><<<g, 256, 0, st>>>(a, v);
		break;
	default:
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

template <bool TOF, bool S64>
static hipError_t ratioT(const KArgs& a, int lpe, dim3 g, float* v, hipStream_t st)
{
	if (a.pr.type == 1)
	{
		if (a.rays != nullptr)
			kEmRatioDD<TOF, S64, true><<<g, 256, a.pr.psf.ldsBytes, st>>>(a, v);
		else
			kEmRatioDD<TOF, S64, false><<<g, 256, a.pr.psf.ldsBytes, st>>>(a, v);
		return hipGetLastError();
	}
	return a.rays != nullptr ? ratioLpe<TOF, S64, true>(a, lpe, g, v, st)
	                         : ratioLpe<TOF, S64, false>(a, lpe, g, v, st);
}

__global__ void __launch_bounds__(256)
    kToBricks(const float* __restrict__ x, float* __restrict__ xb, int nx, int ny, int nz,
              int64_t nvox, int64_t bvox, int64_t total)
{
	const int64_t stride = (int64_t)gridDim.x * blockDim.x;
	const int bnx = brickNx(nx), bny = brickNy(ny);
	for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += stride)
	{
		const int64_t f = o / bvox;
		const int64_t r = o - f * bvox;
		const int64_t b = r >> 5;
		const int w = (int)(r & 31);
		const int64_t t = b / bnx;
		const int vx = (int)(b - t * bnx) * 4 + (w & 3);
		const int vy = (int)(t % bny) * 4 + ((w >> 2) & 3);
		const int vz = (int)(t / bny) * 2 + (w >> 4);
		xb[o] = (vx < nx && vy < ny && vz < nz)
		            ? x[f * nvox + ((int64_t)vz * ny + vy) * nx + vx]
		            : 0.f;
	}
}

hipError_t launchToBricks(const DImage& im, const float* x, float* xb, hipStream_t st)
{
	const int64_t bvox = brickFrameVoxels(im.nx, im.ny, im.nz);
	const int64_t total = bvox * im.nt;
	const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1), 65536);
	kToBricks<<<g, 256, 0, st>>>(x, xb, im.nx, im.ny, im.nz, im.nvox, bvox, total);
	return hipGetLastError();
}

hipError_t launchEmRatio(const KArgs& a, int lpe, float* v, hipStream_t st)
{
	if (a.pr.type != 1 && (a.xb == nullptr || a.bvox != brickFrameVoxels(a.im.nx, a.im.ny, a.im.nz)))
		return hipErrorInvalidValue;
	if (a.pr.type == 1)
		lpe = 1;
	const int64_t threads = a.L.n * lpe;
	int64_t nb = (threads + 255) / 256;
	if (nb > 0x7ffffff0)
		nb = 0x7ffffff0;
	nb = (nb + kXcds - 1) / kXcds * kXcds;  // xcdRange is a bijection on [0, nb)
	if (nb < kXcds)
		nb = kXcds;
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
