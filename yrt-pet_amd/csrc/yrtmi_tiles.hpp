// Tile plans: Siddon walks and distance-driven footprints binned by image
// tile, and the kernels that build and use them (host-side launchers).
#pragma once

#include "yrtmi_rays.hpp"

#include <algorithm>

namespace yrtmi
{

// Image tiles held in LDS by one workgroup, by shape (64-bit fixed-point
// voxels, FP32 for shape 2):
//   0: 16 x 16 x 16 (32 KB, 256 threads)   1: 32 x 32 x 16 (128 KB, 512)
//   2: 32 x 32 x 32 (128 KB FP32, 1024)    3: 32 x 32 x 8 (64 KB, 512)
constexpr int kNumTileShapes = 4;
__host__ __device__ constexpr int tileDim(int shape, int axis)
{
	return shape == 0 ? 16 : (shape == 2 ? 32 : (axis < 2 ? 32 : (shape == 1 ? 16 : 8)));
}
__host__ __device__ constexpr int tileThreads(int shape)
{
	return shape == 0 ? 256 : (shape == 2 ? 1024 : 512);
}
// Binning keeps a counter (count pass) and a counter plus a 32-bit base
// (fill pass; a plan holds fewer than 2^32 records) per (tile, length
// class) in LDS; length classes are added while tiles x classes <=
// kMaxBins / 2.
constexpr int kMaxBins = 16384;
constexpr int kMaxClasses = 16;

// Entry record of a ray in a tile: voxel within the tile (5 bits per axis),
// the ray's plane shift per axis (sidPlaneShift, 4 bits each, offset 8) and
// whether the tile is where the ray's walk starts (bit 31).
constexpr int kMaxShift = 7;
__host__ __device__ inline uint32_t packEntry(int lx, int ly, int lz, int dx, int dy,
                                              int dz, bool first)
{
	return (uint32_t)lx | ((uint32_t)ly << 5) | ((uint32_t)lz << 10) |
	       ((uint32_t)(dx + 8) << 15) | ((uint32_t)(dy + 8) << 19) |
	       ((uint32_t)(dz + 8) << 23) | (first ? 0x80000000u : 0u);
}

// A record's fine length class (0 .. 15), kept in entry-word bits that
// the record's fields leave free (Siddon: 27..30, packEntry; DD: 12..15,
// above the slice range) for the plan's per-chunk sort (kChunkSort).
// Siddon: its segments in the tile in 16 equal steps of the most a tile
// piece can hold (tx + ty + tz); DD: its footprint voxels per slice.
__host__ __device__ inline uint32_t withFineClass(bool dd, uint32_t pv, int segs, int L)
{
	if (dd)
	{
		const int ns = (int)((pv >> 6) & 63u) - (int)(pv & 63u) + 1;
		const int c = std::min(15, std::max(0, (segs + ns - 1) / std::max(ns, 1)));
		return pv | ((uint32_t)c << 12);
	}
	const int c = std::min(15, std::max(0, segs * 16 / std::max(L, 1)));
	return pv | ((uint32_t)c << 27);
}

// Words per plan record: Siddon key, alpha, entry; DD key, entry.
__host__ __device__ constexpr int recWords(bool dd) { return dd ? 2 : 3; }

__host__ __device__ inline int fineClassOf(bool dd, uint32_t pv)
{
	return (int)((pv >> (dd ? 12 : 27)) & 15u);
}

// Tiles of an image with nt frames: ntx * nty * ntz per frame, numbered
// ((f * ntz + tz) * nty + ty) * ntx + tx; tiles of tx * ty * tz voxels.
struct DTiles
{
	int32_t ntx, nty, ntz, nTiles;
	int32_t tx, ty, tz, shape;
	int32_t lx, ly, lz;  // log2 of tx, ty, tz (tile sides are powers of two)
};

// A chunk of one tile's records, processed by one workgroup.
struct TileChunk
{
	int64_t begin;
	int32_t tile;
	int32_t count;
};

// Plan being built.  Records are binned by (tile, length class): a record's
// class counts the cuts at or below its number of segments in the tile (DD:
// its footprint voxels in the tile), and
// within a tile the records are stored class by class, so that the lanes of
// a wave walk pieces of similar length.  Bin b = tile * C + class.
struct DPlan
{
	int32_t R;                  // rays per LOR
	int32_t C;                  // length classes (1 .. kMaxClasses)
	int32_t cut[kMaxClasses - 1];  // class k starts at cut[k-1] segments
	unsigned long long* count;  // [nTiles * C] records per bin (count pass)
	unsigned int* blockCnt;     // [grid][nTiles * C] records per workgroup and bin
	unsigned int* cursor;       // [nTiles * C] fill cursor within the bin
	const int64_t* tileOff;     // [nTiles * C] first record of the bin
	// records of recWords(DD) words: the key (lor * R + ray; DD: lor), the
	// alpha where the walk enters the tile (Siddon, its bits), the entry
	// word (packEntry; DD: the record's slices in the slab, binRayDD) with
	// its fine length class
	uint32_t* rec;
	const float4* ray;          // ray table (rayEntry + sidStore / ddStore, kRayPrep)
	int* err;                   // set when a tile range would overflow
};

// Ray table of a Siddon plan, rayTabVec(S64) float4 per (LOR, ray) t = lor *
// R + ray.  The first two: the ray's end points after multi-ray offsets,
// motion and the image offset (q1, q2, exactly as the walk sets it up), its
// TOF value and its frame (-1: skipped); the backprojection reads them
// instead of resolving the LOR.  With the FP32 setup two more hold the
// ray's set-up (sidStore), from which a walk re-entered at a tile record
// starts without redoing it.
constexpr int rayTabVec(bool S64)
{
	return S64 ? 2 : 4;
}

__device__ __forceinline__ void rayEntry(float4* tab, int vec, uint32_t t, float3 q1, float3 q2,
                                         float tof, int32_t frame)
{
	tab[vec * (uint64_t)t] = make_float4(q1.x, q1.y, q1.z, q2.x);
	tab[vec * (uint64_t)t + 1] = make_float4(q2.y, q2.z, tof, __int_as_float(frame));
}

// Plan in use.  queue != null: persistent launch, workgroups take chunks
// from queue[0] (queue[1] counts finished workgroups; the last one resets
// both, so the next launch starts from zero without a memset).
struct DPlanRead
{
	int32_t R;
	const TileChunk* chunks;
	int64_t nChunks;
	const uint32_t* rec;  // records, recWords(DD) words each (DPlan::rec)
	const float4* ray;    // rayEntry + sidStore (Siddon) or ddStore (DD) table
	unsigned int* queue;
	// binade E of the median nonzero |v| of the launch (kVExpHist + kPickExp
	// before each launch): sets the fixed-point scale of the tile
	// accumulation
	const int* vexp;
};

// LDS of the fill pass: an int counter and a 32-bit base per bin.
inline size_t binLdsBytes(int nB)
{
	return (size_t)nB * (sizeof(int) + sizeof(uint32_t));
}
// The plan's ray table for a.L (before the binning passes).
hipError_t launchRayPrep(const KArgs& a, float4* tab, hipStream_t st);
// pass 0: count; pass 1: fill (same grid in both).
hipError_t launchBin(int pass, const KArgs& a, const DTiles& T, const DPlan& pl,
                     int grid, hipStream_t st);
// Each chunk's records (nChunks chunks, device table) ordered by their fine
// length class, in place (kChunkSort).
hipError_t launchChunkSort(bool dd, const DPlan& pl, const TileChunk* chunks, int64_t nChunks,
                           hipStream_t st);
// y += backprojection of v (per LOR) through the plan's records: one
// workgroup per chunk, or `persistent` workgroups on the chunk queue.
hipError_t launchTileBack(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                          const float* v, float* y, int persistent, hipStream_t st);
// xb (nt frames of brickFrameVoxels floats) = bricked copy of x (nt frames
// of the image's voxels), padding voxels 0.
hipError_t launchToBricks(const DImage& im, const float* x, float* xb, hipStream_t st);
// v[e] = EM ratio of LOR e (0 where skipped), Siddon or DD (a.pr.type);
// Siddon reads a.xb (launchToBricks of a.x).
hipError_t launchEmRatio(const KArgs& a, int lpe, float* v, hipStream_t st);
// v[e] = sensitivity value of LOR e (factor x attenuation).
hipError_t launchSensValue(const KArgs& a, float* v, hipStream_t st);

}  // namespace yrtmi
