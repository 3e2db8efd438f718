// Tile plans: Siddon walks binned by image tile, and the kernels that build
// and use them (host-side launchers).
#pragma once

#include "yrtmi_rays.hpp"

namespace yrtmi
{

// Image tile held in LDS by one workgroup: 32 x 32 x 16 FP32 voxels, 64 KB.
constexpr int kTileX = 32, kTileY = 32, kTileZ = 16;
constexpr int kTileVox = kTileX * kTileY * kTileZ;
constexpr int kTileThreads = 512;
// Binning keeps two int counters per tile in LDS (64 KB at most).
constexpr int kMaxTiles = 8192;

// Entry record of a ray in a tile: voxel within the tile (5 + 5 + 4 bits),
// the ray's plane shift per axis (sidPlaneShift, 4 bits each, offset 8) and
// whether the tile is where the ray's walk starts (bit 31).
static_assert(kTileX <= 32 && kTileY <= 32 && kTileZ <= 16, "entry packing");
constexpr int kMaxShift = 7;
__host__ __device__ inline uint32_t packEntry(int lx, int ly, int lz, int dx, int dy,
                                              int dz, bool first)
{
	return (uint32_t)lx | ((uint32_t)ly << 5) | ((uint32_t)lz << 10) |
	       ((uint32_t)(dx + 8) << 14) | ((uint32_t)(dy + 8) << 18) |
	       ((uint32_t)(dz + 8) << 22) | (first ? 0x80000000u : 0u);
}

// Tiles of an image with nt frames: ntx * nty * ntz per frame, numbered
// ((f * ntz + tz) * nty + ty) * ntx + tx.
struct DTiles
{
	int32_t ntx, nty, ntz, nTiles;
};

// A chunk of one tile's records, processed by one workgroup.
struct TileChunk
{
	int64_t begin;
	int32_t tile;
	int32_t count;
};

// Plan being built.
struct DPlan
{
	int32_t R;                  // rays per LOR
	unsigned long long* count;  // [nTiles] records per tile (count pass)
	unsigned int* cursor;       // [nTiles] fill cursor within the tile
	const int64_t* tileOff;     // [nTiles] first record of the tile
	uint32_t* key;              // lor * R + ray
	float* alpha;               // alpha where the walk enters the tile
	uint32_t* vox;              // packed entry voxel (packEntry)
	int* err;                   // set when a tile range would overflow
};

// Plan in use.
struct DPlanRead
{
	int32_t R;
	const TileChunk* chunks;
	const uint32_t* key;
	const float* alpha;
	const uint32_t* vox;
};

// pass 0: count; pass 1: fill.  Dynamic LDS 8 * nTiles bytes.
hipError_t launchBin(int pass, const KArgs& a, const DTiles& T, const DPlan& pl,
                     int grid, hipStream_t st);
// y += backprojection of v (per LOR) through the plan's records.
hipError_t launchTileBack(const KArgs& a, const DTiles& T, const DPlanRead& pl,
                          int64_t nChunks, const float* v, float* y,
                          hipStream_t st);
// v[e] = EM ratio of LOR e (0 where skipped), Siddon.
hipError_t launchEmRatio(const KArgs& a, int lpe, float* v, hipStream_t st);
// v[e] = sensitivity value of LOR e (factor x attenuation).
hipError_t launchSensValue(const KArgs& a, float* v, hipStream_t st);

}  // namespace yrtmi
