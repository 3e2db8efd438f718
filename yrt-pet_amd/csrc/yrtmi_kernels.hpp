// Launchers of the yrtmi kernels (host side).
#pragma once

#include <algorithm>

#include "yrtmi_rays.hpp"

namespace yrtmi
{

// mode: MODE_FWD / MODE_BACK / MODE_EM / MODE_SENS / MODE_COUNT.
// lpe: lanes per event (1, 2, 4, 8; forced to 1 for DD).  grid <= 0 sizes
// the grid from the event count; otherwise the kernel grid-strides.
hipError_t launchEvents(int mode, const KArgs& a, int lpe, int block,
                        int64_t grid, hipStream_t st);
hipError_t launchAttFactors(const KArgs& a, hipStream_t st);
hipError_t launchEmUpdate(float* x, const float* u, const float* s,
                          int64_t nvox, int64_t nt, int64_t sFrames, float thr,
                          hipStream_t st);
hipError_t launchThreshold(float* x, const float* m, int64_t nvox, int64_t nt,
                           int64_t mFrames, float thr, float les, float leo,
                           float gts, float gto, hipStream_t st);
hipError_t launchFill(float* x, int64_t n, float v, hipStream_t st);
// Histogram values of detector pairs (list-mode events).
struct PairGeom
{
	uint32_t dpr, nrings, ndet, numDOI, numDOIPoss, numR, numPhi, numZBinDiff;
	int32_t maxRingDiff;
};
hipError_t launchGatherPairs(const float* histo, const uint32_t* d1, const uint32_t* d2,
                             int64_t n, const int32_t* ringMap, const PairGeom& g, float* out,
                             int* err, hipStream_t st);
hipError_t launchPairBins(const uint32_t* d1, const uint32_t* d2, int64_t n,
                          const int32_t* ringMap, const PairGeom& g, uint32_t* out, int* err,
                          hipStream_t st);
hipError_t launchFillI32(int32_t* x, int64_t n, int32_t v, hipStream_t st);

// Image-space PSF: one axis (0 x, 1 y, 2 z) of the separable circular
// convolution of every frame (nx * ny * nz * frames = total < 2^31 voxels),
// taps passed by value.
constexpr int kMaxPsfTaps = 63;
struct PsfTaps
{
	float k[kMaxPsfTaps];
	int n;  // odd, <= kMaxPsfTaps
};
hipError_t launchPsfAxis(int axis, const float* in, float* out, const PsfTaps& t, int nx,
                         int ny, int nz, int total, hipStream_t st);

// Time-averaged moved image: out[v] += sum over frames of weight *
// trilinear(in, inverse transform * centre of v), for one frame of
// nx * ny * nz < 2^31 voxels.  `frames` holds 4 float4 per frame: the rows
// (r_i0, r_i1, r_i2, t_i) of the inverse transform, then {weight, 0, 0, 0}.
struct MoveGeom
{
	float v[3], len[3], off[3];  // voxel size (length / n), length, offset
	float origin[3], inv[3];     // centre of voxel 0, 1 / voxel size
	int nx, ny, nz;
};
hipError_t launchTimeAverageMove(const float* in, float* out, const float4* frames, int n,
                                 const MoveGeom& g, hipStream_t st);

// Histogram subset geometry for the bin-list kernels.
struct SubsetGeom
{
	uint32_t numPhi, numR, nps, S, s;
};
constexpr int kCompactRounds = 8;
constexpr int kCompactPerBlock = 256 * kCompactRounds;
// Nonzero bins of a histogram subset, in subset order: per-block counts,
// then (after an exclusive scan of the counts) the writes.
hipError_t launchCountNonzero(const float* v, const SubsetGeom& g, int64_t n,
                              int* blockCount, int64_t nBlocks, hipStream_t st);
hipError_t launchWriteNonzero(const float* v, const SubsetGeom& g, int64_t n,
                              const int64_t* blockOff, int64_t nBlocks,
                              uint32_t* bins, float* meas, hipStream_t st);
// dst[i] = src[s + i*S], i < n, elements of elemBytes bytes.
hipError_t launchInterleaved(const void* src, int elemBytes, int64_t n, int64_t S,
                             int64_t s, void* dst, hipStream_t st);
hipError_t launchGather(const float* src, int64_t srcN, const uint32_t* idx,
                        int64_t n, int invert, float* out, hipStream_t st);
hipError_t launchSubsetGather(const float* h, const SubsetGeom& g, int64_t n,
                              int invert, float* out, hipStream_t st);
hipError_t launchAxpby(int64_t n, float a, const float* x, float b, float* y,
                       hipStream_t st);
// y += sum of the nk per-XCD partial images (row stride `stride`), which are
// left zeroed.
hipError_t launchFoldPartials(float* y, float* part, int64_t stride, int nk,
                              int64_t n, hipStream_t st);

}  // namespace yrtmi
