// Launchers of the yrtmi kernels (host side).
#pragma once

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
// y += sum of the nk per-XCD partial images (row stride `stride`), which are
// left zeroed.
hipError_t launchFoldPartials(float* y, float* part, int64_t stride, int nk,
                              int64_t n, hipStream_t st);

}  // namespace yrtmi
