// Locality order of LOR sets and permutation of per-LOR arrays (host-side
// launchers).
#pragma once

#include "yrtmi_device.hpp"

namespace yrtmi
{

// perm = 0..n-1 sorted stably by each LOR's locality key (frame, axial
// midpoint, tilt, transaxial angle, offset).  Allocates its temporaries and
// synchronises `st`.
hipError_t lorOrder(const DScanner& S, const DLors& L, const DImage& im, uint32_t* perm,
                    hipStream_t st);
// dst[i] = src[perm[i]], elements of 1, 2, 4 or 8 bytes.
hipError_t launchPermute(const void* src, int elemBytes, const uint32_t* perm, int64_t n,
                         void* dst, hipStream_t st);

}  // namespace yrtmi
