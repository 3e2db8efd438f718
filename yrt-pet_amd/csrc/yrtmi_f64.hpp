// FP64 Siddon projector (host-side launcher).
#pragma once

#include "yrtmi_device.hpp"

namespace yrtmi
{

// fwd: out[e] = sum over rays of the FP64 line integrals of x / rays;
// else y += the FP64 backprojection of vals.
hipError_t launchProj64(bool fwd, const KArgs& a, const double* x, double* y,
                        const double* vals, double* out, hipStream_t st);

}  // namespace yrtmi
