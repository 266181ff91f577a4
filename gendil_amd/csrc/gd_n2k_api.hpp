// Host-side interface of the two-node Kronecker apply (gd_n2k.hpp): 4D-6D
// uniform plans of order 1 with constant beta, sigma and inflow data 0.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {
namespace n2k {

// variant code of this kernel in the plan's tuning table
constexpr int kVariant = 64;

template <typename T>
bool available(const gd_plan* plan, bool adv, bool mass);

// GD_ERR_UNSUPPORTED when the plan or coefficients do not fit the kernel
template <typename T>
int apply(const gd_plan* plan, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s);

}  // namespace n2k
}  // namespace gd
