// Lookup of the register-resident 3D kernel instances.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {
namespace k3 {

template <typename T>
struct K3Entry {
  int N, lanes;
  int (*adv)(const Params<T>&, const T*, T*, hipStream_t);
  int (*mass)(const Params<T>&, const T*, T*, hipStream_t);
};

// nullptr if (N, lanes) is not compiled
template <typename T>
const K3Entry<T>* find(int N, int lanes);

}  // namespace k3
}  // namespace gd
