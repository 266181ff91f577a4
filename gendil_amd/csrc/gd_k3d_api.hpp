// Registry and launch of the register-resident 3D kernel instances.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {
namespace k3 {

template <typename T>
struct K3Params;

template <typename T>
using K3Launch = int (*)(const K3Params<T>&, const T*, T*, hipStream_t);

template <typename T>
struct K3Entry {
  int N, lanes, adv, mass, bcg, sig;
  K3Launch<T> fn;
};

template <typename T>
struct K3List {
  const K3Entry<T>* e;
  int n;
};

K3List<double> instances_f64_n2();
K3List<double> instances_f64_n3();
K3List<double> instances_f64_n4();
K3List<float> instances_f32_n2();
K3List<float> instances_f32_n3();
K3List<float> instances_f32_n4();

// Apply y = mass_coef M x + adv_coef A x with the register-resident kernel
// if an instance exists for this plan and lane count; returns
// GD_ERR_UNSUPPORTED (without touching the GPU) otherwise.
template <typename T>
int apply(const gd_plan* plan, int lanes, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s);

// Whether apply() can serve this plan for the given lane count.
template <typename T>
bool available(const gd_plan* plan, int lanes, bool adv, bool mass);

}  // namespace k3
}  // namespace gd
