// Host-side interface of the Kronecker-factored 3D kernel (gd_k3k.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"

namespace gd {
namespace k3k {

template <typename T>
struct KParams;

template <typename T>
using KLaunch = int (*)(const KParams<T>&, const T*, T*, hipStream_t);

template <typename T>
struct KEntry {
  int N, lanes, adv, mass, sig;
  KLaunch<T> fn;
};

template <typename T>
struct KList {
  const KEntry<T>* e;
  int n;
};

KList<double> instances_f64_n2();
KList<double> instances_f64_n3();
KList<double> instances_f64_n4();
KList<float> instances_f32_n2();
KList<float> instances_f32_n3();
KList<float> instances_f32_n4();

// Plan-side cache: per-position factor tables (both precisions).
struct State;
void destroy_state(State* s);

// Rebuild the factor tables for the plan's current coefficients (allocates
// and copies synchronously).  Called by plan creation and by every
// coefficient setter, after the device has drained, so applies never
// allocate or copy.
int refresh(const gd_plan* plan);

// y = mass_coef M x + adv_coef A x if an instance serves this plan and lane
// count; GD_ERR_UNSUPPORTED (nothing launched) otherwise.
template <typename T>
int apply(const gd_plan* plan, int lanes, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s);

template <typename T>
bool available(const gd_plan* plan, int lanes, bool adv, bool mass);

// y += the interior-face terms of the advection operator (GenDiL's
// AdvectionFaceOperator over the interior faces; no cell terms, no faces
// across a periodic wrap or on the boundary) by the streaming kernel;
// GD_ERR_UNSUPPORTED (nothing launched) where it does not serve the plan.
template <typename T>
int apply_faces(const gd_plan* plan, const T* x, T* y, hipStream_t s);

// Lane count of the register-resident Kronecker variant for n nodes per
// dimension, and the code (32..39) of the default LDS-staged variant (-1: none
// compiled for n).
int default_lanes(int n, size_t elem_bytes);
int default_staged(int n, size_t elem_bytes);

}  // namespace k3k
}  // namespace gd
