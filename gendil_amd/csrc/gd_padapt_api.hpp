// Host-side interface of the p-adaptive apply (gd_padapt.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

struct gd_plan;

namespace gd {
namespace pa {

// Threads per element for the n = 4, 3, 2 buckets (64 = one wave, 256 = one
// block); a block of 256 threads holds 256 / tpe elements of one bucket.
struct Variant {
  int tpe4, tpe3, tpe2;
};

int num_variants(int D);
Variant variant(int D, int v);

// Launch one instantiation.  pp: device PParams<T>; it0/it1: number of work
// items of the n=4 and n=3 buckets; grid: total items.
template <typename T>
int launch(int D, bool adv, int v, int grid, int it0, int it1, const void* pp, T mass_coef, T adv_coef, const T* x,
           T* y, hipStream_t s);

// Launch the Kronecker 6D kernel (gd_pa6k.hpp): it0 blocks of one n=4 element,
// it1 blocks of four n=3 elements, then blocks of four n=2 elements.
template <typename T>
int launch_kron(bool adv, int grid, int it0, int it1, const void* pp, T mass_coef, T adv_coef, const T* x, T* y,
                hipStream_t s);

// Variant an apply with variant knob v (-1: default) resolves to.
int resolved_variant(const gd_plan* plan, int v);

// Plan-side state (element buckets, device parameter copies).
struct State;
void destroy_state(State* s);

// Build the element buckets (once) and upload the parameter blocks of both
// precisions for the plan's current coefficients (synchronous copies).
// Called by plan creation and by every coefficient setter, after the device
// has drained, so applies never allocate or copy.
int refresh(const gd_plan* plan);

// Is the plan served by this path?  On false, *why says what is missing.
bool eligible(const gd_plan* plan, const char** why);

template <typename T>
int apply(const gd_plan* plan, int v, T mass_coef, T adv_coef, const T* x, T* y, hipStream_t s);

}  // namespace pa
}  // namespace gd
