// Instance lists of the register-resident 3D kernel, one translation unit per
// (precision, nodes) so they compile in parallel.
#pragma once

#include "gd_k3d.hpp"
#include "gd_k3d_api.hpp"

namespace gd {
namespace k3 {

template <typename T, int N, int LANES, bool ADV, bool MASS, bool BCG, int SIG>
int launch(const K3Params<T>& prm_in, const T* x, T* y, hipStream_t s)
{
  constexpr int NT = 256;
  constexpr int EPB = NT / LANES;
  K3Params<T> prm = prm_in;
  const long long nb = ((long long)prm.E + EPB - 1) / EPB;
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL((k3d_apply<T, N, LANES, ADV, MASS, BCG, SIG, NT>), dim3((unsigned)nb), dim3(NT), 0, s, prm, x,
                     y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

// one-axis signatures compiled: the benchmark field (beta_x(y), beta_y(z),
// beta_z(x)) and all-constant fields; sigma constant
constexpr int kSigCyclic = sig(1, 2, 0, -1);
constexpr int kSigConst = sig(-1, -1, -1, -1);

// every one-axis signature with run-time axes (the default lane count of
// each (precision, nodes) only): advection with and without mass and boundary
// data, and mass alone
#define GD_K3_ANY(T, N, L)                                                                                 \
  GD_K3_ADV(T, N, L, 0, 0, kSigAny), GD_K3_ADV(T, N, L, 1, 0, kSigAny), GD_K3_ADV(T, N, L, 0, 1, kSigAny), \
      GD_K3_ADV(T, N, L, 1, 1, kSigAny), K3Entry<T>{N, L, 0, 1, 0, kSigAny, &launch<T, N, L, false, true, false, kSigAny>}

#define GD_K3_ADV(T, N, L, MASS, BCG, SIG) \
  K3Entry<T> { N, L, 1, MASS, BCG, SIG, &launch<T, N, L, true, MASS, BCG, SIG> }
#define GD_K3_FOR_LANES(T, N, L)                                                                            \
  GD_K3_ADV(T, N, L, 0, 0, kSigCyclic), GD_K3_ADV(T, N, L, 1, 0, kSigCyclic), GD_K3_ADV(T, N, L, 0, 1, kSigCyclic), \
      GD_K3_ADV(T, N, L, 1, 1, kSigCyclic), GD_K3_ADV(T, N, L, 0, 0, kSigConst),                                  \
      GD_K3_ADV(T, N, L, 1, 0, kSigConst), GD_K3_ADV(T, N, L, 0, 1, kSigConst), GD_K3_ADV(T, N, L, 1, 1, kSigConst), \
      K3Entry<T>{N, L, 0, 1, 0, kSigConst, &launch<T, N, L, false, true, false, kSigConst>}

}  // namespace k3
}  // namespace gd
