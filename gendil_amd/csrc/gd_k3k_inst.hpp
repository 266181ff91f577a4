// Instance lists of the Kronecker-factored 3D kernel, one translation unit per
// (precision, nodes) so they compile in parallel.
#pragma once

#include "gd_k3k.hpp"
#include "gd_k3k_api.hpp"

namespace gd {
namespace k3k {

template <typename T, int N, int LANES, bool ADV, bool MASS, int SIG>
int launch(const KParams<T>& prm_in, const T* x, T* y, hipStream_t s)
{
  constexpr int NT = 256;
  constexpr int EPB = NT / LANES;
  KParams<T> prm = prm_in;
  const long long nb = ((long long)prm.E + EPB - 1) / EPB;
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL((k3k_apply<T, N, LANES, ADV, MASS, SIG, NT>), dim3((unsigned)nb), dim3(NT), 0, s, prm, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

// signatures compiled: the benchmark field (beta_x(y), beta_y(z), beta_z(x))
// and constant fields
constexpr int kSigCyclic = ksig(1, 2, 0);
constexpr int kSigConst = ksig(-1, -1, -1);

#define GD_K3K_ENTRY(T, N, L, ADV, MASS, SIG) \
  KEntry<T> { N, L, ADV, MASS, SIG, &launch<T, N, L, ADV, MASS, SIG> }
#define GD_K3K_FOR_LANES(T, N, L)                                                                          \
  GD_K3K_ENTRY(T, N, L, 1, 0, kSigCyclic), GD_K3K_ENTRY(T, N, L, 1, 1, kSigCyclic),                         \
      GD_K3K_ENTRY(T, N, L, 1, 0, kSigConst), GD_K3K_ENTRY(T, N, L, 1, 1, kSigConst),                       \
      GD_K3K_ENTRY(T, N, L, 0, 1, kSigConst)

}  // namespace k3k
}  // namespace gd
