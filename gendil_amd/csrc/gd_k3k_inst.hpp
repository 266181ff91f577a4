// Instance lists of the Kronecker-factored 3D kernel, one translation unit per
// (precision, nodes) so they compile in parallel.
#pragma once

#include "gd_k3k.hpp"
#include "gd_k3k_api.hpp"
#include "gd_k3s.hpp"

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

// LDS-staged kernel: lanes code 32 + (1: outputs staged through LDS) + (2:
// blocks of 128 threads instead of 256) + (4: one term at a time)
template <typename T, int N, int CODE, bool ADV, bool MASS, int SIG>
int launch_staged(const KParams<T>& prm_in, const T* x, T* y, hipStream_t s)
{
  constexpr int NT = (CODE & 2) ? 128 : 256;
  constexpr int EPB = NT / N;
  KParams<T> prm = prm_in;
  const long long nb = ((long long)prm.E + EPB - 1) / EPB;
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL((k3k_staged<T, N, ADV, MASS, SIG, EPB, (CODE & 1) != 0, (CODE & 4) != 0>), dim3((unsigned)nb),
                     dim3(NT), 0, s, prm, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

// Streaming kernel (lanes code 39 + NB: 40 one tile buffer per wave, two
// waves per SIMD; 41 two buffers, one wave per SIMD; 42 one buffer, y +=
// the operator): one wave per block, persistent over tiles of one x-row, as
// many waves as fit on the device at once.
template <typename T, int N, bool ADV, bool MASS, int SIG, int NB, bool ACC = false>
int launch_stream(const KParams<T>& prm_in, const T* x, T* y, hipStream_t s)
{
  using G = StreamGeom<T, N, ADV, SIG>;
  if (prm_in.M0 % G::TW != 0) return GD_ERR_UNSUPPORTED;
  const int ntiles = prm_in.E / G::TW;
  if (ntiles == 0) return GD_OK;
  static int waves = 0;  // resident waves on the device (per instance)
  if (waves == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    GD_HIP(hipGetDevice(&dev));
    GD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    GD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k3k_stream<T, N, ADV, MASS, SIG, NB, ACC>), 64, 0));
    waves = cus * (per_cu > 0 ? per_cu : 1);
  }
  KParams<T> prm = prm_in;
  const int grid = ntiles < waves ? ntiles : waves;
  prm.num_blocks = grid;
  hipLaunchKernelGGL((k3k_stream<T, N, ADV, MASS, SIG, NB, ACC>), dim3((unsigned)grid), dim3(64), 0, s, prm, x, y,
                     ntiles);
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

#define GD_K3K_STAGED_ENTRY(T, N, CODE, ADV, MASS, SIG) \
  KEntry<T> { N, 32 + CODE, ADV, MASS, SIG, &launch_staged<T, N, CODE, ADV, MASS, SIG> }
#define GD_K3K_STAGED(T, N, CODE)                                                                             \
  GD_K3K_STAGED_ENTRY(T, N, CODE, 1, 0, kSigCyclic), GD_K3K_STAGED_ENTRY(T, N, CODE, 1, 1, kSigCyclic),        \
      GD_K3K_STAGED_ENTRY(T, N, CODE, 1, 0, kSigConst), GD_K3K_STAGED_ENTRY(T, N, CODE, 1, 1, kSigConst),      \
      GD_K3K_STAGED_ENTRY(T, N, CODE, 0, 1, kSigConst)
#define GD_K3K_STREAM_ENTRY(T, N, NB, ADV, MASS, SIG) \
  KEntry<T> { N, 39 + NB, ADV, MASS, SIG, &launch_stream<T, N, ADV, MASS, SIG, NB> }
#define GD_K3K_STREAM(T, N, NB)                                                                                \
  GD_K3K_STREAM_ENTRY(T, N, NB, 1, 0, kSigCyclic), GD_K3K_STREAM_ENTRY(T, N, NB, 1, 1, kSigCyclic),             \
      GD_K3K_STREAM_ENTRY(T, N, NB, 1, 0, kSigConst), GD_K3K_STREAM_ENTRY(T, N, NB, 1, 1, kSigConst),           \
      GD_K3K_STREAM_ENTRY(T, N, NB, 0, 1, kSigConst)
// y += the operator, one buffer (the face terms alone: the face operator)
#define GD_K3K_STREAM_ACC(T, N)                                                                  \
  KEntry<T>{N, 42, 1, 0, kSigCyclic, &launch_stream<T, N, true, false, kSigCyclic, 1, true>},    \
      KEntry<T> { N, 42, 1, 0, kSigConst, &launch_stream<T, N, true, false, kSigConst, 1, true> }
#define GD_K3K_STAGED_ALL(T, N)                                                                        \
  GD_K3K_STAGED(T, N, 0), GD_K3K_STAGED(T, N, 1), GD_K3K_STAGED(T, N, 2), GD_K3K_STAGED(T, N, 3),          \
      GD_K3K_STAGED(T, N, 4), GD_K3K_STAGED(T, N, 5), GD_K3K_STAGED(T, N, 6), GD_K3K_STAGED(T, N, 7)

}  // namespace k3k
}  // namespace gd
