// Instances of the register-resident 3D kernel (GLL, Q == N) and their launch.
#include "gd_k3d.hpp"
#include "gd_k3d_api.hpp"

namespace gd {
namespace k3 {

template <typename T, int N, int LANES, bool ADV>
int launch(const Params<T>& prm_in, const T* x, T* y, hipStream_t s)
{
  constexpr int NT = 256;
  constexpr int EPB = NT / LANES;
  Params<T> prm = prm_in;
  const long long nb = (prm.E + EPB - 1) / EPB;
  if (nb > 0x7fffffffLL) return fail(GD_ERR_INVALID, "too many elements for one launch");
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL((k3d_apply<T, N, LANES, ADV, NT>), dim3((unsigned)nb), dim3(NT), 0, s, prm, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

#define GD_K3(T, N, L) \
  {N, L, &launch<T, N, L, true>, &launch<T, N, L, false>}

static const K3Entry<double> k64[] = {GD_K3(double, 2, 1), GD_K3(double, 3, 1), GD_K3(double, 4, 1),
                                      GD_K3(double, 4, 2), GD_K3(double, 4, 4)};
static const K3Entry<float> k32[] = {GD_K3(float, 2, 1), GD_K3(float, 3, 1), GD_K3(float, 4, 1),
                                     GD_K3(float, 4, 2), GD_K3(float, 4, 4)};

template <>
const K3Entry<double>* find<double>(int N, int lanes)
{
  for (const auto& k : k64)
    if (k.N == N && k.lanes == lanes) return &k;
  return nullptr;
}

template <>
const K3Entry<float>* find<float>(int N, int lanes)
{
  for (const auto& k : k32)
    if (k.N == N && k.lanes == lanes) return &k;
  return nullptr;
}

}  // namespace k3
}  // namespace gd
