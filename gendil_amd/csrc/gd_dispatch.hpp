// Registry of compiled (dim, nodes, quadrature points) kernel instances.
#pragma once

#include <hip/hip_runtime.h>

#include "gd_internal.hpp"
#include "gd_kernels.hpp"

namespace gd {

template <typename T>
using LaunchFn = int (*)(const Params<T>&, const T*, T*, hipStream_t);

struct KernelSet {
  int D, N, Q;
  LaunchFn<double> fwd64_adv, fwd64_mass, trn64, grad64;
  LaunchFn<float> fwd32_adv, fwd32_mass;
};

struct KernelList {
  const KernelSet* sets;
  int count;
};

KernelList kernels_d1();
KernelList kernels_d2();
KernelList kernels_d3();
KernelList kernels_d4();
KernelList kernels_d5();
KernelList kernels_d6();

const KernelSet* find_kernels(int D, int N, int Q);

template <class C, bool ADV>
int launch(const Params<typename C::T>& prm_in, const typename C::T* x, typename C::T* y, hipStream_t s)
{
  using T = typename C::T;
  auto kern = dev::dg_kernel<C, ADV>;
  const size_t lds = dev::lds_bytes<C>();
  static bool attr_done = false;
  if (lds + sizeof(dev::ElemInfo<C::D>) * C::EPB > 160 * 1024)
    return fail(GD_ERR_UNSUPPORTED, "element tensors of this order/precision exceed the 160 KB LDS");
  if (lds > 64 * 1024 && !attr_done) {
    GD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    attr_done = true;
  }
  Params<T> prm = prm_in;
  const long long nb = (prm.E + C::EPB - 1) / C::EPB;
  if (nb > 0x7fffffffLL) return fail(GD_ERR_INVALID, "too many elements for one launch");
  prm.num_blocks = (int)nb;
  if (nb == 0) return GD_OK;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(C::NT), lds, s, prm, x, y);
  GD_HIP(hipGetLastError());
  return GD_OK;
}

#define GD_KSET(D, N, Q)                                                                     \
  KernelSet                                                                                  \
  {                                                                                          \
    D, N, Q, &launch<dev::Cfg<double, D, N, Q, dev::kForward>, true>,                        \
        &launch<dev::Cfg<double, D, N, Q, dev::kForward>, false>,                            \
        &launch<dev::Cfg<double, D, N, Q, dev::kTranspose>, true>,                           \
        &launch<dev::Cfg<double, D, N, Q, dev::kGradBeta>, true>,                            \
        &launch<dev::Cfg<float, D, N, Q, dev::kForward>, true>,                              \
        &launch<dev::Cfg<float, D, N, Q, dev::kForward>, false>                              \
  }

}  // namespace gd
