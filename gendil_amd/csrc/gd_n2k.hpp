// Two-node (p = 1) Kronecker apply for 4D-6D Cartesian meshes with
// constant coefficients: the mass term sigma M and the upwind advection
// operator of a constant field beta, any basis and rule of the plan (GenDiL's
// 6D benchmark operator, benchmark-advection-6d.cpp: Legendre p = 1, three
// Gauss points, beta = 1).
//
// With constant beta and sigma the element operator of the quadrature kernel
// (advectionoperator.hpp:137-227 and the face terms of :596-676,
// massoperator.hpp:39-89) is a sum of Kronecker products of 2 x 2 matrices
// built from the rule's 1D tables (the tensor rule factorises every
// integral):
//   Km M x ... x M                       (mass, Km = mass_coef detJ sigma)
//   M x .. x C_d x .. x M                (cell term of component d and its own
//                                         outflow face, Kt_d = adv_coef beta_d
//                                         detJ / h_d: C_d = -Kt_d D + the
//                                         face's outer product)
//   M x .. x F_d x .. x M  on the upwind neighbour along d
// with M = B^T W B, D = Gb^T W B.  All terms share the M factors: factored
// through M^{-1} along the term's own axis,
//   y = (M x ... x M) [ Km u + sum_d (M^{-1} C_d)_d u + sum_d (M^{-1} F_d)_d u_nb(d) ],
// one 2 x 2 contraction per axis and term, then D contractions with M.
//
// Layout: one lane per node, 2^D lanes per element (node bits i_0 .. i_{D-1}
// are the lane bits), 64 / 2^D elements per wave; a contraction along axis d
// pairs lane l with lane l ^ (1 << d).  Every load and store moves an
// element's contiguous 2^D values.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace gd {
namespace n2k {

template <typename T>
struct N2Params {
  int E;
  int M[6];       // elements per axis
  int stride[6];  // element-index stride of each axis
  int periodic;   // bit d: axis d periodic
  int negmask;    // bit d: beta_d < 0 (inflow through the + face)
  int bc_zero;    // boundary faces carry inflow data g = 0 (else they are skipped)
  // The lane coefficients, laid out for the lane that reads them: per axis d
  // (12 values from d * 12) the row (b, b), (b, 1 - b) of M^{-1} C_d
  // without (own = 0) and with (own = 1) the own outflow face at
  // 4 own + 2 b, then the row of M^{-1} F_d at 8 + 2 b; from 72 the rows of
  // M at 72 + 2 b, and Km at 76.
  static constexpr int kTab = 77;
  T tab[kTab];
};

// lane ^ MASK of a 64-lane wave (DPP quad permutes for 1 and 2)
template <int MASK>
__device__ __forceinline__ float xshfl(float v)
{
  if constexpr (MASK == 1)
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xf, 0xf, false));
  else if constexpr (MASK == 2)
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xf, 0xf, false));
  else
    return __shfl_xor(v, MASK, 64);
}
template <int MASK>
__device__ __forceinline__ double xshfl(double v)
{
  if constexpr (MASK == 1 || MASK == 2) {
    constexpr int ctl = MASK == 1 ? 0xB1 : 0x4E;
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, ctl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctl, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
  } else {
    return __shfl_xor(v, MASK, 64);
  }
}

template <int I, int NN, typename F>
__device__ __forceinline__ void sfor(F&& f)
{
  if constexpr (I < NN) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, NN>(f);
  }
}

// One element (or the lanes past the last element, which load in bounds and
// store nothing): e is this lane's element, node its node.
template <typename T, int D>
__device__ __forceinline__ void apply_one(const N2Params<T>& p, const T* tab, const T* __restrict__ x,
                                          T* __restrict__ y, int e, int node)
{
  constexpr int ND = 1 << D;
  const bool valid = e < p.E;
  const int ec = valid ? e : p.E - 1;  // lanes past the last element load in bounds, store nothing
  const T u = x[(long long)ec * ND + node];
  T w = tab[76] * u;
  int rem = ec;
  sfor<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    const int Md = p.M[d];
    const int sd = rem % Md;
    rem /= Md;
    const int b = (node >> d) & 1;
    const bool per = (p.periodic >> d) & 1, neg = (p.negmask >> d) & 1;
    const bool has_lo = Md == 1 ? per : (sd > 0 || per);
    const bool has_hi = Md == 1 ? per : (sd < Md - 1 || per);
    const bool own = neg ? (has_lo || p.bc_zero) : (has_hi || p.bc_zero);
    const bool inflow = neg ? has_hi : has_lo;
    // cell term and own outflow face
    const T* cr = tab + d * 12 + (own ? 4 : 0) + 2 * b;
    const T cs = cr[0], co = cr[1];
    const T up = xshfl<(1 << d)>(u);
    w += cs * u + co * up;
    // inflow face: the upwind neighbour's pencil along d (wrapping on a
    // periodic axis); without one, the element's own values with zero weights
    const int st = p.stride[d];
    int nb = ec;
    if (inflow) {
      if (!neg) nb = sd > 0 ? ec - st : ec + (Md - 1) * st;
      else nb = sd < Md - 1 ? ec + st : ec - (Md - 1) * st;
    }
    const T nv = x[(long long)nb * ND + node];
    const T np = xshfl<(1 << d)>(nv);
    const T* fr = tab + d * 12 + 8 + 2 * b;
    const T fs = inflow ? fr[0] : T(0), fo = inflow ? fr[1] : T(0);
    w += fs * nv + fo * np;
  });
  sfor<0, D>([&](auto d_) {
    constexpr int d = decltype(d_)::value;
    const T* mr = tab + 72 + 2 * ((node >> d) & 1);
    const T ms = mr[0], mo = mr[1];
    const T wp = xshfl<(1 << d)>(w);
    w = ms * w + mo * wp;
  });
  if (valid) y[(long long)e * ND + node] = w;
}

template <typename T, int D>
__global__ void __launch_bounds__(256) n2k_kernel(N2Params<T> p, const T* __restrict__ x, T* __restrict__ y)
{
  constexpr int ND = 1 << D, EPW = 64 / ND;
  // the coefficient table in LDS: each lane reads its own rows there (held in
  // scalar registers the 77 values would not fit)
  __shared__ T tab[N2Params<T>::kTab];
  for (int i = (int)threadIdx.x; i < N2Params<T>::kTab; i += (int)blockDim.x) tab[i] = p.tab[i];
  __syncthreads();
  const int lane = (int)threadIdx.x & 63, node = lane & (ND - 1);
  // Blocks stay resident and stride over the mesh (the table is loaded once
  // per block, not once per four elements); within a round, blocks in launch
  // order.  (One contiguous run of elements per XCD halves the fetched bytes
  // on a 10^6 mesh but measured 15 % slower there, and equal at 5^6.)
  // the wave's first element: wave-uniform, so every lane of a wave runs the
  // same iterations (the lane exchanges need the whole wave)
  const int wave_e = ((int)blockIdx.x * ((int)blockDim.x / 64) + (int)threadIdx.x / 64) * EPW;
  const int estride = (int)gridDim.x * ((int)blockDim.x / 64) * EPW;
  for (int w0 = wave_e; w0 < p.E; w0 += estride) {
    int e = w0 + lane / ND;
    if constexpr (EPW == 1) e = __builtin_amdgcn_readfirstlane(e);  // one element per wave: scalar index work
    apply_one<T, D>(p, tab, x, y, e, node);
  }
}

template <typename T, int D>
int launch(const N2Params<T>& prm, const T* x, T* y, hipStream_t s)
{
  constexpr int EPB = 4 * (64 >> D);  // four waves per block
  // at most the blocks the device holds at once (queried once per process:
  // CUs x resident blocks of this instance); the kernel strides over the
  // rest, so no block starts late with a full share of the mesh
  static int resident = 0;
  if (resident == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, n2k_kernel<T, D>, 256, 0) != hipSuccess)
      return -1;
    resident = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
  }
  const int need = (prm.E + EPB - 1) / EPB;
  const int nb = need < resident ? need : resident;
  if (nb == 0) return 0;
  hipLaunchKernelGGL((n2k_kernel<T, D>), dim3((unsigned)nb), dim3(256), 0, s, prm, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace n2k
}  // namespace gd
