// Streaming form of the LDS-staged Kronecker 3D kernel (gd_k3k.hpp): the same
// operator and the same per-element arithmetic, organised so that memory and
// compute overlap.
//
// * One wave per block, persistent: it walks a sequence of tiles of TW = 64/N
//   consecutive elements of one x-row (N lanes per element, lane h producing
//   output z-plane h).
// * A tile's element chunks, upwind neighbour slabs and beta-dependent factor
//   tables move into LDS by the gfx950 LDS DMA (global_load_lds_dwordx4:
//   every lane sends one 16-byte chunk from its own global address to the
//   buffer at base + 16 * lane), with no registers held and no LDS stores
//   executed by the wave.  Each element occupies CPE + 1 chunks (one pad
//   chunk, so the 16 elements a wave reads at one offset spread over the
//   banks); lanes that fall on a pad chunk or past the tile copy an in-bounds
//   chunk that is never read.
// * NB = 1 (two waves per SIMD): one buffer, refilled in two parts: its
//   element region as soon as the tile's element planes have been read (the
//   first phase), its second region (slabs and tables) once the outputs,
//   staged there, are in registers; the stores follow the copies.  NB = 2
//   (one wave per SIMD): while tile t computes from one buffer, tile t+1's
//   copies fill the other.
// * Wave-local: no block barrier at all.  The wave waits for a tile's copies
//   with s_waitcnt vmcnt (vector memory operations complete in issue order),
//   leaving the previous tile's output stores (and, with two buffers, the
//   next tile's copies) in flight; no other vector memory operation is
//   issued in the loop.
// * Outputs are staged through LDS (NB = 2: the tile's element region; NB = 1:
//   the second region) and leave as coalesced 16-byte non-temporal stores.
// Each XCD walks one contiguous chunk of tiles (TileWalk), so the y and z
// neighbours of a tile were read by the same XCD shortly before and are
// still in its L2.
#pragma once

#include <array>
#include <climits>

#include "gd_k3k.hpp"

namespace gd {
namespace k3k {

// 16-byte global -> LDS copy of every lane (global_load_lds_dwordx4): lane i
// sends the 16 bytes at src to LDS byte address lds + 16 * i (lds in M0,
// wave-uniform).  Written as one asm statement that sets M0 itself, so the
// compiler neither keeps a value of its own in M0 across it nor inserts
// waits for it: the kernel waits for these copies with wait_vmcnt, and the
// "memory" clobber keeps every LDS access in program order around them.
// The address operand is read when the instruction issues.  The M0 clobber
// marks the write for the reader; the compiler treats M0 as reserved and
// warns about it, and the kernels here use M0 for nothing else.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma16(const void* src, unsigned lds)
{
  __asm__ volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(src), "s"(lds) : "memory", "m0");
}
#pragma clang diagnostic pop

// Lane K of this lane's quad (lanes 4i .. 4i+3): a DPP quad_perm broadcast,
// no LDS traffic
template <int K>
__device__ __forceinline__ float quad_bcast(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), K * 0x55, 0xf, 0xf, false));
}
template <int K>
__device__ __forceinline__ double quad_bcast(double v)
{
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, K * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), K * 0x55, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// s_waitcnt vmcnt(K), the other counters left alone (gfx9 encoding: vmcnt
// bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8), then a compiler fence so no
// LDS read is scheduled above it
template <int K>
__device__ __forceinline__ void wait_vmcnt()
{
  static_assert(K >= 0 && K < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((K & 15) | ((K >> 4) << 14) | (7 << 4) | (15 << 8));
  __asm__ volatile("" ::: "memory");
}

template <typename T, int N, bool ADV, int SIG>
struct StreamGeom {
  static constexpr int N2 = N * N, N3 = N2 * N;
  static constexpr int TW = 64 / N;             // elements per tile
  static constexpr int VW = 16 / (int)sizeof(T);
  static constexpr int CPE = N3 / VW;           // chunks per element
  static constexpr int UCS = CPE + 1;           // element stride in chunks
  static constexpr int PC = N2 / VW;            // chunks per z-plane of an element
  // Element chunk c (z-plane c / PC) sits at position swz(c) of its LDS
  // slot: the chunk index within the plane is XORed with a function of the
  // plane, so the N lanes of an element, which write N different planes at
  // one chunk index, and the lanes reading the x-neighbour slab (one value of
  // every plane) fall on different banks.  XOR within a plane: an
  // involution, swz(swz(c)) = c.
  static constexpr int PX(int plane)
  {
    if constexpr (N != 4) return 0;
    else return VW == 2 ? 2 * plane : 2 * (plane >> 1);
  }
  static constexpr int swz(int c) { return c ^ PX(c / PC); }
  // the same XOR in units of x-rows (N values): plane p's row s holds row s ^ QX(p)
  static constexpr int QX(int plane) { return PX(plane) * VW / N; }
  static_assert(N != 4 || (PX(1) * VW) % N == 0, "the swizzle moves whole x-rows");
  // the tile's elements, then one halo slot: the upwind x neighbour of the
  // tile's first (beta_x >= 0) or last (beta_x <= 0) element
  static constexpr int U_INS = ((TW + 1) * UCS + 63) / 64;
  static constexpr int SCH = PC;                // chunks per slab / table
  static constexpr int SCS = SCH + 1;
  static constexpr int AX[3] = {ksig_axis(SIG, 0), ksig_axis(SIG, 1), ksig_axis(SIG, 2)};
  static constexpr bool TIN[3] = {ADV && (AX[0] == 0 || AX[0] == 1), ADV && (AX[1] == 0 || AX[1] == 1),
                                  ADV && (AX[2] == 0 || AX[2] == 1)};
  static constexpr bool TZ[3] = {ADV && AX[0] == 2, ADV && AX[1] == 2, ADV && AX[2] == 2};
  static constexpr int NTB = (TIN[0] ? 1 : 0) + (TIN[1] ? 1 : 0) + (TIN[2] ? 1 : 0);
  static constexpr int TBI[3] = {0, TIN[0] ? 1 : 0, (TIN[0] ? 1 : 0) + (TIN[1] ? 1 : 0)};
  static constexpr int slot_comp(int s)
  {
    for (int d = 0, k = 0; d < 3; ++d)
      if (TIN[d] && k++ == s) return d;
    return 0;
  }
  static constexpr int SC[3] = {slot_comp(0), slot_comp(1), slot_comp(2)};
  static constexpr int SA[3] = {AX[SC[0]] < 0 ? 0 : AX[SC[0]], AX[SC[1]] < 0 ? 0 : AX[SC[1]],
                                AX[SC[2]] < 0 ? 0 : AX[SC[2]]};
  static constexpr int NTZ = (TZ[0] ? 1 : 0) + (TZ[1] ? 1 : 0) + (TZ[2] ? 1 : 0);
  static constexpr int ZC = TZ[0] ? 0 : (TZ[1] ? 1 : 2);
  // second region: y slabs, z slabs, the in-plane tables, then the tile's
  // z-factor table.  A table that varies along x takes one padded entry per
  // element (a run of TW), one that varies along y is the same for every
  // element of the tile's row (a single entry).
  static constexpr int NRUN = (ADV ? 2 : 0) + NTB;
  static constexpr int ADVR = ADV ? 2 : 0;      // the slab runs
  static constexpr bool run_shared(int r) { return r >= ADVR && SA[r - ADVR] != 0; }
  static constexpr int run_len(int r) { return run_shared(r) ? SCH : TW * SCS; }
  static constexpr int run0(int r)
  {
    int o = 0;
    for (int k = 0; k < r; ++k) o += run_len(k);
    return o;
  }
  static constexpr int ZT0 = run0(NRUN);        // first chunk of the z-factor table
  static constexpr int S_INS = (ZT0 + NTZ * SCH + 63) / 64;
  static constexpr int INS = U_INS + S_INS;     // copies per tile and lane
  static constexpr int ST = TW * CPE / 64;      // output stores per tile and lane
  static constexpr int BUF_CHUNKS = INS * 64;
  // one-buffer form: the outputs are staged in the second region, which
  // therefore spans at least TW * UCS chunks
  static constexpr int SREG = S_INS * 64 > TW * UCS ? S_INS * 64 : (TW * UCS + 63) / 64 * 64;
  static constexpr int BUF1_CHUNKS = U_INS * 64 + SREG;
  static_assert(N3 % VW == 0 && N2 % VW == 0, "element planes must be whole 16-byte chunks");
  static_assert((TW * CPE) % 64 == 0, "a tile's outputs are whole wave-wide stores");
  static_assert(INS + ST < 64, "vmcnt counts to 63");
  static_assert(NTZ <= 1, "at most one component depends on z");
  static_assert(!(ADV && (AX[0] == 0 || AX[1] == 1 || AX[2] == 2)),
                "components varying along their own axis read face samples from memory: not served here");
};

// Copy source of one lane for one copy instruction, in two parts: `kind`
// selects a per-tile base (0 the tile's elements, 1 the halo element, 2 the
// y slabs, 3 the z slabs, 4 + s table slot s, 7 the z-factor table, 8 none)
// and `off` is the lane's fixed offset from it, in values.
struct StreamSrc {
  int kind, off;
};

template <typename G>
__host__ __device__ constexpr StreamSrc stream_src(int L, bool second)
{
  constexpr int N = G::N3 / G::N2;
  constexpr int N2 = G::N2, N3 = G::N3, TW = G::TW, VW = G::VW, CPE = G::CPE, UCS = G::UCS, SCS = G::SCS;
  constexpr int SCH = G::SCH;
  if (!second) {
    const int ee = L / UCS, c = L - ee * UCS;
    if (c >= CPE || ee > TW) return {8, 0};
    if (ee == TW) return {1, G::swz(c) * VW};
    return {0, ee * N3 + G::swz(c) * VW};
  }
  if (L >= G::ZT0) {
    const int zc = L - G::ZT0;
    return zc < G::NTZ * SCH ? StreamSrc{7, zc * VW} : StreamSrc{8, 0};
  }
  int run = 0;
  while (run + 1 < G::NRUN && L >= G::run0(run + 1)) ++run;
  const int rr = L - G::run0(run);
  if (G::run_shared(run)) return rr < SCH ? StreamSrc{4 + (run - G::ADVR), rr * VW} : StreamSrc{8, 0};
  const int ee = rr / SCS, c = rr - ee * SCS;
  if (c >= SCH) return {8, 0};
  const int m = c * VW;
  if (G::ADVR == 2 && run == 0) return {2, ee * N3 + (m / N) * N2 + m % N};  // y slab (iz, ix)
  if (G::ADVR == 2 && run == 1) return {3, ee * N3 + m};                      // z slab
  return {4 + (run - G::ADVR), ee * N2 + m};                                 // a table varying along x
}

// Bit k set: some lane of copy instruction j takes its source from base k.
template <typename G>
constexpr unsigned stream_kinds(int j)
{
  unsigned m = 0;
  const bool second = j >= G::U_INS;
  for (int lane = 0; lane < 64; ++lane) m |= 1u << stream_src<G>((second ? j - G::U_INS : j) * 64 + lane, second).kind;
  return m;
}

// The copy instructions J.. of one tile: lane sources from the per-tile
// bases b0..b7 by the lane's fixed kind (indices fixed at compile time: a
// run-time index into a local array would place it in scratch memory, and a
// scratch read would wait for every copy in flight).
template <typename G, typename T, int J, int END = G::INS>
__device__ __forceinline__ void stream_issue(const int (&kind)[G::INS], const int (&off)[G::INS], unsigned base,
                                             const T* b0, const T* b1, const T* b2, const T* b3, const T* b4,
                                             const T* b5, const T* b6, const T* b7)
{
  if constexpr (J < END) {
    constexpr unsigned ks = stream_kinds<G>(J);
    // the lane's kind is re-read every tile (the empty asm), so its
    // comparisons are not kept as loop-invariant masks in scalar registers
    int kd = kind[J];
    __asm__ volatile("" : "+v"(kd));
    const T* src = b0;  // pads: any in-bounds chunk
    if constexpr ((ks >> 1) & 1) src = kd == 1 ? b1 : src;
    if constexpr ((ks >> 2) & 1) src = kd == 2 ? b2 : src;
    if constexpr ((ks >> 3) & 1) src = kd == 3 ? b3 : src;
    if constexpr ((ks >> 4) & 1) src = kd == 4 ? b4 : src;
    if constexpr ((ks >> 5) & 1) src = kd == 5 ? b5 : src;
    if constexpr ((ks >> 6) & 1) src = kd == 6 ? b6 : src;
    if constexpr ((ks >> 7) & 1) src = kd == 7 ? b7 : src;
    dma16(src + off[J], base + J * 1024);
    stream_issue<G, T, J + 1, END>(kind, off, base, b0, b1, b2, b3, b4, b5, b6, b7);
  }
}

// The tiles one wave walks.  XCD x (blocks go round-robin over the 8 XCDs)
// owns one contiguous chunk of tiles, in proportion to its waves, and its wx
// waves walk it wx consecutive tiles at a time, so the upwind z-neighbour
// plane of a tile was read by the same XCD one plane earlier and is still in
// its L2 (`rev`: the walk runs downwards, for beta_z <= 0).  Each XCD starts
// at a staggered point of its chunk and wraps round, so the eight concurrent
// streams do not sit a power-of-two distance apart in memory.  Position jj
// of the chunk is tile(jj); a wave visits j0, next(j0), ... (count tiles).
struct TileWalk {
  int c0, len, wx, count, j0;
  bool rev;
  __host__ __device__ TileWalk(int block, int nw, int ntiles, bool down)
  {
    const int xcd = block % 8, r = block / 8;
    const int per = nw / 8, rem = nw % 8;
    wx = per + (xcd < rem ? 1 : 0);
    const int w0 = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    c0 = (int)((long long)ntiles * w0 / nw);
    len = (int)((long long)ntiles * (w0 + wx) / nw) - c0;
    count = r < len ? (len - r + wx - 1) / wx : 0;
    j0 = count > 0 ? (int)((r + (long long)xcd * (len / 8 + 37)) % len) : 0;
    rev = down;
  }
  __host__ __device__ int tile(int jj) const { return rev ? c0 + len - 1 - jj : c0 + jj; }
  // a wave with a next tile has wx < len, so one wrap suffices
  __host__ __device__ int next(int jj) const { return jj + wx >= len ? jj + wx - len : jj + wx; }
};

// Elements are dealt in tiles of TW consecutive elements of one x-row; the
// host requires sizes[0] to be a multiple of TW (so every tile is full) and
// no component varying along its own axis.  ACC (one buffer only): y += the
// operator (the tile's old outputs come into a third LDS region with the
// tile's other copies) instead of y = the operator.
template <typename T, int N, bool ADV, bool MASS, int SIG, int NB, bool ACC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NB == 1 ? 2 : 1)))
k3k_stream(KParams<T> p, const T* __restrict__ x, T* __restrict__ y, int ntiles)
{
  using G = StreamGeom<T, N, ADV, SIG>;
  constexpr int N2 = G::N2, N3 = G::N3, TW = G::TW, VW = G::VW, CPE = G::CPE, UCS = G::UCS, SCS = G::SCS;
  // (the constant arrays are named through G inside the lambdas below: a
  // local copy captured by reference would be materialised at run time)
  constexpr bool MSH = ADV && MASS && G::AX[0] != 2;
  using KT = KTab<N>;
  using C = Chunk<T, VW>;
  static_assert(NB == 1 || NB == 2, "one or two tile buffers");
  static_assert(!ACC || NB == 1, "accumulation runs with one buffer");
  static_assert(G::INS + 2 * G::ST < 64, "vmcnt counts to 63");
  // one buffer: element region, second region, then (ACC) the old outputs
  constexpr int YREG0 = G::BUF1_CHUNKS;
  __shared__ C lds[NB == 1 ? G::BUF1_CHUNKS + (ACC ? G::ST * 64 : 0) : 2 * G::BUF_CHUNKS];
  const unsigned lds_base = (unsigned)(size_t)(__attribute__((address_space(3))) void*)lds;

  const int lane = (int)threadIdx.x;
  const int h = lane % N;
  const int el = lane / N;
  const bool negx = (p.negmask & 1) != 0, negy = (p.negmask & 2) != 0, negz = (p.negmask & 4) != 0;
  const int ny = negy ? 0 : N - 1, nz = negz ? 0 : N - 1;  // node of the upwind neighbour slab

  const TileWalk walk((int)blockIdx.x, (int)gridDim.x, ntiles, negz);
  int j = walk.j0;

  // this lane's fixed copy offsets (values) and bases, one per instruction
  int kind[G::INS], off[G::INS];
#pragma unroll
  for (int j = 0; j < G::INS; ++j) {
    const bool second = j >= G::U_INS;
    const StreamSrc s = stream_src<G>((second ? j - G::U_INS : j) * 64 + lane, second);
    kind[j] = s.kind;
    off[j] = s.off + (s.kind == 2 ? ny * N : (s.kind == 3 ? nz * N2 : 0));
  }

  // this lane's rows of the constant z factors (kept in registers: selected
  // per lane inside the loop they would become loads from a constant table)
  T rowM[N], rowD[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    rowM[j] = lrow<N, N, 1, T>(KT::M, h, 0, j);
    rowD[j] = lrow<N, N, 1, T>(KT::D, h, 0, j);
  }
  pin<N>(rowM);
  pin<N>(rowD);

  // position of a tile: first element, its x index, the tile's y and z
  struct TilePos {
    int e0, sx0, sy, sz;
  };
  auto tile_pos = [&](int t) __attribute__((always_inline)) {
    TilePos tp;
    tp.e0 = t * TW;
    const unsigned q = (unsigned)tp.e0 / (unsigned)p.M0;
    tp.sx0 = tp.e0 - (int)q * p.M0;
    tp.sz = (int)(q / (unsigned)p.M1);
    tp.sy = (int)q - tp.sz * p.M1;
    return tp;
  };
  // upwind neighbour of a face of element position s along axis d (Ma cells,
  // stride st), or -1 where there is none
  auto upwind_off = [&](int sa, int Ma, int st, bool periodic, bool neg) __attribute__((always_inline)) {
    if (Ma == 1) return periodic ? 0 : INT_MIN;
    if (!neg) return sa > 0 ? -st : (periodic ? (Ma - 1) * st : INT_MIN);
    return sa < Ma - 1 ? st : (periodic ? -(Ma - 1) * st : INT_MIN);
  };

  // ---- the copies of one tile into buffer b ----
  // part 0: every copy of the tile; 1: the element region (elements and
  // halo, U_INS copies); 2: the second region (slabs and tables)
  auto issue = [&](int t, int b, int part) __attribute__((always_inline)) {
    const unsigned base = lds_base + (unsigned)(b * G::BUF_CHUNKS * (int)sizeof(C));
    const TilePos tp = tile_pos(t);
    // per-tile bases (uniform), selected per lane by its fixed kind
    const T* b0 = x + (long long)tp.e0 * N3;
    const T* b1;
    {
      const int xo = upwind_off(negx ? tp.sx0 + TW - 1 : tp.sx0, p.M0, 1, p.periodic & 1, negx);
      const int halo = (negx ? tp.e0 + TW - 1 : tp.e0) + (xo == INT_MIN ? 0 : xo);
      b1 = x + (long long)halo * N3;
    }
    const T* b2 = b0;
    const T* b3 = b0;
    if constexpr (ADV) {
      const int yo = upwind_off(tp.sy, p.M1, p.M0, (p.periodic >> 1) & 1, negy);
      const int zo = upwind_off(tp.sz, p.M2, p.M0 * p.M1, (p.periodic >> 2) & 1, negz);
      b2 = x + (long long)(tp.e0 + (yo == INT_MIN ? 0 : yo)) * N3;
      b3 = x + (long long)(tp.e0 + (zo == INT_MIN ? 0 : zo)) * N3;
    }
    const T* bt[3] = {x, x, x};
#pragma unroll
    for (int sl = 0; sl < G::NTB; ++sl) {
      const int ax = G::SA[sl];
      const int pos = ax == 0 ? tp.sx0 : (ax == 1 ? tp.sy : tp.sz);
      bt[sl] = p.tab[G::SC[sl]] + (long long)pos * N2;
    }
    const T* b7 = G::NTZ > 0 ? p.tab[G::ZC] + (long long)tp.sz * N2 : b0;
    if (part == 1)
      stream_issue<G, T, 0, G::U_INS>(kind, off, base, b0, b1, b2, b3, bt[0], bt[1], bt[2], b7);
    else if (part == 2)
      stream_issue<G, T, G::U_INS, G::INS>(kind, off, base, b0, b1, b2, b3, bt[0], bt[1], bt[2], b7);
    else
      stream_issue<G, T, 0>(kind, off, base, b0, b1, b2, b3, bt[0], bt[1], bt[2], b7);
    if constexpr (ACC) {
      // the tile's old outputs, chunk i * 64 + lane to the same LDS chunk
      if (part != 1) {
        const T* yt = y + (long long)tp.e0 * N3;
#pragma unroll
        for (int i = 0; i < G::ST; ++i)
          dma16(yt + (i * 64 + lane) * VW, lds_base + (unsigned)((YREG0 + i * 64) * (int)sizeof(C)));
      }
    }
  };

  // ---- one tile: compute from buffer b while the next tile moves into b^1 ----
  auto step = [&](int t, int tn, int b, bool first) __attribute__((always_inline)) {
    C* cur = lds + b * G::BUF_CHUNKS;
    const TilePos tp = tile_pos(t);
    const int e0 = tp.e0;
    // face coefficients: y and z faces are the same for the whole tile
    FaceCoefs<T> fc;
    fc.negmask = p.negmask;
    fc.nsl[0] = negx ? 0 : N - 1;
    int sub[3] = {tp.sx0 + el, tp.sy, tp.sz};
    if constexpr (ADV) {
      const int Ms[3] = {p.M0, p.M1, p.M2};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int Ma = Ms[d], sa = sub[d];
        const bool per = (p.periodic >> d) & 1;
        const bool neg = (p.negmask >> d) & 1;
        const bool has_lo = Ma == 1 ? per : (sa > 0 || per);
        const bool has_hi = Ma == 1 ? per : (sa < Ma - 1 || per);
        if (!neg) {
          fc.cown[d] = (has_hi || p.bc_zero) ? p.Kt[d] : T(0);
          fc.cnb[d] = has_lo ? -p.Kt[d] : T(0);
        } else {
          fc.cown[d] = (has_lo || p.bc_zero) ? -p.Kt[d] : T(0);
          fc.cnb[d] = has_hi ? p.Kt[d] : T(0);
        }
      }
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) fc.cown[d] = fc.cnb[d] = T(0);
    }

    // In flight, oldest first: this tile's copies, the previous tile's output
    // stores (none for a wave's first tile), the next tile's copies.  Wait for
    // this tile's copies only: the stores and the next copies stay in flight.
    if constexpr (NB == 2) {
      if (tn >= 0) {
        issue(tn, b ^ 1, 0);
        if (first) wait_vmcnt<G::INS>();
        else wait_vmcnt<G::INS + G::ST>();
      } else {
        if (first) wait_vmcnt<0>();
        else wait_vmcnt<G::ST>();
      }
    } else {
      // one buffer: in flight, oldest first, this tile's element copies
      // (issued halfway through the previous tile), its slab and table
      // copies, and the previous tile's stores (issued after them)
      if (first) wait_vmcnt<0>();
      else wait_vmcnt<G::ST>();
    }

    const C* ue = cur + el * UCS;
    const T* sbase = reinterpret_cast<const T*>(cur + G::U_INS * 64);
    const T* sye = sbase + (0 * TW + el) * SCS * VW;
    const T* sze = sbase + (1 * TW + el) * SCS * VW;
    auto ftab = [&](int d) -> const T* {
      const int r = G::ADVR + G::TBI[d];
      return sbase + G::run0(r) * VW + (G::run_shared(r) ? 0 : el * SCS * VW);
    };
    // this lane's row of the beta-dependent z factor
    T zt[N];
#pragma unroll
    for (int j = 0; j < N; ++j) zt[j] = T(0);
    if constexpr (G::NTZ > 0) {
      const T* zr = sbase + G::ZT0 * VW + h * N;
#pragma unroll
      for (int j = 0; j < N; ++j) zt[j] = zr[j];
    }
    // x-neighbour slab row h: the neighbour's slot in the tile, or the halo
    T vx[N];
#pragma unroll
    for (int j = 0; j < N; ++j) vx[j] = T(0);
    if constexpr (ADV) {
      const int xs = negx ? el + 1 : el - 1;
      const int slot = (xs >= 0 && xs < TW) ? xs : TW;
      // the neighbour's plane h, row s holds y-row s ^ QX(h) (the swizzle
      // permutes whole rows of x values): vx[s] = value (x = nsl, y = s ^ QX(h), z = h)
      const T* un = reinterpret_cast<const T*>(cur + slot * UCS + h * G::PC) + fc.nsl[0];
#pragma unroll
      for (int j = 0; j < N; ++j) vx[j] = un[j * N];
    }

    auto coefX = [&](int j) -> T {
      if constexpr (G::TZ[0]) return zt[j];
      else return rowM[j];
    };
    auto coefY = [&](int j) -> T {
      if constexpr (G::TZ[1]) return zt[j];
      else return rowM[j];
    };
    auto coefZ = [&](int j) -> T {
      T dz;
      if constexpr (G::TZ[2]) dz = zt[j];
      else dz = rowD[j];
      return -p.Kc[2] * dz + ((h == N - 1 && j == N - 1) ? fc.cPown(2) : T(0)) +
             ((h == 0 && j == 0) ? fc.cMown(2) : T(0));
    };
    auto coefM = [&](int j) -> T { return rowM[j]; };

    T out[N2];
#pragma unroll
    for (int m = 0; m < N2; ++m) out[m] = T(0);
    {
      // all terms' z factors in one stream over the element's planes
      T accX[N2], accY[N2], accZ[N2], accM[(MASS && !MSH) ? N2 : 1];
#pragma unroll
      for (int m = 0; m < N2; ++m) {
        accX[m] = accY[m] = accZ[m] = T(0);
        if constexpr (MASS && !MSH) accM[m] = T(0);
      }
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const T cx = (ADV || MSH) ? coefX(j) : T(0);
        const T cy = ADV ? coefY(j) : T(0);
        const T cz = ADV ? coefZ(j) : T(0);
        const T cm = (MASS && !MSH) ? coefM(j) : T(0);
        if (j > 0) {  // one plane in flight at a time
          if constexpr (ADV || MSH) pin<N2>(accX);
          if constexpr (ADV) {
            pin<N2>(accY);
            pin<N2>(accZ);
          }
          if constexpr (MASS && !MSH) pin<N2>(accM);
        }
#pragma unroll
        for (int v = 0; v < N2 / VW; ++v) {
          const C c = ue[G::swz(j * G::PC + v)];
#pragma unroll
          for (int l = 0; l < VW; ++l) {
            const int m = v * VW + l;
            if constexpr (ADV || MSH) accX[m] += cx * c.v[l];
            if constexpr (ADV) {
              accY[m] += cy * c.v[l];
              accZ[m] += cz * c.v[l];
            }
            if constexpr (MASS && !MSH) accM[m] += cm * c.v[l];
          }
        }
      }
      if constexpr (NB == 1) {
        // the element region (planes and x halo) has been read for this
        // tile: once every read has landed (lgkmcnt(0)), refill it with the
        // next tile's elements while the rest of the tile computes from
        // registers and the second region
        __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4) | (0 << 8));
        __builtin_amdgcn_wave_barrier();
        if (tn >= 0) issue(tn, 0, 1);
      }
      if constexpr (ADV) {
        // x-neighbour slab through this lane's z-factor row: the element's
        // lanes exchange their rows
        T sx[N];
#pragma unroll
        for (int j = 0; j < N; ++j) sx[j] = T(0);
        if constexpr (N == 4) {
          // the element's four lanes are one quad
          const T c0 = coefX(0), c1 = coefX(1), c2 = coefX(2), c3 = coefX(3);
#pragma unroll
          for (int j = 0; j < N; ++j) sx[j] += c0 * quad_bcast<0>(vx[j ^ G::QX(0)]);
#pragma unroll
          for (int j = 0; j < N; ++j) sx[j] += c1 * quad_bcast<1>(vx[j ^ G::QX(1)]);
#pragma unroll
          for (int j = 0; j < N; ++j) sx[j] += c2 * quad_bcast<2>(vx[j ^ G::QX(2)]);
#pragma unroll
          for (int j = 0; j < N; ++j) sx[j] += c3 * quad_bcast<3>(vx[j ^ G::QX(3)]);
        } else {
#pragma unroll
          for (int k = 0; k < N; ++k) {
            const T ck = coefX(k);
#pragma unroll
            for (int j = 0; j < N; ++j) sx[j] += ck * __shfl(vx[j], (lane & ~(N - 1)) + k, 64);
          }
        }
        term_x<T, N, SIG>(p, fc, ftab(0), accX, sx, out);
      }
      if constexpr (MSH) term_plane<T, N, SIG, -1>(p, nullptr, accX, out);
      if constexpr (MASS && !MSH) term_plane<T, N, SIG, -1>(p, nullptr, accM, out);
      pin<N2>(out);
      if constexpr (ADV) {
        T sy[N];
#pragma unroll
        for (int j = 0; j < N; ++j) sy[j] = T(0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const T ck = coefY(k);
#pragma unroll
          for (int j = 0; j < N; ++j) sy[j] += ck * sye[k * N + j];
        }
        term_y<T, N, SIG>(p, fc, ftab(1), accY, sy, out);
        pin<N2>(out);
        // upwind neighbour across a z face: its slab enters output plane 0
        // (beta_z >= 0) or N-1 (beta_z <= 0)
        const T czn = (h == 0 ? fc.cMnb(2) : T(0)) + (h == N - 1 ? fc.cPnb(2) : T(0));
#pragma unroll
        for (int m = 0; m < N2; ++m) accZ[m] += czn * sze[m];
        term_plane<T, N, SIG, 2>(p, ftab(2), accZ, out);
      }
    }

    // outputs staged through LDS: with two buffers through the tile's
    // element region, with one through the second region (the element
    // region already holds the next tile's copies); every lane has finished
    // reading it (the values were consumed), so the wave may overwrite it
    C* const stg = NB == 1 ? cur + G::U_INS * 64 : cur;
    static_assert(NB == 2 || TW * UCS <= G::SREG, "the outputs fit the second region");
    __builtin_amdgcn_wave_barrier();
    {
      // PX is even: chunks 2k and 2k+1 stay neighbours, at (2k) ^ px
      static_assert(G::PC % 2 == 0, "pairs of chunks per plane");
      C* ow = stg + el * UCS + h * G::PC;
      const int px = G::PX(h);
#pragma unroll
      for (int k = 0; k < G::PC / 2; ++k) {
        C* w2 = ow + ((2 * k) ^ px);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          C c;
#pragma unroll
          for (int l = 0; l < VW; ++l) c.v[l] = out[(2 * k + i) * VW + l];
          w2[i] = c;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    C* dst = reinterpret_cast<C*>(y) + (long long)e0 * CPE;
    if constexpr (NB == 2) {
#pragma unroll
      for (int i = 0; i < G::ST; ++i) {
        const int cidx = i * 64 + lane;
        const int ee = cidx / CPE;
        store_nt(&dst[cidx], cur[ee * UCS + G::swz(cidx - ee * CPE)]);
      }
    } else {
      // one buffer: take the staged outputs into registers, let every LDS
      // read finish (lgkmcnt(0)), then refill the second region with the
      // next tile's slabs and tables before the stores go out
      C st[G::ST];
#pragma unroll
      for (int i = 0; i < G::ST; ++i) {
        const int cidx = i * 64 + lane;
        const int ee = cidx / CPE;
        st[i] = stg[ee * UCS + G::swz(cidx - ee * CPE)];
        if constexpr (ACC) {
          // plus the old outputs, read before the next tile's copies land
          const C yo = lds[YREG0 + cidx];
#pragma unroll
          for (int l = 0; l < VW; ++l) st[i].v[l] += yo.v[l];
        }
      }
      __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4) | (0 << 8));
      __builtin_amdgcn_wave_barrier();
      if (tn >= 0) issue(tn, 0, 2);
#pragma unroll
      for (int i = 0; i < G::ST; ++i) store_nt(&dst[i * 64 + lane], st[i]);
    }
    __builtin_amdgcn_wave_barrier();
  };

  if (walk.count > 0) issue(walk.tile(j), 0, 0);
  for (int k = 0, b = 0; k < walk.count; ++k, b = NB == 2 ? b ^ 1 : 0) {
    const int jn = walk.next(j);
    step(walk.tile(j), k + 1 < walk.count ? walk.tile(jn) : -1, b, k == 0);
    j = jn;
  }
}

}  // namespace k3k
}  // namespace gd
