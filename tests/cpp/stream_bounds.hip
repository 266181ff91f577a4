// Host-side check of the streaming kernel's copy addresses (gd_k3s.hpp): for
// every tile of a set of meshes, every periodicity and every upwind direction,
// the source of every lane of every global->LDS copy instruction lies inside
// its buffer (the input vector or a factor table) and is 16-byte aligned, and
// each instruction's lanes take their bases from the kinds it was compiled
// for.  The per-tile bases restate the kernel's `issue` lambda; the per-lane
// sources are the kernel's own stream_src.  Also: the element region's chunk
// layout (StreamGeom::swz) against those sources, and the kernel's TileWalk
// visiting every tile once.  Runs on the CPU only.
#include "gd_k3s.hpp"
#include <cstdio>
#include <cstdlib>
#include <climits>

using namespace gd::k3k;

static int upwind_off(int sa, int Ma, int st, bool periodic, bool neg)
{
  if (Ma == 1) return periodic ? 0 : INT_MIN;
  if (!neg) return sa > 0 ? -st : (periodic ? (Ma - 1) * st : INT_MIN);
  return sa < Ma - 1 ? st : (periodic ? -(Ma - 1) * st : INT_MIN);
}

template <typename T, int N, bool ADV, int SIG>
long check(int M0, int M1, int M2, int periodic, int negmask)
{
  using G = StreamGeom<T, N, ADV, SIG>;
  constexpr int N2 = G::N2, N3 = G::N3, TW = G::TW, VW = G::VW;
  if (M0 % TW) return 0;
  const long E = (long)M0 * M1 * M2;
  const int ntiles = (int)(E / TW);
  const bool negx = negmask & 1, negy = negmask & 2, negz = negmask & 4;
  const int ny = negy ? 0 : N - 1, nz = negz ? 0 : N - 1;
  long nbad = 0, nchk = 0;
  for (int j = 0; j < G::INS; ++j) {
    const unsigned ks = stream_kinds<G>(j);
    for (int lane = 0; lane < 64; ++lane) {
      const bool second = j >= G::U_INS;
      const StreamSrc s = stream_src<G>((second ? j - G::U_INS : j) * 64 + lane, second);
      if (!((ks >> s.kind) & 1)) { printf("kind mask mismatch j=%d lane=%d kind=%d\n", j, lane, s.kind); ++nbad; }
    }
  }
  const int Ms[3] = {M0, M1, M2};
  for (int t = 0; t < ntiles; ++t) {
    const int e0 = t * TW;
    const unsigned q = (unsigned)e0 / (unsigned)M0;
    const int sx0 = e0 - (int)q * M0, sz = (int)(q / (unsigned)M1), sy = (int)q - sz * M1;
    long base[9];
    long len[9];
    for (int k = 0; k < 9; ++k) { base[k] = (long)e0 * N3; len[k] = E * N3; }
    {
      const int xo = upwind_off(negx ? sx0 + TW - 1 : sx0, M0, 1, periodic & 1, negx);
      base[1] = (long)((negx ? e0 + TW - 1 : e0) + (xo == INT_MIN ? 0 : xo)) * N3;
    }
    if (ADV) {
      const int yo = upwind_off(sy, M1, M0, (periodic >> 1) & 1, negy);
      const int zo = upwind_off(sz, M2, M0 * M1, (periodic >> 2) & 1, negz);
      base[2] = (long)(e0 + (yo == INT_MIN ? 0 : yo)) * N3;
      base[3] = (long)(e0 + (zo == INT_MIN ? 0 : zo)) * N3;
    }
    for (int sl = 0; sl < G::NTB; ++sl) {
      const int ax = G::SA[sl];
      const int pos = ax == 0 ? sx0 : (ax == 1 ? sy : sz);
      base[4 + sl] = (long)pos * N2;
      len[4 + sl] = (long)Ms[ax] * N2;
    }
    if (G::NTZ > 0) { base[7] = (long)sz * N2; len[7] = (long)M2 * N2; }
    for (int j = 0; j < G::INS; ++j)
      for (int lane = 0; lane < 64; ++lane) {
        const bool second = j >= G::U_INS;
        const StreamSrc s = stream_src<G>((second ? j - G::U_INS : j) * 64 + lane, second);
        const int off = s.off + (s.kind == 2 ? ny * N : (s.kind == 3 ? nz * N2 : 0));
        const int k = s.kind == 8 ? 0 : s.kind;
        const long a = base[k] + off;
        ++nchk;
        if (a < 0 || a + VW > len[k] || (a * (long)sizeof(T)) % 16) {
          if (nbad < 10) printf("OOB t=%d j=%d lane=%d kind=%d a=%ld len=%ld\n", t, j, lane, s.kind, a, len[k]);
          ++nbad;
        }
      }
  }
  if (nbad)
    printf("T=%zu N=%d ADV=%d SIG=%d M=(%d,%d,%d) per=%d neg=%d: %ld of %ld sources out of bounds\n", sizeof(T), N,
           (int)ADV, SIG, M0, M1, M2, periodic, negmask, nbad, nchk);
  return nbad;
}

// The element region of a tile buffer: LDS position ee * UCS + swz(c) of an
// element slot receives chunk c of that element (and the halo slot chunk c of
// the halo element), every chunk exactly once, as the kernel's reads assume.
template <typename T, int N, bool ADV, int SIG>
long check_layout()
{
  using G = StreamGeom<T, N, ADV, SIG>;
  long nbad = 0;
  for (int c = 0; c < G::CPE; ++c)
    if (G::swz(G::swz(c)) != c || G::swz(c) / G::PC != c / G::PC) ++nbad;
  for (int ee = 0; ee <= G::TW; ++ee) {
    int seen[G::CPE] = {};
    for (int p = 0; p < G::UCS; ++p) {
      const StreamSrc s = stream_src<G>(ee * G::UCS + p, false);
      if (p >= G::CPE) {
        if (s.kind != 8) ++nbad;
        continue;
      }
      const int want_kind = ee == G::TW ? 1 : 0;
      const int base = ee == G::TW ? 0 : ee * G::N3;
      if (s.kind != want_kind || (s.off - base) % G::VW) { ++nbad; continue; }
      const int c = (s.off - base) / G::VW;
      if (c < 0 || c >= G::CPE || G::swz(c) != p) { ++nbad; continue; }
      ++seen[c];
    }
    for (int c = 0; c < G::CPE; ++c) nbad += seen[c] != 1;
  }
  if (nbad) printf("T=%zu N=%d ADV=%d SIG=%d: element layout wrong in %ld places\n", sizeof(T), N, (int)ADV, SIG, nbad);
  return nbad;
}

// Every tile is walked by exactly one wave, whatever the grid and direction.
long check_walks()
{
  long nbad = 0;
  const int cases[][2] = {{1, 1}, {7, 3}, {16, 16}, {100, 37}, {4096, 2048}, {16384, 2048}, {16384, 1024},
                          {5000, 999}, {262144, 2048}, {9, 8}, {2049, 2048}};
  for (auto& cs : cases)
    for (int down = 0; down < 2; ++down) {
      const int ntiles = cs[0], nw = cs[1];
      int* hits = (int*)calloc(ntiles, sizeof(int));
      for (int b = 0; b < nw; ++b) {
        const TileWalk w(b, nw, ntiles, down != 0);
        int j = w.j0;
        for (int k = 0; k < w.count; ++k, j = w.next(j)) {
          const int t = w.tile(j);
          if (j < 0 || j >= w.len || t < 0 || t >= ntiles) { ++nbad; break; }
          ++hits[t];
        }
      }
      for (int t = 0; t < ntiles; ++t) nbad += hits[t] != 1;
      free(hits);
      if (nbad) { printf("tile walk wrong: ntiles=%d waves=%d down=%d\n", ntiles, nw, down); return nbad; }
    }
  return nbad;
}

template <typename T, int N, bool ADV, int SIG>
long sweep()
{
  long bad = 0;
  const int sizes[][3] = {{16, 1, 1}, {16, 2, 3}, {32, 3, 2}, {64, 8, 8}, {16, 4, 4}, {48, 5, 1}};
  for (auto& m : sizes)
    for (int per = 0; per < 8; ++per)
      for (int neg = 0; neg < 8; ++neg) bad += check<T, N, ADV, SIG>(m[0], m[1], m[2], per, ADV ? neg : 0);
  return bad;
}

int main()
{
  constexpr int cyc = ksig(1, 2, 0), cst = ksig(-1, -1, -1);
  long bad = check_walks();
  bad += check_layout<double, 4, true, cyc>();
  bad += check_layout<float, 4, true, cyc>();
  bad += check_layout<double, 4, false, cst>();
  bad += sweep<double, 4, true, cyc>();
  bad += sweep<double, 4, true, cst>();
  bad += sweep<double, 4, false, cst>();
  bad += sweep<float, 4, true, cyc>();
  bad += sweep<float, 4, true, cst>();
  bad += sweep<float, 4, false, cst>();
  printf("TOTAL BAD %ld\n", bad);
  return bad != 0;
}
