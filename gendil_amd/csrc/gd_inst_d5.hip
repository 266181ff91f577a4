// Kernel instances for 5D meshes: (nodes per dim, quadrature points per dim).
// GenDiL's 5D range benchmarks on the device: advection with q = n
// (benchmarks/range-benchmark-advection-5d.cpp:121-141, p = 0..3) and mass with
// q = n + 1 (range-benchmark-mass-5d.cpp:117-141, p = 0..2).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(5, 1, 1),
    GD_KSET(5, 1, 2),
    GD_KSET(5, 2, 2),
    GD_KSET(5, 2, 3),
    GD_KSET(5, 3, 3),
    GD_KSET(5, 3, 4),
    GD_KSET(5, 4, 4)};

KernelList kernels_d5() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
