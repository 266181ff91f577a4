// Kernel instances for 3D meshes: (nodes per dim, quadrature points per dim),
// including GenDiL's 3D range benchmarks (q = n + 1, p = 0..6:
// benchmarks/range-benchmark-advection-3d.cpp:95-137, range-benchmark-mass-3d.cpp).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(3, 1, 1),
    GD_KSET(3, 1, 2),
    GD_KSET(3, 1, 3),
    GD_KSET(3, 2, 2),
    GD_KSET(3, 2, 3),
    GD_KSET(3, 2, 4),
    GD_KSET(3, 3, 3),
    GD_KSET(3, 3, 4),
    GD_KSET(3, 3, 5),
    GD_KSET(3, 4, 4),
    GD_KSET(3, 4, 5),
    GD_KSET(3, 4, 6),
    GD_KSET(3, 5, 5),
    GD_KSET(3, 5, 6),
    GD_KSET(3, 5, 7),
    GD_KSET(3, 6, 7),
    GD_KSET(3, 7, 8)};

KernelList kernels_d3() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
