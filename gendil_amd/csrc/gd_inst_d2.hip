// Kernel instances for 2D meshes: (nodes per dim, quadrature points per dim).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(2, 2, 2),
    GD_KSET(2, 3, 3),
    GD_KSET(2, 3, 4),
    GD_KSET(2, 4, 4),
    GD_KSET(2, 4, 5)};

KernelList kernels_d2() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
