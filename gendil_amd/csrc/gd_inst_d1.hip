// Kernel instances for 1D meshes: (nodes per dim, quadrature points per dim).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(1, 2, 2),
    GD_KSET(1, 3, 3),
    GD_KSET(1, 4, 4),
    GD_KSET(1, 4, 5)};

KernelList kernels_d1() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
