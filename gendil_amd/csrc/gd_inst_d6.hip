// Kernel instances for 6D meshes: (nodes per dim, quadrature points per dim).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(6, 2, 2),
    GD_KSET(6, 2, 3),
    GD_KSET(6, 3, 3),
    GD_KSET(6, 4, 4)};

KernelList kernels_d6() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
