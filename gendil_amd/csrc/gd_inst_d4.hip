// Kernel instances for 4D meshes: (nodes per dim, quadrature points per dim).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(4, 2, 2),
    GD_KSET(4, 2, 3),
    GD_KSET(4, 3, 3),
    GD_KSET(4, 3, 4)};

KernelList kernels_d4() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
