// SIPDG diffusion kernel instances for 3D meshes: (nodes, points) of GenDiL's
// diffusion tests (GLL, q = p + 3: test-diffusion-3d-generic.cpp:78-100), of
// q = p + 2 (examples/poisson) and of q = p + 1.
#include "gd_diff.hpp"

namespace gd {

static const DiffSet kSets[] = {GD_DSET(3, 2, 3), GD_DSET(3, 2, 4), GD_DSET(3, 3, 3), GD_DSET(3, 3, 4),
                                GD_DSET(3, 3, 5), GD_DSET(3, 4, 4), GD_DSET(3, 4, 5), GD_DSET(3, 4, 6)};

DiffList diff_kernels_d3() { return DiffList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
