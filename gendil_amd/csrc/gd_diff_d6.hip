// SIPDG diffusion kernel instance for 6D meshes: Legendre p = 1, q = p + 2
// (examples/poisson/poisson-convergence-6d.cpp).
#include "gd_diff.hpp"

namespace gd {

static const DiffSet kSets[] = {GD_DSET(6, 2, 3)};

DiffList diff_kernels_d6() { return DiffList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
