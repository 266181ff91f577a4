// SIPDG diffusion kernel instances for 1D and 2D meshes: (nodes, points) of
// GenDiL's diffusion tests and Poisson examples (GLL q = p + 3,
// test-diffusion-symmetry-1d-minimal.cpp; Legendre q = p + 2,
// examples/poisson/poisson-convergence-{1d,2d}.cpp) and q = p + 1.
#include "gd_diff.hpp"

namespace gd {

static const DiffSet kSets[] = {
    GD_DSET(1, 2, 3), GD_DSET(1, 2, 4), GD_DSET(1, 3, 4), GD_DSET(1, 4, 5), GD_DSET(1, 5, 6), GD_DSET(1, 4, 4),
    GD_DSET(2, 2, 3), GD_DSET(2, 2, 4), GD_DSET(2, 3, 4), GD_DSET(2, 3, 5), GD_DSET(2, 4, 4), GD_DSET(2, 4, 5),
    GD_DSET(2, 4, 6)};

DiffList diff_kernels_d12() { return DiffList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
