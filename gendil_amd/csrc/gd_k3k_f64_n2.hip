// Kronecker-factored 3D kernel instances: double, 2 nodes per dimension.
#include "gd_k3k_inst.hpp"

namespace gd {
namespace k3k {

static const KEntry<double> kList[] = {GD_K3K_FOR_LANES(double, 2, 1), GD_K3K_FOR_LANES(double, 2, 2), GD_K3K_STAGED_ALL(double, 2)};

KList<double> instances_f64_n2() { return KList<double>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3k
}  // namespace gd
