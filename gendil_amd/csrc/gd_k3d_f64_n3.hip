// Register-resident 3D kernel instances: double, 3 nodes per dimension.
#include "gd_k3d_inst.hpp"

namespace gd {
namespace k3 {

static const K3Entry<double> kList[] = {GD_K3_FOR_LANES(double, 3, 1), GD_K3_ANY(double, 3, 1)};

K3List<double> instances_f64_n3() { return K3List<double>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3
}  // namespace gd
