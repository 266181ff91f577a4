// Register-resident 3D kernel instances: double, 4 nodes per dimension.
#include "gd_k3d_inst.hpp"

namespace gd {
namespace k3 {

static const K3Entry<double> kList[] = {GD_K3_FOR_LANES(double, 4, 1), GD_K3_FOR_LANES(double, 4, 2), GD_K3_FOR_LANES(double, 4, 4), GD_K3_ANY(double, 4, 2), GD_K3_ANY(double, 4, 4)};

K3List<double> instances_f64_n4() { return K3List<double>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3
}  // namespace gd
