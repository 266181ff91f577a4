// Register-resident 3D kernel instances: float, 2 nodes per dimension.
#include "gd_k3d_inst.hpp"

namespace gd {
namespace k3 {

static const K3Entry<float> kList[] = {GD_K3_FOR_LANES(float, 2, 1), GD_K3_ANY(float, 2, 1)};

K3List<float> instances_f32_n2() { return K3List<float>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3
}  // namespace gd
