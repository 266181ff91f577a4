// Register-resident 3D kernel instances: float, 3 nodes per dimension.
#include "gd_k3d_inst.hpp"

namespace gd {
namespace k3 {

static const K3Entry<float> kList[] = {GD_K3_FOR_LANES(float, 3, 1), GD_K3_ANY(float, 3, 1)};

K3List<float> instances_f32_n3() { return K3List<float>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3
}  // namespace gd
