// Register-resident 3D kernel instances: float, 4 nodes per dimension.
#include "gd_k3d_inst.hpp"

namespace gd {
namespace k3 {

static const K3Entry<float> kList[] = {GD_K3_FOR_LANES(float, 4, 1), GD_K3_FOR_LANES(float, 4, 2), GD_K3_FOR_LANES(float, 4, 4), GD_K3_ANY(float, 4, 2), GD_K3_ANY(float, 4, 4)};

K3List<float> instances_f32_n4() { return K3List<float>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3
}  // namespace gd
