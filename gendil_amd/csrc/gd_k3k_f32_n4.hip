// Kronecker-factored 3D kernel instances: float, 4 nodes per dimension.
#include "gd_k3k_inst.hpp"

namespace gd {
namespace k3k {

static const KEntry<float> kList[] = {GD_K3K_FOR_LANES(float, 4, 2), GD_K3K_FOR_LANES(float, 4, 4), GD_K3K_STAGED_ALL(float, 4), GD_K3K_STREAM(float, 4, 2), GD_K3K_STREAM(float, 4, 1)};

KList<float> instances_f32_n4() { return KList<float>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3k
}  // namespace gd
