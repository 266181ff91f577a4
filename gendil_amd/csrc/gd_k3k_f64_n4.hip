// Kronecker-factored 3D kernel instances: double, 4 nodes per dimension.
#include "gd_k3k_inst.hpp"

namespace gd {
namespace k3k {

static const KEntry<double> kList[] = {GD_K3K_FOR_LANES(double, 4, 2), GD_K3K_FOR_LANES(double, 4, 4), GD_K3K_STAGED_ALL(double, 4), GD_K3K_STREAM(double, 4, 2), GD_K3K_STREAM(double, 4, 1), GD_K3K_STREAM_ACC(double, 4)};

KList<double> instances_f64_n4() { return KList<double>{kList, (int)(sizeof(kList) / sizeof(kList[0]))}; }

}  // namespace k3k
}  // namespace gd
