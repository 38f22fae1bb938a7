// 3-D step kernels, FP64 instantiations.
#include "k3d_impl.hpp"
namespace k3d {
K3D_INSTANTIATE(double)
}
