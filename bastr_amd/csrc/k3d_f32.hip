// 3-D step kernels, FP32 instantiations (FP64 kept in every reduction).
#include "k3d_impl.hpp"
namespace k3d {
K3D_INSTANTIATE(float)
}
