// 2-D step kernels, FP32 instantiations (FP64 kept in every reduction).
#include "k2d_impl.hpp"
namespace k2d {
K2D_INSTANTIATE(float)
}
