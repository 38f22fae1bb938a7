// 2-D step kernels, FP64 instantiations.
#include "k2d_impl.hpp"
namespace k2d {
K2D_INSTANTIATE(double)
}
