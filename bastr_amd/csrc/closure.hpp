// closure.hpp -- spectral helper kernels of the learned subgrid closure
// (closure.hip): the coarse fields the network reads, the sharp filter from a
// fine grid to a coarse one, and the subgrid force of a training pair.
#pragma once
#include "k2d.hpp"

namespace clo {

// Half spectra of one 2-D grid (n0 x n1, pitch hp), as the step keeps them
struct Half {
    int n0, n1, nh, hp;
};

// theta_hat = i (k1 u1 + k2 u2), omega_hat = i (k1 u2 - k2 u1) (odd powers of a
// wavenumber vanish at that axis's Nyquist index, as in every pass)
template <typename T>
int launch_grad(Half g, const k2d::Gate& gate, const cx<T>* u1, const cx<T>* u2, cx<T>* th, cx<T>* om, hipStream_t s);
// dst = src on the modes |k|^2 < cut^2 of the coarse grid, 0 elsewhere; nf fields;
// src is the fine grid's half spectrum (same normalisation, 1/N^d forward)
template <typename T>
int launch_filter(Half fine, Half coarse, int cut, const cx<T>* const* src, cx<T>* const* dst, int nf, hipStream_t s);
// out_c = a_c - b_c on |k|^2 < cut^2, 0 elsewhere (two fields)
template <typename T>
int launch_tau(Half g, int cut, const cx<T>* a0, const cx<T>* a1, const cx<T>* b0, const cx<T>* b1, cx<T>* o0,
               cx<T>* o1, hipStream_t s);
// x_c += y_c over n entries (two fields)
template <typename T>
int launch_add(const k2d::Gate& gate, cx<T>* x0, cx<T>* x1, const cx<T>* y0, const cx<T>* y1, size_t n, hipStream_t s);
// dst = (float)src
template <typename T>
int launch_to_float(const T* src, float* dst, size_t n, hipStream_t s);

}  // namespace clo
