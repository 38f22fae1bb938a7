// pp.hpp -- post-processing of a field on the device: Bastrpp's spectra2d and
// transfer2d/3d (Bastrpp.F90:174-599) on the half spectrum.
#pragma once
#include "k2d.hpp"

namespace pp {

using k2d::Gate;
using k2d::Geo;

// x pass of transfer: from the y- (and z-) inverse half rows of u and its
// derivatives, the physical products, and their x-forward transforms.
//   2-D hin: u1, u2, d2 u1, d2 u2;   hout: tA1 tA2 tC1 tC2 S11 S12 S22 (+ scratch)
//   3-D hin: u_c, d2 u_c, d3 u_c;    hout: S11 S12 S13 S22 S23 S33 tA1 tA2 tA3 tC1 tC2 tC3
// tA_i = -u_j d_j u_i, tC_i = u_i theta, S_ij = u_i u_j.
template <typename T>
struct RowArgs {
    Geo<T> g;
    int n2;              // rows per plane times planes: n1 * n2 rows in all
    Gate gate;
    const cx<T>* hin[9];
    cx<T>* hout[12];
    const cx<T>* tw;
};

// shell sums over the half spectrum of T, T1, T2 (per-block partials
// part[(b * 3 + q) * gridDim.x + block], b = shell 0..K).
template <typename T>
struct BinArgs {
    int nd, n0, n1, n2, nh, hp;
    Gate gate;
    const cx<T>* U[3];   // u_hat
    const cx<T>* F[12];  // fft of the products, in the order of RowArgs.hout
    int allkmax;
    double* part;
};

template <typename T> int launch_row2(int n0, const RowArgs<T>& a, hipStream_t s);
template <typename T> int launch_row3(int n0, const RowArgs<T>& a, hipStream_t s);
template <typename T> int launch_bins(const BinArgs<T>& a, int nblk, hipStream_t s);

}  // namespace pp
