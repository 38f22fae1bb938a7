// 2-D step: precision-independent kernels and launch-shape queries.
#include <climits>
#include <math.h>
#include "k2d.hpp"

namespace k2d {

// Reduce the per-workgroup partials of one iteration into its diagnostics
// slot, in a fixed order (spectra_compute2D :991-1000, velgrad :1158-1180),
// and raise the NaN stop (solution.F90:69-81).
__global__ __launch_bounds__(256) void finalize_kernel(FinArgs a) {
    if (a.gate.closed()) return;
    const int K = a.allkmax, nb = K + 1;
    const int tid = threadIdx.x;
    for (int b = 1 + tid; b <= K; b += blockDim.x) {
        double es = 0, ed = 0;
        for (int w = 0; w < a.nblk_spec; ++w) {
            es += a.part_bins[((size_t)w * nb + b) * 2 + 0];
            ed += a.part_bins[((size_t)w * nb + b) * 2 + 1];
        }
        a.spec[K + b - 1] = es;
        a.spec[2 * K + b - 1] = ed;
    }
    if (tid == 0) {
        double t[4] = {0, 0, 0, 0};
        for (int w = 0; w < a.nblk_spec; ++w)
            for (int q = 0; q < 4; ++q) t[q] += a.part_tot[(size_t)w * 4 + q];
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        double eta = (double)(2.0f * 3.14f);
        for (int w = 0; w < a.nblk_vg; ++w) {
            for (int q = 0; q < 7; ++q) v[q] += a.part_vg[(size_t)w * 8 + q];
            eta = fmin(eta, a.part_vg[(size_t)w * 8 + 7]);
        }
        const double N = a.npts, nu = a.nu;
        const double dim = a.ndims == 2 ? 2.0 : 3.0;
        const double urms = sqrt(v[4] / N), dudx2 = v[5] / N, eps = v[6] / N;
        const double ufluc = urms / sqrt(dim);
        const double tay = ufluc / sqrt(dudx2 / dim);
        bastr_diag* d = a.diag;
        d->Esspe = t[0];
        d->Edspe = t[1];
        d->k2Edspe = t[2];
        d->kMEdspe = t[3];
        d->umumtheta2 = v[0] / N;
        d->umumijji = v[1] / N;
        d->u2theta = v[2] / N;
        d->dissp = v[3] / N;
        d->eta_min = eta;
        d->Taylength = tay;
        d->ReTay = ufluc * tay / nu;
        d->Kollength = sqrt(sqrt(nu * nu * nu / eps));
        if (isnan(t[0] + t[1])) {
            d->nan_flag = 1;
            if (*a.status > a.gate.it) *a.status = a.gate.it;
        }
    }
}

int launch_finalize(const FinArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, s, a);
    return 0;
}

template <typename T>
static int cb_of(int n) {
    switch (n) {
#define X(NN) case NN: return ColCfg<T, NN>::CB;
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return -1;
}
template <typename T>
static int l_of(int n) {
    switch (n) {
#define X(NN) case NN: return RowCfg<T, NN>::L;
        X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#undef X
    }
    return -1;
}

int col_block_width(int n, int prec) { return prec == BASTR_PREC_F32 ? cb_of<float>(n) : cb_of<double>(n); }
int row_block_rows(int n, int prec) { return prec == BASTR_PREC_F32 ? l_of<float>(n) : l_of<double>(n); }
int velgrad_blocks(int n0, int n1, int prec) {
    const int L = row_block_rows(n0, prec);
    return L > 0 ? (n1 + L - 1) / L : -1;
}

}  // namespace k2d
