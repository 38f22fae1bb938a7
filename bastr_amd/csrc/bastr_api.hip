// bastr_api.hip -- C-ABI of include/bastr.h over the 2-D HIP step kernels.
//
// One handle = one periodic field resident in HBM, one HIP stream.  The
// spectral state is the pair of normalised half spectra u1_hat, u2_hat; the
// physical field exists only when bastr_get_fields asks for it.  One loop
// iteration (solution.F90:54-97) is the launch sequence in run_iteration().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/bastr.h"
#include "../../include/bastr_closure.h"
#include "../../include/bastr_prof.h"
#include "closure.hpp"
#include "k2d.hpp"
#include "k3d.hpp"
#include "init_rp.hpp"
#include "mlp.hpp"
#include "pp.hpp"

using namespace k2d;

namespace {

thread_local char g_err[512] = "";

// Graph capture (bastr_step) is kept apart from the library's blocking
// runtime calls (allocations, synchronous copies) made on other threads for
// other handles: the runtime invalidates a capture in progress when such a
// call runs meanwhile.  Stream-ordered work (launches, async copies, stream and
// event waits) does not take the lock.
std::recursive_mutex g_sync_mu;
using SyncLock = std::lock_guard<std::recursive_mutex>;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(BASTR_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(call)                                                                    \
    do {                                                                                   \
        if ((call) != 0) return fail(BASTR_EINVAL, "launch shape rejected: %s", #call);    \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess) return fail(BASTR_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

__global__ void set_ints_kernel(int* p, int n, int v) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

__global__ void fill_diag_kernel(bastr_diag* d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        bastr_diag z = {};
        z.force_factor = z.force_E = z.force_energy = z.force_Fen = __builtin_nan("");
        d[i] = z;
    }
}

template <typename A, typename B>
__global__ void convert_kernel(const A* __restrict__ in, B* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (B)in[i];
}

// 16-byte copy for bastr_prof_copy_gbps: one element per thread, blocks in
// address order (the fastest plain copy measured here; a grid-stride loop
// reaches about three quarters of its rate)
__global__ void copy16_kernel(const double2* __restrict__ in, double2* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Device allocation, zero-filled.  The fill runs on the null stream, which the
// handles' non-blocking streams do not wait for: it is waited for here, or a
// copy or kernel queued next on a handle's stream could run before it and be
// overwritten with zeros (seen as a bastr_autotune_step snapshot that came
// back zeroed now and then).
void* dalloc(size_t bytes, bool* ok) {
    SyncLock lk(g_sync_mu);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) { *ok = false; return nullptr; }
    if (hipMemset(p, 0, bytes) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) { *ok = false; }
    return p;
}

}  // namespace

// A subgrid-closure network: FP32 master parameters, their BF16 operand form,
// Adam state and the scales of its inputs and outputs.
struct bastr_closure {
    hipStream_t stream = nullptr;
    float* params = nullptr;   // [NPARAM] FP32 master copy
    float* am = nullptr;       // Adam first and second moments
    float* av = nullptr;
    float* grad = nullptr;
    mlp::Packed* packed = nullptr;
    float scales[BASTR_CLOSURE_NSCALE] = {1, 1, 1, 1, 1, 1};
    mlp::Scales sc{};
    uint64_t version = 1;      // bumped when the scales change (captured graphs hold them)
    int t = 0;                 // Adam steps taken
    float* part = nullptr;     // per-wavefront partial gradients
    float* part_loss = nullptr;
    int part_w = 0;
    double* dloss = nullptr;
};

struct bastr_handle {
    bastr_params p{};
    int nd = 2, n0 = 0, n1 = 0, prec = 0, allkmax = 0;
    double nu = 0;
    int nh = 0, hp = 0, cut2 = 0;
    // Dealias-limited passes.  The step's spectra are zero from |k| = cut on
    // along every axis (the entry dealias zeroes them and every tendency is
    // dealiased), so the step's passes neither read nor write those modes:
    // x passes stop at kx = cut, column passes run cols columns (cut rounded
    // up to 16, or every column when that exceeds hp) and skip |k| >= cut
    // along the pass, and the ky rows |ky| < cut are rows_n rows, row r being
    // ky index r, or r + rows_skip from r = rows_at on.
    bool prune = true;
    int cut = 0, cols = 0, rows_n = 0, rows_at = 0, rows_skip = 0;
    size_t npts = 0, nhalf = 0;
    size_t esz = 8;  // sizeof(T)
    hipStream_t stream = nullptr;
    // 2-D overlap: the spectra and the diagnostics reduction run on `side`
    // beside the main chain (forked and joined with the events fork[0..2]
    // each iteration)
    bool overlap = false;
    hipStream_t side = nullptr;
    hipEvent_t fork[3] = {};
    void* U[3] = {};
    void* UB[3] = {};
    void* KA[3] = {};
    void* W[13] = {};            // 2-D: 7 y-inverse outputs; 3-D: 13 y-inverse outputs
    void* Z[10] = {};            // 3-D: z-inverse outputs
    void* NX[3] = {};            // x-pass outputs (3-D: aliases Z[0..2])
    void* NY[3] = {};            // 3-D y-forward outputs (aliases Z[3..5])
    void* ureal[3] = {};         // physical fields (3-D: aliases W[0..2])
    int n2 = 1;
    double* dstage = nullptr;  // FP64 host-exchange staging (FP32 mode)
    void* tw0 = nullptr;
    void* tw1 = nullptr;
    double* tw0d = nullptr;    // FP64 twiddles for the c2c entry points in FP32 mode
    double* tw1d = nullptr;
    void* cbuf = nullptr;      // complex128 work array of the c2c entry points
    double* part_bins = nullptr;
    double* part_tot = nullptr;
    double* part_vg = nullptr;
    double* red = nullptr;
    double* part_en = nullptr;   // energy partials (forcing methods 1, 5)
    double* dfactor = nullptr;   // forcing method 1 factor
    int nblk_spec = 0, nblk_vg = 0, nblk_en = 0;
    int* status = nullptr;       // [nmem]
    // Ensemble: nmem independent 2-D members of this grid, stacked member-major
    // in every field buffer, partial-sum buffer and diagnostic slot array;
    // every launch of the step covers all of them (blockIdx.z = member).  lw is
    // the current launch width: nmem, or 1 while a MemberView points the
    // buffers at one member.
    int nmem = 1, lw = 1;
    int dev = 0;                      // device current at bastr_create
    int law = BASTR_LAW_HITGEN;       // spectrum law of the random-phase generator
    // learned subgrid closure (2-D one-field RK3 handles): the attached network,
    // the graph's view of its scales, and the buffers of its force: theta_hat and
    // omega_hat (CW), the y-inverses of u1, u2, theta, omega (CY), those fields in
    // physical space (CR), the force (CF) and its x-forward (CX)
    bastr_closure* net = nullptr;
    uint64_t net_version = 0;
    void* CW[2] = {};
    void* CY[4] = {};
    void* CR[4] = {};
    void* CF[2] = {};
    void* CX[2] = {};
    bool clo_per_step = false;  // the closure force of stage A held for stages B and C
    int* mforce = nullptr;            // [nmem] device: member is band-forced (null: all are)
    std::vector<char> mset;           // member's field has been set
    std::vector<int64_t> mnstep;      // per-member nstep / time (ensembles)
    std::vector<double> mtime;
    bastr_diag* ddiag = nullptr;
    double* dspec = nullptr;
    int cap = 0;
    std::vector<double> kn;
    std::vector<int64_t> counts;
    int64_t nstep = 0;
    double time = 0;
    bool prologue_pending = false, have_fields = false;
    // per-kernel event timing (bastr_prof.h)
    uint32_t prof_mask = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_pending;
    int64_t prof_launch[BASTR_K_NCLASS] = {};
    double prof_ms[BASTR_K_NCLASS] = {};
    // the last unprofiled step call's iteration sequence, captured as a HIP
    // graph and replayed while the step count stays the same
    hipGraphExec_t gexec = nullptr;
    int gsteps = 0;
    // identity for graphs captured over several handles (bastr_step_group):
    // gen is unique per handle in the process, epoch counts the events that
    // invalidate captured launches (new buffers, closure scales, attach)
    uint64_t gen = 0, epoch = 0;
    uint64_t plan_epoch = 0;  // column radix plans the captured graph was made with
    bool use_graph = true;
    bool tune_pending = false;  // BASTR_TUNE=1: tune the column plans at the first bastr_step
    // post-processing arena (bastr_pp_*), allocated on first use
    void* pp[30] = {};
    int* pp_status = nullptr;  // always INT_MAX: the post-processing gate stays open
    double* pp_part = nullptr;
    int pp_nblk = 0;
};

namespace {

hipEvent_t pool_event(bastr_handle* h) {
    if (h->ev_used == h->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->ev_pool.push_back(e);
    }
    return h->ev_pool[h->ev_used++];
}

// Bracket one launch of class `cls` with events when that class is enabled.
struct ProfScope {
    bastr_handle* h;
    int cls;
    bool on;
    hipStream_t s;
    ProfScope(bastr_handle* h_, int c, hipStream_t st = nullptr)
        : h(h_), cls(c), on((h_->prof_mask >> c) & 1u), s(st ? st : h_->stream) {
        if (on) {
            hipEvent_t e = pool_event(h);
            if (!e) { on = false; return; }
            h->ev_pending.push_back({cls, h->ev_used - 1});
            (void)hipEventRecord(e, s);
        }
    }
    ~ProfScope() {
        if (on) {
            hipEvent_t e = pool_event(h);
            if (e) (void)hipEventRecord(e, s);
            else h->ev_pending.pop_back();
        }
    }
};

// After a stream synchronisation: fold the recorded pairs into the totals.
void prof_collect(bastr_handle* h) {
    for (auto& pc : h->ev_pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, h->ev_pool[pc.second], h->ev_pool[pc.second + 1]) == hipSuccess) {
            h->prof_launch[pc.first] += 1;
            h->prof_ms[pc.first] += ms;
        }
    }
    h->ev_pending.clear();
    h->ev_used = 0;
}

template <typename T>
void make_twiddles(int n, std::vector<cx<T>>& tw) {
    tw.resize(n);
    for (int m = 0; m < n; ++m) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        tw[m].x = (T)cosl(a);
        tw[m].y = (T)(-sinl(a));
    }
}

template <typename T>
int upload_twiddles(int n, void** out) {
    SyncLock lk(g_sync_mu);
    std::vector<cx<T>> tw;
    make_twiddles<T>(n, tw);
    HIPCHK(hipMalloc(out, sizeof(cx<T>) * n));
    HIPCHK(hipMemcpy(*out, tw.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
Geo<T> geo(const bastr_handle* h) {
    Geo<T> g;
    g.n0 = h->n0;
    g.n1 = h->n1;
    g.nh = h->nh;
    g.hp = h->hp;
    g.cut2 = h->cut2;
    g.lproject = h->p.lproject;
    g.kxlim = h->hp;
    g.nu = h->nu;
    g.dt = h->p.deltat;
    g.inv_n = 1.0 / ((double)h->n0 * (double)h->n1);
    g.hstride = (long)h->nhalf;
    g.rstride = (long)h->npts;
    g.nmem = h->lw;
    return g;
}

// every member's NaN gate open (status = INT_MAX)
int reset_status(bastr_handle* h) {
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(256), 0, h->stream, h->status, h->lw, INT_MAX);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Points a handle's per-member buffers at member m of an ensemble and
// narrows its launches to that member, for the one-member host operations
// (set/get fields, the random-phase generator); restores them on exit.
struct MemberView {
    bastr_handle* h;
    void *U[2], *UB[2], *KA[2], *NX[2], *ureal[2], *W[7];
    int* status;
    double *part_en, *dfactor;
    int lw;
    MemberView(bastr_handle* h_, int m) : h(h_) {
        const size_t co = (size_t)m * h->nhalf * 2 * h->esz, ro = (size_t)m * h->npts * h->esz;
        auto mv = [](void*& p, size_t o) { p = (void*)((char*)p + o); };
        for (int c = 0; c < 2; ++c) {
            U[c] = h->U[c]; UB[c] = h->UB[c]; KA[c] = h->KA[c]; NX[c] = h->NX[c]; ureal[c] = h->ureal[c];
            mv(h->U[c], co); mv(h->UB[c], co); mv(h->KA[c], co); mv(h->NX[c], co); mv(h->ureal[c], ro);
        }
        for (int q = 0; q < 7; ++q) {
            W[q] = h->W[q];
            mv(h->W[q], co);
        }
        status = h->status;
        part_en = h->part_en;
        dfactor = h->dfactor;
        lw = h->lw;
        h->status += m;
        h->part_en += (size_t)m * h->nblk_en;
        h->dfactor += m;
        h->lw = 1;
    }
    ~MemberView() {
        for (int c = 0; c < 2; ++c) {
            h->U[c] = U[c]; h->UB[c] = UB[c]; h->KA[c] = KA[c]; h->NX[c] = NX[c]; h->ureal[c] = ureal[c];
        }
        for (int q = 0; q < 7; ++q) h->W[q] = W[q];
        h->status = status;
        h->part_en = part_en;
        h->dfactor = dfactor;
        h->lw = lw;
    }
};

// x-pass limit of the step's passes (kx < cut when dealias-limited)
int step_kxlim(const bastr_handle* h) { return h->prune ? h->cut : h->hp; }

template <typename T>
cx<T>* C(void* p) { return (cx<T>*)p; }

k3d::ColJob job1(int in, int mult, int out) {
    k3d::ColJob j{};
    j.nterm = 1;
    j.in[0] = in;
    j.mult[0] = mult;
    j.out = out;
    return j;
}

// 2-D y-inverse of the half spectra src[0..1] into W:
//   mode 0: W0 = u1, W1 = u2
//   mode 1: + W2 = i k2 u1, W3 = i k2 u2                 (compute_ut2D :825-830)
//   mode 2: + W4 = Lap u1, W5 = Lap u2 = -(k1^2 + k2^2) u_i, W6 = d2 Lap u2
//           = -i k2 (k1^2 + k2^2) u2, the y part of Lap theta = d1 Lap u1 +
//           d2 Lap u2 (velgrad_calculate2D :1125-1140); k1 is the column's kx
template <typename T>
int run_yinv2(bastr_handle* h, const Gate& g, const void* const* src, int mode, bool prune) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    c.in[0] = C<T>((void*)src[0]);
    c.in[1] = C<T>((void*)src[1]);
    for (int q = 0; q < 7; ++q) c.out[q] = C<T>(h->W[q]);
    int nj = 0;
    c.job[nj++] = job1(0, k3d::M_ONE, 0);
    c.job[nj++] = job1(1, k3d::M_ONE, 1);
    if (mode >= 1) {
        c.job[nj++] = job1(0, k3d::M_IK, 2);
        c.job[nj++] = job1(1, k3d::M_IK, 3);
    }
    if (mode == 2) {
        c.job[nj++] = job1(0, k3d::M_LAP, 4);
        c.job[nj++] = job1(1, k3d::M_LAP, 5);
        c.job[nj++] = job1(1, k3d::M_LAP_IK, 6);
        c.kc_div = 1;  // the column offset within a row is kx
        c.kc_n = h->n0;
    }
    c.njob = nj;
    c.pitch = h->hp;
    c.ncols = prune && h->prune ? h->cols : h->hp;
    c.kcut = prune && h->prune ? h->cut : 0;
    c.plane_stride = 0;
    c.nplanes = 1;
    c.scale = 1.0;
    c.tw = C<T>(h->tw1);
    c.mstride = (long)h->nhalf;
    c.nmem = h->lw;
    LAUNCHCHK(k3d::launch_colx<T>(h->n1, true, c, h->stream));
    return 0;
}

// 2-D y-forward of nf half-spectrum fields src -> dst, times `scale`
// (prune: only the columns and rows the dealiased update reads)
template <typename T>
int run_yfwd2(bastr_handle* h, const Gate& g, void* const* src, void* const* dst, int nf, double scale, int pitch,
              const void* tw, bool prune = false) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < nf; ++q) {
        c.in[q] = C<T>(src[q]);
        c.out[q] = C<T>(dst[q]);
        c.job[q] = job1(q, k3d::M_ONE, q);
    }
    c.njob = nf;
    c.pitch = pitch;
    c.ncols = prune && h->prune ? h->cols : pitch;
    c.kcut = prune && h->prune ? h->cut : 0;
    c.plane_stride = 0;
    c.nplanes = 1;
    c.scale = scale;
    c.tw = (const cx<T>*)tw;
    c.mstride = (long)h->nhalf;
    c.nmem = h->lw;
    LAUNCHCHK(k3d::launch_colx<T>(h->n1, false, c, h->stream));
    return 0;
}

// spectral update (dealias/stage combination) from the y-forward outputs in
// W[0..1] (prune: the modes inside the dealias limits)
template <typename T>
int run_update2(bastr_handle* h, const Gate& g, int mode, bool first, bool prune, void* const* src = nullptr) {
    ColFwdArgs<T> c{};
    c.g = geo<T>(h);
    c.gate = g;
    c.nx[0] = C<T>(src ? src[0] : h->W[0]);
    c.nx[1] = C<T>(src ? src[1] : h->W[1]);
    for (int q = 0; q < 2; ++q) {
        c.U[q] = C<T>(h->U[q]);
        c.UB[q] = C<T>(h->UB[q]);
        c.KA[q] = C<T>(h->KA[q]);
    }
    c.pitch = h->hp;
    c.cn_first = first ? 1 : 0;
    if (prune && h->prune) {
        c.nkx = h->cut;
        c.nky = h->rows_n;
        c.ky_skip_at = h->rows_at;
        c.ky_skip = h->rows_skip;
    } else {
        c.nkx = h->nh;
        c.nky = h->n1;
        c.ky_skip_at = h->n1;
        c.ky_skip = 0;
    }
    LAUNCHCHK(launch_update<T>(mode, c, h->stream));
    return 0;
}

// y-forward of the x-pass outputs NX fused with the spectral update of a stage
template <typename T>
int run_cf2(bastr_handle* h, const Gate& g, int mode, bool first) {
    ColFwdArgs<T> c{};
    c.g = geo<T>(h);
    c.gate = g;
    c.nx[0] = C<T>(h->NX[0]);
    c.nx[1] = C<T>(h->NX[1]);
    for (int q = 0; q < 2; ++q) {
        c.U[q] = C<T>(h->U[q]);
        c.UB[q] = C<T>(h->UB[q]);
        c.KA[q] = C<T>(h->KA[q]);
    }
    c.pitch = h->hp;
    c.cn_first = first ? 1 : 0;
    c.nkx = h->prune ? h->cols : h->hp;
    c.kcut = h->prune ? h->cut : 0;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_cf2<T>(h->n1, mode, c, h->stream));
    return 0;
}

// u_hat = dealias(fft2d(u)) from the real fields in ureal (mainloop2D :41-52)
template <typename T>
int run_prologue(bastr_handle* h) {
    const Gate open{h->status, 0, 0};
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.gate = open;
    r.rin[0] = (const T*)h->ureal[0];
    r.rin[1] = (const T*)h->ureal[1];
    r.hout[0] = C<T>(h->NX[0]);
    r.hout[1] = C<T>(h->NX[1]);
    r.tw = C<T>(h->tw0);
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    LAUNCHCHK(launch_row<T>(h->n0, R_FWD_PACK, r, h->n1, h->stream));
    void* nx[2] = {h->NX[0], h->NX[1]};
    void* w[2] = {h->W[0], h->W[1]};
    int rc = run_yfwd2<T>(h, open, nx, w, 2, 1.0 / ((double)h->n0 * (double)h->n1), h->hp, h->tw1);
    if (rc) return rc;
    return run_update2<T>(h, open, UP_PROLOGUE, false, false);
}

// u = ifft2d(u_hat) into ureal
template <typename T>
int run_to_physical(bastr_handle* h) {
    const Gate open{h->status, 0, 0};
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    const void* Us[2] = {h->U[0], h->U[1]};
    int rc = run_yinv2<T>(h, open, Us, 0, false);
    if (rc) return rc;
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.gate = open;
    r.hin[0] = C<T>(h->W[0]);
    r.hin[1] = C<T>(h->W[1]);
    r.rout[0] = (T*)h->ureal[0];
    r.rout[1] = (T*)h->ureal[1];
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(launch_row<T>(h->n0, R_INV_PACK, r, h->n1, h->stream));
    return 0;
}

// The closure's coarse fields u1, u2, theta, omega in physical space (CR) from
// the half spectra Uin (prune: the step's dealias-limited passes)
template <typename T>
int closure_fields(bastr_handle* h, const Gate& g, const void* const* Uin, bool prune) {
    const clo::Half hf{h->n0, h->n1, h->nh, h->hp};
    LAUNCHCHK(clo::launch_grad<T>(hf, g, C<T>((void*)Uin[0]), C<T>((void*)Uin[1]), C<T>(h->CW[0]), C<T>(h->CW[1]),
                                  h->stream));
    const bool pr = prune && h->prune;
    k3d::ColXArgs<T> c{};
    c.gate = g;
    c.in[0] = C<T>((void*)Uin[0]);
    c.in[1] = C<T>((void*)Uin[1]);
    c.in[2] = C<T>(h->CW[0]);
    c.in[3] = C<T>(h->CW[1]);
    for (int q = 0; q < 4; ++q) {
        c.out[q] = C<T>(h->CY[q]);
        c.job[q] = job1(q, k3d::M_ONE, q);
    }
    c.njob = 4;
    c.pitch = h->hp;
    c.ncols = pr ? h->cols : h->hp;
    c.kcut = pr ? h->cut : 0;
    c.plane_stride = 0;
    c.nplanes = 1;
    c.scale = 1.0;
    c.tw = C<T>(h->tw1);
    c.mstride = (long)h->nhalf;
    c.nmem = 1;
    LAUNCHCHK(k3d::launch_colx<T>(h->n1, true, c, h->stream));
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.g.kxlim = pr ? h->cut : h->hp;
    r.gate = g;
    r.tw = C<T>(h->tw0);
    for (int q = 0; q < 4; q += 2) {
        r.hin[0] = C<T>(h->CY[q]);
        r.hin[1] = C<T>(h->CY[q + 1]);
        r.rout[0] = (T*)h->CR[q];
        r.rout[1] = (T*)h->CR[q + 1];
        LAUNCHCHK(launch_row<T>(h->n0, R_INV_PACK, r, h->n1, h->stream));
    }
    return 0;
}

// The attached network's force on the stage field Uin, added to the stage's
// x-forwarded nonlinear term NX (before its y-forward, dealiasing and projection)
template <typename T>
int run_closure(bastr_handle* h, const Gate& g, const void* const* Uin) {
    int rc = closure_fields<T>(h, g, Uin, true);
    if (rc) return rc;
    mlp::Fields<T> F{};
    for (int q = 0; q < 4; ++q) F.f[q] = (const T*)h->CR[q];
    F.n = h->n0;
    F.shift = __builtin_ctz((unsigned)h->n0);
    LAUNCHCHK(mlp::launch_forward<T>(h->net->packed, h->net->sc, F, (T*)h->CF[0], (T*)h->CF[1], h->stream));
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.gate = g;
    r.rin[0] = (const T*)h->CF[0];
    r.rin[1] = (const T*)h->CF[1];
    r.hout[0] = C<T>(h->CX[0]);
    r.hout[1] = C<T>(h->CX[1]);
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(launch_row<T>(h->n0, R_FWD_PACK, r, h->n1, h->stream));
    LAUNCHCHK(clo::launch_add<T>(g, C<T>(h->NX[0]), C<T>(h->NX[1]), C<T>(h->CX[0]), C<T>(h->CX[1]), h->nhalf,
                                 h->stream));
    return 0;
}

// One RK3 stage (stage 0..2) or CN sweep (stage 3 = middle sweep, 4 = last
// sweep; cn_first marks sweep 1): y-inverse (derivatives), x pass (nonlinear
// term), y-forward + update.  The first stage/sweep's y-inverse and x pass ran
// in the diagnostic phase (shared with velgrad).
template <typename T>
int run_stage(bastr_handle* h, int stage, const Gate& gstep, const void* const* Uin, bool first = false) {
    if (!first) {
        ProfScope ps(h, BASTR_K_COL_INV);
        const int rc = run_yinv2<T>(h, gstep, Uin, 1, true);
        if (rc) return rc;
    }
    // (stage A's y-inverse ran in the diagnostic phase with the velgrad outputs)
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.g.kxlim = step_kxlim(h);
    r.gate = gstep;
    for (int q = 0; q < 4; ++q) r.hin[q] = C<T>(h->W[q]);
    r.hout[0] = C<T>(h->NX[0]);
    r.hout[1] = C<T>(h->NX[1]);
    r.tw = C<T>(h->tw0);
    if (!first) {  // the first x pass ran fused with velgrad in the diagnostic phase
        ProfScope ps(h, BASTR_K_ROW_STAGE);
        LAUNCHCHK(launch_row<T>(h->n0, R_STAGE, r, h->n1, h->stream));
    }
    if (h->net) {  // subgrid force of the stage field (RK3 only: Uin is the stage input)
        ProfScope ps(h, BASTR_K_CLOSURE);
        if (h->clo_per_step && stage > 0) {
            // once per step: stage A's x-transformed force, still in CX
            LAUNCHCHK(clo::launch_add<T>(gstep, C<T>(h->NX[0]), C<T>(h->NX[1]), C<T>(h->CX[0]), C<T>(h->CX[1]),
                                         h->nhalf, h->stream));
        } else {
            const int rc = run_closure<T>(h, gstep, Uin);
            if (rc) return rc;
        }
    }
    {
        static const int modes[5] = {UP_A, UP_B, UP_C, UP_CN_MID, UP_CN_LAST};
        ProfScope ps(h, BASTR_K_COL_FWD_A + (stage < 3 ? stage : 2));
        int rc;
        if (k3d::cf2_full_rows<T>(h->n1)) {
            rc = run_cf2<T>(h, gstep, modes[stage], first);
        } else {
            // the fused kernel would need narrower column blocks here: y-forward, then the update
            void* nx[2] = {h->NX[0], h->NX[1]};
            void* w[2] = {h->W[0], h->W[1]};
            rc = run_yfwd2<T>(h, gstep, nx, w, 2, 1.0 / ((double)h->n0 * (double)h->n1), h->hp, h->tw1, true);
            if (!rc) rc = run_update2<T>(h, gstep, modes[stage], first, true, w);
        }
        if (rc) return rc;
    }
    return 0;
}

// One iteration of mainloop2D's body (solution.F90:54-97).
// With h->overlap the diagnostics that only feed the log run on h->side:
//   main: fork0 | col_inv_vg, velgrad (stage A's x pass) | fork1 | stage A | wait fork2 | stages B, C, forcing
//   side: wait fork0 | spectra | wait fork1 | reduction | fork2
// Nothing writes u_hat before the join; a NaN found by the reduction stops
// stages B and C and the forcing (stage A writes only its scratch KA, UB), as
// in the one-stream order.
template <typename T>
int run_iteration(bastr_handle* h, int it) {
    const Gate gdiag{h->status, it, 0};
    const Gate gstep{h->status, it, 1};
    const bool ov = h->overlap;
    hipStream_t ss = ov ? h->side : h->stream;
    if (ov) {
        HIPCHK(hipEventRecord(h->fork[0], h->stream));
        HIPCHK(hipStreamWaitEvent(h->side, h->fork[0], 0));
    }
    // spectra_compute2D on u_hat
    SpecArgs<T> sa{};
    sa.g = geo<T>(h);
    sa.gate = gdiag;
    sa.U[0] = C<T>(h->U[0]);
    sa.U[1] = C<T>(h->U[1]);
    sa.allkmax = h->allkmax;
    sa.part_bins = h->part_bins;
    sa.part_tot = h->part_tot;
    if (h->prune) {  // the modes inside the dealias limits (the rest are zeros)
        sa.nkx = h->cut;
        sa.nky = h->rows_n;
        sa.ky_at = h->rows_at;
        sa.ky_skip = h->rows_skip;
    }
    {
        ProfScope ps(h, BASTR_K_SPECTRA, ss);
        LAUNCHCHK(launch_spectra<T>(sa, h->nblk_spec, ss));
    }
    // velgrad_calculate2D: derivative fields shared with RK stage A
    {
        ProfScope ps(h, BASTR_K_COL_INV_VG);
        const void* Us[2] = {h->U[0], h->U[1]};
        const int rc = run_yinv2<T>(h, gdiag, Us, 2, true);
        if (rc) return rc;
    }
    VgArgs<T> va{};
    va.g = geo<T>(h);
    va.g.kxlim = step_kxlim(h);
    va.gate = gdiag;
    for (int q = 0; q < 7; ++q) va.w[q] = C<T>(h->W[q]);
    va.part = h->part_vg;
    va.hout[0] = C<T>(h->NX[0]);
    va.hout[1] = C<T>(h->NX[1]);
    va.tw = C<T>(h->tw0);
    {
        ProfScope ps(h, BASTR_K_VELGRAD);
        LAUNCHCHK(launch_velgrad<T>(h->n0, va, h->n1, h->stream, &h->nblk_vg));
    }
    if (ov) {
        HIPCHK(hipEventRecord(h->fork[1], h->stream));
        HIPCHK(hipStreamWaitEvent(h->side, h->fork[1], 0));
    }
    FinArgs fa{};
    fa.gate = gdiag;
    fa.nblk_spec = h->nblk_spec;
    fa.nblk_vg = h->nblk_vg;
    fa.allkmax = h->allkmax;
    fa.ndims = 2;
    fa.nu = h->nu;
    fa.npts = (double)h->npts;
    fa.part_bins = h->part_bins;
    fa.part_tot = h->part_tot;
    fa.part_vg = h->part_vg;
    fa.red = h->red;
    fa.diag = h->ddiag + it;
    fa.spec = h->dspec + (size_t)it * 3 * h->allkmax;
    fa.status = h->status;
    fa.nmem = h->lw;
    fa.dstride = h->cap;
    fa.sstride = (long)h->cap * 3 * h->allkmax;
    {
        ProfScope ps(h, BASTR_K_FINALIZE, ss);
        LAUNCHCHK(launch_finalize(fa, ss));
    }
    if (ov) HIPCHK(hipEventRecord(h->fork[2], h->side));
    const void* Ua[2] = {h->U[0], h->U[1]};
    const void* Ub[2] = {h->UB[0], h->UB[1]};
    int rc;
    if (h->p.timemethod == 1) {
        // RK32D (solution.F90:172-203)
        if ((rc = run_stage<T>(h, 0, gstep, Ua, true))) return rc;
        if (ov) HIPCHK(hipStreamWaitEvent(h->stream, h->fork[2], 0));
        if ((rc = run_stage<T>(h, 1, gstep, Ub))) return rc;
        if ((rc = run_stage<T>(h, 2, gstep, Ub))) return rc;
    } else {
        // CN2D (solution.F90:205-231): six sweeps u = uold + dt*F(0.5*(u + uold))
        for (int sweep = 1; sweep <= 6; ++sweep) {
            if ((rc = run_stage<T>(h, sweep == 6 ? 4 : 3, gstep, Ub, sweep == 1))) return rc;
            if (sweep == 1 && ov) HIPCHK(hipStreamWaitEvent(h->stream, h->fork[2], 0));
        }
    }
    // forcing2D (solution.F90:303-540)
    const int fm = h->p.forcemethod;
    if (fm == 1 || fm == 2 || fm == 5) {
        ProfScope ps(h, BASTR_K_FORCE);
        if (fm == 1 || fm == 5) {
            EnergyArgs<T> e{};
            e.g = geo<T>(h);
            e.gate = gstep;
            e.U[0] = C<T>(h->U[0]);
            e.U[1] = C<T>(h->U[1]);
            e.part_en = h->part_en;
            LAUNCHCHK(launch_energy<T>(e, h->nblk_en, h->stream));
        }
        ForceArgs<T> f{};
        f.g = geo<T>(h);
        f.gate = gstep;
        f.U[0] = C<T>(h->U[0]);
        f.U[1] = C<T>(h->U[1]);
        f.forcemethod = fm;
        f.forcek = h->p.forcek;
        f.target_energy = h->p.target_energy;
        f.part_en = h->part_en;
        f.nblk_en = h->nblk_en;
        f.factor_out = h->dfactor;
        f.diag = h->ddiag + it;
        f.dstride = h->cap;
        f.mforce = h->mforce;
        LAUNCHCHK(launch_force_band<T>(f, h->stream));
        if (fm == 1) {
            ScaleArgs<T> sc{};
            sc.g = geo<T>(h);
            sc.gate = gstep;
            sc.U[0] = C<T>(h->U[0]);
            sc.U[1] = C<T>(h->U[1]);
            sc.factor = h->dfactor;
            sc.mforce = h->mforce;
            LAUNCHCHK(launch_scale<T>(sc, h->stream));
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// 3-D launch sequence (mainloop3D body, solution.F90:124-169)
// ---------------------------------------------------------------------------
template <typename T>
k3d::Geo3<T> geo3(const bastr_handle* h) {
    k3d::Geo3<T> g;
    g.n0 = h->n0;
    g.n1 = h->n1;
    g.n2 = h->n2;
    g.nh = h->nh;
    g.hp = h->hp;
    g.cut2 = h->cut2;
    g.lproject = h->p.lproject;
    g.kxlim = h->hp;
    g.nu = h->nu;
    g.dt = h->p.deltat;
    g.inv_n = 1.0 / ((double)h->n0 * (double)h->n1 * (double)h->n2);
    return g;
}

// z-inverse of src[3]: Z_c, Z3_c = i k3 (and with vg: ZL_c = -(k1^2 + k2^2 + k3^2) u_c
// and ZD = -i (k1^2 + k2^2 + k3^2)(k2 u2 + k3 u3), Lap(u_c) and d2 Lap(u2) +
// d3 Lap(u3); k1 and k2 are the column's own wavenumbers)
template <typename T>
int run_zinv3(bastr_handle* h, const Gate& g, void* const* src, int mode /*0 plain 1 stage 2 vg*/, bool prune) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < 3; ++q) c.in[q] = C<T>(src[q]);
    for (int q = 0; q < 10; ++q) c.out[q] = C<T>(h->Z[q]);
    int nj = 0;
    for (int q = 0; q < 3; ++q) c.job[nj++] = job1(q, k3d::M_ONE, q);
    if (mode >= 1)
        for (int q = 0; q < 3; ++q) c.job[nj++] = job1(q, k3d::M_IK, 3 + q);
    if (mode == 2) {
        for (int q = 0; q < 3; ++q) c.job[nj++] = job1(q, k3d::M_LAP, 6 + q);
        k3d::ColJob d{};
        d.nterm = 2;
        d.in[0] = 1; d.mult[0] = k3d::M_LAP_IKC;
        d.in[1] = 2; d.mult[1] = k3d::M_LAP_IK;
        d.out = 9;
        c.job[nj++] = d;
        c.kc_div = h->hp;  // the column offset is ky hp + kx
        c.kc_n = h->n1;
        c.kx_mod = h->hp;
    }
    c.njob = nj;
    c.pitch = h->hp * h->n1;
    if (prune && h->prune && h->cols < h->hp) {
        // one plane per kept ky row, its first cols kx columns
        c.ncols = h->cols;
        c.plane_stride = h->hp;
        c.nplanes = h->rows_n;
        c.pskip_at = h->rows_at;
        c.pskip = h->rows_skip;
    } else {
        c.ncols = h->hp * h->n1;
        c.plane_stride = 0;
        c.nplanes = 1;
    }
    c.kcut = prune && h->prune ? h->cut : 0;
    c.scale = 1.0;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_colx<T>(h->n2, true, c, h->stream));
    return 0;
}

// y-inverse of the z-inverse outputs into W: Y_c, Y2_c = i k2 Y, Y3_c (and with
// vg: L_c = Lap(u_c), D = d2 Lap(u2) + d3 Lap(u3), the y-inverses of ZL_c and ZD)
template <typename T>
int run_yinv3(bastr_handle* h, const Gate& g, int mode /*0 plain 1 stage 2 vg*/, bool prune, int out0 = 0) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < 10; ++q) c.in[q] = C<T>(h->Z[q]);
    for (int q = 0; q < 13; ++q) c.out[q] = C<T>(h->W[q]);
    int nj = 0;
    for (int q = 0; q < 3; ++q) c.job[nj++] = job1(q, k3d::M_ONE, out0 + q);
    if (mode >= 1) {
        for (int q = 0; q < 3; ++q) c.job[nj++] = job1(q, k3d::M_IK, 3 + q);
        for (int q = 0; q < 3; ++q) c.job[nj++] = job1(3 + q, k3d::M_ONE, 6 + q);
    }
    if (mode == 2) {
        for (int q = 0; q < 3; ++q) c.job[nj++] = job1(6 + q, k3d::M_ONE, 9 + q);
        c.job[nj++] = job1(9, k3d::M_ONE, 12);
    }
    c.njob = nj;
    c.pitch = h->hp;
    c.ncols = prune && h->prune ? h->cols : h->hp;
    c.kcut = prune && h->prune ? h->cut : 0;
    c.plane_stride = (long)h->hp * h->n1;
    c.nplanes = h->n2;
    c.scale = 1.0;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_colx<T>(h->n1, true, c, h->stream));
    return 0;
}

// y-forward of the x-pass outputs NX -> NY
template <typename T>
int run_yfwd3(bastr_handle* h, const Gate& g, bool prune) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < 3; ++q) {
        c.in[q] = C<T>(h->NX[q]);
        c.out[q] = C<T>(h->NY[q]);
        c.job[q] = job1(q, k3d::M_ONE, q);
    }
    c.njob = 3;
    c.pitch = h->hp;
    c.ncols = prune && h->prune ? h->cols : h->hp;
    c.kcut = prune && h->prune ? h->cut : 0;
    c.plane_stride = (long)h->hp * h->n1;
    c.nplanes = h->n2;
    c.scale = 1.0;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_colx<T>(h->n1, false, c, h->stream));
    return 0;
}

// 3-D z-forward of nf fields src -> dst, times 1/(ia*ja*ka) (the fft3d normalisation)
template <typename T>
int run_zfwd3(bastr_handle* h, const Gate& g, void* const* src, void* const* dst, int nf) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < nf; ++q) {
        c.in[q] = C<T>(src[q]);
        c.out[q] = C<T>(dst[q]);
        c.job[q] = job1(q, k3d::M_ONE, q);
    }
    c.njob = nf;
    c.pitch = h->hp * h->n1;
    c.ncols = h->hp * h->n1;
    c.plane_stride = 0;
    c.nplanes = 1;
    c.scale = 1.0 / ((double)h->n0 * (double)h->n1 * (double)h->n2);
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_colx<T>(h->n2, false, c, h->stream));
    return 0;
}

// z-forward of NY fused with the spectral update of the state
template <typename T>
int run_cf3(bastr_handle* h, const Gate& g, int mode, bool first, bool prune) {
    k3d::Cf3Args<T> c{};
    c.g = geo3<T>(h);
    c.gate = g;
    for (int q = 0; q < 3; ++q) {
        c.nx[q] = C<T>(h->NY[q]);
        c.U[q] = C<T>(h->U[q]);
        c.UB[q] = C<T>(h->UB[q]);
        c.KA[q] = C<T>(h->KA[q]);
    }
    c.cn_first = first ? 1 : 0;
    if (prune && h->prune && h->cols < h->hp) {
        c.nkx = h->cols;
        c.nky = h->rows_n;
        c.ky_skip_at = h->rows_at;
        c.ky_skip = h->rows_skip;
    } else {  // every column, as one flat row
        c.nkx = h->hp * h->n1;
        c.nky = 1;
        c.ky_skip_at = 1;
        c.ky_skip = 0;
    }
    c.kzcut = prune && h->prune ? h->cut : 0;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_cf3<T>(h->n2, mode, c, h->stream));
    return 0;
}

template <typename T>
int run_x3(bastr_handle* h, const Gate& g) {
    k3d::Row3Args<T> r{};
    r.g = geo3<T>(h);
    r.g.kxlim = step_kxlim(h);
    r.gate = g;
    for (int q = 0; q < 9; ++q) r.hin[q] = C<T>(h->W[q]);
    for (int q = 0; q < 3; ++q) r.hout[q] = C<T>(h->NX[q]);
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(k3d::launch_row3<T>(h->n0, k3d::R3_STAGE, r, h->n1 * h->n2, h->stream));
    return 0;
}

template <typename T>
int run_prologue3(bastr_handle* h) {
    const Gate open{h->status, 0, 0};
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    k3d::Row3Args<T> r{};
    r.g = geo3<T>(h);
    r.gate = open;
    for (int q = 0; q < 3; ++q) {
        r.rin[q] = (const T*)h->ureal[q];
        r.hout[q] = C<T>(h->NX[q]);
    }
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(k3d::launch_row3<T>(h->n0, k3d::R3_FWD_PACK, r, h->n1 * h->n2, h->stream));
    int rc;
    if ((rc = run_yfwd3<T>(h, open, false))) return rc;
    return run_cf3<T>(h, open, k3d::C3_PROLOGUE, false, false);
}

template <typename T>
int run_to_physical3(bastr_handle* h) {
    const Gate open{h->status, 0, 0};
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    int rc;
    if ((rc = run_zinv3<T>(h, open, h->U, 0, false))) return rc;
    if ((rc = run_yinv3<T>(h, open, 0, false, 3))) return rc;  // into W[3..5]; ureal aliases W[0..2]
    k3d::Row3Args<T> r{};
    r.g = geo3<T>(h);
    r.gate = open;
    for (int q = 0; q < 3; ++q) {
        r.hin[q] = C<T>(h->W[3 + q]);
        r.rout[q] = (T*)h->ureal[q];
    }
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(k3d::launch_row3<T>(h->n0, k3d::R3_INV_PACK, r, h->n1 * h->n2, h->stream));
    return 0;
}

// stage: 0..2 RK3 stages, 3 CN middle sweep, 4 CN last sweep
template <typename T>
int run_stage3(bastr_handle* h, int stage, const Gate& g, bool first) {
    int rc;
    if (!first) {
        {
            ProfScope ps(h, BASTR_K_COL_INV);
            if ((rc = run_zinv3<T>(h, g, h->UB, 1, true))) return rc;
            if ((rc = run_yinv3<T>(h, g, 1, true))) return rc;
        }
    }
    {
        ProfScope ps(h, BASTR_K_ROW_STAGE);
        if ((rc = run_x3<T>(h, g))) return rc;
    }
    static const int modes[5] = {k3d::C3_A, k3d::C3_B, k3d::C3_C, k3d::C3_CN_MID, k3d::C3_CN_LAST};
    ProfScope ps(h, BASTR_K_COL_FWD_A + (stage < 3 ? stage : 2));
    if ((rc = run_yfwd3<T>(h, g, true))) return rc;
    return run_cf3<T>(h, g, modes[stage], first, true);
}

template <typename T>
int run_iteration3(bastr_handle* h, int it) {
    const Gate gdiag{h->status, it, 0};
    const Gate gstep{h->status, it, 1};
    int rc;
    {
        ProfScope ps(h, BASTR_K_SPECTRA);
        k3d::Spec3Args<T> sa{};
        sa.g = geo3<T>(h);
        sa.gate = gdiag;
        for (int q = 0; q < 3; ++q) sa.U[q] = C<T>(h->U[q]);
        sa.allkmax = h->allkmax;
        sa.part_bins = h->part_bins;
        sa.part_tot = h->part_tot;
        if (h->prune) {  // the modes inside the dealias limits (the rest are zeros)
            sa.nkx = h->cut;
            sa.nky = sa.nkz = h->rows_n;
            sa.ky_at = sa.kz_at = h->rows_at;
            sa.ky_skip = sa.kz_skip = h->rows_skip;
        }
        LAUNCHCHK(k3d::launch_spec3<T>(sa, h->nblk_spec, h->stream));
    }
    {
        ProfScope ps(h, BASTR_K_COL_INV_VG);
        if ((rc = run_zinv3<T>(h, gdiag, h->U, 2, true))) return rc;
        if ((rc = run_yinv3<T>(h, gdiag, 2, true))) return rc;
    }
    {
        ProfScope ps(h, BASTR_K_VELGRAD);
        k3d::Vg3Args<T> va{};
        va.g = geo3<T>(h);
        va.g.kxlim = step_kxlim(h);
        va.gate = gdiag;
        for (int q = 0; q < 13; ++q) va.w[q] = C<T>(h->W[q]);
        va.part = h->part_vg;
        va.tw = C<T>(h->tw0);
        LAUNCHCHK(k3d::launch_vg3<T>(h->n0, va, h->n1 * h->n2, h->stream, h->nblk_vg));
    }
    {
        FinArgs fa{};
        fa.gate = gdiag;
        fa.nblk_spec = h->nblk_spec;
        fa.nblk_vg = h->nblk_vg;
        fa.allkmax = h->allkmax;
        fa.ndims = 3;
        fa.nu = h->nu;
        fa.npts = (double)h->npts;
        fa.part_bins = h->part_bins;
        fa.part_tot = h->part_tot;
        fa.part_vg = h->part_vg;
        fa.red = h->red;
        fa.diag = h->ddiag + it;
        fa.spec = h->dspec + (size_t)it * 3 * h->allkmax;
        fa.status = h->status;
        ProfScope ps(h, BASTR_K_FINALIZE);
        LAUNCHCHK(launch_finalize(fa, h->stream));
    }
    if (h->p.timemethod == 1) {
        // RK33D (solution.F90:233-270)
        if ((rc = run_stage3<T>(h, 0, gstep, true))) return rc;
        if ((rc = run_stage3<T>(h, 1, gstep, false))) return rc;
        if ((rc = run_stage3<T>(h, 2, gstep, false))) return rc;
    } else {
        // CN3D (solution.F90:272-301)
        for (int sweep = 1; sweep <= 6; ++sweep)
            if ((rc = run_stage3<T>(h, sweep == 6 ? 4 : 3, gstep, sweep == 1))) return rc;
    }
    const int fm = h->p.forcemethod;
    if (fm == 1 || fm == 2 || fm == 5) {
        ProfScope ps(h, BASTR_K_FORCE);
        if (fm == 1 || fm == 5) {
            k3d::Energy3Args<T> e{};
            e.g = geo3<T>(h);
            e.gate = gstep;
            for (int q = 0; q < 3; ++q) e.U[q] = C<T>(h->U[q]);
            e.part_en = h->part_en;
            LAUNCHCHK(k3d::launch_energy3<T>(e, h->nblk_en, h->stream));
        }
        k3d::Force3Args<T> f{};
        f.g = geo3<T>(h);
        f.gate = gstep;
        for (int q = 0; q < 3; ++q) f.U[q] = C<T>(h->U[q]);
        f.forcemethod = fm;
        f.forcek = h->p.forcek;
        f.target_energy = h->p.target_energy;
        f.part_en = h->part_en;
        f.nblk_en = h->nblk_en;
        f.factor_out = h->dfactor;
        f.diag = h->ddiag + it;
        LAUNCHCHK(k3d::launch_force3<T>(f, h->stream));
        if (fm == 1) {
            k3d::Scale3Args<T> sc{};
            sc.g = geo3<T>(h);
            sc.gate = gstep;
            for (int q = 0; q < 3; ++q) sc.U[q] = C<T>(h->U[q]);
            sc.factor = h->dfactor;
            LAUNCHCHK(k3d::launch_scale3<T>(sc, h->stream));
        }
    }
    return 0;
}

void drop_graph(bastr_handle* h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->gsteps = 0;
    h->epoch += 1;  // group graphs holding this handle's launches are stale too
}

// longest step call replayed as one graph (about 25 nodes per iteration)
constexpr int kMaxGraphSteps = 256;
// largest grid (points over all members) whose step calls replay a graph by default
constexpr double kGraphMaxPoints = 1u << 21;

int ensure_capacity(bastr_handle* h, int n) {
    if (n <= h->cap) return 0;
    SyncLock lk(g_sync_mu);
    drop_graph(h);  // the captured launches point at the old buffers
    if (h->ddiag) (void)hipFree(h->ddiag);
    if (h->dspec) (void)hipFree(h->dspec);
    h->ddiag = nullptr;
    h->dspec = nullptr;
    const size_t slots = (size_t)n * h->nmem;  // member-major: member m's slots start at m*n
    HIPCHK(hipMalloc(&h->ddiag, sizeof(bastr_diag) * slots));
    HIPCHK(hipMalloc(&h->dspec, sizeof(double) * 3 * (size_t)(h->allkmax > 0 ? h->allkmax : 1) * slots));
    HIPCHK(hipMemsetAsync(h->dspec, 0, sizeof(double) * 3 * (size_t)(h->allkmax > 0 ? h->allkmax : 1) * slots,
                          h->stream));
    h->cap = n;
    return 0;
}

// Ensemble results of a step call: member m's diagnostics go to
// diags[m*nsteps + it] and its spectra to spectra[(m*nsteps + it)*3*allkmax];
// a member whose energy went NaN at iteration s has s+1 iterations of
// diagnostics (the rest keep the fill values) and advances s iterations.
int collect_ensemble(bastr_handle* h, int nsteps, bastr_diag* diags, double* spectra) {
    const int M = h->nmem, K = h->allkmax;
    std::vector<int> status(M, INT_MAX);
    HIPCHK(hipMemcpyAsync(status.data(), h->status, sizeof(int) * M, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    prof_collect(h);
    bool any_nan = false;
    for (int m = 0; m < M; ++m) {
        if (diags)
            HIPCHK(hipMemcpyAsync(diags + (size_t)m * nsteps, h->ddiag + (size_t)m * h->cap,
                                  sizeof(bastr_diag) * nsteps, hipMemcpyDeviceToHost, h->stream));
        if (spectra && K > 0)
            HIPCHK(hipMemcpyAsync(spectra + (size_t)m * nsteps * 3 * K, h->dspec + (size_t)m * h->cap * 3 * K,
                                  sizeof(double) * 3 * K * (size_t)nsteps, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int m = 0; m < M; ++m) {
        const int done = status[m] == INT_MAX ? nsteps : status[m] + 1;
        any_nan = any_nan || status[m] != INT_MAX;
        if (spectra && K > 0) {
            double* sp = spectra + (size_t)m * nsteps * 3 * K;
            for (int it = 0; it < nsteps; ++it)
                for (int b = 1; b <= K; ++b) sp[(size_t)it * 3 * K + b - 1] = h->kn[b];
        }
        for (int it = 0; it < done; ++it) {
            if (diags) {
                diags[(size_t)m * nsteps + it].nstep = h->mnstep[m];
                diags[(size_t)m * nsteps + it].time = h->mtime[m];
            }
            if (it == status[m]) break;
            h->mnstep[m] = h->mnstep[m] + 1;
            h->mtime[m] = h->mtime[m] + h->p.deltat;
        }
    }
    return any_nan ? BASTR_STOP_NAN : BASTR_OK;
}

// The per-call launches before a step call's iterations: the pending entry
// dealias, diagnostic slots, the NaN gate and the fill values
template <typename T>
int step_prepare(bastr_handle* h, int nsteps) {
    int rc = 0;
    if (h->prologue_pending) {
        if ((rc = (h->nd == 3 ? run_prologue3<T>(h) : run_prologue<T>(h)))) return rc;
        h->prologue_pending = false;
    }
    if ((rc = ensure_capacity(h, nsteps))) return rc;
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    {
        const int nfill = h->nmem > 1 ? h->cap * h->nmem : nsteps;
        hipLaunchKernelGGL(fill_diag_kernel, dim3((nfill + 255) / 256), dim3(256), 0, h->stream, h->ddiag, nfill);
        HIPCHK(hipGetLastError());
    }
    if (h->net && h->net_version != h->net->version) {
        drop_graph(h);  // the captured closure launches hold the network's scales
        h->net_version = h->net->version;
    }
    if (h->plan_epoch != k3d::colx_plan_epoch()) {
        drop_graph(h);  // the captured column passes are those of the old radix plan
        h->plan_epoch = k3d::colx_plan_epoch();
    }
    return 0;
}

// Time the radix plans of the single-term column passes of length n
// (precision prec) on the current device, in the shape of a real pass: ncols
// columns of each of nplanes planes, with the pruned passes' cut along the
// pass (n/3), launched as an RK3 iteration of a 2-D step launches them (one
// 7-job inverse, two 4-job inverses, three 2-job forwards).  Every launch
// reads and writes arrays at other addresses than the launch before it (slots
// of two buffers of at least 768 MiB), so the passes stream from HBM as a
// step's do rather than from the 256 MB Infinity Cache, and each launch has the
// real pass's workgroup count (a pass over more planes than the real one hides
// how evenly its workgroups fill the CUs: 2048-point FP64 2-D forwards are 344
// workgroups).  The mix runs `reps` times between HIP events after a warm-up
// mix; ms[v] = milliseconds per mix for the plans (points, width) = (16, 1),
// (8, 1), (16, 2), (8, 2), NaN where a plan does not exist.
template <typename T>
int time_plans(int n, int ncols, int nplanes, double* ms) {
    const int reps = 5;
    // one array of the pass; at most 2^25 elements so that 8 slots fit 2^28
    size_t F = (size_t)ncols * n;
    nplanes = std::max(1, std::min<int>(nplanes, (int)(((size_t)1 << 25) / F)));
    F *= (size_t)nplanes;
    const size_t fbytes = sizeof(cx<T>) * F;
    const size_t nslot = std::max<size_t>(8, std::min<size_t>((((size_t)768 << 20) + fbytes - 1) / fbytes,
                                                              ((size_t)1 << 28) / F));
    const size_t elems = nslot * F;
    SyncLock lk(g_sync_mu);
    void *in = nullptr, *out = nullptr, *tw = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        if (s) (void)hipStreamSynchronize(s);
        for (void* p : {in, out, tw})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    };
    bool ok = hipMalloc(&in, sizeof(cx<T>) * elems) == hipSuccess && hipMalloc(&out, sizeof(cx<T>) * elems) == hipSuccess;
    ok = ok && hipMemset(in, 0, sizeof(cx<T>) * elems) == hipSuccess && hipStreamSynchronize(nullptr) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    if (ok) {
        std::vector<cx<T>> twh;
        make_twiddles<T>(n, twh);
        ok = hipMalloc(&tw, sizeof(cx<T>) * n) == hipSuccess &&
             hipMemcpy(tw, twh.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice) == hipSuccess;
    }
    int* st = nullptr;
    ok = ok && hipMalloc(&st, sizeof(int)) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, st, 1, INT_MAX);
        ok = hipGetLastError() == hipSuccess;
    }
    const int prec = sizeof(T) == 4 ? 1 : 0;
    const int keep_p = k3d::colx_points(n, prec), keep_w = k3d::colx_width(n, prec);
    // the mix: (jobs, inverse) per launch
    const int mix[6][2] = {{7, 1}, {4, 1}, {4, 1}, {2, 0}, {2, 0}, {2, 0}};
    size_t slot = 0;
    auto launch = [&](int njob, bool inv) {
        k3d::ColXArgs<T> c{};
        c.gate = Gate{st, 0, 0};
        c.in[0] = C<T>(in) + (slot % nslot) * F;
        c.in[1] = C<T>(in) + ((slot + 1) % nslot) * F;
        for (int q = 0; q < njob; ++q) {
            c.out[q] = C<T>(out) + ((slot + q) % nslot) * F;
            c.job[q] = job1(q & 1, inv && (q / 2) % 2 == 1 ? k3d::M_IK : k3d::M_ONE, q);
        }
        slot += 7;
        c.njob = njob;
        c.pitch = ncols;
        c.ncols = ncols;
        c.plane_stride = (long)ncols * n;
        c.nplanes = nplanes;
        c.scale = 1.0;
        c.kcut = n / 3 > 0 ? n / 3 : 0;
        c.tw = C<T>(tw);
        c.mstride = 0;
        c.nmem = 1;
        return k3d::launch_colx<T>(n, inv, c, s) == 0;
    };
    for (int v = 0; v < 4 && ok; ++v) {
        ms[v] = std::nan("");
        const int pts = (v & 1) ? 8 : 16, wid = (v & 2) ? 2 : 1;
        if (!k3d::colx_plan_ok(n, prec, pts, wid) || ncols % k3d::colx_plan_cb(n, prec, pts, wid) != 0) continue;
        k3d::set_colx_plan(n, prec, pts, wid);
        float t0 = 0.f;
        for (int r = 0; r <= reps && ok; ++r) {
            if (r == 1) ok = hipEventRecord(e0, s) == hipSuccess;
            for (int m = 0; m < 6 && ok; ++m) ok = launch(mix[m][0], mix[m][1] != 0);
        }
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&t0, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess;
        ms[v] = (double)t0 / reps;
    }
    k3d::set_colx_plan(n, prec, keep_p, keep_w);
    if (st) (void)hipFree(st);
    cleanup();
    if (!ok) rc = fail(BASTR_EDEVICE, "bastr_autotune: timing of the column plans failed");
    return rc;
}

bool graph_ok(const bastr_handle* h, int nsteps) {
    return h->use_graph && h->prof_mask == 0 && nsteps <= kMaxGraphSteps;
}

template <typename T>
int step_launch(bastr_handle* h, int nsteps) {
    int rc = step_prepare<T>(h, nsteps);
    if (rc) return rc;
    // The ~25 launches of each iteration are replayed as one graph; the capture
    // is kept for the next call with the same step count (every argument is
    // fixed per handle: buffers, the iteration index within the call, the
    // device-side NaN gate).  Profiled calls launch directly: their timing
    // events stay out of the capture.
    if (graph_ok(h, nsteps)) {
        if (!h->gexec || h->gsteps != nsteps) {
            SyncLock lk(g_sync_mu);
            drop_graph(h);
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < nsteps && !rc; ++it)
                rc = h->nd == 3 ? run_iteration3<T>(h, it) : run_iteration<T>(h, it);
            const hipError_t ec = hipStreamEndCapture(h->stream, &g);
            if (rc || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : fail(BASTR_EDEVICE, "bastr_step: graph capture failed: %s", hipGetErrorString(ec));
            }
            const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                h->gexec = nullptr;
                return fail(BASTR_EDEVICE, "bastr_step: graph instantiation failed: %s", hipGetErrorString(ei));
            }
            h->gsteps = nsteps;
        }
        HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else {
        for (int it = 0; it < nsteps; ++it)
            if ((rc = (h->nd == 3 ? run_iteration3<T>(h, it) : run_iteration<T>(h, it)))) return rc;
    }
    return 0;
}

// Wait for a step call's launches and return its results (the one-field
// layout, or the ensemble layout of collect_ensemble).
int step_collect(bastr_handle* h, int nsteps, bastr_diag* diags, double* spectra) {
    if (h->nmem > 1) return collect_ensemble(h, nsteps, diags, spectra);
    int status = INT_MAX;
    HIPCHK(hipMemcpyAsync(&status, h->status, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    prof_collect(h);
    const int done = status == INT_MAX ? nsteps : status + 1;  // iterations with diagnostics
    const int K = h->allkmax;
    if (diags)
        HIPCHK(hipMemcpyAsync(diags, h->ddiag, sizeof(bastr_diag) * done, hipMemcpyDeviceToHost, h->stream));
    if (spectra && K > 0)
        HIPCHK(hipMemcpyAsync(spectra, h->dspec, sizeof(double) * 3 * K * (size_t)done, hipMemcpyDeviceToHost,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (spectra && K > 0) {
        for (int it = 0; it < done; ++it)
            for (int b = 1; b <= K; ++b) spectra[(size_t)it * 3 * K + b - 1] = h->kn[b];
    }
    // nstep/time bookkeeping exactly as the reference accumulates them
    for (int it = 0; it < done; ++it) {
        if (diags) {
            diags[it].nstep = h->nstep;
            diags[it].time = h->time;
        }
        if (it == status) break;  // the NaN iteration does not advance
        h->nstep = h->nstep + 1;
        h->time = h->time + h->p.deltat;
    }
    return status == INT_MAX ? BASTR_OK : BASTR_STOP_NAN;
}

template <typename T>
int step_impl(bastr_handle* h, int nsteps, bastr_diag* diags, double* spectra) {
    const int rc = step_launch<T>(h, nsteps);
    return rc ? rc : step_collect(h, nsteps, diags, spectra);
}

// The column plans timed on the handle's own step (bastr_autotune_step): each
// plan valid for every column length of the handle runs `iters` iterations
// from the handle's state, once untimed (graph capture, caches) and once
// between HIP events on its stream, in two interleaved rounds (the faster
// round counts); the spectral state is copied aside before and back after
// every run, and nstep, time and the diagnostics of later calls are those of
// a handle that was never tuned.  Keeps the default (16, 1) unless another
// plan is at least 3 % faster.  ms[v]: milliseconds per iteration.
template <typename T>
int autotune_step_impl(bastr_handle* h, int iters, double* ms, int* best_out) {
    int rc = 0;
    if (h->prologue_pending) {
        if ((rc = (h->nd == 3 ? run_prologue3<T>(h) : run_prologue<T>(h)))) return rc;
        h->prologue_pending = false;
    }
    const int nd = h->nd, prec = h->prec;
    const int lens[2] = {h->n1, nd == 3 ? h->n2 : 0};
    const size_t bytes = h->nhalf * 2 * h->esz * (size_t)h->nmem;
    bool ok = true;
    void* snap[3] = {};
    for (int c = 0; c < nd; ++c) snap[c] = dalloc(bytes, &ok);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    int keep_p[2], keep_w[2];
    for (int i = 0; i < 2; ++i)
        if (lens[i]) keep_p[i] = k3d::colx_points(lens[i], prec), keep_w[i] = k3d::colx_width(lens[i], prec);
    const uint32_t prof = h->prof_mask;
    h->prof_mask = 0;  // the tuning runs stay out of the per-kernel timings
    auto copy_state = [&](bool save) {
        for (int c = 0; c < nd && ok; ++c)
            ok = hipMemcpyAsync(save ? snap[c] : h->U[c], save ? h->U[c] : snap[c], bytes, hipMemcpyDeviceToDevice,
                                h->stream) == hipSuccess;
    };
    copy_state(true);
    const bool saved = ok;  // the snapshots hold the state (restore only then)
    std::vector<int> st(h->nmem);
    bool nan_seen = false;
    for (int v = 0; v < 4; ++v) ms[v] = std::nan("");
    for (int round = 0; round < 2 && ok && !rc && !nan_seen; ++round) {
        for (int v = 0; v < 4 && ok && !rc && !nan_seen; ++v) {
            const int pts = (v & 1) ? 8 : 16, wid = (v & 2) ? 2 : 1;
            bool valid = true;
            for (int i = 0; i < 2; ++i)
                if (lens[i] && !k3d::colx_plan_ok(lens[i], prec, pts, wid)) valid = false;
            if (!valid) continue;
            for (int i = 0; i < 2; ++i)
                if (lens[i]) k3d::set_colx_plan(lens[i], prec, pts, wid);
            float tm = 0.f;
            for (int run = 0; run < 2 && ok && !rc; ++run) {
                copy_state(false);
                if (run == 1) ok = ok && hipEventRecord(e0, h->stream) == hipSuccess;
                if (ok) rc = step_launch<T>(h, iters);
                if (run == 1) ok = ok && hipEventRecord(e1, h->stream) == hipSuccess;
            }
            ok = ok && !rc && hipMemcpyAsync(st.data(), h->status, sizeof(int) * h->nmem, hipMemcpyDeviceToHost,
                                             h->stream) == hipSuccess;
            ok = ok && hipStreamSynchronize(h->stream) == hipSuccess && hipEventElapsedTime(&tm, e0, e1) == hipSuccess;
            for (int m = 0; m < h->nmem && ok; ++m) nan_seen = nan_seen || st[m] != INT_MAX;
            const double per = (double)tm / iters;
            if (ok && !rc && !nan_seen && (std::isnan(ms[v]) || per < ms[v])) ms[v] = per;
        }
    }
    {
        // put the state back whatever failed above (a run may have advanced it)
        bool rok = true;
        for (int c = 0; c < nd && saved; ++c)
            rok = rok && hipMemcpyAsync(h->U[c], snap[c], bytes, hipMemcpyDeviceToDevice, h->stream) == hipSuccess;
        rok = rok && hipStreamSynchronize(h->stream) == hipSuccess;
        ok = ok && rok;
    }
    h->prof_mask = prof;
    for (int c = 0; c < nd; ++c)
        if (snap[c]) {
            SyncLock lk(g_sync_mu);
            (void)hipFree(snap[c]);
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    int best = 0;
    if (ok && !rc && !nan_seen && !std::isnan(ms[0]))
        for (int v = 1; v < 4; ++v)
            if (!std::isnan(ms[v]) && ms[v] < 0.97 * ms[0] && ms[v] < ms[best]) best = v;
    const bool good = ok && !rc && !nan_seen && !std::isnan(ms[0]);
    for (int i = 0; i < 2; ++i)
        if (lens[i]) {
            if (good) k3d::set_colx_plan(lens[i], prec, (best & 1) ? 8 : 16, (best & 2) ? 2 : 1);
            else k3d::set_colx_plan(lens[i], prec, keep_p[i], keep_w[i]);
        }
    if (rc) return rc;
    if (!ok) return fail(BASTR_EDEVICE, "bastr_autotune_step: a device call failed while timing the plans");
    if (nan_seen)
        return fail(BASTR_EINVAL, "bastr_autotune_step: the field reached a NaN iteration within %d iterations "
                                  "(tune on an earlier state or with fewer iterations)", iters);
    if (std::isnan(ms[0])) return fail(BASTR_EDEVICE, "bastr_autotune_step: the default plan was not timed");
    *best_out = best;
    return 0;
}

// ---------------------------------------------------------------------------
// One graph over several handles (bastr_step_group): the iterations of every
// handle, each on its own stream forked from and joined to the first handle's,
// captured once and replayed as a single graph launch.  Entries are keyed by
// the handles' (gen, epoch) and the step count: a handle's new buffers, a
// changed closure or its destruction never match an old entry.
// ---------------------------------------------------------------------------
struct GroupGraph {
    std::vector<std::pair<uint64_t, uint64_t>> key;
    int nsteps = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;  // [0] fork, [1..nh-1] joins (capture); [nh..2nh-1] prepared, [2nh] done
    uint64_t last_use = 0;
};
std::vector<GroupGraph> g_groups;
std::mutex g_groups_mu;
uint64_t g_group_clock = 0;
constexpr size_t kMaxGroupGraphs = 4;

void group_entry_free(GroupGraph& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    for (hipEvent_t x : e.ev)
        if (x) (void)hipEventDestroy(x);
    e.exec = nullptr;
    e.ev.clear();
}

// drop every group graph that holds handle `gen`
void group_forget(uint64_t gen) {
    std::lock_guard<std::mutex> lk(g_groups_mu);
    for (size_t i = 0; i < g_groups.size();) {
        bool has = false;
        for (auto& k : g_groups[i].key) has = has || k.first == gen;
        if (has) {
            group_entry_free(g_groups[i]);
            g_groups.erase(g_groups.begin() + (long)i);
        } else {
            ++i;
        }
    }
}

template <typename T>
int group_iterations(bastr_handle* h, int nsteps) {
    int rc = 0;
    for (int it = 0; it < nsteps && !rc; ++it) rc = h->nd == 3 ? run_iteration3<T>(h, it) : run_iteration<T>(h, it);
    return rc;
}

int group_body(bastr_handle* h, int nsteps) {
    return h->prec == BASTR_PREC_F32 ? group_iterations<float>(h, nsteps) : group_iterations<double>(h, nsteps);
}

// Launch the iterations of every handle as one graph (after each handle's
// step_prepare); the caller then collects each handle on its own stream.
int group_launch(bastr_handle* const* hs, int nh, int nsteps) {
    std::vector<std::pair<uint64_t, uint64_t>> key(nh);
    for (int i = 0; i < nh; ++i) key[i] = {hs[i]->gen, hs[i]->epoch};
    std::lock_guard<std::mutex> lk(g_groups_mu);
    GroupGraph* e = nullptr;
    for (auto& x : g_groups)
        if (x.nsteps == nsteps && x.key == key) e = &x;
    if (!e) {
        if (g_groups.size() >= kMaxGroupGraphs) {  // evict the least recently used
            size_t lru = 0;
            for (size_t i = 1; i < g_groups.size(); ++i)
                if (g_groups[i].last_use < g_groups[lru].last_use) lru = i;
            group_entry_free(g_groups[lru]);
            g_groups.erase(g_groups.begin() + (long)lru);
        }
        GroupGraph n;
        n.key = key;
        n.nsteps = nsteps;
        n.ev.assign(2 * nh + 1, nullptr);
        for (auto& x : n.ev)
            if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) {
                group_entry_free(n);
                return fail(BASTR_EDEVICE, "bastr_step_group: event creation failed");
            }
        SyncLock sl(g_sync_mu);
        hipStream_t s0 = hs[0]->stream;
        hipGraph_t g = nullptr;
        int rc = 0;
        hipError_t ec = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
        if (ec == hipSuccess) {
            // every other handle's stream joins the capture through the fork event
            ec = hipEventRecord(n.ev[0], s0);
            for (int i = 1; i < nh && ec == hipSuccess; ++i) ec = hipStreamWaitEvent(hs[i]->stream, n.ev[0], 0);
            for (int i = 0; i < nh && ec == hipSuccess && !rc; ++i) rc = group_body(hs[i], nsteps);
            for (int i = 1; i < nh && ec == hipSuccess && !rc; ++i) {
                ec = hipEventRecord(n.ev[i], hs[i]->stream);
                if (ec == hipSuccess) ec = hipStreamWaitEvent(s0, n.ev[i], 0);
            }
            const hipError_t ee = hipStreamEndCapture(s0, &g);
            if (ec == hipSuccess) ec = ee;
        }
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            group_entry_free(n);
            return rc ? rc : fail(BASTR_EDEVICE, "bastr_step_group: graph capture failed: %s", hipGetErrorString(ec));
        }
        ec = hipGraphInstantiate(&n.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ec != hipSuccess) {
            n.exec = nullptr;
            group_entry_free(n);
            return fail(BASTR_EDEVICE, "bastr_step_group: graph instantiation failed: %s", hipGetErrorString(ec));
        }
        g_groups.push_back(n);
        e = &g_groups.back();
    }
    e->last_use = ++g_group_clock;
    hipStream_t s0 = hs[0]->stream;
    // each handle's prepared state (entry dealias, gates, slots) before the graph
    for (int i = 1; i < nh; ++i) {
        HIPCHK(hipEventRecord(e->ev[nh + i], hs[i]->stream));
        HIPCHK(hipStreamWaitEvent(s0, e->ev[nh + i], 0));
    }
    HIPCHK(hipGraphLaunch(e->exec, s0));
    // each handle's result copies after the graph
    HIPCHK(hipEventRecord(e->ev[2 * nh], s0));
    for (int i = 1; i < nh; ++i) HIPCHK(hipStreamWaitEvent(hs[i]->stream, e->ev[2 * nh], 0));
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Step without allocation (bastr_step_noalloc).  The step's per-call device
// scratch -- the NaN status word, the partial sums of the spectra, the
// velocity-gradient moments and the forcing energy, the reduced sums, the
// forcing factor and the per-iteration diagnostic and spectrum slots, which
// bastr_step keeps in handle buffers (the slots grow with the step count by
// hipMalloc) -- is carved out of this fixed array instead.  Each device holds
// its own copy for the life of the process; its address on a device is
// resolved once, with that device current, when the first handle is created
// there.  Calls on one device share the array, so they are serialised by a
// per-device lock (a call is synchronous and holds it until its results are
// copied out).
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_init_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // scratch's device address per device
std::mutex g_scratch_mu[kMaxDevices];     // one noalloc call per device at a time

// Resolves scratch's address on the current device (once per device).
int scratch_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return fail(BASTR_EDEVICE, "device %d beyond the %d the library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    if (g_scratch_ptr[dev]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(BASTR_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
    return 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of a one-field handle's step scratch for a call of nsteps
// iterations (the sizes of the handle's own buffers, create_impl / ensure_capacity).
struct ScratchLayout {
    size_t status, bins, tot, vg, red, en, factor, diag, spec, total;
};

ScratchLayout scratch_layout(const bastr_handle* h, int nsteps) {
    const size_t K = (size_t)h->allkmax;
    ScratchLayout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) { const size_t at = o; o = align256(o + bytes); return at; };
    L.status = take(sizeof(int));
    L.bins = take(sizeof(double) * (size_t)h->nblk_spec * (K + 1) * 2);
    L.tot = take(sizeof(double) * (size_t)h->nblk_spec * 4);
    L.vg = take(sizeof(double) * (size_t)h->nblk_vg * 8);
    L.red = take(sizeof(double) * (2 * K + 12));
    L.en = take(sizeof(double) * (size_t)h->nblk_en);
    L.factor = take(sizeof(double));
    L.diag = take(sizeof(bastr_diag) * (size_t)nsteps);
    L.spec = take(sizeof(double) * 3 * (K > 0 ? K : 1) * (size_t)nsteps);
    L.total = o;
    return L;
}

// Points the handle's step scratch at the array for one call: direct launches
// (no graph instantiation), no profiling events, slot capacity = nsteps.
struct ScratchView {
    bastr_handle* h;
    int* status;
    double *part_bins, *part_tot, *part_vg, *red, *part_en, *dfactor, *dspec;
    bastr_diag* ddiag;
    int cap;
    bool use_graph;
    uint32_t prof_mask;
    ScratchView(bastr_handle* h_, char* base, const ScratchLayout& L, int nsteps) : h(h_) {
        status = h->status; part_bins = h->part_bins; part_tot = h->part_tot; part_vg = h->part_vg;
        red = h->red; part_en = h->part_en; dfactor = h->dfactor; ddiag = h->ddiag; dspec = h->dspec;
        cap = h->cap; use_graph = h->use_graph; prof_mask = h->prof_mask;
        h->status = (int*)(base + L.status);
        h->part_bins = (double*)(base + L.bins);
        h->part_tot = (double*)(base + L.tot);
        h->part_vg = (double*)(base + L.vg);
        h->red = (double*)(base + L.red);
        h->part_en = (double*)(base + L.en);
        h->dfactor = (double*)(base + L.factor);
        h->ddiag = (bastr_diag*)(base + L.diag);
        h->dspec = (double*)(base + L.spec);
        h->cap = nsteps;
        h->use_graph = false;
        h->prof_mask = 0;
    }
    ~ScratchView() {
        h->status = status; h->part_bins = part_bins; h->part_tot = part_tot; h->part_vg = part_vg;
        h->red = red; h->part_en = part_en; h->dfactor = dfactor; h->ddiag = ddiag; h->dspec = dspec;
        h->cap = cap; h->use_graph = use_graph; h->prof_mask = prof_mask;
    }
};

__global__ void zero_doubles_kernel(double* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace

extern "C" {

const char* bastr_last_error(void) { return g_err; }

static int create_impl(const bastr_params* p, int nmem, const int* forced, bastr_handle** out) {
    if (!p || !out) return fail(BASTR_EINVAL, "bastr_create: NULL argument");
    *out = nullptr;
    const int nd = p->ka == 0 ? 2 : 3;
    if (nmem < 1 || nmem > 65535) return fail(BASTR_EINVAL, "bastr_create_ensemble: nmembers %d not in [1, 65535]", nmem);
    if (nmem > 1 && nd != 2) return fail(BASTR_EINVAL, "bastr_create_ensemble: ensembles are 2-D (ka = 0)");
    if (forced)
        for (int m = 0; m < nmem; ++m)
            if (forced[m] != 0 && forced[m] != 1)
                return fail(BASTR_EINVAL, "bastr_create_ensemble: forced[%d] = %d is not 0 or 1", m, forced[m]);
    if (nd == 2 && (!pow2_in(p->ia, 16, 4096) || !pow2_in(p->ja, 16, 4096)))
        return fail(BASTR_EINVAL, "bastr_create: ia, ja must be powers of two in [16, 4096] (got %d, %d)", p->ia, p->ja);
    if (nd == 3 && (!pow2_in(p->ia, 16, 512) || p->ja != p->ia || p->ka != p->ia))
        return fail(BASTR_EINVAL, "bastr_create: 3-D grids must be cubic, a power of two in [16, 512] (got %d %d %d)",
                    p->ia, p->ja, p->ka);
    if (p->timemethod != 1 && p->timemethod != 2)
        return fail(BASTR_EINVAL, "Undefined time method %d (1 = RK3, 2 = CN)", p->timemethod);
    if (p->forcemethod < 0 || p->forcemethod > 5 || p->forcemethod == 3 || p->forcemethod == 4)
        return fail(BASTR_EINVAL,
                    "bastr_create: forcemethod %d not supported (0, 1, 2, 5; methods 3 and 4 draw from the "
                    "reference's unseeded Fortran RNG)", p->forcemethod);
    if (p->precision != BASTR_PREC_F64 && p->precision != BASTR_PREC_F32)
        return fail(BASTR_EINVAL, "bastr_create: bad precision %d", p->precision);
    if (p->kmax < 1) return fail(BASTR_EINVAL, "bastr_create: kmax must be >= 1");
    if ((p->forcemethod == 2 || p->forcemethod == 5) && 2 * p->forcek + 2 > p->ja)
        return fail(BASTR_EINVAL, "bastr_create: forcek %d too large for ja=%d", p->forcek, p->ja);

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(BASTR_EDEVICE, "bastr_create: hipGetDevice failed");
    if (int rs = scratch_init(dev)) return rs;
    bastr_handle* h = new bastr_handle();
    static std::atomic<uint64_t> next_gen{1};
    h->gen = next_gen.fetch_add(1);
    h->dev = dev;
    h->p = *p;
    h->nd = nd;
    h->nmem = h->lw = nmem;
    h->mset.assign(nmem, 0);
    h->mnstep.assign(nmem, 0);
    h->mtime.assign(nmem, 0.0);
    h->n0 = p->ia;
    h->n1 = p->ja;
    h->n2 = nd == 3 ? p->ka : 1;
    h->prec = p->precision;
    h->esz = p->precision == BASTR_PREC_F32 ? 4 : 8;
    {
        // Graph replay where the launches' own cost shows: grids of at most
        // kGraphMaxPoints points over all members.  Larger grids launch
        // directly (measured at 2048^2 FP64: 0.515-0.520 ms per iteration
        // direct against 0.538-0.547 ms replayed; 256^2: 0.092 direct against
        // 0.087-0.095 replayed).  BASTR_GRAPH "0" / "1" forces either.
        const char* gv = std::getenv("BASTR_GRAPH");
        const double pts = (double)p->ia * p->ja * (nd == 3 ? p->ka : 1) * nmem;
        h->use_graph = gv && std::strcmp(gv, "0") == 0 ? false
                     : gv && std::strcmp(gv, "1") == 0 ? true
                     : pts <= kGraphMaxPoints;
    }
    h->allkmax = p->ia / 3 < p->kmax ? p->ia / 3 : p->kmax;
    h->nu = p->reynolds == 0.0 ? 0.0 : 1.0 / p->reynolds;
    const int cutoff = (int)std::floor((double)p->ia / 3.0);
    h->cut2 = cutoff * cutoff;
    h->nh = h->n0 / 2 + 1;
    const int cb = col_block_width(h->n1, h->prec);
    h->hp = (h->nh + cb - 1) / cb * cb;
    {
        const char* pv = std::getenv("BASTR_PRUNE");  // "0": the step's passes cover every mode
        h->prune = !(pv && std::strcmp(pv, "0") == 0);
        h->cut = cutoff;
        const int c16 = (cutoff + 15) / 16 * 16;  // a multiple of every column block width
        h->cols = c16 <= h->hp ? c16 : h->hp;
        if (2 * cutoff - 1 < h->n1) {
            h->rows_n = 2 * cutoff - 1;
            h->rows_at = cutoff;
            h->rows_skip = h->n1 - (2 * cutoff - 1);
        } else {
            h->rows_n = h->n1;
            h->rows_at = h->n1;
            h->rows_skip = 0;
        }
    }
    h->npts = (size_t)h->n0 * h->n1 * h->n2;
    h->nhalf = (size_t)h->hp * h->n1 * h->n2;
    {
        // "1": the first bastr_step call tunes the column plans of this handle's
        // lengths on its own step (bastr_autotune_step), once per length and
        // precision in the process
        const char* tv = std::getenv("BASTR_TUNE");
        h->tune_pending = tv && std::strcmp(tv, "1") == 0;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(BASTR_EDEVICE, "bastr_create: hipStreamCreate failed");
    }
    if (nd == 2) {
        // "1": the two-stream order (measured: graph replay at 2048^2 FP64 0.575 -> 0.553 ms
        // per iteration; direct launches and small grids slower, so not the default)
        const char* ov = std::getenv("BASTR_OVERLAP");
        h->overlap = ov && std::strcmp(ov, "1") == 0;
    }
    if (h->overlap) {
        bool oks = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) == hipSuccess;
        for (int i = 0; i < 3 && oks; ++i)
            oks = hipEventCreateWithFlags(&h->fork[i], hipEventDisableTiming) == hipSuccess;
        if (!oks) {
            bastr_destroy(h);
            return fail(BASTR_EDEVICE, "bastr_create: second stream or its events failed");
        }
    }
    bool ok = true;
    const size_t cbytes = h->nhalf * 2 * h->esz;
    if (nd == 2) {
        const size_t M = (size_t)nmem;  // member-major stacks of every field
        for (int c = 0; c < 2; ++c) {
            h->U[c] = dalloc(cbytes * M, &ok);
            h->UB[c] = dalloc(cbytes * M, &ok);
            h->KA[c] = dalloc(cbytes * M, &ok);
            h->NX[c] = dalloc(cbytes * M, &ok);
            h->ureal[c] = dalloc(h->npts * h->esz * M, &ok);
        }
        for (int q = 0; q < 7; ++q) h->W[q] = dalloc(cbytes * M, &ok);
    } else {
        for (int c = 0; c < 3; ++c) {
            h->U[c] = dalloc(cbytes, &ok);
            h->UB[c] = dalloc(cbytes, &ok);
            h->KA[c] = dalloc(cbytes, &ok);
        }
        for (int q = 0; q < 10; ++q) h->Z[q] = dalloc(cbytes, &ok);
        for (int q = 0; q < 13; ++q) h->W[q] = dalloc(cbytes, &ok);
        for (int c = 0; c < 3; ++c) {
            h->NX[c] = h->Z[c];
            h->NY[c] = h->Z[3 + c];
            h->ureal[c] = h->W[c];  // a real n0*n1*n2 field fits in one half-spectrum buffer
        }
    }
    h->dstage = (double*)dalloc(h->npts * sizeof(double), &ok);
    const int nbx = (h->nh + 63) / 64;
    const long ntask = (long)h->n1 * h->n2 * nbx;
    if (nd == 2) {
        // 64-mode tasks the (dealias-limited) spectra pass sums; each block
        // writes (allkmax+1)*2 partial bins, so a block takes at least
        // (allkmax+1)/8 tasks: its partials stay under a quarter of its reads
        const long ntask2 = h->prune ? (long)h->rows_n * ((h->cut + 63) / 64) : ntask;
        // (small FP64 grids: 4 tasks per block, so the latency-bound pass
        // spreads over more CUs; 256^2 FP64 0.0615 -> 0.0603 ms per iteration,
        // while the 96-member FP32 ensemble measured 1.641 -> 1.679 ms with it.
        // The rule ignores the member count: an ensemble member sums its
        // spectra in the order of the one-field run of its grid)
        const bool small = ntask2 <= 1024 && h->prec == BASTR_PREC_F64;
        const long per = small ? 4L : std::max(8L, (long)(h->allkmax + 1) / 8);
        h->nblk_spec = (int)std::max(1L, std::min(512L, ntask2 / per));
    } else {
        h->nblk_spec = (int)std::max(1L, std::min(512L, ntask / 8));
    }
    {
        const int L = vg_block_rows(h->n0, h->prec, nd);
        h->nblk_vg = (h->n1 * h->n2 + L - 1) / L;
    }
    const size_t M = (size_t)nmem;
    h->part_bins = (double*)dalloc(sizeof(double) * (size_t)h->nblk_spec * (h->allkmax + 1) * 2 * M, &ok);
    h->part_tot = (double*)dalloc(sizeof(double) * (size_t)h->nblk_spec * 4 * M, &ok);
    h->part_vg = (double*)dalloc(sizeof(double) * (size_t)h->nblk_vg * 8 * M, &ok);
    h->status = (int*)dalloc(sizeof(int) * M, &ok);
    h->red = (double*)dalloc(sizeof(double) * (2 * (size_t)h->allkmax + 12) * M, &ok);
    h->nblk_en = 512;
    h->part_en = (double*)dalloc(sizeof(double) * h->nblk_en * M, &ok);
    h->dfactor = (double*)dalloc(sizeof(double) * M, &ok);
    if (forced) h->mforce = (int*)dalloc(sizeof(int) * M, &ok);
    if (!ok) {
        bastr_destroy(h);
        return fail(BASTR_ENOMEM, "bastr_create: device allocation failed");
    }
    int rc = h->prec == BASTR_PREC_F32 ? upload_twiddles<float>(h->n0, &h->tw0) : upload_twiddles<double>(h->n0, &h->tw0);
    if (!rc) rc = h->prec == BASTR_PREC_F32 ? upload_twiddles<float>(h->n1, &h->tw1) : upload_twiddles<double>(h->n1, &h->tw1);
    if (!rc && h->mforce) {
        SyncLock lk(g_sync_mu);
        if (hipMemcpy(h->mforce, forced, sizeof(int) * nmem, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(BASTR_EDEVICE, "bastr_create_ensemble: copy of the forcing flags failed");
    }
    if (rc) {
        bastr_destroy(h);
        return rc;
    }
    // static shell data: Ecount and kn = mean |k| per shell (solution.F90:974-996)
    const int K = h->allkmax;
    h->kn.assign(K + 1, 0.0);
    h->counts.assign(K + 1, 0);
    for (int k = 0; k < h->n2; ++k)
        for (int j = 0; j < h->n1; ++j)
            for (int i = 0; i < h->n0; ++i) {
                const int k1 = i <= h->n0 / 2 ? i : i - h->n0;
                const int k2 = j <= h->n1 / 2 ? j : j - h->n1;
                const int k3 = nd == 3 ? (k <= h->n2 / 2 ? k : k - h->n2) : 0;
                const double kk = std::sqrt((double)k1 * k1 + (double)k2 * k2 + (double)k3 * k3);
                const long sh = std::lround(kk);
                if (sh <= K) {
                    h->counts[sh] += 1;
                    h->kn[sh] += kk;
                }
            }
    for (int b = 1; b <= K; ++b) h->kn[b] = h->kn[b] / (double)h->counts[b];
    *out = h;
    return BASTR_OK;
}

int bastr_create(const bastr_params* p, bastr_handle** out) { return create_impl(p, 1, nullptr, out); }

int bastr_create_ensemble(const bastr_params* p, int32_t nmembers, const int32_t* forced, bastr_handle** out) {
    return create_impl(p, nmembers, forced, out);
}

int bastr_nmembers(const bastr_handle* h) { return h ? h->nmem : -1; }

int bastr_destroy(bastr_handle* h) {
    if (!h) return BASTR_OK;
    group_forget(h->gen);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->side) (void)hipStreamSynchronize(h->side);
    for (hipEvent_t e : h->fork)
        if (e) (void)hipEventDestroy(e);
    if (h->side) (void)hipStreamDestroy(h->side);
    void* bufs[] = {h->dstage, h->tw0, h->tw1, h->tw0d, h->tw1d, h->cbuf, h->part_bins, h->part_tot,
                    h->part_vg, h->red, h->part_en, h->dfactor, h->status, h->ddiag, h->dspec, h->mforce};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (void* b : {h->CW[0], h->CW[1], h->CY[0], h->CY[1], h->CY[2], h->CY[3], h->CR[0], h->CR[1], h->CR[2],
                    h->CR[3], h->CF[0], h->CF[1], h->CX[0], h->CX[1]})
        if (b) (void)hipFree(b);
    for (int c = 0; c < 3; ++c) {
        if (h->U[c]) (void)hipFree(h->U[c]);
        if (h->UB[c]) (void)hipFree(h->UB[c]);
        if (h->KA[c]) (void)hipFree(h->KA[c]);
        if (h->nd == 2) {  // 3-D NX/ureal alias other buffers
            if (h->NX[c]) (void)hipFree(h->NX[c]);
            if (h->ureal[c]) (void)hipFree(h->ureal[c]);
        }
    }
    for (int q = 0; q < 13; ++q)
        if (h->W[q]) (void)hipFree(h->W[q]);
    for (int q = 0; q < 10; ++q)
        if (h->Z[q]) (void)hipFree(h->Z[q]);
    for (int q = 0; q < 30; ++q)
        if (h->pp[q]) (void)hipFree(h->pp[q]);
    if (h->pp_status) (void)hipFree(h->pp_status);
    if (h->pp_part) (void)hipFree(h->pp_part);
    drop_graph(h);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return BASTR_OK;
}

int bastr_allkmax(const bastr_handle* h) { return h ? h->allkmax : -1; }

int bastr_set_fields(bastr_handle* h, const double* u1, const double* u2, const double* u3) {
    if (!h || !u1 || !u2 || (h && h->nd == 3 && !u3)) return fail(BASTR_EINVAL, "bastr_set_fields: NULL argument");
    if (h->lw > 1) return fail(BASTR_EINVAL, "bastr_set_fields: an ensemble handle takes the bastr_*_member entry points");
    const double* src[3] = {u1, u2, u3};
    SyncLock lk(g_sync_mu);
    for (int c = 0; c < h->nd; ++c) {
        if (h->prec == BASTR_PREC_F64) {
            HIPCHK(hipMemcpy(h->ureal[c], src[c], sizeof(double) * h->npts, hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(h->dstage, src[c], sizeof(double) * h->npts, hipMemcpyHostToDevice));
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3(1024), dim3(256), 0, h->stream, h->dstage,
                               (float*)h->ureal[c], h->npts);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(h->stream));
        }
    }
    h->prologue_pending = true;
    h->have_fields = true;
    return BASTR_OK;
}

static int prologue_any(bastr_handle* h) {
    if (h->nd == 3) return h->prec == BASTR_PREC_F32 ? run_prologue3<float>(h) : run_prologue3<double>(h);
    return h->prec == BASTR_PREC_F32 ? run_prologue<float>(h) : run_prologue<double>(h);
}

int bastr_get_fields(bastr_handle* h, double* u1, double* u2, double* u3) {
    if (!h || !u1 || !u2 || (h && h->nd == 3 && !u3)) return fail(BASTR_EINVAL, "bastr_get_fields: NULL argument");
    if (h->lw > 1) return fail(BASTR_EINVAL, "bastr_get_fields: an ensemble handle takes the bastr_*_member entry points");
    if (!h->have_fields) return fail(BASTR_EINVAL, "bastr_get_fields: fields not set");
    int rc;
    if (h->prologue_pending) {
        if ((rc = prologue_any(h))) return rc;
        h->prologue_pending = false;
    }
    if (h->nd == 3) rc = h->prec == BASTR_PREC_F32 ? run_to_physical3<float>(h) : run_to_physical3<double>(h);
    else rc = h->prec == BASTR_PREC_F32 ? run_to_physical<float>(h) : run_to_physical<double>(h);
    if (rc) return rc;
    double* dst[3] = {u1, u2, u3};
    for (int c = 0; c < h->nd; ++c) {
        if (h->prec == BASTR_PREC_F64) {
            HIPCHK(hipMemcpyAsync(dst[c], h->ureal[c], sizeof(double) * h->npts, hipMemcpyDeviceToHost, h->stream));
        } else {
            hipLaunchKernelGGL((convert_kernel<float, double>), dim3(1024), dim3(256), 0, h->stream,
                               (const float*)h->ureal[c], h->dstage, h->npts);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(dst[c], h->dstage, sizeof(double) * h->npts, hipMemcpyDeviceToHost, h->stream));
        }
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return BASTR_OK;
}

}  // extern "C"

template <typename T>
static int init_random_phase_impl(bastr_handle* h, double k0, uint64_t seed, double energy) {
    const Gate open{h->status, 0, 0};
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    // 1. noise in the first physical field, zeros in the others
    LAUNCHCHK(rp::launch_noise<T>(h->ureal[0], h->npts, seed, h->stream));
    for (int c = 1; c < h->nd; ++c) LAUNCHCHK(rp::launch_zero<T>(h->ureal[c], h->npts, h->stream));
    // 2. its normalised spectrum into U[0]
    if (h->nd == 2) {
        RowArgs<T> r{};
        r.g = geo<T>(h);
        r.gate = open;
        r.rin[0] = (const T*)h->ureal[0];
        r.rin[1] = (const T*)h->ureal[1];
        r.hout[0] = C<T>(h->NX[0]);
        r.hout[1] = C<T>(h->NX[1]);
        r.tw = C<T>(h->tw0);
        LAUNCHCHK(launch_row<T>(h->n0, R_FWD_PACK, r, h->n1, h->stream));
        void* src[1] = {h->NX[0]};
        void* dst[1] = {h->U[0]};
        int rc = run_yfwd2<T>(h, open, src, dst, 1, 1.0 / ((double)h->n0 * (double)h->n1), h->hp, h->tw1);
        if (rc) return rc;
    } else {
        k3d::Row3Args<T> r{};
        r.g = geo3<T>(h);
        r.gate = open;
        for (int q = 0; q < 3; ++q) {
            r.rin[q] = (const T*)h->ureal[q];
            r.hout[q] = C<T>(h->NX[q]);
        }
        r.tw = C<T>(h->tw0);
        LAUNCHCHK(k3d::launch_row3<T>(h->n0, k3d::R3_FWD_PACK, r, h->n1 * h->n2, h->stream));
        int rc = run_yfwd3<T>(h, open, false);
        if (rc) return rc;
        void* src[1] = {h->NY[0]};
        void* dst[1] = {h->U[0]};
        if ((rc = run_zfwd3<T>(h, open, src, dst, 1))) return rc;
    }
    // 3. the spectrum law with the noise phases
    rp::RpArgs b{};
    for (int c = 0; c < 3; ++c) b.U[c] = h->U[c];
    b.nd = h->nd;
    b.n0 = h->n0;
    b.n1 = h->n1;
    b.n2 = h->n2;
    b.nh = h->nh;
    b.hp = h->hp;
    b.k0 = k0;
    b.law = h->law;
    LAUNCHCHK(rp::launch_build<T>(b, h->stream));
    // 4. rescale to <sum u_i^2> = energy (Parseval over the half spectrum)
    std::vector<double> part(h->nblk_en);
    if (h->nd == 2) {
        EnergyArgs<T> e{};
        e.g = geo<T>(h);
        e.gate = open;
        e.U[0] = C<T>(h->U[0]);
        e.U[1] = C<T>(h->U[1]);
        e.part_en = h->part_en;
        LAUNCHCHK(launch_energy<T>(e, h->nblk_en, h->stream));
    } else {
        k3d::Energy3Args<T> e{};
        e.g = geo3<T>(h);
        e.gate = open;
        for (int q = 0; q < 3; ++q) e.U[q] = C<T>(h->U[q]);
        e.part_en = h->part_en;
        LAUNCHCHK(k3d::launch_energy3<T>(e, h->nblk_en, h->stream));
    }
    HIPCHK(hipMemcpyAsync(part.data(), h->part_en, sizeof(double) * h->nblk_en, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double e = 0;
    for (double x : part) e += x;
    if (!(e > 0)) return fail(BASTR_EINVAL, "bastr_init_random_phase: the generated field has zero energy");
    const double factor = std::sqrt(energy / e);
    HIPCHK(hipMemcpyAsync(h->dfactor, &factor, sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (h->nd == 2) {
        ScaleArgs<T> sc{};
        sc.g = geo<T>(h);
        sc.gate = open;
        sc.U[0] = C<T>(h->U[0]);
        sc.U[1] = C<T>(h->U[1]);
        sc.factor = h->dfactor;
        LAUNCHCHK(launch_scale<T>(sc, h->stream));
    } else {
        k3d::Scale3Args<T> sc{};
        sc.g = geo3<T>(h);
        sc.gate = open;
        for (int q = 0; q < 3; ++q) sc.U[q] = C<T>(h->U[q]);
        sc.factor = h->dfactor;
        LAUNCHCHK(k3d::launch_scale3<T>(sc, h->stream));
    }
    // 5. physical field u = real(ifft(u_hat)): the file hitgen would write
    int rc = h->nd == 3 ? run_to_physical3<T>(h) : run_to_physical<T>(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

extern "C" {

int bastr_init_random_phase(bastr_handle* h, double k0, uint64_t seed, double energy) {
    if (!h) return fail(BASTR_EINVAL, "bastr_init_random_phase: NULL handle");
    if (h->lw > 1) return fail(BASTR_EINVAL, "bastr_init_random_phase: an ensemble handle takes the bastr_*_member entry points");
    if (!(k0 > 0) || !(energy > 0)) return fail(BASTR_EINVAL, "bastr_init_random_phase: k0 and energy must be > 0");
    int rc = h->prec == BASTR_PREC_F32 ? init_random_phase_impl<float>(h, k0, seed, energy)
                                       : init_random_phase_impl<double>(h, k0, seed, energy);
    if (rc) return rc;
    h->prologue_pending = true;
    h->have_fields = true;
    h->nstep = 0;
    h->time = 0.0;
    return BASTR_OK;
}

int bastr_set_spectrum_law(bastr_handle* h, int32_t law) {
    if (!h) return fail(BASTR_EINVAL, "bastr_set_spectrum_law: NULL handle");
    if (law != BASTR_LAW_HITGEN && law != BASTR_LAW_PURE)
        return fail(BASTR_EINVAL, "bastr_set_spectrum_law: law %d is neither BASTR_LAW_HITGEN nor BASTR_LAW_PURE", law);
    h->law = law;
    return BASTR_OK;
}

int bastr_set_time(bastr_handle* h, int64_t nstep, double time) {
    if (!h) return fail(BASTR_EINVAL, "bastr_set_time: NULL handle");
    h->nstep = nstep;
    h->time = time;
    for (int m = 0; m < h->nmem; ++m) {
        h->mnstep[m] = nstep;
        h->mtime[m] = time;
    }
    return BASTR_OK;
}

int bastr_get_time(const bastr_handle* h, int64_t* nstep, double* time) {
    if (!h) return fail(BASTR_EINVAL, "bastr_get_time: NULL handle");
    if (nstep) *nstep = h->nmem > 1 ? h->mnstep[0] : h->nstep;
    if (time) *time = h->nmem > 1 ? h->mtime[0] : h->time;
    return BASTR_OK;
}

// One member of an ensemble handle through the one-field entry points: the
// handle's buffers point at the member for the call, and the member's spectral
// state is formed at once (the entry dealias of mainloop2D :41-52).
static int member_check(bastr_handle* h, int member, const char* what) {
    if (!h) return fail(BASTR_EINVAL, "%s: NULL handle", what);
    if (member < 0 || member >= h->nmem)
        return fail(BASTR_EINVAL, "%s: member %d not in [0, %d)", what, member, h->nmem);
    return 0;
}

// have_fields of an ensemble is "every member is set"; the one-field calls a
// member operation makes set it on their own, so it is recomputed after each
// member operation, and a failed one leaves that member unset (its buffers may
// hold part of the new field)
static void member_done(bastr_handle* h, int member, bool ok = true) {
    h->mset[member] = ok ? 1 : 0;
    h->have_fields = std::all_of(h->mset.begin(), h->mset.end(), [](char c) { return c != 0; });
}

int bastr_set_member_fields(bastr_handle* h, int32_t member, const double* u1, const double* u2, const double* u3) {
    int rc = member_check(h, member, "bastr_set_member_fields");
    if (rc) return rc;
    if (h->nmem == 1) return bastr_set_fields(h, u1, u2, u3);
    {
        MemberView v(h, member);
        rc = bastr_set_fields(h, u1, u2, u3);
        if (!rc) rc = prologue_any(h);
        h->prologue_pending = false;
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess)
            rc = fail(BASTR_EDEVICE, "bastr_set_member_fields: stream synchronisation failed");
    }
    member_done(h, member, rc == 0);
    if (rc) return rc;
    return BASTR_OK;
}

int bastr_get_member_fields(bastr_handle* h, int32_t member, double* u1, double* u2, double* u3) {
    int rc = member_check(h, member, "bastr_get_member_fields");
    if (rc) return rc;
    if (h->nmem == 1) return bastr_get_fields(h, u1, u2, u3);
    if (!h->mset[member]) return fail(BASTR_EINVAL, "bastr_get_member_fields: member %d not set", member);
    MemberView v(h, member);
    const bool hf = h->have_fields;
    h->have_fields = true;
    rc = bastr_get_fields(h, u1, u2, u3);
    h->have_fields = hf;
    return rc;
}

int bastr_init_random_phase_member(bastr_handle* h, int32_t member, double k0, uint64_t seed, double energy) {
    int rc = member_check(h, member, "bastr_init_random_phase_member");
    if (rc) return rc;
    if (h->nmem == 1) return bastr_init_random_phase(h, k0, seed, energy);
    {
        MemberView v(h, member);
        rc = bastr_init_random_phase(h, k0, seed, energy);
        if (!rc) rc = prologue_any(h);
        h->prologue_pending = false;
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess)
            rc = fail(BASTR_EDEVICE, "bastr_init_random_phase_member: stream synchronisation failed");
    }
    member_done(h, member, rc == 0);
    if (rc) return rc;
    h->mnstep[member] = 0;
    h->mtime[member] = 0.0;
    return BASTR_OK;
}

int bastr_get_member_time(const bastr_handle* h, int32_t member, int64_t* nstep, double* time) {
    if (!h) return fail(BASTR_EINVAL, "bastr_get_member_time: NULL handle");
    if (member < 0 || member >= h->nmem)
        return fail(BASTR_EINVAL, "bastr_get_member_time: member %d not in [0, %d)", member, h->nmem);
    if (h->nmem == 1) return bastr_get_time(h, nstep, time);
    if (nstep) *nstep = h->mnstep[member];
    if (time) *time = h->mtime[member];
    return BASTR_OK;
}

int bastr_step(bastr_handle* h, int32_t nsteps, bastr_diag* diags, double* spectra) {
    if (!h) return fail(BASTR_EINVAL, "bastr_step: NULL handle");
    if (nsteps < 0) return fail(BASTR_EINVAL, "bastr_step: nsteps < 0");
    if (!h->have_fields) return fail(BASTR_EINVAL, "bastr_step: fields not set");
    if (nsteps == 0) return BASTR_OK;
    if (h->tune_pending) {
        // BASTR_TUNE=1: once per column length and precision in the process; a
        // failed tuning (a NaN within its iterations) leaves the plans as they were
        static std::mutex tuned_mu;
        static std::vector<int> tuned;
        h->tune_pending = false;
        std::lock_guard<std::mutex> tl(tuned_mu);
        const int key = h->n1 * 2 + h->prec;
        if (std::find(tuned.begin(), tuned.end(), key) == tuned.end()) {
            int32_t pts, wid;
            double ms[4];
            if (bastr_autotune_step(h, h->nd == 3 ? 2 : 8, &pts, &wid, ms) == BASTR_OK) tuned.push_back(key);
        }
    }
    return h->prec == BASTR_PREC_F32 ? step_impl<float>(h, nsteps, diags, spectra)
                                     : step_impl<double>(h, nsteps, diags, spectra);
}

int bastr_step_noalloc(bastr_handle* h, int32_t nsteps, bastr_diag* diags, double* spectra) {
    if (!h) return fail(BASTR_EINVAL, "bastr_step_noalloc: NULL handle");
    if (nsteps < 0) return fail(BASTR_EINVAL, "bastr_step_noalloc: nsteps < 0");
    if (h->nmem > 1) return fail(BASTR_EINVAL, "bastr_step_noalloc: ensemble handles take bastr_step");
    if (!h->have_fields) return fail(BASTR_EINVAL, "bastr_step_noalloc: fields not set");
    if (nsteps == 0) return BASTR_OK;
    void* base = g_scratch_ptr[h->dev];
    if (!base) return fail(BASTR_EDEVICE, "bastr_step_noalloc: scratch address of device %d not resolved", h->dev);
    const ScratchLayout L = scratch_layout(h, nsteps);
    if (L.total > kScratchBytes)
        return fail(BASTR_EINVAL,
                    "bastr_step_noalloc: %d iterations need %zu bytes of scratch, more than its %zu "
                    "(call with fewer iterations)", nsteps, L.total, kScratchBytes);
    std::lock_guard<std::mutex> lk(g_scratch_mu[h->dev]);
    int rc;
    {
        ScratchView v(h, (char*)base, L, nsteps);
        // the spectrum slots start at zero, as ensure_capacity leaves them
        const size_t nspec = 3 * (size_t)(h->allkmax > 0 ? h->allkmax : 1) * nsteps;
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(64), dim3(256), 0, h->stream, h->dspec, nspec);
        rc = hipGetLastError() == hipSuccess ? 0 : fail(BASTR_EDEVICE, "bastr_step_noalloc: zero launch failed");
        if (!rc) rc = h->prec == BASTR_PREC_F32 ? step_launch<float>(h, nsteps) : step_launch<double>(h, nsteps);
        if (!rc) rc = step_collect(h, nsteps, diags, spectra);
        if (rc < 0) (void)hipStreamSynchronize(h->stream);  // nothing still reads the array
    }
    return rc;
}

int bastr_step_group(bastr_handle* const* hs, int32_t nh, int32_t nsteps, bastr_diag* const* diags,
                     double* const* spectra) {
    if (!hs || nh < 1) return fail(BASTR_EINVAL, "bastr_step_group: no handles");
    if (nsteps < 0) return fail(BASTR_EINVAL, "bastr_step_group: nsteps < 0");
    for (int i = 0; i < nh; ++i) {
        if (!hs[i]) return fail(BASTR_EINVAL, "bastr_step_group: NULL handle %d", i);
        if (!hs[i]->have_fields) return fail(BASTR_EINVAL, "bastr_step_group: fields of handle %d not set", i);
        for (int j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(BASTR_EINVAL, "bastr_step_group: handle %d given twice", i);
    }
    if (nsteps == 0) return BASTR_OK;
    bool one_graph = nh > 1;
    for (int i = 0; i < nh; ++i) one_graph = one_graph && graph_ok(hs[i], nsteps) && hs[i]->dev == hs[0]->dev;
    if (one_graph) {
        // every handle's iterations in one captured graph, replayed by one launch
        for (int i = 0; i < nh; ++i) {
            bastr_handle* h = hs[i];
            const int rc = h->prec == BASTR_PREC_F32 ? step_prepare<float>(h, nsteps) : step_prepare<double>(h, nsteps);
            if (rc) {
                for (int j = 0; j < nh; ++j) (void)hipStreamSynchronize(hs[j]->stream);
                return rc;
            }
        }
        const int rc = group_launch(hs, nh, nsteps);
        if (rc) {
            for (int j = 0; j < nh; ++j) (void)hipStreamSynchronize(hs[j]->stream);
            return rc;
        }
    } else {
        // every handle's launches go to its own stream first, so they overlap on
        // the device; then each is waited for and its results collected
        for (int i = 0; i < nh; ++i) {
            bastr_handle* h = hs[i];
            const int rc = h->prec == BASTR_PREC_F32 ? step_launch<float>(h, nsteps) : step_launch<double>(h, nsteps);
            if (rc) {
                for (int j = 0; j < i; ++j) (void)hipStreamSynchronize(hs[j]->stream);
                return rc;
            }
        }
    }
    int ret = BASTR_OK;
    for (int i = 0; i < nh; ++i) {
        const int rc = step_collect(hs[i], nsteps, diags ? diags[i] : nullptr, spectra ? spectra[i] : nullptr);
        if (rc < 0) {
            for (int j = i + 1; j < nh; ++j) (void)hipStreamSynchronize(hs[j]->stream);
            return rc;
        }
        if (rc == BASTR_STOP_NAN) ret = BASTR_STOP_NAN;
    }
    return ret;
}

int bastr_autotune(int32_t n, int32_t precision, int32_t ncols, int32_t nplanes, int32_t* points, int32_t* width,
                   double* ms) {
    if (!pow2_in(n, 16, 4096)) return fail(BASTR_EINVAL, "bastr_autotune: n = %d is not a power of two in [16, 4096]", n);
    if (precision != BASTR_PREC_F64 && precision != BASTR_PREC_F32)
        return fail(BASTR_EINVAL, "bastr_autotune: bad precision %d", precision);
    if (ncols < 16 || ncols % 16 || ncols > 4096 || nplanes < 1 || (size_t)ncols * n * nplanes > ((size_t)1 << 28))
        return fail(BASTR_EINVAL, "bastr_autotune: ncols must be a multiple of 16 in [16, 4096] and "
                                  "ncols * n * nplanes at most 2^28 (got %d, %d)", ncols, nplanes);
    double t[4] = {std::nan(""), std::nan(""), std::nan(""), std::nan("")};
    const int rc = precision == BASTR_PREC_F32 ? time_plans<float>(n, ncols, nplanes, t)
                                               : time_plans<double>(n, ncols, nplanes, t);
    if (rc) return rc;
    // the default plan unless another is at least 3 % faster (box-to-box spread
    // of these timings is a few per cent; a smaller gain is not a gain)
    int best = 0;
    for (int v = 1; v < 4; ++v)
        if (!std::isnan(t[v]) && t[v] < 0.97 * t[0] && t[v] < t[best]) best = v;
    const int pts = (best & 1) ? 8 : 16, wid = (best & 2) ? 2 : 1;
    k3d::set_colx_plan(n, precision, pts, wid);
    if (points) *points = pts;
    if (width) *width = wid;
    if (ms)
        for (int v = 0; v < 4; ++v) ms[v] = t[v];
    return BASTR_OK;
}

int bastr_autotune_step(bastr_handle* h, int32_t iters, int32_t* points, int32_t* width, double* ms) {
    if (!h) return fail(BASTR_EINVAL, "bastr_autotune_step: NULL handle");
    if (!h->have_fields) return fail(BASTR_EINVAL, "bastr_autotune_step: fields not set");
    if (iters < 1 || iters > kMaxGraphSteps)
        return fail(BASTR_EINVAL, "bastr_autotune_step: iters = %d not in [1, %d]", iters, kMaxGraphSteps);
    double t[4];
    int best = 0;
    const int rc = h->prec == BASTR_PREC_F32 ? autotune_step_impl<float>(h, iters, t, &best)
                                             : autotune_step_impl<double>(h, iters, t, &best);
    if (rc) return rc;
    if (points) *points = (best & 1) ? 8 : 16;
    if (width) *width = (best & 2) ? 2 : 1;
    if (ms)
        for (int v = 0; v < 4; ++v) ms[v] = t[v];
    return BASTR_OK;
}

int bastr_tune_set(int32_t n, int32_t precision, int32_t points, int32_t width) {
    if (!pow2_in(n, 16, 4096) || (precision != BASTR_PREC_F64 && precision != BASTR_PREC_F32))
        return fail(BASTR_EINVAL, "bastr_tune_set: bad length or precision");
    if (!k3d::colx_plan_ok(n, precision, points, width))
        return fail(BASTR_EINVAL, "bastr_tune_set: no plan of %d points per thread at width x%d for n = %d in this "
                                  "precision", points, width, n);
    k3d::set_colx_plan(n, precision, points, width);
    return BASTR_OK;
}

int bastr_tune_get(int32_t n, int32_t precision, int32_t* points, int32_t* width) {
    if (points) *points = k3d::colx_points(n, precision);
    if (width) *width = k3d::colx_width(n, precision);
    return BASTR_OK;
}

int bastr_shell_counts(const bastr_handle* h, int64_t* counts) {
    if (!h || !counts) return fail(BASTR_EINVAL, "bastr_shell_counts: NULL argument");
    for (int b = 1; b <= h->allkmax; ++b) counts[b - 1] = h->counts[b];
    return BASTR_OK;
}

static int fft_c2c3(bastr_handle* h, bool inverse, const cx<double>* tw) {
    const Gate open{h->status, 0, 0};
    cx<double>* z = (cx<double>*)h->cbuf;
    Geo<double> g = geo<double>(h);
    g.n1 = h->n1 * h->n2;  // rows of the x pass
    RowArgs<double> r{};
    r.g = g;
    r.gate = open;
    r.cin = z;
    r.cout = z;
    r.tw = tw;
    k3d::ColXArgs<double> y{};
    y.gate = open;
    y.in[0] = z;
    y.out[0] = z;
    y.job[0] = job1(0, k3d::M_ONE, 0);
    y.njob = 1;
    y.pitch = h->n0;
    y.ncols = h->n0;
    y.plane_stride = (long)h->n0 * h->n1;
    y.nplanes = h->n2;
    y.scale = 1.0;
    y.tw = tw;
    k3d::ColXArgs<double> zc = y;
    zc.pitch = h->n0 * h->n1;
    zc.ncols = h->n0 * h->n1;
    zc.plane_stride = 0;
    zc.nplanes = 1;
    if (!inverse) {
        LAUNCHCHK(launch_row<double>(h->n0, R_C2C_FWD, r, h->n1 * h->n2, h->stream));
        LAUNCHCHK(k3d::launch_colx<double>(h->n1, false, y, h->stream));
        zc.scale = 1.0 / ((double)h->n0 * h->n1 * h->n2);  // fft3d normalisation (fftwlink.F90:347)
        LAUNCHCHK(k3d::launch_colx<double>(h->n2, false, zc, h->stream));
    } else {
        LAUNCHCHK(k3d::launch_colx<double>(h->n2, true, zc, h->stream));
        LAUNCHCHK(k3d::launch_colx<double>(h->n1, true, y, h->stream));
        LAUNCHCHK(launch_row<double>(h->n0, R_C2C_INV, r, h->n1 * h->n2, h->stream));
    }
    return 0;
}

static int fft_c2c(bastr_handle* h, double* inout, bool inverse) {
    if (h && h->nmem > 1) return fail(BASTR_EINVAL, "bastr_fft_*: not available on an ensemble handle");
    SyncLock lk(g_sync_mu);
    if (!h || !inout) return fail(BASTR_EINVAL, "bastr_fft: NULL argument");
    const size_t bytes = sizeof(double) * 2 * h->npts;
    if (!h->cbuf) HIPCHK(hipMalloc(&h->cbuf, bytes));
    const cx<double>* tw0 = (const cx<double>*)h->tw0;
    const cx<double>* tw1 = (const cx<double>*)h->tw1;
    if (h->prec != BASTR_PREC_F64) {
        if (!h->tw0d) {
            int rc = upload_twiddles<double>(h->n0, (void**)&h->tw0d);
            if (!rc) rc = upload_twiddles<double>(h->n1, (void**)&h->tw1d);
            if (rc) return rc;
        }
        tw0 = (const cx<double>*)h->tw0d;
        tw1 = (const cx<double>*)h->tw1d;
    }
    HIPCHK(hipMemcpyAsync(h->cbuf, inout, bytes, hipMemcpyHostToDevice, h->stream));
    if (reset_status(h)) return fail(BASTR_EDEVICE, "status reset launch failed");
    const Gate open{h->status, 0, 0};
    Geo<double> g = geo<double>(h);
    cx<double>* z = (cx<double>*)h->cbuf;
    if (h->nd == 3) {
        int rc = fft_c2c3(h, inverse, tw0);
        if (rc) return rc;
    } else if (!inverse) {
        RowArgs<double> r{};
        r.g = g;
        r.gate = open;
        r.cin = z;
        r.cout = z;
        r.tw = tw0;
        LAUNCHCHK(launch_row<double>(h->n0, R_C2C_FWD, r, h->n1, h->stream));
        if (col_block_width(h->n1, BASTR_PREC_F64) > h->n0) return fail(BASTR_EINVAL, "bastr_fft: ia too small for ja");
        void* zz[1] = {z};
        const int rc = run_yfwd2<double>(h, open, zz, zz, 1, 1.0 / ((double)h->n0 * (double)h->n1), h->n0, tw1);
        if (rc) return rc;
    } else {
        k3d::ColXArgs<double> c{};
        c.gate = open;
        c.in[0] = z;
        c.out[0] = z;
        c.job[0] = job1(0, k3d::M_ONE, 0);
        c.njob = 1;
        c.pitch = h->n0;
        c.ncols = h->n0;
        c.plane_stride = 0;
        c.nplanes = 1;
        c.scale = 1.0;
        c.tw = tw1;
        if (col_block_width(h->n1, BASTR_PREC_F64) > h->n0) return fail(BASTR_EINVAL, "bastr_fft: ia too small for ja");
        LAUNCHCHK(k3d::launch_colx<double>(h->n1, true, c, h->stream));
        RowArgs<double> r{};
        r.g = g;
        r.gate = open;
        r.cin = z;
        r.cout = z;
        r.tw = tw0;
        LAUNCHCHK(launch_row<double>(h->n0, R_C2C_INV, r, h->n1, h->stream));
    }
    HIPCHK(hipMemcpyAsync(inout, h->cbuf, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return BASTR_OK;
}

int bastr_fft_forward(bastr_handle* h, double* inout) { return fft_c2c(h, inout, false); }

int bastr_prof_enable(bastr_handle* h, uint32_t mask) {
    if (!h) return fail(BASTR_EINVAL, "bastr_prof_enable: NULL handle");
    h->prof_mask = mask;
    for (int c = 0; c < BASTR_K_NCLASS; ++c) {
        h->prof_launch[c] = 0;
        h->prof_ms[c] = 0;
    }
    return BASTR_OK;
}

int bastr_prof_read(const bastr_handle* h, int c, int64_t* launches, double* total_ms, double* bytes) {
    if (!h || c < 0 || c >= BASTR_K_NCLASS) return fail(BASTR_EINVAL, "bastr_prof_read: bad argument");
    // Compulsory traffic, in units of one half-spectrum field F = nh*n1*n2
    // complex: every array a class must read or write, once, over the modes
    // its passes touch.  With the dealias limits, an array indexed by kx
    // counts kx < cut (the fraction fc of F), one still indexed by ky (not
    // yet y-transformed, or y-transformed again) also |ky| < cut (fy), and
    // by kz |kz| < cut (fz).  3-D classes cover a stage's z and y passes.
    const bool lim = h->prune;
    const double fc = lim ? (double)h->cut / h->nh : 1.0;
    const double fy = lim ? (double)h->rows_n / h->n1 : 1.0;
    const double fz = h->nd == 3 ? fy : 1.0;
    double units[BASTR_K_NCLASS] = {};
    if (h->nd == 2) {
        units[BASTR_K_SPECTRA] = 2 * fc * fy;
        units[BASTR_K_COL_INV_VG] = fc * (2 * fy + 7);   // y-inverse: u_hat -> 7 fields
        units[BASTR_K_VELGRAD] = 9 * fc;                 // stage A's x pass and every moment: r7 w2
        units[BASTR_K_COL_INV] = fc * (2 * fy + 4);      // y-inverse: 2 -> 4
        units[BASTR_K_ROW_STAGE] = 6 * fc;               // r4 w2
        units[BASTR_K_COL_FWD_A] = 2 * fc + 6 * fc * fy; // y-forward of 2, update r U w KA UB
        units[BASTR_K_COL_FWD_B] = 2 * fc + 8 * fc * fy; // r U KA w UB KA
        units[BASTR_K_COL_FWD_C] = 2 * fc + 8 * fc * fy; // r U KA UB w U
        // closure force of a stage: theta/omega spectra r2 w2, y-inverse r4 w4,
        // x-inverse r4 w4 (real), network r4 w2 (real), x-forward r2 w2, add r4 w2
        units[BASTR_K_CLOSURE] = 36;
    } else {
        const double fyz = fy * fz;
        units[BASTR_K_SPECTRA] = 3 * fc * fyz;
        units[BASTR_K_COL_INV_VG] = 3 * fc * fyz + 20 * fc * fy + 13 * fc;  // z: 3 -> 10, y: 10 -> 13
        units[BASTR_K_VELGRAD] = 13 * fc;
        units[BASTR_K_COL_INV] = 3 * fc * fyz + 12 * fc * fy + 9 * fc;      // z: 3 -> 6, y: 6 -> 9
        units[BASTR_K_ROW_STAGE] = 12 * fc;                                  // r9 w3
        units[BASTR_K_COL_FWD_A] = 3 * fc + 6 * fc * fy + 9 * fc * fyz;    // y-fwd 3 -> 3, z-fwd + update
        units[BASTR_K_COL_FWD_B] = 3 * fc + 6 * fc * fy + 12 * fc * fyz;
        units[BASTR_K_COL_FWD_C] = 3 * fc + 6 * fc * fy + 12 * fc * fyz;
    }
    // (an ensemble launch covers every member)
    const double half = (double)h->nh * h->n1 * h->n2 * 2.0 * h->esz * h->nmem;
    if (launches) *launches = h->prof_launch[c];
    if (total_ms) *total_ms = h->prof_ms[c];
    if (bytes) *bytes = units[c] * half;
    return BASTR_OK;
}

int bastr_prof_copy_gbps(bastr_handle* h, uint64_t bytes, int reps, double* gbps) {
    if (!h || !gbps || reps < 1 || bytes < 16 || bytes % 16 || bytes / 16 > ((uint64_t)1 << 39))
        return fail(BASTR_EINVAL, "bastr_prof_copy_gbps: bad argument");
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        if (a) (void)hipFree(a);
        return fail(BASTR_ENOMEM, "bastr_prof_copy_gbps: cannot allocate 2 x %llu bytes", (unsigned long long)bytes);
    }
    const size_t n = bytes / 16;
    int rc = BASTR_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMemsetAsync(a, 0, bytes, h->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(BASTR_EDEVICE, "bastr_prof_copy_gbps: setup failed");
    } else {
        for (int r = 0; r < reps + 2; ++r) {
            if (r == 2) (void)hipEventRecord(e0, h->stream);
            hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                               (const double2*)a, (double2*)b, n);
        }
        (void)hipEventRecord(e1, h->stream);
        float ms = 0;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0)
            rc = fail(BASTR_EDEVICE, "bastr_prof_copy_gbps: timing failed");
        else
            *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

const char* bastr_prof_name(int c) {
    static const char* names[BASTR_K_NCLASS] = {"spectra", "col_inv_vg", "velgrad", "finalize", "col_inv",
                                                "row_stage", "col_fwd_a", "col_fwd_b", "col_fwd_c", "force",
                                                "vg_dis", "closure"};
    return (c >= 0 && c < BASTR_K_NCLASS) ? names[c] : "";
}
int bastr_prof_graph(const bastr_handle* h) { return h && h->use_graph ? 1 : 0; }
int bastr_fft_backward(bastr_handle* h, double* inout) { return fft_c2c(h, inout, true); }

}  // extern "C"

// ---------------------------------------------------------------------------
// post-processing of a field: Bastrpp's spectra2d and transfer2d/3d
// (Bastrpp.F90:174-599).  A separate arena of half-spectrum buffers, so the
// solver's state, its scratch and a pending prologue stay as they are:
//   2-D: [0,1] u_hat, [2..5] y-inverse outputs, [6..13] products
//   3-D: [0..2] u_hat, [3..8] z-inverse outputs, [9..17] y-inverse outputs,
//        [18..29] products
// The real input fields are staged in the product buffers.
// ---------------------------------------------------------------------------
namespace {

int pp_base(const bastr_handle* h) { return h->nd == 2 ? 6 : 18; }

int pp_arena(bastr_handle* h) {
    if (h->pp_status) return 0;
    SyncLock lk(g_sync_mu);
    const int nbuf = h->nd == 2 ? 14 : 30;
    const size_t cbytes = h->nhalf * 2 * h->esz;
    const long nbx = (h->nh + 63) / 64;
    const long ntask = (long)h->n1 * h->n2 * nbx;
    h->pp_nblk = (int)std::max(1L, std::min(1024L, ntask / 4));
    const size_t K1 = (size_t)h->allkmax + 1;
    const size_t npart = std::max((size_t)h->pp_nblk * K1 * 3, (size_t)h->nblk_spec * (K1 * 2 + 4));
    bool ok = true;
    for (int q = 0; q < nbuf && ok; ++q) h->pp[q] = dalloc(cbytes, &ok);
    if (ok) h->pp_part = (double*)dalloc(sizeof(double) * npart, &ok);
    int* st = ok ? (int*)dalloc(sizeof(int), &ok) : nullptr;
    const int open = INT_MAX;
    if (ok && hipMemcpy(st, &open, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) ok = false;
    if (!ok) {
        for (int q = 0; q < 30; ++q) {
            if (h->pp[q]) (void)hipFree(h->pp[q]);
            h->pp[q] = nullptr;
        }
        if (h->pp_part) (void)hipFree(h->pp_part);
        if (st) (void)hipFree(st);
        h->pp_part = nullptr;
        return fail(BASTR_ENOMEM, "post-processing: device allocation of %d half fields failed", nbuf);
    }
    h->pp_status = st;
    return 0;
}

// column transforms along y (axis 1) or z (axis 2) of the arena buffers
template <typename T>
int pp_cols(bastr_handle* h, const Gate& g, const int* in, int nin, const int* out, int nout,
            const k3d::ColJob* jobs, int nj, bool inverse, int axis, double scale) {
    k3d::ColXArgs<T> c{};
    c.gate = g;
    for (int q = 0; q < nin; ++q) c.in[q] = C<T>(h->pp[in[q]]);
    for (int q = 0; q < nout; ++q) c.out[q] = C<T>(h->pp[out[q]]);
    for (int q = 0; q < nj; ++q) c.job[q] = jobs[q];
    c.njob = nj;
    if (axis == 1) {
        c.pitch = h->hp;
        c.ncols = h->hp;
        c.plane_stride = (long)h->hp * h->n1;
        c.nplanes = h->n2;
    } else {
        c.pitch = h->hp * h->n1;
        c.ncols = h->hp * h->n1;
        c.plane_stride = 0;
        c.nplanes = 1;
    }
    c.scale = scale;
    c.tw = C<T>(h->tw1);
    LAUNCHCHK(k3d::launch_colx<T>(axis == 1 ? h->n1 : h->n2, inverse, c, h->stream));
    return 0;
}

// plain forward transforms, in place, of arena buffers first..first+nf-1
template <typename T>
int pp_fwd_inplace(bastr_handle* h, const Gate& g, int first, int nf, int axis, double scale) {
    for (int b = 0; b < nf; b += 6) {
        const int m = std::min(6, nf - b);
        int io[6];
        k3d::ColJob j[6];
        for (int q = 0; q < m; ++q) {
            io[q] = first + b + q;
            j[q] = job1(q, k3d::M_ONE, q);
        }
        int rc = pp_cols<T>(h, g, io, m, io, m, j, m, false, axis, scale);
        if (rc) return rc;
    }
    return 0;
}

// u_hat = fft(u) (normalised, no dealias) into arena [0..nd-1]
template <typename T>
int pp_load(bastr_handle* h, const double* const* src) {
    SyncLock lk(g_sync_mu);
    const int rb = pp_base(h);
    for (int c = 0; c < h->nd; ++c) {
        if (h->prec == BASTR_PREC_F64) {
            HIPCHK(hipMemcpy(h->pp[rb + c], src[c], sizeof(double) * h->npts, hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(h->dstage, src[c], sizeof(double) * h->npts, hipMemcpyHostToDevice));
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3(1024), dim3(256), 0, h->stream, h->dstage,
                               (float*)h->pp[rb + c], h->npts);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(h->stream));
        }
    }
    const Gate open{h->pp_status, 0, 0};
    const double inv_n = 1.0 / ((double)h->n0 * (double)h->n1 * (double)h->n2);
    k3d::ColJob j[3] = {job1(0, k3d::M_ONE, 0), job1(1, k3d::M_ONE, 1), job1(2, k3d::M_ONE, 2)};
    if (h->nd == 2) {
        RowArgs<T> r{};
        r.g = geo<T>(h);
        r.gate = open;
        r.rin[0] = (const T*)h->pp[rb];
        r.rin[1] = (const T*)h->pp[rb + 1];
        r.hout[0] = C<T>(h->pp[2]);
        r.hout[1] = C<T>(h->pp[3]);
        r.tw = C<T>(h->tw0);
        LAUNCHCHK(launch_row<T>(h->n0, R_FWD_PACK, r, h->n1, h->stream));
        const int in[2] = {2, 3}, out[2] = {0, 1};
        return pp_cols<T>(h, open, in, 2, out, 2, j, 2, false, 1, inv_n);
    }
    k3d::Row3Args<T> r{};
    r.g = geo3<T>(h);
    r.gate = open;
    for (int c = 0; c < 3; ++c) {
        r.rin[c] = (const T*)h->pp[rb + c];
        r.hout[c] = C<T>(h->pp[9 + c]);
    }
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(k3d::launch_row3<T>(h->n0, k3d::R3_FWD_PACK, r, h->n1 * h->n2, h->stream));
    const int y[3] = {9, 10, 11}, z[3] = {3, 4, 5}, u[3] = {0, 1, 2};
    int rc = pp_cols<T>(h, open, y, 3, z, 3, j, 3, false, 1, 1.0);
    if (rc) return rc;
    return pp_cols<T>(h, open, z, 3, u, 3, j, 3, false, 2, inv_n);
}

template <typename T>
int pp_spectra_impl(bastr_handle* h, const double* const* src, double* shells, double* totals) {
    int rc = pp_load<T>(h, src);
    if (rc) return rc;
    const int nb = h->nblk_spec, K = h->allkmax;
    SpecArgs<T> sa{};
    sa.g = geo<T>(h);
    sa.gate = Gate{h->pp_status, 0, 0};
    sa.U[0] = C<T>(h->pp[0]);
    sa.U[1] = C<T>(h->pp[1]);
    sa.allkmax = K;
    sa.part_bins = h->pp_part;
    sa.part_tot = h->pp_part + (size_t)nb * (K + 1) * 2;
    LAUNCHCHK(launch_spectra<T>(sa, nb, h->stream));
    std::vector<double> part((size_t)nb * ((K + 1) * 2 + 4));
    HIPCHK(hipMemcpyAsync(part.data(), h->pp_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // the device bins carry spectra_compute2D's 1/2; spectra2d has none
    for (int b = 1; b <= K; ++b)
        for (int c = 0; c < 2; ++c) {
            double acc = 0;
            for (int k = 0; k < nb; ++k) acc += part[((size_t)b * 2 + c) * nb + k];
            shells[(size_t)c * K + b - 1] = 2.0 * acc;
        }
    const double* tot = part.data() + (size_t)nb * (K + 1) * 2;
    for (int q = 0; q < 4; ++q) {
        double acc = 0;
        for (int k = 0; k < nb; ++k) acc += tot[(size_t)q * nb + k];
        totals[q] = 2.0 * acc;
    }
    return BASTR_OK;
}

template <typename T>
int pp_transfer_impl(bastr_handle* h, const double* const* src, double* out) {
    int rc = pp_load<T>(h, src);
    if (rc) return rc;
    const Gate open{h->pp_status, 0, 0};
    const int nd = h->nd, rb = pp_base(h), K = h->allkmax;
    const double inv_n = 1.0 / ((double)h->n0 * (double)h->n1 * (double)h->n2);
    pp::RowArgs<T> ra{};
    ra.g = geo<T>(h);
    ra.n2 = h->n2;
    ra.gate = open;
    ra.tw = C<T>(h->tw0);
    if (nd == 2) {
        // y-inverse: u1, u2, i k2 u1, i k2 u2
        const int in[2] = {0, 1}, yo[4] = {2, 3, 4, 5};
        const k3d::ColJob j[4] = {job1(0, k3d::M_ONE, 0), job1(1, k3d::M_ONE, 1), job1(0, k3d::M_IK, 2),
                                  job1(1, k3d::M_IK, 3)};
        if ((rc = pp_cols<T>(h, open, in, 2, yo, 4, j, 4, true, 1, 1.0))) return rc;
        for (int q = 0; q < 4; ++q) ra.hin[q] = C<T>(h->pp[2 + q]);
        for (int q = 0; q < 8; ++q) ra.hout[q] = C<T>(h->pp[rb + q]);
        LAUNCHCHK(pp::launch_row2<T>(h->n0, ra, h->stream));
        if ((rc = pp_fwd_inplace<T>(h, open, rb, 7, 1, inv_n))) return rc;
    } else {
        // z-inverse: u_c, i k3 u_c; y-inverse: u_c, i k2 u_c, (i k3 u_c)
        const int in[3] = {0, 1, 2}, zo[6] = {3, 4, 5, 6, 7, 8};
        k3d::ColJob j[9];
        for (int q = 0; q < 3; ++q) {
            j[q] = job1(q, k3d::M_ONE, q);
            j[3 + q] = job1(q, k3d::M_IK, 3 + q);
        }
        if ((rc = pp_cols<T>(h, open, in, 3, zo, 6, j, 6, true, 2, 1.0))) return rc;
        const int yo[9] = {9, 10, 11, 12, 13, 14, 15, 16, 17};
        for (int q = 0; q < 3; ++q) j[6 + q] = job1(3 + q, k3d::M_ONE, 6 + q);
        if ((rc = pp_cols<T>(h, open, zo, 6, yo, 9, j, 9, true, 1, 1.0))) return rc;
        for (int q = 0; q < 9; ++q) ra.hin[q] = C<T>(h->pp[9 + q]);
        for (int q = 0; q < 12; ++q) ra.hout[q] = C<T>(h->pp[rb + q]);
        LAUNCHCHK(pp::launch_row3<T>(h->n0, ra, h->stream));
        if ((rc = pp_fwd_inplace<T>(h, open, rb, 12, 1, 1.0))) return rc;
        if ((rc = pp_fwd_inplace<T>(h, open, rb, 12, 2, inv_n))) return rc;
    }
    pp::BinArgs<T> ba{};
    ba.nd = nd;
    ba.n0 = h->n0;
    ba.n1 = h->n1;
    ba.n2 = h->n2;
    ba.nh = h->nh;
    ba.hp = h->hp;
    ba.gate = open;
    for (int c = 0; c < nd; ++c) ba.U[c] = C<T>(h->pp[c]);
    for (int q = 0; q < (nd == 2 ? 8 : 12); ++q) ba.F[q] = C<T>(h->pp[rb + q]);
    ba.allkmax = K;
    ba.part = h->pp_part;
    LAUNCHCHK(pp::launch_bins<T>(ba, h->pp_nblk, h->stream));
    const int nb = h->pp_nblk;
    std::vector<double> part((size_t)nb * (K + 1) * 3);
    HIPCHK(hipMemcpyAsync(part.data(), h->pp_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int b = 1; b <= K; ++b)
        for (int q = 0; q < 3; ++q) {
            double acc = 0;
            for (int k = 0; k < nb; ++k) acc += part[((size_t)b * 3 + q) * nb + k];
            out[(size_t)q * K + b - 1] = acc;
        }
    return BASTR_OK;
}

}  // namespace

extern "C" {

int bastr_pp_spectra(bastr_handle* h, const double* u1, const double* u2, const double* u3, double* shells,
                     double* totals) {
    (void)u3;
    if (!h || !u1 || !u2 || !shells || !totals) return fail(BASTR_EINVAL, "bastr_pp_spectra: NULL argument");
    if (h->nmem > 1) return fail(BASTR_EINVAL, "bastr_pp_spectra: not available on an ensemble handle");
    if (h->nd != 2) return fail(BASTR_EINVAL, "bastr_pp_spectra: spectra2d is 2-D only");
    int rc = pp_arena(h);
    if (rc) return rc;
    const double* src[3] = {u1, u2, nullptr};
    return h->prec == BASTR_PREC_F32 ? pp_spectra_impl<float>(h, src, shells, totals)
                                     : pp_spectra_impl<double>(h, src, shells, totals);
}

int bastr_pp_transfer(bastr_handle* h, const double* u1, const double* u2, const double* u3, double* out) {
    if (!h || !u1 || !u2 || !out || (h->nd == 3 && !u3)) return fail(BASTR_EINVAL, "bastr_pp_transfer: NULL argument");
    if (h->nmem > 1) return fail(BASTR_EINVAL, "bastr_pp_transfer: not available on an ensemble handle");
    int rc = pp_arena(h);
    if (rc) return rc;
    const double* src[3] = {u1, u2, u3};
    return h->prec == BASTR_PREC_F32 ? pp_transfer_impl<float>(h, src, out) : pp_transfer_impl<double>(h, src, out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned subgrid closure (bastr_closure.h): the sharp filter between grids,
// training pairs, the network's host side and its attachment to a handle
// ---------------------------------------------------------------------------
struct bastr_pairs {
    int n = 0, shift = 0, cap = 0, count = 0;
    float* fld = nullptr;          // [cap][4][n*n] u1, u2, theta, omega
    float* tgt = nullptr;          // [cap][2][n*n] tau1, tau2
    bastr_handle* coarse = nullptr;  // n x n, the fine handles' precision (made by the first add)
    void* FA[2] = {};              // the filtered fine nonlinear term, coarse half spectra
};

namespace {

clo::Half half_of(const bastr_handle* h) { return clo::Half{h->n0, h->n1, h->nh, h->hp}; }

// the closure buffers of a 2-D one-field handle (allocated once)
int closure_buffers(bastr_handle* h) {
    if (h->CR[0]) return 0;
    bool ok = true;
    const size_t cbytes = h->nhalf * 2 * h->esz, rbytes = h->npts * h->esz;
    for (int q = 0; q < 2; ++q) {
        h->CW[q] = dalloc(cbytes, &ok);
        h->CF[q] = dalloc(rbytes, &ok);
        h->CX[q] = dalloc(cbytes, &ok);
    }
    for (int q = 0; q < 4; ++q) {
        h->CY[q] = dalloc(cbytes, &ok);
        h->CR[q] = dalloc(rbytes, &ok);
    }
    return ok ? 0 : fail(BASTR_ENOMEM, "closure buffers: device allocation failed");
}

bool square_pow2(const bastr_handle* h, int lo, int hi) {
    return h->nd == 2 && h->n0 == h->n1 && pow2_in(h->n0, lo, hi);
}

int check_fine_coarse(const bastr_handle* f, const bastr_handle* c, const char* what) {
    if (!f || !c) return fail(BASTR_EINVAL, "%s: NULL handle", what);
    if (f->nmem != 1 || c->nmem != 1) return fail(BASTR_EINVAL, "%s: one-field handles only", what);
    if (!square_pow2(f, 32, 4096) || !square_pow2(c, 32, 4096))
        return fail(BASTR_EINVAL, "%s: square 2-D grids, powers of two in [32, 4096]", what);
    if (c->n0 > f->n0) return fail(BASTR_EINVAL, "%s: the coarse grid (%d) is finer than the fine one (%d)", what, c->n0, f->n0);
    if (f->prec != c->prec) return fail(BASTR_EINVAL, "%s: the two handles differ in precision", what);
    if (!f->have_fields) return fail(BASTR_EINVAL, "%s: fine fields not set", what);
    return 0;
}

// fine's normalised spectrum (its pending entry dealias first) and its open gate
template <typename T>
int fine_spectrum(bastr_handle* f) {
    if (f->prologue_pending) {
        const int rc = run_prologue<T>(f);
        if (rc) return rc;
        f->prologue_pending = false;
    }
    return reset_status(f) ? fail(BASTR_EDEVICE, "status reset launch failed") : 0;
}

// nonlinear term N(u) = -u_j d_j u_i of a handle's state, normalised half
// spectra in W[0..1] (the step's stage pass without the update; modes inside
// the dealias limits)
template <typename T>
int nonlinear_term(bastr_handle* h) {
    const Gate open{h->status, 0, 0};
    const void* Us[2] = {h->U[0], h->U[1]};
    int rc = run_yinv2<T>(h, open, Us, 1, true);
    if (rc) return rc;
    RowArgs<T> r{};
    r.g = geo<T>(h);
    r.g.kxlim = step_kxlim(h);
    r.gate = open;
    for (int q = 0; q < 4; ++q) r.hin[q] = C<T>(h->W[q]);
    r.hout[0] = C<T>(h->NX[0]);
    r.hout[1] = C<T>(h->NX[1]);
    r.tw = C<T>(h->tw0);
    LAUNCHCHK(launch_row<T>(h->n0, R_STAGE, r, h->n1, h->stream));
    void* nx[2] = {h->NX[0], h->NX[1]};
    void* w[2] = {h->W[0], h->W[1]};
    return run_yfwd2<T>(h, open, nx, w, 2, 1.0 / ((double)h->n0 * (double)h->n1), h->hp, h->tw1, true);
}

// coarse state := sharp filter of fine's (plus up to two more fine half
// spectra filtered into extra[] of coarse size)
template <typename T>
int filter_impl(bastr_handle* f, bastr_handle* c, const void* const* more, void* const* extra, int nmore) {
    const cx<T>* src[4] = {C<T>(f->U[0]), C<T>(f->U[1]), nullptr, nullptr};
    cx<T>* dst[4] = {C<T>(c->U[0]), C<T>(c->U[1]), nullptr, nullptr};
    for (int q = 0; q < nmore; ++q) {
        src[2 + q] = C<T>((void*)more[q]);
        dst[2 + q] = C<T>(extra[q]);
    }
    LAUNCHCHK(clo::launch_filter<T>(half_of(f), half_of(c), c->cut, src, dst, 2 + nmore, f->stream));
    HIPCHK(hipStreamSynchronize(f->stream));
    c->prologue_pending = false;
    c->have_fields = true;
    c->nstep = f->nstep;
    c->time = f->time;
    return 0;
}

template <typename T>
int pairs_add_impl(bastr_pairs* ps, bastr_handle* f) {
    bastr_handle* c = ps->coarse;
    int rc = fine_spectrum<T>(f);
    if (!rc) rc = nonlinear_term<T>(f);
    const void* more[2] = {f->W[0], f->W[1]};
    if (!rc) rc = filter_impl<T>(f, c, more, ps->FA, 2);
    if (rc) return rc;
    if (reset_status(c)) return fail(BASTR_EDEVICE, "status reset launch failed");
    const Gate open{c->status, 0, 0};
    // D_c N(u_bar) in c->W[0..1]; tau = P_c N(u) - D_c N(u_bar) in KA, then physical (ureal)
    if ((rc = nonlinear_term<T>(c))) return rc;
    LAUNCHCHK(clo::launch_tau<T>(half_of(c), c->cut, C<T>(ps->FA[0]), C<T>(ps->FA[1]), C<T>(c->W[0]), C<T>(c->W[1]),
                                 C<T>(c->KA[0]), C<T>(c->KA[1]), c->stream));
    const void* K[2] = {c->KA[0], c->KA[1]};
    if ((rc = run_yinv2<T>(c, open, K, 0, false))) return rc;
    RowArgs<T> r{};
    r.g = geo<T>(c);
    r.gate = open;
    r.hin[0] = C<T>(c->W[0]);
    r.hin[1] = C<T>(c->W[1]);
    r.rout[0] = (T*)c->ureal[0];
    r.rout[1] = (T*)c->ureal[1];
    r.tw = C<T>(c->tw0);
    LAUNCHCHK(launch_row<T>(c->n0, R_INV_PACK, r, c->n1, c->stream));
    // the network's input fields of u_bar
    const void* Uc[2] = {c->U[0], c->U[1]};
    if ((rc = closure_fields<T>(c, open, Uc, false))) return rc;
    const size_t n2 = c->npts;
    for (int q = 0; q < 4; ++q)
        LAUNCHCHK(clo::launch_to_float<T>((const T*)c->CR[q], ps->fld + ((size_t)ps->count * 4 + q) * n2, n2, c->stream));
    for (int q = 0; q < 2; ++q)
        LAUNCHCHK(clo::launch_to_float<T>((const T*)c->ureal[q], ps->tgt + ((size_t)ps->count * 2 + q) * n2, n2,
                                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ps->count += 1;
    return 0;
}

mlp::Data pairs_data(const bastr_pairs* ps) {
    mlp::Data d{};
    d.fld = ps->fld;
    d.tgt = ps->tgt;
    d.n = ps->n;
    d.shift = ps->shift;
    d.nsnap = ps->count;
    return d;
}

void set_scales(bastr_closure* net, const float* s) {
    for (int q = 0; q < BASTR_CLOSURE_NSCALE; ++q) net->scales[q] = s[q];
    for (int q = 0; q < mlp::NFLD; ++q) net->sc.inv_in[q] = 1.0f / s[q];
    for (int q = 0; q < mlp::NO; ++q) net->sc.out[q] = s[mlp::NFLD + q];
    net->version += 1;
}

int net_part(bastr_closure* net, int nw) {
    if (nw <= net->part_w) return 0;
    SyncLock lk(g_sync_mu);
    if (net->part) (void)hipFree(net->part);
    if (net->part_loss) (void)hipFree(net->part_loss);
    net->part = nullptr;
    net->part_loss = nullptr;
    net->part_w = 0;
    HIPCHK(hipMalloc(&net->part, sizeof(float) * (size_t)nw * mlp::NPARAM));
    HIPCHK(hipMalloc(&net->part_loss, sizeof(float) * (size_t)nw));
    net->part_w = nw;
    return 0;
}

// forward + backward over the selected tiles: gradient in net->grad, loss in *dloss
int loss_grad_launch(bastr_closure* net, const bastr_pairs* ps, const mlp::TileSel& sel, double npts, double* dloss) {
    const int nw = mlp::train_waves(sel.ntiles);
    int rc = net_part(net, nw);
    if (rc) return rc;
    LAUNCHCHK(mlp::launch_train(net->packed, net->sc, pairs_data(ps), sel, npts, net->part, net->part_loss, net->stream));
    LAUNCHCHK(mlp::launch_reduce(net->part, net->part_loss, nw, net->grad, dloss, net->stream));
    return 0;
}

}  // namespace

extern "C" {

int bastr_filter_to(bastr_handle* fine, bastr_handle* coarse) {
    int rc = check_fine_coarse(fine, coarse, "bastr_filter_to");
    if (rc) return rc;
    if (fine == coarse) return fail(BASTR_EINVAL, "bastr_filter_to: the same handle twice");
    rc = fine->prec == BASTR_PREC_F32 ? fine_spectrum<float>(fine) : fine_spectrum<double>(fine);
    if (rc) return rc;
    return fine->prec == BASTR_PREC_F32 ? filter_impl<float>(fine, coarse, nullptr, nullptr, 0)
                                        : filter_impl<double>(fine, coarse, nullptr, nullptr, 0);
}

int bastr_pairs_create(int32_t n, int32_t capacity, bastr_pairs** out) {
    if (!out) return fail(BASTR_EINVAL, "bastr_pairs_create: NULL argument");
    *out = nullptr;
    if (!pow2_in(n, 32, 4096)) return fail(BASTR_EINVAL, "bastr_pairs_create: n = %d is not a power of two in [32, 4096]", n);
    if (capacity < 1) return fail(BASTR_EINVAL, "bastr_pairs_create: capacity must be >= 1");
    bastr_pairs* ps = new bastr_pairs();
    ps->n = n;
    ps->shift = __builtin_ctz((unsigned)n);
    ps->cap = capacity;
    SyncLock lk(g_sync_mu);
    const size_t n2 = (size_t)n * n;
    if (hipMalloc(&ps->fld, sizeof(float) * n2 * 4 * (size_t)capacity) != hipSuccess ||
        hipMalloc(&ps->tgt, sizeof(float) * n2 * 2 * (size_t)capacity) != hipSuccess) {
        bastr_pairs_destroy(ps);
        return fail(BASTR_ENOMEM, "bastr_pairs_create: device allocation failed");
    }
    *out = ps;
    return BASTR_OK;
}

int bastr_pairs_add(bastr_pairs* ps, bastr_handle* fine) {
    if (!ps || !fine) return fail(BASTR_EINVAL, "bastr_pairs_add: NULL argument");
    if (ps->count >= ps->cap) return fail(BASTR_EINVAL, "bastr_pairs_add: all %d snapshots used", ps->cap);
    if (!ps->coarse) {
        bastr_params p{};
        p.ia = p.ja = ps->n;
        p.ka = 0;
        p.reynolds = fine->p.reynolds;
        p.deltat = fine->p.deltat;
        p.timemethod = 1;
        p.forcemethod = 0;
        p.forcek = 1;
        p.target_energy = 0;
        p.lproject = 0;
        p.kmax = 1;
        p.precision = fine->prec;
        int rc = bastr_create(&p, &ps->coarse);
        if (rc) return rc;
        bool ok = true;
        const size_t cbytes = ps->coarse->nhalf * 2 * ps->coarse->esz;
        ps->FA[0] = dalloc(cbytes, &ok);
        ps->FA[1] = dalloc(cbytes, &ok);
        if (!ok || closure_buffers(ps->coarse)) return fail(BASTR_ENOMEM, "bastr_pairs_add: device allocation failed");
    }
    int rc = check_fine_coarse(fine, ps->coarse, "bastr_pairs_add");
    if (rc) return rc;
    return fine->prec == BASTR_PREC_F32 ? pairs_add_impl<float>(ps, fine) : pairs_add_impl<double>(ps, fine);
}

int bastr_pairs_count(const bastr_pairs* ps) { return ps ? ps->count : -1; }

int bastr_pairs_get(const bastr_pairs* ps, int32_t i, float* fields, float* target) {
    if (!ps) return fail(BASTR_EINVAL, "bastr_pairs_get: NULL handle");
    if (i < 0 || i >= ps->count) return fail(BASTR_EINVAL, "bastr_pairs_get: snapshot %d not in [0, %d)", i, ps->count);
    SyncLock lk(g_sync_mu);
    const size_t n2 = (size_t)ps->n * ps->n;
    if (fields) HIPCHK(hipMemcpy(fields, ps->fld + (size_t)i * 4 * n2, sizeof(float) * 4 * n2, hipMemcpyDeviceToHost));
    if (target) HIPCHK(hipMemcpy(target, ps->tgt + (size_t)i * 2 * n2, sizeof(float) * 2 * n2, hipMemcpyDeviceToHost));
    return BASTR_OK;
}

int bastr_pairs_destroy(bastr_pairs* ps) {
    if (!ps) return BASTR_OK;
    {
        SyncLock lk(g_sync_mu);
        if (ps->fld) (void)hipFree(ps->fld);
        if (ps->tgt) (void)hipFree(ps->tgt);
        for (void* p : ps->FA)
            if (p) (void)hipFree(p);
    }
    if (ps->coarse) bastr_destroy(ps->coarse);
    delete ps;
    return BASTR_OK;
}

int bastr_closure_create(uint64_t seed, bastr_closure** out) {
    if (!out) return fail(BASTR_EINVAL, "bastr_closure_create: NULL argument");
    *out = nullptr;
    // Glorot-uniform weights from a counter-based draw of seed, zero biases
    std::vector<float> p(mlp::NPARAM, 0.0f);
    auto draw = [seed](uint64_t i) {
        uint64_t x = seed * 0x100000001B3ull + i + 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x = x ^ (x >> 31);
        return (double)(x >> 11) * 0x1.0p-53;  // [0, 1)
    };
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; ++i) p[off + i] = (float)((2.0 * draw(off + i) - 1.0) * a);
    };
    fill(mlp::OFF_W1, mlp::NH, mlp::NF);
    fill(mlp::OFF_W2, mlp::NH, mlp::NH);
    fill(mlp::OFF_W3, mlp::NO, mlp::NH);
    bastr_closure* net = new bastr_closure();
    {
        SyncLock lk(g_sync_mu);
        bool ok = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipMalloc(&net->params, sizeof(float) * mlp::NPARAM) == hipSuccess;
        ok = ok && hipMalloc(&net->am, sizeof(float) * mlp::NPARAM) == hipSuccess;
        ok = ok && hipMalloc(&net->av, sizeof(float) * mlp::NPARAM) == hipSuccess;
        ok = ok && hipMalloc(&net->grad, sizeof(float) * mlp::NPARAM) == hipSuccess;
        ok = ok && hipMalloc(&net->packed, sizeof(mlp::Packed)) == hipSuccess;
        ok = ok && hipMalloc(&net->dloss, sizeof(double) * 4096) == hipSuccess;
        if (!ok) {
            bastr_closure_destroy(net);
            return fail(BASTR_ENOMEM, "bastr_closure_create: device allocation failed");
        }
    }
    const float ones[BASTR_CLOSURE_NSCALE] = {1, 1, 1, 1, 1, 1};
    const int rc = bastr_closure_set_params(net, p.data(), ones);
    if (rc) {
        bastr_closure_destroy(net);
        return rc;
    }
    *out = net;
    return BASTR_OK;
}

int bastr_closure_destroy(bastr_closure* net) {
    if (!net) return BASTR_OK;
    SyncLock lk(g_sync_mu);
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (void* q : {(void*)net->params, (void*)net->am, (void*)net->av, (void*)net->grad, (void*)net->packed,
                    (void*)net->part, (void*)net->part_loss, (void*)net->dloss})
        if (q) (void)hipFree(q);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return BASTR_OK;
}

int bastr_closure_set_params(bastr_closure* net, const float* params, const float* scales) {
    if (!net || !params || !scales) return fail(BASTR_EINVAL, "bastr_closure_set_params: NULL argument");
    for (int q = 0; q < BASTR_CLOSURE_NSCALE; ++q)
        if (!(scales[q] > 0.0f) || !std::isfinite(scales[q]))
            return fail(BASTR_EINVAL, "bastr_closure_set_params: scale %d = %g is not a positive number", q, (double)scales[q]);
    SyncLock lk(g_sync_mu);
    HIPCHK(hipMemcpy(net->params, params, sizeof(float) * mlp::NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(net->am, 0, sizeof(float) * mlp::NPARAM));
    HIPCHK(hipMemset(net->av, 0, sizeof(float) * mlp::NPARAM));
    HIPCHK(hipStreamSynchronize(nullptr));  // the fills, before net->stream's next kernels
    net->t = 0;
    set_scales(net, scales);
    LAUNCHCHK(mlp::launch_pack(net->params, net->packed, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    return BASTR_OK;
}

int bastr_closure_get_params(const bastr_closure* net, float* params, float* scales) {
    if (!net) return fail(BASTR_EINVAL, "bastr_closure_get_params: NULL handle");
    SyncLock lk(g_sync_mu);
    if (params) HIPCHK(hipMemcpy(params, net->params, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost));
    if (scales)
        for (int q = 0; q < BASTR_CLOSURE_NSCALE; ++q) scales[q] = net->scales[q];
    return BASTR_OK;
}

int bastr_closure_fit_scales(bastr_closure* net, const bastr_pairs* ps) {
    if (!net || !ps) return fail(BASTR_EINVAL, "bastr_closure_fit_scales: NULL argument");
    if (ps->count < 1) return fail(BASTR_EINVAL, "bastr_closure_fit_scales: no pairs");
    const int nb = mlp::sumsq_blocks();
    std::vector<double> sums((size_t)nb * 6);
    {
        SyncLock lk(g_sync_mu);
        double* d = nullptr;
        HIPCHK(hipMalloc(&d, sizeof(double) * sums.size()));
        const int rl = mlp::launch_sumsq(pairs_data(ps), d, net->stream);
        const hipError_t e1 = hipGetLastError();
        const hipError_t e2 = hipStreamSynchronize(net->stream);
        const hipError_t e3 = hipMemcpy(sums.data(), d, sizeof(double) * sums.size(), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        if (rl || e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
            return fail(BASTR_EDEVICE, "bastr_closure_fit_scales: reduction failed");
    }
    float s[BASTR_CLOSURE_NSCALE];
    const double npts = (double)ps->count * ps->n * ps->n;
    double rms[6];
    for (int q = 0; q < 6; ++q) {
        double a = 0.0;
        for (int b = 0; b < nb; ++b) a += sums[(size_t)b * 6 + q];
        rms[q] = std::sqrt(a / npts);
    }
    // fields come in pairs of one unit (u1 u2, theta omega, tau1 tau2); a field
    // at roundoff level beside its partner (omega of a potential flow, which
    // stays potential under Burgers) takes the partner's scale, so that its
    // noise enters the network near zero instead of at unit size
    for (int q = 0; q < 6; ++q) {
        const double other = rms[q ^ 1];
        double v = rms[q] < 1e-6 * other ? other : rms[q];
        if (!(v > 0.0)) v = 1.0;  // both identically zero: unit scale
        s[q] = (float)v;
    }
    set_scales(net, s);
    return BASTR_OK;
}

int bastr_closure_loss_grad(bastr_closure* net, const bastr_pairs* ps, int32_t first, int32_t count, double* loss,
                            float* grad) {
    if (!net || !ps) return fail(BASTR_EINVAL, "bastr_closure_loss_grad: NULL argument");
    if (first < 0 || count < 1 || first + count > ps->count)
        return fail(BASTR_EINVAL, "bastr_closure_loss_grad: pairs [%d, %d) not in [0, %d)", first, first + count, ps->count);
    mlp::TileSel sel{};
    sel.random = 0;
    sel.s0 = first;
    sel.ns = count;
    sel.ntiles = (long)count * ps->n * (ps->n / 32);
    int rc = loss_grad_launch(net, ps, sel, (double)count * ps->n * ps->n, net->dloss);
    if (rc) return rc;
    SyncLock lk(g_sync_mu);
    HIPCHK(hipStreamSynchronize(net->stream));
    if (loss) HIPCHK(hipMemcpy(loss, net->dloss, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) HIPCHK(hipMemcpy(grad, net->grad, sizeof(float) * mlp::NPARAM, hipMemcpyDeviceToHost));
    return BASTR_OK;
}

int bastr_closure_train(bastr_closure* net, const bastr_pairs* ps, int32_t iters, int32_t batch, float lr,
                        uint64_t seed, double* losses) {
    if (!net || !ps) return fail(BASTR_EINVAL, "bastr_closure_train: NULL argument");
    if (ps->count < 1) return fail(BASTR_EINVAL, "bastr_closure_train: no pairs");
    if (iters < 0 || batch < 1 || !(lr > 0.0f)) return fail(BASTR_EINVAL, "bastr_closure_train: bad iters/batch/lr");
    const int chunk = 4096;  // losses kept on the device per chunk of steps
    for (int i0 = 0; i0 < iters; i0 += chunk) {
        const int ni = std::min(chunk, iters - i0);
        for (int i = 0; i < ni; ++i) {
            mlp::TileSel sel{};
            sel.random = 1;
            sel.seed = seed * 0x9E3779B97F4A7C15ull + (uint64_t)(net->t + 1);
            sel.ntiles = batch;
            int rc = loss_grad_launch(net, ps, sel, (double)batch * 32.0, net->dloss + i);
            if (rc) return rc;
            net->t += 1;
            LAUNCHCHK(mlp::launch_adam(net->params, net->am, net->av, net->grad, lr, net->t, net->stream));
            LAUNCHCHK(mlp::launch_pack(net->params, net->packed, net->stream));
        }
        SyncLock lk(g_sync_mu);
        HIPCHK(hipStreamSynchronize(net->stream));
        if (losses) HIPCHK(hipMemcpy(losses + i0, net->dloss, sizeof(double) * ni, hipMemcpyDeviceToHost));
    }
    SyncLock lk(g_sync_mu);
    HIPCHK(hipStreamSynchronize(net->stream));
    return BASTR_OK;
}

int bastr_closure_apply(bastr_closure* net, const bastr_pairs* ps, int32_t i, float* force) {
    if (!net || !ps || !force) return fail(BASTR_EINVAL, "bastr_closure_apply: NULL argument");
    if (i < 0 || i >= ps->count) return fail(BASTR_EINVAL, "bastr_closure_apply: snapshot %d not in [0, %d)", i, ps->count);
    const size_t n2 = (size_t)ps->n * ps->n;
    SyncLock lk(g_sync_mu);
    float* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(float) * 2 * n2));
    mlp::Fields<float> F{};
    for (int q = 0; q < 4; ++q) F.f[q] = ps->fld + ((size_t)i * 4 + q) * n2;
    F.n = ps->n;
    F.shift = ps->shift;
    const int rl = mlp::launch_forward<float>(net->packed, net->sc, F, d, d + n2, net->stream);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(net->stream);
    const hipError_t e3 = hipMemcpy(force, d, sizeof(float) * 2 * n2, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (rl || e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return fail(BASTR_EDEVICE, "bastr_closure_apply: inference failed");
    return BASTR_OK;
}

int bastr_prof_closure(bastr_closure* net, const bastr_pairs* ps, int reps, int batch, double* train_ms,
                       double* forward_ms) {
    if (!net || !ps || reps < 1 || batch < 1) return fail(BASTR_EINVAL, "bastr_prof_closure: bad argument");
    if (ps->count < 1) return fail(BASTR_EINVAL, "bastr_prof_closure: no pairs");
    mlp::TileSel sel{};
    sel.random = 1;
    sel.seed = 12345;
    sel.ntiles = batch;
    const int nw = mlp::train_waves(sel.ntiles);
    int rc = net_part(net, nw);
    if (rc) return rc;
    const size_t n2 = (size_t)ps->n * ps->n;
    SyncLock lk(g_sync_mu);
    float* d = nullptr;
    hipEvent_t e[4] = {};
    HIPCHK(hipMalloc(&d, sizeof(float) * 2 * n2));
    bool ok = true;
    for (hipEvent_t& ev : e) ok = ok && hipEventCreate(&ev) == hipSuccess;
    mlp::Fields<float> F{};
    for (int q = 0; q < 4; ++q) F.f[q] = ps->fld + (size_t)q * n2;
    F.n = ps->n;
    F.shift = ps->shift;
    const mlp::Data dd = pairs_data(ps);
    const double npts = (double)batch * 32.0;
    // one untimed launch of each, then the timed ones
    ok = ok && mlp::launch_train(net->packed, net->sc, dd, sel, npts, net->part, net->part_loss, net->stream) == 0;
    ok = ok && mlp::launch_forward<float>(net->packed, net->sc, F, d, d + n2, net->stream) == 0;
    ok = ok && hipEventRecord(e[0], net->stream) == hipSuccess;
    for (int r = 0; r < reps && ok; ++r)
        ok = mlp::launch_train(net->packed, net->sc, dd, sel, npts, net->part, net->part_loss, net->stream) == 0;
    ok = ok && hipEventRecord(e[1], net->stream) == hipSuccess;
    ok = ok && hipEventRecord(e[2], net->stream) == hipSuccess;
    for (int r = 0; r < reps && ok; ++r)
        ok = mlp::launch_forward<float>(net->packed, net->sc, F, d, d + n2, net->stream) == 0;
    ok = ok && hipEventRecord(e[3], net->stream) == hipSuccess;
    ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(net->stream) == hipSuccess;
    float t1 = 0, t2 = 0;
    ok = ok && hipEventElapsedTime(&t1, e[0], e[1]) == hipSuccess && hipEventElapsedTime(&t2, e[2], e[3]) == hipSuccess;
    for (hipEvent_t ev : e)
        if (ev) (void)hipEventDestroy(ev);
    (void)hipFree(d);
    if (!ok) return fail(BASTR_EDEVICE, "bastr_prof_closure: timing failed");
    if (train_ms) *train_ms = t1 / reps;
    if (forward_ms) *forward_ms = t2 / reps;
    return BASTR_OK;
}

int bastr_closure_attach(bastr_handle* h, bastr_closure* net) {
    if (!h) return fail(BASTR_EINVAL, "bastr_closure_attach: NULL handle");
    if (net) {
        if (h->nmem != 1 || !square_pow2(h, 32, 4096))
            return fail(BASTR_EINVAL, "bastr_closure_attach: one-field square 2-D handles, n a power of two in [32, 4096]");
        if (h->p.timemethod != 1) return fail(BASTR_EINVAL, "bastr_closure_attach: RK3 (timemethod 1) only");
        if (h->overlap) return fail(BASTR_EINVAL, "bastr_closure_attach: not with the two-stream order (BASTR_OVERLAP)");
        const int rc = closure_buffers(h);
        if (rc) return rc;
    }
    drop_graph(h);  // captured step graphs do not hold the closure's launches (or still do)
    h->net = net;
    h->net_version = net ? net->version : 0;
    return BASTR_OK;
}

int bastr_closure_per_step(bastr_handle* h, int32_t per_step) {
    if (!h) return fail(BASTR_EINVAL, "bastr_closure_per_step: NULL handle");
    if (per_step != 0 && per_step != 1) return fail(BASTR_EINVAL, "bastr_closure_per_step: per_step must be 0 or 1");
    if (h->clo_per_step != (per_step == 1)) drop_graph(h);  // the captured stages launch the other cadence
    h->clo_per_step = per_step == 1;
    return BASTR_OK;
}

}  // extern "C"
