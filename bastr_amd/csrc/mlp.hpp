// mlp.hpp -- the per-point MLP subgrid closure on BF16 MFMA (gfx950).
//
// Network: 36 features -> 64 -> 64 -> 2, GELU after both hidden layers, in
// its tanh form (Hendrycks and Gimpel, 2016): gelu(z) = z (1 + tanh u) / 2,
// u = sqrt(2/pi) (z + 0.044715 z^3).  Features of a point are the 3x3 periodic stencil of four coarse
// physical fields (u1, u2, theta = d1 u1 + d2 u2, omega = d1 u2 - d2 u1),
// feature f = 4*pos + field, pos = 3*(dy+1) + (dx+1), each field divided by
// its scale; the two outputs times their scales are the subgrid force.
//
// Arithmetic: weights W1, W2, W3 rounded to BF16 (round to nearest even) from
// an FP32 master copy; the layer inputs of the two MFMA layers (features,
// hidden activations) rounded to BF16; FP32 accumulation; biases, GELU and the
// output layer (64 -> 2, on the vector units) in FP32.  Backward: the MFMA
// operands dZ2, dZ1, H1 and the features are BF16, accumulation FP32.
//
// MFMA formulation (v_mfma_f32_32x32x8_bf16, D = A B + C, A 32x8, B 8x32):
// rows M are hidden units, columns N are points, so a wavefront advances 32
// points at a time, D[unit][point] = sum_k W[unit][k] in[k][point].  Lane l
// holds B[4*(l/32) + t][l%32] and D[8*(r/4) + 4*(l/32) + r%4][l%32],
// r = 0..15: the accumulator block of one layer is the B operand of the next
// (k-step j of the next layer is accumulators 4*(j%4)..+3 of block j/4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlp {

constexpr int NFLD = 4;   // feature fields
constexpr int NF = 36;    // features (9 stencil points x 4 fields)
constexpr int NH = 64;    // hidden width
constexpr int NO = 2;     // outputs
constexpr int KS1 = 6;    // k-steps of layer 1 (36 features padded to 48)
constexpr int KS2 = 8;    // k-steps of layer 2 (64)

// FP32 parameter vector, in this order
constexpr int OFF_W1 = 0;                  // W1[h][f], 64 x 36
constexpr int OFF_B1 = OFF_W1 + NH * NF;   // b1[h]
constexpr int OFF_W2 = OFF_B1 + NH;        // W2[h2][h1], 64 x 64
constexpr int OFF_B2 = OFF_W2 + NH * NH;   // b2[h2]
constexpr int OFF_W3 = OFF_B2 + NH;        // W3[c][h2], 2 x 64
constexpr int OFF_B3 = OFF_W3 + NO * NH;   // b3[c]
constexpr int NPARAM = OFF_B3 + NO;        // 6658

// Packed device form of the network (written by launch_pack from the master copy)
struct Packed {
    uint16_t A1[2 * KS1 * 64 * 4];   // layer-1 A operands [block][k-step][lane][4], W1 rows
    uint16_t A2[2 * KS2 * 64 * 4];   // layer-2 A operands, W2[h2][h1]
    uint16_t A2T[2 * KS2 * 64 * 4];  // backward dH1 = W2^T dZ2: operands W2[h2][h1] with rows h1
    float W3[NO * NH];               // BF16-rounded W3
    float b1[NH], b2[NH], b3[NO];
};

struct Scales {
    float inv_in[NFLD];   // 1 / feature-field scale
    float out[NO];        // output scale
};

// Feature fields of one grid (n x n, n a power of two, x fastest)
template <typename T>
struct Fields {
    const T* f[NFLD];
    int n, shift;  // n = 1 << shift
};

// Training data: nsnap snapshots of one n x n coarse grid, fields[s][c][n*n],
// targets[s][c][n*n] (FP32)
struct Data {
    const float* fld;
    const float* tgt;
    int n, shift, nsnap;
};

// How a training launch picks its 32-point tiles (a tile is 32 consecutive x of
// one row of one snapshot): random (counter-based hash of seed, tile index) or
// every tile of snapshots [s0, s0 + ns) in order.
struct TileSel {
    int random;
    uint64_t seed;
    int s0, ns;
    long ntiles;      // tiles this launch processes
};

int launch_pack(const float* params, Packed* out, hipStream_t s);
template <typename T>
int launch_forward(const Packed* net, Scales sc, Fields<T> in, T* out0, T* out1, hipStream_t s);
// forward + backward over the selected tiles; per-wavefront partial gradients
// part[w * NPARAM + j] and partial losses part_loss[w] for w < nwaves(); loss =
// sum over points and outputs of (out - target/scale)^2 / (2 * npts_total)
int train_waves(long ntiles);
int launch_train(const Packed* net, Scales sc, Data d, TileSel sel, double npts_total, float* part, float* part_loss,
                 hipStream_t s);
// grad[j] = sum_w part[w * NPARAM + j] (fixed order); loss = sum_w part_loss[w]
int launch_reduce(const float* part, const float* part_loss, int nw, float* grad, double* loss, hipStream_t s);
// Adam step t (1-based) on the master copy
int launch_adam(float* params, float* m, float* v, const float* grad, float lr, int t, hipStream_t s);
// sum of squares of each feature field and target component over the data
// (partials per block, reduced on the host): sums[blk * 6 + q]
int sumsq_blocks();
int launch_sumsq(Data d, double* sums, hipStream_t s);

}  // namespace mlp
