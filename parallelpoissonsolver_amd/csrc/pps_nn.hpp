// pps_nn.hpp -- learned multilevel preconditioner: network layout, the MFMA
// convolution kernels' launchers and the preconditioner interface.
//
// Network (3 levels, 16 channels, 3^3 kernels, ReLU; BF16 weights and
// activations, FP32 accumulation):
//   L0  A0 = relu(lift(r))            1 -> 16
//       A1 = relu(a1(A0))             16 -> 16   (skip S0)
//   L1  B0 = relu(b0(pool(A1)))       16 -> 16
//       B1 = relu(b1(B0))             16 -> 16   (skip S1)
//   L2  C0 = relu(c0(pool(B1)))       16 -> 16
//       C1..C5 = relu(ci(C_{i-1}))    16 -> 16
//   L1  E0 = relu(e0(up(C5) + S1))    16 -> 16
//       E1 = relu(e1(E0))             16 -> 16
//   L0  F0 = relu(f0(up(E1) + S0))    16 -> 16
//       z  = out(F0)                  16 -> 1   (linear)
// pool = 2x2x2 average, up = nearest-neighbour copy; both are fused into the
// input stage of the convolution that consumes them.  Activations are stored
// channels-last (16 BF16 = 32 bytes per voxel), dense per level.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pps {

struct Ctx;

struct NNPrecond {
    virtual ~NNPrecond() = default;
    // out = M(in) on the solver region (zero elsewhere); kernels gated on *gate
    virtual int apply(Ctx& C, void* out, int out_prec, void* in, int in_prec, const int* gate) = 0;
    virtual int fallback_switches() const { return 0; }
};

namespace nn {

constexpr int kC = 16;        // channels
constexpr int kTaps = 27;
constexpr int kSteps = 14;    // K = 28 taps x 16 channels in steps of 32 (tap 27 is zero)
constexpr int kTileX = 16, kTileY = 4, kTileZ = 4;  // output tile of one conv block
constexpr int kNumConv = 14;  // lift, a1, b0, b1, c0..c5, e0, e1, f0, out

// where a convolution reads its input from
enum InMode {
    IN_PLAIN = 0,   // act at this level
    IN_POOL = 1,    // 2x2x2 average of act at the finer level
    IN_UPADD = 2,   // act at the coarser level (nearest) + skip at this level
    IN_MASK = 3,    // act * (mask > 0): ReLU gradient of the consumer's output
    IN_FIELD = 4    // one-channel FP32/FP64 field (strided), times a device scale
};
enum OutMode {
    OUT_RELU = 0,   // bf16 act, + bias, relu
    OUT_LINEAR = 1, // bf16 act, no bias, no relu (backward data)
    OUT_FIELD = 2   // channel 0 (+ bias) into a strided FP32/FP64 field
};

// a strided one-channel field (solver field or dense training sample)
struct FieldRef {
    void* p;           // float* or double*
    int is64;
    long long off;     // element offset of voxel (0,0,0) of sample 0
    long long sx, sy, sz, sb;  // strides in elements
    const double* scale_sum;   // if set: scale = 1/sqrt(*scale_sum / scale_count)
    double scale_count;
    float scale;       // used when scale_sum is null
};

struct ActRef {
    uint16_t* p;       // bf16 bits, [b][z][y][x][16]
    int B, D, H, W;
};

struct ConvArgs {
    int mode_in, mode_out;
    ActRef in, in2;     // in2: skip (UPADD) or ReLU mask (MASK); fine level for POOL is `in`
    FieldRef fin, fout;
    ActRef out;          // output activation (D,H,W of this level)
    const uint16_t* wl;  // lane-arranged bf16 weights [kSteps][64][8]
    const float* bias;   // [16] or null
    const int* gate;
    const int* gate2;
};

void launch_conv(const ConvArgs& a, hipStream_t s);

struct DwArgs {
    int mode_in;        // layer input as in the forward pass
    ActRef in, in2;
    FieldRef fin;
    int dz_mode;        // 0: dZ = dY (bf16 act) masked by Y>0; 1: dZ = dY unmasked bf16; 2: dZ = FP32 field
    ActRef dy, y;
    FieldRef fdy;
    int D, H, W, B;     // output geometry of the layer
    float* part;        // [nblocks][27*16*16 + 16]
    int tiles_per_block;
};
int dw_blocks(const DwArgs& a);
void launch_dw(const DwArgs& a, hipStream_t s);
// sum the partials: g[27*256 + 16] = sum over blocks
void launch_dw_reduce(const float* part, int nblocks, float* g, hipStream_t s);

// gradient routing of the fused input stages
// dfine[v] (+)= dcoarse[v>>1] / 8   (pool backward)
void launch_pool_back(const ActRef& dcoarse, const ActRef& dfine, int accumulate, hipStream_t s);
// dskip (+)= dx ; dcoarse = sum of the 8 children of dx   (up + skip backward)
void launch_upadd_back(const ActRef& dx, const ActRef& dskip, int accumulate_skip, const ActRef& dcoarse, hipStream_t s);

// lane-arranged weights from FP32 master weights W[27][16 ci][16 co] (ci < cin, co < cout)
// transposed = 1: the flipped, transposed filter of the backward-data pass
void launch_pack_weights(const float* W, int cin, int cout, int transposed, uint16_t* wl, hipStream_t s);

// Adam update of n parameters
void launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                 float bc1, float bc2, hipStream_t s);

// loss = mean over samples of ||y - t||^2 / ||t||^2 ; dy = 2 (y - t) / ||t||^2 / B
void launch_loss(const float* y, const float* t, const float* tnorm2, int B, long long n, float* dy, double* loss_partial,
                 hipStream_t s);

}  // namespace nn
}  // namespace pps
