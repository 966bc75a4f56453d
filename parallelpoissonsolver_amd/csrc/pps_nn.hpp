// pps_nn.hpp -- learned multilevel preconditioner (BF16 MFMA convolutions).
//
// Network: 3 levels, 16 channels, 3^3 kernels, ReLU; BF16 weights and
// activations, FP32 accumulation, FP32 master weights (Adam).
//   layer   name    level  input                      out
//     0     lift      0    field r (1 channel)        act0   (skip of level 0)
//     1     b0        1    pool(act0)                 act1
//     2     b1        1    act1                       act2   (skip of level 1)
//     3     c0        2    pool(act2)                 act3
//   4..10   c1..c7    2    act(i-1)                   act4..act10
//    11     e0        1    up(act10) + act2           act11
//    12     e1        1    act11                      act12
//    13     out       0    up(act12) + act0           field z (16 -> 1, linear)
// Level 0 (the full grid) carries only the lift and the output convolution;
// the other twelve 16 -> 16 convolutions run at 1/8 and 1/64 of the voxels.
// pool = 2x2x2 average, up = nearest neighbour; both run inside the input
// stage of the convolution that consumes them.  Activations are dense,
// channels-last (16 BF16 = 32 bytes per voxel), [b][z][y][x][c].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace pps {

// solver region of a pitched field: element (i,j,k) of the region at
// off + i + sy*j + sz*k, region extent n0 x n1 x n2
struct NNFieldGeom {
    long long off, sy, sz;
    int n0, n1, n2;
};

// residual update folded into the preconditioner's input read:
// r_cur = r_prev - alpha ap_prev (r_cur = r_prev while *skip != 0), written to
// r_cur on the region; all three fields share the input's precision and layout
struct NNResUpdate {
    const void* r_prev;
    const void* ap_prev;
    void* r_cur;
    const double* alpha;
    const int* skip;
};

// learned preconditioner as seen by the Krylov stages
struct NNPrecond {
    virtual ~NNPrecond() = default;
    // out = M(in) on the region (other points untouched).  The input is scaled
    // by 1/RMS with RMS^2 = *rr / count; every kernel returns at once when
    // *gate != 0 or *gate2 != 0.
    virtual int apply(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                      const int* gate, const int* gate2, const double* rr) = 0;
    // apply() that also sums out . dot over the region (dot: a field of out's
    // precision and layout): per-block FP64 partials go to dpart[0 .. *nblocks)
    // when they fit in dcap; otherwise *nblocks = 0 and nothing is summed.
    // upd (or null): the input `in` is first formed as upd->r_cur from the
    // previous residual (see NNResUpdate); *rr must then already hold ||in||^2
    virtual int apply_dot(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                          const int* gate, const int* gate2, const double* rr, const void* dot, double* dpart, int dcap,
                          int* nblocks, const NNResUpdate* upd) = 0;
    virtual const char* error() const = 0;
};

namespace nn {

constexpr int kC = 16;
constexpr int kTaps = 27;
constexpr int kOps = 15;          // weight operands per layer: 5 tap pairs x 3 dy (K = 2 taps x 16 channels)
constexpr int kLayers = 14;
constexpr int kActs = 13;         // activations act0..act12
constexpr int kLayerParams = kTaps * kC * kC + kC;  // W[tap][ci][co] then bias[co]
constexpr int kParams = kLayers * kLayerParams;
constexpr int kPackedHalves = kOps * 64 * 8;        // lane-arranged BF16 weights of one layer

enum InMode { IN_PLAIN = 0, IN_POOL = 1, IN_UPADD = 2, IN_MASK = 3, IN_FIELD = 4 };
enum OutMode { OUT_RELU = 0, OUT_LINEAR = 1, OUT_FIELD = 2 };

struct LayerDef {
    int cin, cout, level, mode_in;
    int in_act;    // activation read (finer one for POOL, coarser one for UPADD); -1 = field
    int skip_act;  // UPADD: the skip at this level
    int relu;
};
extern const LayerDef kNet[kLayers];

// one-channel strided field (solver field or dense sample batch)
struct FieldRef {
    void* p = nullptr;
    int is64 = 0;
    long long off = 0, sx = 1, sy = 0, sz = 0, sb = 0;
    const double* scale_sum = nullptr;  // scale = 1/sqrt(*scale_sum / scale_count) when set
    double scale_count = 1.0;
    float scale = 1.0f;
};

struct ActRef {
    uint16_t* p = nullptr;
    int B = 0, D = 0, H = 0, W = 0;
};

struct ConvArgs {
    int mode_in, mode_out;
    ActRef in, in2;   // in2: skip (UPADD) or ReLU mask (MASK); POOL reads the finer level through `in`
    FieldRef fin, fout;
    ActRef out;       // output activation; its extent is the level's (p unused for OUT_FIELD)
    ActRef pout;      // lift: 2x2x2-pooled copy of the output (p = null: none)
    const uint16_t* wl;
    const float* bias;
    const int* gate;
    const int* gate2;
    int fused;               // output layer with the lift recomputed (inference)
    const uint16_t* wl2;     // fused: lift operand
    const float* bias2;      // fused: lift bias
    const void* dot;         // fused: field of fout's precision and layout; z . dot per block into dpart
    double* dpart;
    // lift with the residual update: fin is r_prev; the staged value is
    // fin - alpha fap (fin while *skip), also written to fcur on the lift's own tile
    const void* fap;
    void* fcur;
    const double* alpha;
    const int* skip;
};

struct BufSet {
    int B = 0, n[3] = {0, 0, 0};  // batch, level-0 extent (x, y, z)
    bool train = false;
    int dims[3][3] = {};          // per level: D, H, W
    std::vector<uint16_t*> act, dact;
    uint16_t* dx[3] = {nullptr, nullptr, nullptr};  // backward-data scratch per level
    std::vector<uint16_t*> pooled;                  // pool(act i), written by act i's producer (or null)
    float *in = nullptr, *y = nullptr, *dy = nullptr, *ky = nullptr;
    double *g = nullptr, *lpart = nullptr, *coef = nullptr, *lossv = nullptr;
    float* part = nullptr;
    long long part_blocks = 0;
    int lblocks = 0;
    std::vector<void*> owned;
};

struct TrainCfg {
    int batch;
    int steps;
    float lr, lr_final;
    unsigned long long seed;
};

struct TrainStats {
    double loss_first, loss_last;
    int steps;
    double seconds;
};

class Net : public NNPrecond {
public:
    Net() = default;
    ~Net() override;
    int init(unsigned long long seed);  // own stream, He-initialised weights
    int apply(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
              const int* gate, const int* gate2, const double* rr) override;
    int apply_dot(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                  const int* gate, const int* gate2, const double* rr, const void* dot, double* dpart, int dcap,
                  int* nblocks, const NNResUpdate* upd) override;
    const char* error() const override { return err_.c_str(); }
    // average device time of `reps` launches of layer `layer` of the inference
    // pass over the region (run apply() on the same region first); algorithmic
    // HBM bytes and FLOPs of one launch.  dot (a field of out's precision and
    // layout, or null): the fused output layer also sums out . dot, as FCG runs it
    int time_layer(hipStream_t s, const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec,
                   const void* dot, const double* rr, int layer, int reps, double* ms, double* bytes, double* flops);

    int set_params(const float* host, long long n);
    int get_params(float* host, long long n);
    static long long effective_params();
    // dense host batches [B][n2][n1][n0] (float)
    int forward_host(const float* in, int B, const int n[3], float* out);
    int activation_host(int act, float* out);  // act of the last forward_host / loss_grad_host
    int loss_grad_host(const float* in, const double* G, int B, const int n[3], double* loss, float* grad);
    // training set of normalised residuals (device) and their energy norms G
    int dataset_reserve(long long npairs, const int n[3]);
    float* dataset_sample(long long i) const { return data_r_ + i * data_n_; }
    long long sample_size() const { return data_n_; }
    std::vector<double>& dataset_g() { return data_g_; }
    long long dataset_size() const { return data_pairs_; }
    int train(const TrainCfg& c, TrainStats* st);
    int fail(const char* fmt, ...);
    int hip(hipError_t e, const char* what);

private:
    int ensure(BufSet& bs, int B, const int n[3], bool train);
    void release(BufSet& bs);
    void forward(BufSet& bs, const FieldRef& fin, const FieldRef& fout, const int* gate, const int* gate2,
                 const void* dot = nullptr, double* dpart = nullptr, const NNResUpdate* upd = nullptr);
    ConvArgs layer_args(const BufSet& bs, int i, const FieldRef& fin, const FieldRef& fout, const int* gate,
                        const int* gate2) const;
    void launch_layer(const ConvArgs& a, int i);
    static void field_refs(const NNFieldGeom& g, void* out, int out_prec, const void* in, int in_prec, const double* rr,
                           FieldRef& fi, FieldRef& fo);
    void backward(BufSet& bs, const FieldRef& fin);
    void loss(BufSet& bs, int B, long long nvox);
    void pack_all();
    ActRef act(const BufSet& bs, int i) const;
    ActRef dact(const BufSet& bs, int i) const;
    ActRef level(const BufSet& bs, int lv, uint16_t* p) const;

    hipStream_t stream_ = nullptr;  // host-driven work; apply() runs on the caller's stream
    std::string err_;
    float *params_ = nullptr, *grads_ = nullptr, *adam_m_ = nullptr, *adam_v_ = nullptr;
    uint16_t *wl_ = nullptr, *wlt_ = nullptr;  // packed forward / backward-data weights, kLayers each
    uint16_t* wlo_ = nullptr;                  // output layer as tap-partial operands (fused inference)
    BufSet inf_, trn_;
    float* data_r_ = nullptr;
    long long data_n_ = 0, data_pairs_ = 0, data_cap_ = 0;
    int data_dims_[3] = {0, 0, 0};
    std::vector<double> data_g_;
    int adam_t_ = 0;
};

}  // namespace nn
}  // namespace pps
