// pps_nn.hpp -- interface of the learned preconditioner held by a context.
#pragma once

namespace pps {

struct Ctx;

struct NNPrecond {
    virtual ~NNPrecond() = default;
    // out = M(in) on the solver region (zero elsewhere); kernels gated on *gate
    virtual int apply(Ctx& C, void* out, int out_prec, void* in, int in_prec, const int* gate) = 0;
    virtual int fallback_switches() const { return 0; }
};

}  // namespace pps
