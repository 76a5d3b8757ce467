// libacehip.so, host-memory variants of the evaluator-slot calls, for
// callers that hold their configurations in host arrays and have no device
// memory of their own (a Julia ccall with Base arrays: the evaluate,
// grad_config, grad_params, evaluate_d and adjoint_EVAL_D of
// linearmodel.jl:107-134 and evaluator.jl:117-244).  The offsets and species
// are checked on the host, the inputs are copied to the context's staging
// buffers on the context's own stream, the device call runs there, and the
// outputs are copied back before the call returns.  Layouts are those of the
// device entry points.
#include "ace_common.hpp"

namespace {

// Host checks of env_offsets (start at 0, non-decreasing) and species.
int check_host_inputs(ace_ctx* ctx, int64_t nenv, const int32_t* off, const int8_t* species, int64_t* nnbr) {
    if (!ctx) return set_err("null context");
    if (nenv < 0) return set_err("nenv < 0");
    if (nenv > 0 && !off) return set_err("null env_offsets");
    *nnbr = 0;
    if (nenv == 0) return 0;
    if (off[0] != 0) return set_err("env_offsets must start at 0 and be non-negative");
    for (int64_t e = 0; e < nenv; ++e)
        if (off[e + 1] < off[e]) return set_err("env_offsets must be non-decreasing");
    *nnbr = off[nenv];
    if (ctx->nspecies > 0 && *nnbr > 0) {
        if (!species) return set_err("species array required for a categorical 1p basis");
        for (int64_t j = 0; j < *nnbr; ++j)
            if (species[j] < 0 || species[j] >= ctx->nspecies)
                return set_err("species index out of range (unknown category)");
    }
    return 0;
}

// Stage the inputs, run fn(device pointers, stream), copy the outputs back.
template <typename F>
int host_call(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz, const int8_t* species,
              const double* w, double* out1, size_t n1, double* out2, size_t n2, F fn) {
    int64_t nnbr = 0;
    if (check_host_inputs(ctx, nenv, off, species, &nnbr)) return -1;
    if (nenv == 0) return 0;
    if (nnbr > 0 && !xyz) return set_err("null xyz");
    if (!out1 || (n2 > 0 && !out2)) return set_err("null output");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->io_stream) HIP_TRY(hipStreamCreateWithFlags(&ctx->io_stream, hipStreamNonBlocking));
    hipStream_t st = ctx->io_stream;
    if (ctx->io_off.ensure((size_t)nenv + 1) || ctx->io_xyz.ensure(std::max<size_t>(1, 3 * (size_t)nnbr)) ||
        ctx->io_out1.ensure(std::max<size_t>(1, n1)) || (n2 > 0 && ctx->io_out2.ensure(n2)))
        return -1;
    HIP_TRY(hipMemcpyAsync(ctx->io_off.p, off, sizeof(int32_t) * (nenv + 1), hipMemcpyHostToDevice, st));
    if (nnbr > 0)
        HIP_TRY(hipMemcpyAsync(ctx->io_xyz.p, xyz, sizeof(double) * 3 * nnbr, hipMemcpyHostToDevice, st));
    const int8_t* dsp = nullptr;
    if (ctx->nspecies > 0 && nnbr > 0) {
        if (ctx->io_sp.ensure((size_t)nnbr)) return -1;
        HIP_TRY(hipMemcpyAsync(ctx->io_sp.p, species, (size_t)nnbr, hipMemcpyHostToDevice, st));
        dsp = ctx->io_sp.p;
    }
    const double* dw = nullptr;
    if (w && nnbr > 0) {
        if (ctx->io_w.ensure(3 * (size_t)nnbr)) return -1;
        HIP_TRY(hipMemcpyAsync(ctx->io_w.p, w, sizeof(double) * 3 * nnbr, hipMemcpyHostToDevice, st));
        dw = ctx->io_w.p;
    }
    // checked above: skip the device entry points' own host copy of the inputs
    const bool validate = ctx->validate;
    ctx->validate = false;
    const int rc = fn(ctx->io_off.p, ctx->io_xyz.p, dsp, dw, ctx->io_out1.p, n2 > 0 ? ctx->io_out2.p : nullptr,
                      (void*)st);
    ctx->validate = validate;
    if (rc) return -1;
    HIP_TRY(hipMemcpyAsync(out1, ctx->io_out1.p, sizeof(double) * n1, hipMemcpyDeviceToHost, st));
    if (n2 > 0) HIP_TRY(hipMemcpyAsync(out2, ctx->io_out2.p, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int64_t host_nnbr(int64_t nenv, const int32_t* off) { return (nenv > 0 && off) ? off[nenv] : 0; }

}  // namespace

extern "C" int ace_energy_host(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                               const int8_t* species, double* E_out) {
    if (!ctx) return set_err("null context");
    const size_t nc = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    return host_call(ctx, nenv, off, xyz, species, nullptr, E_out, (size_t)nenv * nc, nullptr, 0,
                     [&](const int32_t* o, const double* x, const int8_t* s, const double*, double* E, double*,
                         void* st) { return ace_energy(ctx, nenv, o, x, s, E, st); });
}

extern "C" int ace_energy_forces_host(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                      const int8_t* species, double* E_out, double* dE_dr) {
    if (!ctx) return set_err("null context");
    const size_t nc = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    const size_t nnbr = (size_t)std::max<int64_t>(0, host_nnbr(nenv, off));
    return host_call(ctx, nenv, off, xyz, species, nullptr, E_out, (size_t)nenv * nc, dE_dr, 3 * nc * nnbr,
                     [&](const int32_t* o, const double* x, const int8_t* s, const double*, double* E, double* g,
                         void* st) {
                         // zero neighbours: no gradient buffer to fill
                         return g ? ace_energy_forces(ctx, nenv, o, x, s, E, g, st) : ace_energy(ctx, nenv, o, x, s, E, st);
                     });
}

extern "C" int ace_basis_B_host(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                const int8_t* species, double* B_out) {
    if (!ctx) return set_err("null context");
    const size_t nc = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    return host_call(ctx, nenv, off, xyz, species, nullptr, B_out, (size_t)nenv * ctx->nB * nc, nullptr, 0,
                     [&](const int32_t* o, const double* x, const int8_t* s, const double*, double* B, double*,
                         void* st) { return ace_basis_B(ctx, nenv, o, x, s, B, st); });
}

extern "C" int ace_basis_dB_host(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                 const int8_t* species, double* dB_out) {
    if (!ctx) return set_err("null context");
    const size_t nc = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    int64_t nnbr = 0;
    if (check_host_inputs(ctx, nenv, off, species, &nnbr)) return -1;
    if (nnbr == 0) return 0;  // no neighbours: an empty dB
    return host_call(ctx, nenv, off, xyz, species, nullptr, dB_out, (size_t)nnbr * ctx->nB * 3 * nc, nullptr, 0,
                     [&](const int32_t* o, const double* x, const int8_t* s, const double*, double* dB, double*,
                         void* st) { return ace_basis_dB(ctx, nenv, o, x, s, dB, st); });
}

extern "C" int ace_adjoint_EVAL_D_host(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                       const int8_t* species, const double* w, double* dB_out) {
    if (!ctx) return set_err("null context");
    const int64_t nnbr = host_nnbr(nenv, off);
    if (nnbr > 0 && !w) return set_err("ace_adjoint_EVAL_D_host: null w");
    return host_call(ctx, nenv, off, xyz, species, w, dB_out, (size_t)nenv * ctx->nB, nullptr, 0,
                     [&](const int32_t* o, const double* x, const int8_t* s, const double* dw, double* dB, double*,
                         void* st) { return ace_adjoint_EVAL_D(ctx, nenv, o, x, s, dw, dB, st); });
}
