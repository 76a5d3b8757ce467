// libacehip.so, learned-surrogate support (no counterpart in the
// reference): the order <= 2 invariants of the basis as network inputs, the
// ensemble combination with fallback to plain ACE (the product DAG over the
// listed environments), and the chain rule through the features for forces.
#include "ace_common.hpp"

// ---------------------------------------------------------------------------
// learned-surrogate support: order <= 2 invariants, their gradient, fallback

struct FeatArgs {
    int nfeat, npool;
    const int32_t* feat_ptr;
    const int32_t* feat_f;
    const double* feat_coef;
    const double* cfeat;
    const int32_t* fgrad_ptr;
    const int32_t* fgrad_feat;
    const int32_t* fgrad_f;
    const double* fgrad_coef;
};

__device__ __forceinline__ double2 aplus_val(const double2* Aplus, int64_t estride, int f, int64_t env) {
    if (f < 0) return make_double2(1.0, 0.0);
    double2 v = Aplus[(int64_t)(f & 0x7fff) * estride + env];
    if (f & 0x8000) v.y = -v.y;
    return v;
}

// lanes = environments.  X[f][nenv] (FP64), Xn[env][160] (BF16, normalized),
// E2 = c0 + sum_f c_f X_f.
__global__ __launch_bounds__(64) void k_feat(FeatArgs a, int64_t nenv, const double2* __restrict__ Aplus,
                                             int64_t estride, const double* __restrict__ xmean,
                                             const double* __restrict__ xinvstd, double* __restrict__ X,
                                             __bf16* __restrict__ Xn, double* __restrict__ E2,
                                             const double* __restrict__ e0p) {
    const int64_t env = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (env >= nenv) return;
    double e = *e0p;
    for (int f0 = 0; f0 < 160; f0 += 8) {
        __bf16 xb[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int f = f0 + q;
            double v = 0.0;
            if (f < a.nfeat) {
                for (int z = a.feat_ptr[f]; z < a.feat_ptr[f + 1]; ++z) {
                    const double2 a1 = aplus_val(Aplus, estride, a.feat_f[2 * z], env);
                    const double2 a2 = aplus_val(Aplus, estride, a.feat_f[2 * z + 1], env);
                    v += a.feat_coef[z] * (a1.x * a2.x - a1.y * a2.y);
                }
                e += a.cfeat[f] * v;
                if (X) X[(int64_t)f * nenv + env] = v;
                xb[q] = (__bf16)(float)((v - xmean[f]) * xinvstd[f]);
            } else {
                xb[q] = (__bf16)0.0f;
            }
        }
        if (Xn) {
#pragma unroll
            for (int q = 0; q < 8; ++q) Xn[env * 160 + f0 + q] = xb[q];
        }
    }
    E2[env] = e;
}

__global__ void k_zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// E = E2 + ymean + ystd * mean_m y_m; fallback where the spread exceeds tau
__global__ void k_surr_combine(int64_t nenv, int M, const float* __restrict__ y, double ymean, double ystd,
                               double tau, const double* __restrict__ E2, double* __restrict__ E_out,
                               int32_t* __restrict__ fb_count, int32_t* __restrict__ fb_list) {
    const int64_t env = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (env >= nenv) return;
    double s = 0.0, s2 = 0.0;
    for (int m = 0; m < M; ++m) {
        const double v = y[(int64_t)m * nenv + env];
        s += v;
        s2 += v * v;
    }
    const double mean = s / M;
    const double var = fmax(s2 / M - mean * mean, 0.0);
    E_out[env] = E2[env] + ymean + ystd * mean;
    if (ystd * sqrt(var) > tau) {
        const int pos = atomicAdd(fb_count, 1);
        fb_list[pos] = (int32_t)env;
    }
}

// feature weights w_f = c_f + ystd * xinvstd_f * mean_m dX[m][env][f], then
// G[k] = sum_entries w_f coef value(other)   (lanes = environments)
// Adjoint of the accepted environments:
//   w_f = c_f + ystd * invstd_f * mean_m dy_m/dX_f
//   G[k] = sum_z w_{feat(z)} coef_z value(other_z)
// One workgroup per 64 environments.  dX [M][nenv][160] is read coalesced
// (consecutive threads = consecutive features of one environment) and the
// weights are kept in LDS as [f][65] doubles (padded: the transposed writes
// fall in different banks).  Waves then split the pooled index k; lanes are
// environments and read A+ coalesced from [k][env].
constexpr int FG_PAD = 65;
__global__ __launch_bounds__(256) void k_feat_grad(FeatArgs a, int64_t nenv, const double2* __restrict__ Aplus,
                                                   int64_t estride, int M, const float* __restrict__ dX,
                                                   double ystd, const double* __restrict__ xinvstd,
                                                   const int32_t* __restrict__ gslot_ptr,
                                                   const int32_t* __restrict__ gslot_list,
                                                   double2* __restrict__ Gbuf) {
    extern __shared__ double Wl[];  // [nfeat][FG_PAD]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t env0 = (int64_t)blockIdx.x * 64;
    const int nloc = (int)std::min<int64_t>(64, nenv - env0);
    const double invM = 1.0 / M;
    for (int idx = tid; idx < 64 * 160; idx += 256) {
        const int e = idx / 160, f = idx - e * 160;
        if (f < a.nfeat) {
            double sm = 0.0;
            if (e < nloc)
                for (int m = 0; m < M; ++m) sm += (double)dX[((int64_t)m * nenv + env0 + e) * 160 + f];
            Wl[f * FG_PAD + e] = a.cfeat[f] + ystd * xinvstd[f] * sm * invM;
        }
    }
    __syncthreads();
    const int64_t env = env0 + lane;
    const bool valid = lane < nloc;
    for (int k = wave; k < a.npool; k += 4) {
        double2 acc = make_double2(0.0, 0.0);
        for (int z = a.fgrad_ptr[k]; z < a.fgrad_ptr[k + 1]; ++z) {
            const double w = Wl[a.fgrad_feat[z] * FG_PAD + lane] * a.fgrad_coef[z];
            const double2 v = valid ? aplus_val(Aplus, estride, a.fgrad_f[z], env) : make_double2(0.0, 0.0);
            acc.x += w * v.x;
            acc.y += w * v.y;
        }
        if (valid) {
            const int s0 = gslot_ptr[k], s1 = gslot_ptr[k + 1];
            for (int q = s0; q < s1; ++q)
                Gbuf[(int64_t)gslot_list[q] * estride + env] = q == s0 ? acc : make_double2(0.0, 0.0);
        }
    }
}

// ---- learned surrogate

static FeatArgs feat_args(ace_ctx* ctx) {
    FeatArgs a;
    a.nfeat = ctx->nfeat;
    a.npool = ctx->npool;
    a.feat_ptr = ctx->feat_ptr.p;
    a.feat_f = ctx->feat_f.p;
    a.feat_coef = ctx->feat_coef.p;
    a.cfeat = ctx->cfeat.p;
    a.fgrad_ptr = ctx->fgrad_ptr.p;
    a.fgrad_feat = ctx->fgrad_feat.p;
    a.fgrad_f = ctx->fgrad_f.p;
    a.fgrad_coef = ctx->fgrad_coef.p;
    return a;
}

extern "C" int ace_surr_features(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                 const int8_t* species, const double* xmean, const double* xinvstd, double* X,
                                 void* Xn, double* E2, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (ctx->prop != ACE_PROP_INVARIANT || ctx->nfeat < 1) return set_err("surrogate features need an Invariant basis");
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
    if (nenv == 0) return 0;
    if (!xmean || !xinvstd || !E2) return set_err("ace_surr_features: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    // the neighbour map is written for the force kernel of the following call
    if (ctx->flat_force && nnbr > 0 && ctx->nbr_env.ensure(nnbr)) return -1;
    if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), ctx->flat_force))
        return -1;
    hipLaunchKernelGGL(k_feat, dim3((unsigned)((nenv + 63) / 64)), dim3(64), 0, st, feat_args(ctx), nenv,
                       ctx->Aplus.p, estride, xmean, xinvstd, X, (__bf16*)Xn, E2, ctx->e0v.p);
    HIP_TRY(hipGetLastError());
    ctx->surr_nenv = nenv;
    return 0;
}

extern "C" int ace_surr_energy_forces(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                      const int8_t* species, int nmemb, const float* y, const float* dX,
                                      double ymean, double ystd, const double* xinvstd, double tau,
                                      const double* E2, double* E_out, double* dE, int32_t* fb_count,
                                      int32_t* fb_list, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (ctx->prop != ACE_PROP_INVARIANT || ctx->nfeat < 1) return set_err("surrogate needs an Invariant basis");
    if (nenv == 0) return 0;
    if (ctx->surr_nenv != nenv) return set_err("ace_surr_energy_forces must follow ace_surr_features on the same batch");
    if (nmemb < 1 || nmemb > 16 || !y || !E2 || !E_out || !fb_count || !fb_list || !xinvstd)
        return set_err("ace_surr_energy_forces: bad arguments");
    if (dE && !dX) return set_err("forces need the input gradients dX");
    hipStream_t st = (hipStream_t)stream;
    const int64_t ntile = (nenv + 63) / 64;
    const int64_t estride = ntile * 64;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, fb_count, 1);
    hipLaunchKernelGGL(k_surr_combine, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, nmemb, y, ymean,
                       ystd, tau, E2, E_out, fb_count, fb_list);
    HIP_TRY(hipGetLastError());
    const bool forces = dE != nullptr;
    if (forces) {
        if (ctx->Gbuf.ensure((size_t)std::max(1, ctx->ngslots) * estride)) return -1;
        const size_t fg_lds = (size_t)ctx->nfeat * FG_PAD * sizeof(double);
        set_func_lds((const void*)k_feat_grad, fg_lds);
        hipLaunchKernelGGL(k_feat_grad, dim3((unsigned)ntile), dim3(256), fg_lds, st, feat_args(ctx), nenv,
                           ctx->Aplus.p, estride, nmemb, dX, ystd, xinvstd, ctx->gslot_ptr.p, ctx->gslot_list.p,
                           ctx->Gbuf.p);
        HIP_TRY(hipGetLastError());
    }
    // plain ACE for the listed environments (energy, and G for forces)
    if (launch_dag_listed(ctx, st, nenv, estride, forces, E_out, fb_list, fb_count)) return -1;
    if (forces && nnbr > 0 && launch_force_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), dE))
        return -1;
    return 0;
}
