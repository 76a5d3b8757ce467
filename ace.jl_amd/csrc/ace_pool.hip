// libacehip.so, 1p basis and pooling (product_1pbasis.jl:99-134): per
// neighbour the radial basis (orthpolys.jl:253-263 after
// distancetransforms.jl:73) and the m >= 0 complex spherical harmonics
// (sphericalharmonics.jl:43-51,167-197,379-403), pooled over the neighbours
// of each environment into A+[k][env]; and the directional derivative of the
// pooled values along per-neighbour weights (adjoint_EVAL_D).
#include "ace_common.hpp"

// k_pool: A+[k][env].  EPW environments per wave, one per group of
// LPE = 64 / EPW lanes (EPW = 2 possible when npool <= 32, so that small
// bases keep the lanes busy).  A group stages its environment's neighbours CH
// at a time (CH <= LPE) in its own LDS region ([slot][R_n | pad | Y_lm |
// species], Y at an even offset so that one 16-byte read fetches Y_lm); its
// lanes then turn into pooled functions and sum over the staged neighbours in
// order, holding the sums in registers across chunks.  The summation reads
// LDS once per function for R_n and once per Y_lm it needs: SHY (shared Y),
// the lane's KP functions k = pslot[hl + 64 p] share (species, l, m), so Y_lm
// is read once per neighbour (16 bytes instead of 16 KP; the launcher takes
// it for bases without species, whose summation has no per-function branch);
// otherwise lane hl takes k = hl + LPE p.
template <int L, int EPW, int KP, int CH, bool SHY>
__global__ __launch_bounds__(64, KP > 4 ? 4 : 5) void k_pool(RadialParams rp, int64_t nenv,
                                              const int32_t* __restrict__ off,
                                              const double* __restrict__ xyz, int64_t nnbr,
                                              const int8_t* __restrict__ species,
                                              const int32_t* __restrict__ pool, const int32_t* __restrict__ pslot,
                                              int npool, double2* __restrict__ Aplus, int64_t estride,
                                              int32_t* __restrict__ nbr_env) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int LPE = 64 / EPW;                 // lanes per environment
    static_assert(EPW == 1 || KP == 1, "two environments per wave take one pooled function per lane");
    static_assert(CH <= LPE, "a chunk holds at most one neighbour per lane of the group");
    static_assert(!SHY || EPW == 1, "shared-Y lane tables are laid out for one environment per wave");
    extern __shared__ double smem[];  // [EPW][CH][rstride]
    const int lane = threadIdx.x;
    const int hl = lane % LPE;
    const int nr = rp.nr;
    const int yoff = pool_yoff(nr);
    const int rstride = pool_rstride(nr, NP);  // doubles per neighbour slot
    const int64_t nblk = (nenv + EPW - 1) / EPW;
    const int64_t env = xcd_env(blockIdx.x, nblk) * EPW + lane / LPE;
    const bool valid = env < nenv;
    const int j0 = valid ? off[env] : 0, j1 = valid ? off[env + 1] : 0;
    double* const sgrp = smem + (lane / LPE) * CH * rstride;
    // this lane's pooled functions: index, species, slot offsets of R_n and Y_lm
    int kr[KP], ky[KP], kmu[KP];
    double ar[KP], ai[KP];
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        const int k = SHY ? pslot[hl + 64 * p] : (hl + LPE * p < npool ? hl + LPE * p : -1);
        kr[p] = 0;
        ky[p] = yoff;
        kmu[p] = -2;  // -2: no pooled function in this pass
        if (k >= 0) {
            kmu[p] = pool[4 * k];
            kr[p] = pool[4 * k + 1] - 1;
            ky[p] = yoff + 2 * idxp(pool[4 * k + 2], pool[4 * k + 3]);
        }
        ar[p] = ai[p] = 0.0;
    }
    if (SHY) {
        // a set's first member is valid whenever the lane has one
#pragma unroll
        for (int p = 1; p < KP; ++p) ky[p] = ky[0];
    }
    for (int jb = j0; jb < j1; jb += CH) {
        const int nch = min(CH, j1 - jb);
        if (hl < nch) {
            const int j = jb + hl;
            if (nbr_env) nbr_env[j] = (int32_t)env;  // neighbour -> environment, for the flat force kernel
            const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
            double* slot = sgrp + hl * rstride;
            Sph S = cart2spher(x, y, z);
            // radial basis
            double u = (1.0 + rp.r0) / (1.0 + S.r);
            double t = ipow(u, rp.tp);
            double p1 = rp.A[0] * fcut(rp, t);
            slot[0] = p1;
            double p0 = 0.0;
            if (nr > 1) {
                p0 = p1;
                p1 = (rp.A[1] * t + rp.B[1]) * p0;
                slot[1] = p1;
            }
            for (int n = 2; n < nr; ++n) {
                double p2 = (rp.A[n] * t + rp.B[n]) * p1 + rp.C[n] * p0;
                p0 = p1;
                p1 = p2;
                slot[n] = p2;
            }
            // Y_lm = e^{i m phi}/sqrt2 P_lm, m-outer: P_mm from P_{m-1,m-1}, then
            // the l recursion at fixed m (the operations of alp_eval, in an
            // order that keeps two degrees live), written straight to LDS
            double* ys = slot + yoff;
            const double isq2 = 0.7071067811865475;
            double epr = isq2, epi = 0.0;
            double Pmm = 0.3989422804014327;  // sqrt(1/(2 pi))
#pragma unroll
            for (int m = 0; m <= L; ++m) {
                if (m == 1) {
                    Pmm = -1.224744871391589 * S.sth * Pmm;
                } else if (m >= 2) {
                    Pmm = -sqrt(1.0 + 0.5 / m) * S.sth * Pmm;
                }
                if (m > 0) {
                    const double nr_ = epr * S.cphi - epi * S.sphi;
                    const double ni_ = epr * S.sphi + epi * S.cphi;
                    epr = nr_;
                    epi = ni_;
                }
                double Pa = 0.0, Pb = Pmm;
#pragma unroll
                for (int l = m; l <= L; ++l) {
                    double Pl;
                    if (l == m) {
                        Pl = Pmm;
                    } else if (l == m + 1) {
                        Pl = m == 0 ? S.cth * 1.7320508075688772 * 0.3989422804014327
                                    : S.cth * sqrt(2.0 * m + 3.0) * Pmm;
                    } else {
                        Pl = alp_a(l, m) * (S.cth * Pb + alp_b(l, m) * Pa);
                    }
                    if (l > m) {
                        Pa = Pb;
                        Pb = Pl;
                    }
                    ys[2 * idxp(l, m)] = (m == 0 ? isq2 : epr) * Pl;
                    ys[2 * idxp(l, m) + 1] = m == 0 ? 0.0 : epi * Pl;
                }
            }
            slot[yoff + 2 * NP] = species ? (double)species[j] : -1.0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // lanes over pooled functions, neighbours summed in order.  One
        // species: no branches, lanes past npool (kr = 0, ky = yoff) sum values
        // they never store (KP is the smallest count of passes that covers npool).
        if (species) {
            // pooled functions are grouped by species, so most passes skip a
            // neighbour of another species as a whole wave
            for (int jj = 0; jj < nch; ++jj) {
                const double* slot = sgrp + jj * rstride;
                const int smu = (int)slot[yoff + 2 * NP];
#pragma unroll
                for (int p = 0; p < KP; ++p) {
                    if (kmu[p] == -1 || kmu[p] == smu) {
                        const double R = slot[kr[p]];
                        const double2 Y = *reinterpret_cast<const double2*>(slot + ky[p]);
                        ar[p] = fma(R, Y.x, ar[p]);
                        ai[p] = fma(R, Y.y, ai[p]);
                    }
                }
            }
        } else {
            // two neighbours per iteration, one for 8 passes (registers)
#pragma unroll(KP > 4 ? 1 : 2)
            for (int jj = 0; jj < nch; ++jj) {
                const double* slot = sgrp + jj * rstride;
                double2 Y = *reinterpret_cast<const double2*>(slot + ky[0]);
#pragma unroll
                for (int p = 0; p < KP; ++p) {
                    if (!SHY && p > 0) Y = *reinterpret_cast<const double2*>(slot + ky[p]);
                    const double R = slot[kr[p]];
                    ar[p] = fma(R, Y.x, ar[p]);
                    ai[p] = fma(R, Y.y, ai[p]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!valid) return;
#pragma unroll
    for (int p = 0; p < KP; ++p)
        if (kmu[p] != -2) {
            const int k = SHY ? pslot[hl + 64 * p] : hl + LPE * p;
            Aplus[(int64_t)k * estride + env] = make_double2(ar[p], ai[p]);
        }
}


// k_poolw: dAw+[k][env] = sum_j w_j . grad_j phi_k over the pooled (m >= 0)
// 1p functions: the directional derivative of A+ along per-neighbour weights
// w[3][nnbr] (evaluator.jl:213-223, dAw = sum_j contract(w_j, dA[k, j]),
// restated on A+; m < 0 follows by conjugation since w is real).  One wave
// per environment; neighbours staged POOL_CH at a time in LDS as
// [R_n | w.grad R_n | Y_lm | w.grad Y_lm | species] (radial and Ylm
// gradients as in k_dphi), then lanes = pooled functions sum over them in
// order: dAw+[k] += w.grad R_n Y_lm + R_n w.grad Y_lm.
template <int L, int KP>
__global__ __launch_bounds__(64) void k_poolw(RadialParams rp, int64_t nenv, const int32_t* __restrict__ off,
                                              const double* __restrict__ xyz, const double* __restrict__ w,
                                              int64_t nnbr, const int8_t* __restrict__ species,
                                              const int32_t* __restrict__ pool, int npool,
                                              double2* __restrict__ dAw, int64_t estride) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int64_t env = blockIdx.x;
    const int nr = rp.nr;
    const int rstride = 2 * nr + 4 * NP + 1;
    if (env >= nenv) return;
    const int j0 = off[env], j1 = off[env + 1];
    int kr[KP], ky[KP], kmu[KP];
    double ar[KP], ai[KP];
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        const int k = lane + 64 * p;
        kr[p] = ky[p] = 0;
        kmu[p] = -2;
        if (k < npool) {
            kmu[p] = pool[4 * k];
            kr[p] = pool[4 * k + 1] - 1;
            ky[p] = 2 * nr + 2 * idxp(pool[4 * k + 2], pool[4 * k + 3]);
        }
        ar[p] = ai[p] = 0.0;
    }
    for (int jb = j0; jb < j1; jb += POOL_CH) {
        const int nch = min(POOL_CH, j1 - jb);
        if (lane < nch) {
            const int j = jb + lane;
            const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
            const double wx = w[j], wy = w[nnbr + j], wz = w[2 * nnbr + j];
            double* slot = smem + lane * rstride;
            const Sph S = cart2spher(x, y, z);
            double t, dtdr;
            transform_ed(rp, S.r, t, dtdr);
            const double wr = (wx * x + wy * y + wz * z) / S.r;  // w . rhat
            // R_n and dR_n/dt by the recursion of k_dphi; slot holds R_n and
            // w . grad R_n = dR_n/dt dt/dr (w . rhat)
            double Rm2 = 0.0, dRm2 = 0.0;
            double Rm1 = rp.A[0] * fcut(rp, t), dRm1 = rp.A[0] * fcut_d(rp, t);
            slot[0] = Rm1;
            slot[nr] = dRm1 * dtdr * wr;
            for (int n = 1; n < nr; ++n) {
                const double al = rp.A[n] * t + rp.B[n];
                double Rn = al * Rm1, dRn = al * dRm1 + rp.A[n] * Rm1;
                if (n >= 2) {
                    Rn += rp.C[n] * Rm2;
                    dRn += rp.C[n] * dRm2;
                }
                Rm2 = Rm1;
                dRm2 = dRm1;
                Rm1 = Rn;
                dRm1 = dRn;
                slot[n] = Rn;
                slot[nr + n] = dRn * dtdr * wr;
            }
            double P[NP], dP[NP];
            alp_eval_ed<L>(rp, S, P, dP);
            const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
            // w . grad = (fphi-part) + (ftheta-part) with the Cartesian rows of k_dphi
            const double gph = (-S.sphi * wx + S.cphi * wy) * rinv;
            const double gth = (S.cphi * S.cth * wx + S.sphi * S.cth * wy - S.sth * wz) * rinv;
            double* ys = slot + 2 * nr;
            const double isq2 = 0.7071067811865475;
            for (int l = 0; l <= L; ++l) {
                const int i = idxp(l, 0);
                ys[2 * i] = P[i] * isq2;
                ys[2 * i + 1] = 0.0;
                ys[2 * NP + 2 * i] = gth * dP[i] * isq2;
                ys[2 * NP + 2 * i + 1] = 0.0;
            }
            double epr = isq2, epi = 0.0;
            for (int m = 1; m <= L; ++m) {
                const double a = epr * S.cphi - epi * S.sphi;
                const double b = epr * S.sphi + epi * S.cphi;
                epr = a;
                epi = b;
                for (int l = m; l <= L; ++l) {
                    const int i = idxp(l, m);
                    const double pds = P[i], dpt = dP[i];
                    ys[2 * i] = epr * pds * S.sth;
                    ys[2 * i + 1] = epi * pds * S.sth;
                    // fphi = i m ep P~, ftheta = ep dP~
                    ys[2 * NP + 2 * i] = gph * (-m * epi * pds) + gth * (epr * dpt);
                    ys[2 * NP + 2 * i + 1] = gph * (m * epr * pds) + gth * (epi * dpt);
                }
            }
            slot[2 * nr + 4 * NP] = species ? (double)species[j] : -1.0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int jj = 0; jj < nch; ++jj) {
            const double* slot = smem + jj * rstride;
            const int smu = species ? (int)slot[2 * nr + 4 * NP] : -1;
#pragma unroll
            for (int p = 0; p < KP; ++p) {
                if (kmu[p] == -1 || kmu[p] == smu) {
                    const double Rv = slot[kr[p]], wdR = slot[nr + kr[p]];
                    const double Yr = slot[ky[p]], Yi = slot[ky[p] + 1];
                    const double dYr = slot[2 * NP + ky[p]], dYi = slot[2 * NP + ky[p] + 1];
                    ar[p] += wdR * Yr + Rv * dYr;
                    ai[p] += wdR * Yi + Rv * dYi;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
#pragma unroll
    for (int p = 0; p < KP; ++p)
        if (kmu[p] != -2) dAw[(int64_t)(lane + 64 * p) * estride + env] = make_double2(ar[p], ai[p]);
}

template <int L, int EPW, int KP, int CH, bool SHY>
static void launch_pool_shape(const dim3& grid, size_t lds, hipStream_t st, ace_ctx* ctx, int64_t nenv,
                              const int32_t* off, const double* xyz, int64_t nnbr, const int8_t* sp,
                              const Workspace& ws, int64_t estride, int32_t* nbr_env) {
    set_func_lds((const void*)k_pool<L, EPW, KP, CH, SHY>, lds);
    hipLaunchKernelGGL((k_pool<L, EPW, KP, CH, SHY>), grid, dim3(64), lds, st, radial_params(ctx), nenv, off, xyz,
                       nnbr, sp, ctx->pool.p, SHY ? ctx->pool_slot.p : nullptr, ctx->npool, ws.Aplus, estride,
                       nbr_env);
}

template <int L, int EPW, int KP, bool SHY>
static void launch_pool_ch(int ch, const dim3& grid, size_t lds, hipStream_t st, ace_ctx* ctx, int64_t nenv,
                           const int32_t* off, const double* xyz, int64_t nnbr, const int8_t* sp, const Workspace& ws,
                           int64_t estride, int32_t* nbr_env) {
    if (ch == 8)
        launch_pool_shape<L, EPW, KP, 8, SHY>(grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else if (ch == 32)
        launch_pool_shape<L, EPW, KP, 32, SHY>(grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else
        launch_pool_shape<L, EPW, KP, 16, SHY>(grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
}

template <int L, int KP>
static void launch_pool_kp(bool shy, int ch, const dim3& grid, size_t lds, hipStream_t st, ace_ctx* ctx, int64_t nenv,
                           const int32_t* off, const double* xyz, int64_t nnbr, const int8_t* sp, const Workspace& ws,
                           int64_t estride, int32_t* nbr_env) {
    if (shy)
        launch_pool_ch<L, 1, KP, true>(ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else
        launch_pool_ch<L, 1, KP, false>(ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
}

template <int L>
static int launch_pool(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off,
                       const double* xyz, int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws,
                       bool nbr_map) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    // environments per wave: two when a half-wave covers the pooled functions
    // (unless the context asks for one), else one; neighbours per chunk: the
    // context's choice, 16 by default
    const int epw = ctx->npool <= 32 && ctx->pool_epw != 1 ? 2 : 1;
    const int ch = ctx->pool_ch > 0 ? ctx->pool_ch : POOL_CH;
    const size_t lds = (size_t)epw * ch * pool_rstride(ctx->nr, NP) * sizeof(double);
    int32_t* nbr_env = nullptr;
    if (nbr_map && nnbr > 0) {
        if (!ws.nbr_env || ws.nbr_cap < (size_t)nnbr) return set_err("neighbour map buffer too small");
        nbr_env = ws.nbr_env;
    }
    if (!ws.Aplus) return set_err("no buffer for the pooled values");
    if (ctx->npool > 512) return set_err("pooled 1p basis too large (npool > 512)");
    const dim3 grid((unsigned)((nenv + epw - 1) / epw));
    const int8_t* sp = ctx->nspecies > 0 ? species : nullptr;
    // pooled functions per lane: 1 (two environments per wave), else the
    // fewest passes of 64 that cover npool: 2, 3, 4 or 8; the shared-Y lane
    // sets when the basis has no species and the sets fit the wave
    const int kp = pool_passes(ctx->npool);
    const bool shy = !sp && ctx->pool_shy_kp == kp;
    if (epw == 2)
        launch_pool_ch<L, 2, 1, false>(ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else if (kp == 2)
        launch_pool_kp<L, 2>(shy, ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else if (kp == 3)
        launch_pool_kp<L, 3>(shy, ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else if (kp == 4)
        launch_pool_kp<L, 4>(shy, ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    else
        launch_pool_kp<L, 8>(shy, ch, grid, lds, st, ctx, nenv, off, xyz, nnbr, sp, ws, estride, nbr_env);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int L>
static int launch_poolw(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                         const double* w, int64_t nnbr, const int8_t* species, int64_t estride) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    const size_t lds = (size_t)POOL_CH * (2 * ctx->nr + 4 * NP + 1) * sizeof(double);
    const int8_t* sp = ctx->nspecies > 0 ? species : nullptr;
    if (ctx->npool <= 256) {
        set_func_lds((const void*)k_poolw<L, 4>, lds);
        hipLaunchKernelGGL((k_poolw<L, 4>), dim3((unsigned)nenv), dim3(64), lds, st, radial_params(ctx), nenv, off,
                           xyz, w, nnbr, sp, ctx->pool.p, ctx->npool, ctx->Wplus.p, estride);
    } else {
        if (ctx->npool > 512) return set_err("pooled 1p basis too large (npool > 512)");
        set_func_lds((const void*)k_poolw<L, 8>, lds);
        hipLaunchKernelGGL((k_poolw<L, 8>), dim3((unsigned)nenv), dim3(64), lds, st, radial_params(ctx), nenv, off,
                           xyz, w, nnbr, sp, ctx->pool.p, ctx->npool, ctx->Wplus.p, estride);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_pool_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                    int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws, bool nbr_map) {
    switch (ctx->L) {
        case 0: return launch_pool<0>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 1: return launch_pool<1>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 2: return launch_pool<2>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 3: return launch_pool<3>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 4: return launch_pool<4>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 5: return launch_pool<5>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 6: return launch_pool<6>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 7: return launch_pool<7>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 8: return launch_pool<8>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        case 9: return launch_pool<9>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
        default: return launch_pool<10>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, nbr_map);
    }
}

int launch_poolw_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                     const double* w, int64_t nnbr, const int8_t* species, int64_t estride) {
    switch (ctx->L) {
        case 0: return launch_poolw<0>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 1: return launch_poolw<1>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 2: return launch_poolw<2>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 3: return launch_poolw<3>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 4: return launch_poolw<4>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 5: return launch_poolw<5>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 6: return launch_poolw<6>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 7: return launch_poolw<7>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 8: return launch_poolw<8>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        case 9: return launch_poolw<9>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        default: return launch_poolw<10>(ctx, st, nenv, off, xyz, w, nnbr, species, estride);
    }
}
