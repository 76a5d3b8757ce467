// libacehip.so — MI355X (gfx950) basis-evaluation path of the Atomic Cluster
// Expansion with ACE.jl v0.12.46 semantics.  FP64 throughout.
//
// Pipeline of an energy (+ forces) call:
//   k_pool   : one wave per environment (two, one per half-wave, for
//              bases with at most 32 pooled values), lanes = neighbours.  Radial basis
//              (orthpolys.jl:253-263 after distancetransforms.jl:73) and the
//              m >= 0 complex spherical harmonics (sphericalharmonics.jl:
//              43-51,167-197,379-403) per neighbour, staged in LDS, then
//              lanes = pooled 1p functions sum over neighbours in order
//              (product_1pbasis.jl:99-134) -> A+[k][env] in HBM.
//   k_dag    : 64 environments per workgroup (lanes = environments), the
//              A+ tile staged in LDS, every wave interprets its share of the
//              run programs (depth-first product DAG: prefix products in
//              registers, leaves streamed through scalar loads).  Computes
//              the site energy sum c~ Re AA (evaluator.jl:121-147) and the
//              adjoint polynomials G[k] (evaluator.jl:161-185 restated on A+).
//   k_force  : one wave per environment, lanes = neighbours; recomputes the
//              radial/Ylm values and gradients (orthpolys.jl:287-306,
//              sphericalharmonics.jl:212-267,410-443,60-65) and contracts
//              them with G (evaluator.jl:187-192): +dE/dr_j.
// Basis-level kernels (A, AA, B = A2B AA, dB) follow pibasis.jl:265-275,
// 338-432 and symmbasis.jl:248-336.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/acehip.h"
#include "host_util.hpp"
#include "pack_host.hpp"

#define ACE_LMAX 10
#define ACE_NPMAX ((ACE_LMAX + 1) * (ACE_LMAX + 2) / 2)
#define ACE_NRMAX 32
#define ACE_CONJ 0x8000

// ---------------------------------------------------------------------------
// universal ALP recursion coefficients (sphericalharmonics.jl:146-159)

__constant__ double c_alpA[ACE_NPMAX];
__constant__ double c_alpB[ACE_NPMAX];

__host__ __device__ constexpr int idxp(int l, int m) { return m + l * (l + 1) / 2; }

struct RadialParams {
    const double* A;
    const double* B;
    const double* C;
    double r0, tl, tr;
    int nr, tp, pl, pr;
};

__device__ __forceinline__ double ipow(double x, int p) {
    double y = 1.0;
    for (int i = 0; i < p; ++i) y *= x;
    return y;
}

// t(r) and dt/dr for t = ((1 + r0)/(1 + r))^tp
__device__ __forceinline__ void transform_ed(const RadialParams& rp, double r, double& t,
                                             double& dt) {
    double u = (1.0 + rp.r0) / (1.0 + r);
    double du = -(1.0 + rp.r0) / ((1.0 + r) * (1.0 + r));
    t = ipow(u, rp.tp);
    dt = (double)rp.tp * ipow(u, rp.tp - 1) * du;
}

__device__ __forceinline__ double fcut(const RadialParams& rp, double t) {
    if ((rp.pl > 0 && t < rp.tl) || (rp.pr > 0 && t > rp.tr)) return 0.0;
    return ipow(t - rp.tl, rp.pl) * ipow(t - rp.tr, rp.pr);
}

__device__ __forceinline__ double fcut_d(const RadialParams& rp, double t) {
    if ((rp.pl > 0 && t < rp.tl) || (rp.pr > 0 && t > rp.tr)) return 0.0;
    double d = 0.0;
    if (rp.pl > 0) d += rp.pl * ipow(t - rp.tl, rp.pl - 1) * ipow(t - rp.tr, rp.pr);
    if (rp.pr > 0) d += ipow(t - rp.tl, rp.pl) * rp.pr * ipow(t - rp.tr, rp.pr - 1);
    return d;
}

// Spherical coordinates; cos/sin of the azimuth from x/rho, y/rho.
struct Sph {
    double r, cphi, sphi, cth, sth;
};

__device__ __forceinline__ Sph cart2spher(double x, double y, double z) {
    Sph s;
    s.r = sqrt(x * x + y * y + z * z);
    double rho = sqrt(x * x + y * y);
    if (rho > 0.0) {
        s.cphi = x / rho;
        s.sphi = y / rho;
    } else {
        s.cphi = 1.0;
        s.sphi = 0.0;
    }
    s.cth = z / s.r;
    s.sth = rho / s.r;
    return s;
}

// ALP values, sphericalharmonics.jl:167-197
template <int L>
__device__ __forceinline__ void alp_eval(const Sph& S, double* P) {
    double temp = 0.3989422804014327;  // sqrt(1/(2 pi))
    P[0] = temp;
    if (L == 0) return;
    P[idxp(1, 0)] = S.cth * 1.7320508075688772 * temp;
    temp = -1.224744871391589 * S.sth * temp;
    P[idxp(1, 1)] = temp;
#pragma unroll
    for (int l = 2; l <= L; ++l) {
#pragma unroll
        for (int m = 0; m <= l - 2; ++m)
            P[idxp(l, m)] = c_alpA[idxp(l, m)] *
                            (S.cth * P[idxp(l - 1, m)] + c_alpB[idxp(l, m)] * P[idxp(l - 2, m)]);
        P[idxp(l, l - 1)] = S.cth * sqrt(2.0 * (l - 1) + 3.0) * temp;
        temp = -sqrt(1.0 + 0.5 / l) * S.sth * temp;
        P[idxp(l, l)] = temp;
    }
}

// ALP values (P/sin(theta) for m >= 1) and d/dtheta, sphericalharmonics.jl:212-267
template <int L>
__device__ __forceinline__ void alp_eval_ed(const Sph& S, double* P, double* dP) {
    double temp = 0.3989422804014327;
    P[0] = temp;
    dP[0] = 0.0;
    if (L == 0) return;
    const double s3 = 1.7320508075688772;
    P[idxp(1, 0)] = S.cth * s3 * temp;
    dP[idxp(1, 0)] = -S.sth * s3 * temp;
    double temp1 = -1.224744871391589 * temp;
    double temp_d = -1.224744871391589 * (S.cth * temp);
    P[idxp(1, 1)] = temp1;
    dP[idxp(1, 1)] = temp_d;
    const double st2 = S.sth * S.sth;
#pragma unroll
    for (int l = 2; l <= L; ++l) {
        {
            const int i = idxp(l, 0), i1 = idxp(l - 1, 0), i2 = idxp(l - 2, 0);
            P[i] = c_alpA[i] * (S.cth * P[i1] + c_alpB[i] * P[i2]);
            dP[i] = c_alpA[i] * (-S.sth * P[i1] + S.cth * dP[i1] + c_alpB[i] * dP[i2]);
        }
#pragma unroll
        for (int m = 1; m <= l - 2; ++m) {
            const int i = idxp(l, m), i1 = idxp(l - 1, m), i2 = idxp(l - 2, m);
            P[i] = c_alpA[i] * (S.cth * P[i1] + c_alpB[i] * P[i2]);
            dP[i] = c_alpA[i] * (-st2 * P[i1] + S.cth * dP[i1] + c_alpB[i] * dP[i2]);
        }
        const double c = sqrt(2.0 * (l - 1) + 3.0);
        P[idxp(l, l - 1)] = c * S.cth * temp1;
        dP[idxp(l, l - 1)] = c * (-st2 * temp1 + S.cth * temp_d);
        const double f = sqrt(1.0 + 0.5 / l);
        const double nt1 = -f * S.sth * temp1;
        const double ntd = -f * (S.cth * temp1 * S.sth + S.sth * temp_d);
        temp1 = nt1;
        temp_d = ntd;
        P[idxp(l, l)] = temp1;
        dP[idxp(l, l)] = temp_d;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
// Give every XCD a contiguous range of environments so that neighbouring
// environments, which share cache lines of the [k][env] arrays, meet in the
// same L2.  Bijection on [0, n).
__device__ __forceinline__ int64_t xcd_env(int64_t b, int64_t n) {
    const int64_t x = b & 7, i = b >> 3;
    const int64_t q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------------------
// k_pool: A+[k][env].  EPW environments per wave, one per group of
// LPE = 64 / EPW lanes (EPW = 2 when npool <= 32, so that small bases keep
// the lanes busy).  A group stages its environment's neighbours POOL_CH at a
// time in its own LDS region ([slot][R_n | Y_lm | species]); its lanes then
// turn into pooled functions k = hl + LPE p (p < KP) and sum over the staged
// neighbours in order, holding the sums in registers across chunks.
constexpr int POOL_CH = 16;
constexpr int POOL_KP = 4;

template <int L, int EPW>
__global__ __launch_bounds__(64) void k_pool(RadialParams rp, int64_t nenv,
                                              const int32_t* __restrict__ off,
                                              const double* __restrict__ xyz, int64_t nnbr,
                                              const int8_t* __restrict__ species,
                                              const int32_t* __restrict__ pool, int npool,
                                              double2* __restrict__ Aplus, int64_t estride,
                                              int32_t* __restrict__ nbr_env) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int LPE = 64 / EPW;                 // lanes per environment
    constexpr int KP = EPW == 1 ? POOL_KP : 1;    // pooled functions per lane
    extern __shared__ double smem[];  // [EPW][POOL_CH][rstride]
    const int lane = threadIdx.x;
    const int hl = lane % LPE;
    const int nr = rp.nr;
    const int rstride = nr + 2 * NP + 1;  // doubles per neighbour slot
    const int64_t nblk = (nenv + EPW - 1) / EPW;
    const int64_t env = xcd_env(blockIdx.x, nblk) * EPW + lane / LPE;
    const bool valid = env < nenv;
    const int j0 = valid ? off[env] : 0, j1 = valid ? off[env + 1] : 0;
    double* const sgrp = smem + (lane / LPE) * POOL_CH * rstride;
    // this lane's pooled functions k = hl + LPE p: slot offsets of R_n and Y_lm
    int kr[KP], ky[KP], kmu[KP];
    double ar[KP], ai[KP];
#pragma unroll
    for (int p = 0; p < KP; ++p) {
        const int k = hl + LPE * p;
        kr[p] = ky[p] = 0;
        kmu[p] = -2;  // -2: no pooled function in this pass
        if (k < npool) {
            kmu[p] = pool[4 * k];
            kr[p] = pool[4 * k + 1] - 1;
            ky[p] = nr + 2 * idxp(pool[4 * k + 2], pool[4 * k + 3]);
        }
        ar[p] = ai[p] = 0.0;
    }
    for (int jb = j0; jb < j1; jb += POOL_CH) {
        const int nch = min(POOL_CH, j1 - jb);
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
            double* ys = slot + nr;
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
                        const int i = idxp(l, m);
                        Pl = c_alpA[i] * (S.cth * Pb + c_alpB[i] * Pa);
                    }
                    if (l > m) {
                        Pa = Pb;
                        Pb = Pl;
                    }
                    ys[2 * idxp(l, m)] = (m == 0 ? isq2 : epr) * Pl;
                    ys[2 * idxp(l, m) + 1] = m == 0 ? 0.0 : epi * Pl;
                }
            }
            slot[nr + 2 * NP] = species ? (double)species[j] : -1.0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // lanes over pooled functions, neighbours summed in order
        for (int jj = 0; jj < nch; ++jj) {
            const double* slot = sgrp + jj * rstride;
            const int smu = species ? (int)slot[nr + 2 * NP] : -1;
#pragma unroll
            for (int p = 0; p < KP; ++p) {
                if (kmu[p] == -1 || kmu[p] == smu) {
                    const double R = slot[kr[p]];
                    ar[p] += R * slot[ky[p]];
                    ai[p] += R * slot[ky[p] + 1];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!valid) return;
#pragma unroll
    for (int p = 0; p < KP; ++p)
        if (kmu[p] != -2) Aplus[(int64_t)(hl + LPE * p) * estride + env] = make_double2(ar[p], ai[p]);
}

// ---------------------------------------------------------------------------
// k_dag: run-program interpreter, lanes = environments

struct DagArgs {
    const int32_t* runs;       // [nruns][7]
    const int32_t* leaf_fid;   // [nleaves]
    const double* leaf_coef;   // [nleaves][ncoef]
    const double4* vleaf;      // EuclideanVector: [nleaves][2] decoded leaves (factor id, then 3 complex coefs)
    const int32_t* tasks;      // [ntasks][3]
    const double* task_const;  // [ntasks][2 | 6]
    const int32_t* wtask_ptr;  // [W+1]
    const int32_t* wtask_list;
    int npool;
    int W;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int ENVS>
__device__ __forceinline__ double2 lds_factor(const double2* As, int f, int lane) {
    double2 v = As[(f & 0x7fff) * ENVS + lane];
    if (f & ACE_CONJ) v.y = -v.y;
    return v;
}

// PROP 0: real coefficient, output Re(acc) (energy) or acc (G slot).
// PROP 1: 3 complex coefficients, output Re(acc_p) or acc_p.
// ENVS = 64: lanes = 64 environments, one task stream per wave.
// ENVS = 32: lanes 0-31 and 32-63 are the same 32 environments running two
//            task streams (halves the LDS tile for large pooled bases).
template <int PROP, int ENVS>
__global__ __launch_bounds__(1024) void k_dag(DagArgs a, int64_t nenv,
                                              const double2* __restrict__ Aplus,
                                              int64_t estride, double* __restrict__ E_out,
                                              double2* __restrict__ Gbuf, double e0r,
                                              const double* __restrict__ e0v) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    constexpr int SPW = 64 / ENVS;  // task streams per wave
    extern __shared__ double2 As[];  // [npool][ENVS]
    const int lane = threadIdx.x & 63;
    const int el = lane & (ENVS - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    const int stream = wave * SPW + lane / ENVS;
    const int64_t env = (int64_t)blockIdx.x * ENVS + el;
    const bool valid = env < nenv;
    for (int k = wave * SPW + lane / ENVS; k < a.npool; k += nwaves * SPW)
        As[k * ENVS + el] = valid ? Aplus[(int64_t)k * estride + env] : make_double2(0.0, 0.0);
    __syncthreads();

    double ep[NC];
#pragma unroll
    for (int p = 0; p < NC; ++p) ep[p] = 0.0;

    const int tb = a.wtask_ptr[stream], te = a.wtask_ptr[stream + 1];
    for (int ti = tb; ti < te; ++ti) {
        const int t = a.wtask_list[ti];
        const int rb = a.tasks[5 * t], re = a.tasks[5 * t + 1], slot = a.tasks[5 * t + 2];
        double2 acc[NC];
#pragma unroll
        for (int p = 0; p < NC; ++p)
            acc[p] = make_double2(a.task_const[(2 * NC) * t + 2 * p],
                                  a.task_const[(2 * NC) * t + 2 * p + 1]);
        double2 P1 = make_double2(1.0, 0.0), P2 = P1, P3 = P1;
        for (int r = rb; r < re; ++r) {
            const int* run = a.runs + 7 * r;
            const int d = run[0], j = run[1];
            const int nl = run[5], lo = run[6];
            if (j <= 1 && d >= 1) P1 = lds_factor<ENVS>(As, run[2], el);
            if (j <= 2 && d >= 2) P2 = cmul(P1, lds_factor<ENVS>(As, run[3], el));
            if (j <= 3 && d >= 3) P3 = cmul(P2, lds_factor<ENVS>(As, run[4], el));
            const double2 Pd = d == 0 ? make_double2(1.0, 0.0) : d == 1 ? P1 : d == 2 ? P2 : P3;
            double2 S[NC];
#pragma unroll
            for (int p = 0; p < NC; ++p) S[p] = make_double2(0.0, 0.0);
            if (PROP == 0) {
                for (int i = 0; i < nl; ++i) {
                    const int f = a.leaf_fid[lo + i];
                    const double2 v = lds_factor<ENVS>(As, f, el);
                    const double c = a.leaf_coef[lo + i];
                    S[0].x += c * v.x;
                    S[0].y += c * v.y;
                }
            } else {
                // one 64-byte decoded leaf per iteration: {factor id, 0, c0, c1}, {c2, c3, c4, c5};
                // unrolled by 2 so that consecutive leaf fetches overlap
#pragma unroll 2
                for (int i = 0; i < nl; ++i) {
                    const double4 h = a.vleaf[2 * (int64_t)(lo + i)], q = a.vleaf[2 * (int64_t)(lo + i) + 1];
                    const double2 v = lds_factor<ENVS>(As, (int)__double2loint(h.x), el);
                    const double c[6] = {h.z, h.w, q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        S[p].x += c[2 * p] * v.x - c[2 * p + 1] * v.y;
                        S[p].y += c[2 * p] * v.y + c[2 * p + 1] * v.x;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < NC; ++p) {
                acc[p].x += Pd.x * S[p].x - Pd.y * S[p].y;
                acc[p].y += Pd.x * S[p].y + Pd.y * S[p].x;
            }
        }
        if (slot < 0) {
#pragma unroll
            for (int p = 0; p < NC; ++p) ep[p] += acc[p].x;
        } else if (valid) {
#pragma unroll
            for (int p = 0; p < NC; ++p) Gbuf[((int64_t)slot * NC + p) * estride + env] = acc[p];
        }
    }
    // reduce energy partials over streams (reuse the A tile region)
    __syncthreads();
    double* red = reinterpret_cast<double*>(As);
#pragma unroll
    for (int p = 0; p < NC; ++p) red[(stream * NC + p) * ENVS + el] = ep[p];
    __syncthreads();
    if (wave == 0 && lane < ENVS && valid && E_out) {
#pragma unroll
        for (int p = 0; p < NC; ++p) {
            double sum = PROP == 0 ? e0r : e0v[p];
            for (int w = 0; w < nwaves * SPW; ++w) sum += red[(w * NC + p) * ENVS + el];
            E_out[env * NC + p] = sum;
        }
    }
}

// ---------------------------------------------------------------------------
// k_dagr: the same programs as 64-byte decoded records (Invariant).  A
// record holds one run (or the continuation of a long run): up to four leaves
// with their coefficients and the prefix levels it recomputes, all as ready
// LDS byte offsets, so the loop spends few scalar instructions per record.

struct DagRecArgs {
    const uint32_t* recs;      // [nrec + 1][16]
    const int32_t* tasks;      // [ntasks][5]
    const double* task_const;  // [ntasks][2]
    const int32_t* wtask_ptr;  // [parts * W + 1]
    const int32_t* wtask_list;
    int npool;
    int W;
    int parts;                 // workgroups per tile, each running 1/parts of the tasks
    // tiles [ntile_lo, ntile) run parts + 1 workgroups each, with these task
    // lists: shorter workgroups that fill the grid's last round
    int64_t ntile_lo;
    const int32_t* wtask_ptr_hi;  // [(parts + 1) * W + 1]
    const int32_t* wtask_list_hi;
};

// Streaming accesses: the A+ tile is read once and G written once per
// environment; the non-temporal hint keeps them from evicting the program
// records, which every tile re-reads, from L2.
__device__ __forceinline__ double2 nt_load2(const double2* p) {
    return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ __forceinline__ void nt_store2(double2* p, double2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

__device__ __forceinline__ double flip_sign_if(double c, uint32_t f) {
    // negate c when the conjugation bit (0x8000) of factor id f is set
    return __hiloint2double(__double2hiint(c) ^ (int)((f & 0x8000u) << 16), __double2loint(c));
}

__device__ __forceinline__ double2 lds_conj(const double2* As, uint32_t f, int lane) {
    double2 v = As[(f & 0x7fffu) * 64 + lane];
    v.y = flip_sign_if(v.y, f);
    return v;
}

__device__ __forceinline__ double rec_coef(uint32_t lo, uint32_t hi) {
    return __hiloint2double((int)hi, (int)lo);
}

// One record of a given type (MORE: leaves 2 and 3 present; U1 / U2 / U3:
// prefix levels recomputed): its LDS reads first, then its arithmetic.
template <bool MORE, bool U1, bool U2, bool U3>
__device__ __forceinline__ void dagr_record(const char* lds, uint32_t ctl, const uint4& h0, const uint4& h1,
                                            const uint4& h2, const uint4& h3, double2& acc, double2& P1,
                                            double2& P2, double2& Pc) {
    auto at = [&](uint32_t off) { return *reinterpret_cast<const double2*>(lds + off); };
    const double2 v0 = at(h0.y), v1 = at(h0.z);
    double2 v2, v3, a1, a2, a3;
    if (MORE) {
        v2 = at(h0.w);
        v3 = at(h1.x);
    }
    if (U1) a1 = at(h1.y & 0x7fffffffu);
    if (U2) a2 = at(h1.z & 0x7fffffffu);
    if (U3) a3 = at(h1.w & 0x7fffffffu);
    const double c0 = rec_coef(h2.x, h2.y), c1 = rec_coef(h2.z, h2.w);
    double2 S;
    S.x = c0 * v0.x + c1 * v1.x;
    S.y = c0 * v0.y + c1 * v1.y;
    if (MORE) {
        const double c2 = rec_coef(h3.x, h3.y), c3 = rec_coef(h3.z, h3.w);
        S.x += c2 * v2.x + c3 * v3.x;
        S.y += c2 * v2.y + c3 * v3.y;
    }
    S.y = flip_sign_if(S.y, ctl >> 16);  // conjugated leaves: conj(sum c A) = sum c conj(A)
    if (U1) {
        a1.y = flip_sign_if(a1.y, h1.y >> 16);
        P1 = a1;
        Pc = a1;
    }
    if (U2) {
        a2.y = flip_sign_if(a2.y, h1.z >> 16);
        P2 = cmul(P1, a2);
        Pc = P2;
    }
    if (U3) {
        a3.y = flip_sign_if(a3.y, h1.w >> 16);
        Pc = cmul(P2, a3);
    }
    acc.x += Pc.x * S.x - Pc.y * S.y;
    acc.y += Pc.x * S.y + Pc.y * S.x;
}

__global__ __launch_bounds__(1024) void k_dagr(DagRecArgs a, int64_t nenv,
                                               const double2* __restrict__ Aplus,
                                               int64_t estride, double* __restrict__ E_out,
                                               double2* __restrict__ Gbuf, double e0r,
                                               const int32_t* __restrict__ env_list,
                                               const int32_t* __restrict__ env_count) {
    extern __shared__ double2 As[];  // [npool][64]
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // workgroup = (tile, part): the parts of a tile run disjoint task lists
    const int64_t blo = a.ntile_lo * a.parts;
    const bool hi = (int64_t)blockIdx.x >= blo;
    const int P = hi ? a.parts + 1 : a.parts;
    const int64_t bt = hi ? (int64_t)blockIdx.x - blo : (int64_t)blockIdx.x;
    const int64_t tile = (hi ? a.ntile_lo : 0) + bt / P;
    const int part = (int)(bt % P);
    const int32_t* const wptr = hi ? a.wtask_ptr_hi : a.wtask_ptr;
    const int32_t* const wlist = hi ? a.wtask_list_hi : a.wtask_list;
    int64_t env = tile * 64 + lane;
    bool valid = env < nenv;
    if (env_list) {
        // environments listed by the caller (fallback of the learned surrogate)
        const int64_t cnt = *env_count;
        if (tile * 64 >= cnt) return;  // whole workgroup idle: before any barrier
        valid = env < cnt;
        env = valid ? env_list[env] : 0;
    }
    for (int k = wave; k < a.npool; k += a.W)
        As[k * 64 + lane] = valid ? nt_load2(Aplus + (int64_t)k * estride + env) : make_double2(0.0, 0.0);
    __syncthreads();

    const char* lds = reinterpret_cast<const char*>(As) + 16 * lane;
    double ep = 0.0;
    const int tb = wptr[part * a.W + wave], te = wptr[part * a.W + wave + 1];
    for (int ti = tb; ti < te; ++ti) {
        const int t = wlist[ti];
        const int slot = a.tasks[5 * t + 2], rb = a.tasks[5 * t + 3], re = a.tasks[5 * t + 4];
        double2 acc = make_double2(a.task_const[2 * t], a.task_const[2 * t + 1]);
        // P1, P2: prefix products of levels 1 and 2; Pc: the product of the
        // current record's whole prefix.  Runs are sorted, so every new run
        // recomputes its deepest level and Pc is always the last level computed
        // (1 for the prefix-free runs, which come first).
        double2 P1 = make_double2(1.0, 0.0), P2 = P1, Pc = P1;
        const uint4* rp = reinterpret_cast<const uint4*>(a.recs) + 4 * (int64_t)rb;
        uint4 h0 = rp[0], h1 = rp[1], h2 = rp[2], h3 = rp[3];
        // unrolled by 2: the prefetched and current record headers alternate
        // between two register sets instead of being copied every record
        // (measured 18% faster than the rolled loop)
#pragma unroll 2
        for (int r = rb; r < re; ++r) {
            // next record (record nrec exists and is all zeros)
            const uint4* np = reinterpret_cast<const uint4*>(a.recs) + 4 * (int64_t)(r + 1);
            const uint4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            const uint32_t ctl = h0.x;
#define ACE_REC(T)                                                                                         \
    case T:                                                                                                \
        dagr_record<(T & 1) != 0, (T & 2) != 0, (T & 4) != 0, (T & 8) != 0>(lds, ctl, h0, h1, h2, h3, acc, \
                                                                              P1, P2, Pc);                 \
        break;
            switch ((ctl >> 2) & 15) {
                ACE_REC(0) ACE_REC(1) ACE_REC(2) ACE_REC(3) ACE_REC(4) ACE_REC(5) ACE_REC(6) ACE_REC(7)
                ACE_REC(8) ACE_REC(9) ACE_REC(10) ACE_REC(11) ACE_REC(12) ACE_REC(13) ACE_REC(14) ACE_REC(15)
            }
#undef ACE_REC
            h0 = n0;
            h1 = n1;
            h2 = n2;
            h3 = n3;
        }
        if (slot < 0)
            ep += acc.x;
        else if (valid)
            nt_store2(Gbuf + (int64_t)slot * estride + env, acc);
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(As);
    red[wave * 64 + lane] = ep;
    __syncthreads();
    if (wave == 0 && valid && E_out) {
        // one part: the energy; several: this part's share, summed by k_esum
        double s = a.parts == 1 ? e0r : 0.0;
        for (int w = 0; w < a.W; ++w) s += red[w * 64 + lane];
        E_out[(int64_t)part * estride + env] = s;
    }
}

// E = e0 + sum of the parts' energy shares (fixed order: deterministic).
// Environments from env_hi on belong to tiles with parts + 1 shares.
__global__ void k_esum(int64_t nenv, int parts, int64_t env_hi, const double* __restrict__ Ep, int64_t estride,
                       double e0, double* __restrict__ E_out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    double s = e0;
    const int np = e >= env_hi ? parts + 1 : parts;
    for (int p = 0; p < np; ++p) s += Ep[(int64_t)p * estride + e];
    E_out[e] = s;
}



// ---------------------------------------------------------------------------
// k_force: +dE/dr_j = Re sum_k G[k] grad(R_n Y_lm)(r_j)

template <int L, int PROP>
__global__ __launch_bounds__(64) void k_force_lds(RadialParams rp, int64_t nenv,
                                               const int32_t* __restrict__ off,
                                               const double* __restrict__ xyz, int64_t nnbr,
                                               const int8_t* __restrict__ species,
                                               const int16_t* __restrict__ kidx, int npool,
                                               const int32_t* __restrict__ gslot_ptr,
                                               const int32_t* __restrict__ gslot_list,
                                               const double2* __restrict__ Gbuf, int64_t estride,
                                               double* __restrict__ dE) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int64_t env = xcd_env(blockIdx.x, nenv);
    const int nr = rp.nr;
    double2* Gs = reinterpret_cast<double2*>(smem);
    double* Rs = reinterpret_cast<double*>(Gs + npool * NC);  // [64][2*nr]
    if (env >= nenv) return;
    for (int q = lane; q < npool * NC; q += 64) {
        const int k = q / NC, p = q % NC;
        double2 g = make_double2(0.0, 0.0);
        for (int s = gslot_ptr[k]; s < gslot_ptr[k + 1]; ++s) {
            const double2 v = Gbuf[((int64_t)gslot_list[s] * NC + p) * estride + env];
            g.x += v.x;
            g.y += v.y;
        }
        Gs[q] = g;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int j0 = off[env], j1 = off[env + 1];
    double* myR = Rs + lane * 2 * nr;
    for (int j = j0 + lane; j < j1; j += 64) {
        const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
        const int mu = species ? (int)species[j] : 0;
        Sph S = cart2spher(x, y, z);
        double t, dtdr;
        transform_ed(rp, S.r, t, dtdr);
        {
            double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
            myR[0] = p0;
            myR[nr] = d0;
            if (nr > 1) {
                double al = rp.A[1] * t + rp.B[1];
                double p1 = al * p0, d1 = al * d0 + rp.A[1] * p0;
                myR[1] = p1;
                myR[nr + 1] = d1;
                for (int n = 2; n < nr; ++n) {
                    al = rp.A[n] * t + rp.B[n];
                    const double p2 = al * p1 + rp.C[n] * p0;
                    const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                    p0 = p1;
                    p1 = p2;
                    d0 = d1;
                    d1 = d2;
                    myR[n] = p2;
                    myR[nr + n] = d2;
                }
            }
        }
        double P[NP], dP[NP];
        alp_eval_ed<L>(S, P, dP);
        const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
        const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
        double g[NC][3];
#pragma unroll
        for (int p = 0; p < NC; ++p) g[p][0] = g[p][1] = g[p][2] = 0.0;
        double epr = 0.7071067811865475, epi = 0.0;
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            if (m > 0) {
                const double nr_ = epr * S.cphi - epi * S.sphi;
                const double ni_ = epr * S.sphi + epi * S.cphi;
                epr = nr_;
                epi = ni_;
            }
#pragma unroll
            for (int l = m; l <= L; ++l) {
                const int ip = idxp(l, m);
                // Y and its Cartesian gradient (cYlm_ed!, dspher_to_dcart)
                double Yr, Yi, fpr, fpi, ftr, fti;
                if (m == 0) {
                    Yr = P[ip] * 0.7071067811865475;
                    Yi = 0.0;
                    fpr = fpi = 0.0;
                    ftr = dP[ip] * 0.7071067811865475;
                    fti = 0.0;
                } else {
                    const double pds = P[ip];
                    Yr = epr * pds * S.sth;
                    Yi = epi * pds * S.sth;
                    // dep/dphi = i m ep
                    fpr = -m * epi * pds;
                    fpi = m * epr * pds;
                    ftr = epr * dP[ip];
                    fti = epi * dP[ip];
                }
                const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                       (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                       (-S.sth * ftr) * rinv};
                const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                       (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                       (-S.sth * fti) * rinv};
                const int16_t* kr = kidx + ((int64_t)(mu * (L + 1) + l) * (L + 1) + m) * nr;
                for (int n = 0; n < nr; ++n) {
                    const int k = kr[n];
                    if (k < 0) continue;
                    const double R = myR[n], dR = myR[nr + n] * dtdr;
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        const double2 G = Gs[k * NC + p];
                        // Re(G * (dR rhat Y + R dY))
                        const double gy = G.x * Yr - G.y * Yi;
                        g[p][0] += dR * gy * rx + R * (G.x * dYr[0] - G.y * dYi[0]);
                        g[p][1] += dR * gy * ry + R * (G.x * dYr[1] - G.y * dYi[1]);
                        g[p][2] += dR * gy * rz + R * (G.x * dYr[2] - G.y * dYi[2]);
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < NC; ++p)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) dE[(int64_t)(p * 3 + dd) * nnbr + j] = g[p][dd];
    }
}

// ---------------------------------------------------------------------------
// basis level: A (full), AA, B, dB

__global__ void k_expandA(int64_t nenv, int nA, const int32_t* __restrict__ fid,
                          const double* __restrict__ sgn, const double2* __restrict__ Aplus,
                          int64_t estride, double2* __restrict__ A_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nenv * nA) return;
    const int64_t env = i / nA;
    const int iA = (int)(i % nA);
    const int f = fid[iA];
    double2 v = Aplus[(int64_t)(f & 0x7fff) * estride + env];
    if (f & ACE_CONJ) v.y = -v.y;
    v.x *= sgn[iA];
    v.y *= sgn[iA];
    A_out[i] = v;
}

// AA[env][iAA] = prod_t A[iAA2iA[iAA][t]], left to right (pibasis.jl:265-275)
__global__ void k_AA(int64_t nenv, int nA, int nAA, int maxord, const int32_t* __restrict__ orders,
                     const int32_t* __restrict__ iAA2iA, const double2* __restrict__ A,
                     double2* __restrict__ AA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nenv * nAA) return;
    const int64_t env = i / nAA;
    const int r = (int)(i % nAA);
    const double2* Ae = A + env * nA;
    double2 aa = make_double2(1.0, 0.0);
    const int o = orders[r];
    for (int t = 0; t < o; ++t) aa = cmul(aa, Ae[iAA2iA[r * maxord + t]]);
    AA[i] = aa;
}

// B[env][b][p] = Re sum_{nz in row b} val * AA
template <int PROP>
__global__ void k_B(int64_t nenv, int nB, int nAA, const int32_t* __restrict__ rowptr,
                    const int32_t* __restrict__ cols, const double* __restrict__ vals,
                    const double2* __restrict__ AA, double* __restrict__ B) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nenv * nB) return;
    const int64_t env = i / nB;
    const int b = (int)(i % nB);
    double s[NC];
#pragma unroll
    for (int p = 0; p < NC; ++p) s[p] = 0.0;
    for (int z = rowptr[b]; z < rowptr[b + 1]; ++z) {
        const double2 a = AA[env * nAA + cols[z]];
        if (PROP == 0) {
            s[0] += vals[z] * a.x;  // isrealAA: AA = real(prod A)
        } else {
#pragma unroll
            for (int p = 0; p < NC; ++p) s[p] += vals[6 * z + 2 * p] * a.x - vals[6 * z + 2 * p + 1] * a.y;
        }
    }
#pragma unroll
    for (int p = 0; p < NC; ++p) B[i * NC + p] = s[p];
}

// per-neighbour 1p gradients dphi[j][iA][3] (complex), product_1pbasis.jl:169-221
template <int L>
__global__ __launch_bounds__(64) void k_dphi(RadialParams rp, int64_t nnbr,
                                             const int32_t* __restrict__ nbr_env,
                                             const double* __restrict__ xyz,
                                             const int8_t* __restrict__ species,
                                             const int32_t* __restrict__ aspec4, int nA,
                                             double2* __restrict__ dphi) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnbr) return;
    const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
    const int mu = species ? (int)species[j] : -1;
    Sph S = cart2spher(x, y, z);
    double t, dtdr;
    transform_ed(rp, S.r, t, dtdr);
    double R[ACE_NRMAX], dR[ACE_NRMAX];
    R[0] = rp.A[0] * fcut(rp, t);
    dR[0] = rp.A[0] * fcut_d(rp, t);
    for (int n = 1; n < rp.nr && n < ACE_NRMAX; ++n) {
        const double al = rp.A[n] * t + rp.B[n];
        if (n == 1) {
            R[1] = al * R[0];
            dR[1] = al * dR[0] + rp.A[1] * R[0];
        } else {
            R[n] = al * R[n - 1] + rp.C[n] * R[n - 2];
            dR[n] = al * dR[n - 1] + rp.C[n] * dR[n - 2] + rp.A[n] * R[n - 1];
        }
    }
    double P[NP], dP[NP];
    alp_eval_ed<L>(S, P, dP);
    const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
    const double rh[3] = {x / S.r, y / S.r, z / S.r};
    for (int iA = 0; iA < nA; ++iA) {
        const int amu = aspec4[4 * iA], n = aspec4[4 * iA + 1] - 1, l = aspec4[4 * iA + 2],
                  mm = aspec4[4 * iA + 3];
        double2* out = dphi + (j * nA + iA) * 3;
        if (amu >= 0 && amu != mu) {
            for (int d = 0; d < 3; ++d) out[d] = make_double2(0.0, 0.0);
            continue;
        }
        const int m = mm < 0 ? -mm : mm;
        // e^{i m phi}/sqrt2
        double epr = 0.7071067811865475, epi = 0.0;
        for (int q = 0; q < m; ++q) {
            const double a = epr * S.cphi - epi * S.sphi;
            const double b = epr * S.sphi + epi * S.cphi;
            epr = a;
            epi = b;
        }
        double pds = 0.0, dpt = 0.0;
        for (int ll = 0; ll <= L; ++ll)
            for (int q = 0; q <= ll; ++q)
                if (ll == l && q == m) {
                    pds = P[idxp(ll, q)];
                    dpt = dP[idxp(ll, q)];
                }
        double Yr, Yi, fpr, fpi, ftr, fti;
        if (m == 0) {
            Yr = pds * 0.7071067811865475;
            Yi = 0.0;
            fpr = fpi = 0.0;
            ftr = dpt * 0.7071067811865475;
            fti = 0.0;
        } else {
            Yr = epr * pds * S.sth;
            Yi = epi * pds * S.sth;
            fpr = -m * epi * pds;
            fpi = m * epr * pds;
            ftr = epr * dpt;
            fti = epi * dpt;
        }
        if (mm < 0) {  // Y_l^-m = (-1)^m conj(Y_l^m), same for the gradient
            const double s = (m & 1) ? -1.0 : 1.0;
            Yr *= s;
            Yi *= -s;
            fpr *= s;
            fpi *= -s;
            ftr *= s;
            fti *= -s;
        }
        const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                               (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                               (-S.sth * ftr) * rinv};
        const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                               (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                               (-S.sth * fti) * rinv};
        const double Rv = R[n], dRv = dR[n] * dtdr;
        for (int d = 0; d < 3; ++d)
            out[d] = make_double2(dRv * rh[d] * Yr + Rv * dYr[d], dRv * rh[d] * Yi + Rv * dYi[d]);
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
template <int L>
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
    int kr[POOL_KP], ky[POOL_KP], kmu[POOL_KP];
    double ar[POOL_KP], ai[POOL_KP];
#pragma unroll
    for (int p = 0; p < POOL_KP; ++p) {
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
            alp_eval_ed<L>(S, P, dP);
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
            for (int p = 0; p < POOL_KP; ++p) {
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
    for (int p = 0; p < POOL_KP; ++p)
        if (kmu[p] != -2) dAw[(int64_t)(lane + 64 * p) * estride + env] = make_double2(ar[p], ai[p]);
}

// dAAw[env][iAA] = sum_t dAw_t prod_{s != t} A_s by a left-to-right dual
// product (p, dp) <- (p a, dp a + p da) (evaluator.jl:225-235).
__global__ void k_AAdir(int64_t nenv, int nA, int nAA, int maxord, const int32_t* __restrict__ orders,
                        const int32_t* __restrict__ iAA2iA, const double2* __restrict__ A,
                        const double2* __restrict__ dA, double2* __restrict__ dAA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nenv * nAA) return;
    const int64_t env = i / nAA;
    const int r = (int)(i % nAA);
    const double2* Ae = A + env * nA;
    const double2* dAe = dA + env * nA;
    double2 p = make_double2(1.0, 0.0), dp = make_double2(0.0, 0.0);
    const int o = orders[r];
    for (int t = 0; t < o; ++t) {
        const int v = iAA2iA[r * maxord + t];
        const double2 a = Ae[v], da = dAe[v];
        const double2 x = cmul(dp, a), y = cmul(p, da);
        dp = make_double2(x.x + y.x, x.y + y.y);
        p = cmul(p, a);
    }
    dAA[i] = dp;
}

// One A2B nonzero of the fused basis-value kernel, pre-decoded on the host:
// the AA column's order and pooled factor ids (k | ACE_CONJ), and the A2B
// value times the sign of the m < 0 factors (Invariant; EuclideanVector
// rows read their 3 complex values from the CSR array).
struct BtRec {
    int32_t o;
    int32_t f[4];
    int32_t z;  // CSR index of the nonzero
    double coef;
};
static_assert(sizeof(BtRec) == 32, "BtRec is 8 dwords");

// Product (DIR = false) or dual-number derivative (DIR = true) of the AA
// column of record r, left to right, from the staged tiles.
template <bool DIR, int ENVS>
__device__ __forceinline__ double2 bt_product(const double2* As, const double2* Ws, const BtRec& r, int e) {
    double2 a[4], da[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < r.o) {
            a[q] = lds_factor<ENVS>(As, r.f[q], e);
            if (DIR) da[q] = lds_factor<ENVS>(Ws, r.f[q], e);
        }
    double2 p = make_double2(1.0, 0.0), dp = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < r.o) {
            if (DIR) {
                const double2 x = cmul(dp, a[q]), y = cmul(p, da[q]);
                dp = make_double2(x.x + y.x, x.y + y.y);
            }
            p = cmul(p, a[q]);
        }
    return DIR ? dp : p;
}

// k_Bt: basis values B = A2B . AA (DIR = false; pibasis.jl:265-275,
// symmbasis.jl:248-264) or their directional derivative A2B . dAA along
// per-neighbour weights (DIR = true; evaluator.jl:225-240) for a tile of
// ENVS environments, without materialising AA.  The pooled A+ tile (and
// dA+w for DIR) is staged in LDS as in k_dagr; every A2B nonzero (one BtRec,
// read with one scalar load) recomputes its AA product from it.  A
// workgroup is (tile, chunk of rows); each wave takes rpw consecutive rows
// (with ENVS = 32 the two half-waves take alternate rows).  Workgroups of
// one tile are consecutive on one XCD so the staged tile is read from that
// XCD's L2.
constexpr int BT_NW = 16;  // waves per k_Bt workgroup

template <int PROP, bool DIR, int ENVS>
__global__ __launch_bounds__(64 * BT_NW) void k_Bt(int64_t nenv, int npool, int nB, int nchunk, int rpw,
                                            const int32_t* __restrict__ rowptr, const BtRec* __restrict__ recs,
                                            const double* __restrict__ vals, const double2* __restrict__ Aplus,
                                            const double2* __restrict__ Wplus, int64_t estride,
                                            double* __restrict__ out) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double2 tilebuf[];
    double2* As = tilebuf;
    double2* Ws = tilebuf + (size_t)npool * ENVS;
    const int64_t ntile = (nenv + ENVS - 1) / ENVS;
    const int64_t lin = xcd_env(blockIdx.x, ntile * nchunk);
    const int64_t t = lin / nchunk;
    const int chunk = (int)(lin % nchunk);
    const int64_t e0 = t * ENVS;
    for (int i = threadIdx.x; i < npool * ENVS; i += 64 * BT_NW) {
        const int k = i / ENVS;
        const int64_t env = e0 + (i % ENVS);
        double2 v = make_double2(0.0, 0.0), w = make_double2(0.0, 0.0);
        if (env < nenv) {
            v = Aplus[(int64_t)k * estride + env];
            if (DIR) w = Wplus[(int64_t)k * estride + env];
        }
        As[i] = v;
        if (DIR) Ws[i] = w;
    }
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int e = lane % ENVS;
    const int half = ENVS == 32 ? (lane >> 5) : 0;
    const int64_t env = e0 + e;
    const int b0 = (chunk * BT_NW + wave) * rpw;
    const int b1 = min(b0 + rpw, nB);
    for (int b = b0 + half; b < b1; b += (ENVS == 32 ? 2 : 1)) {
        double s[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) s[q] = 0.0;
        const int z1 = rowptr[b + 1];
        for (int z = rowptr[b]; z < z1; ++z) {
            const BtRec r = recs[z];
            const double2 v = bt_product<DIR, ENVS>(As, Ws, r, e);
            if (PROP == 0) {
                s[0] += r.coef * v.x;
            } else {
                const double sg = r.coef;
#pragma unroll
                for (int q = 0; q < NC; ++q)
                    s[q] += sg * (vals[6 * r.z + 2 * q] * v.x - vals[6 * r.z + 2 * q + 1] * v.y);
            }
        }
        if (env < nenv) {
#pragma unroll
            for (int q = 0; q < NC; ++q) out[(env * nB + b) * NC + q] = s[q];
        }
    }
}

// dAA[j][iAA][3] = sum_a (prod_{t != a} A_t) dphi[j][v_a]  (pibasis.jl:402-432)
__global__ void k_dAA(int64_t nnbr, int nA, int nAA, int maxord, const int32_t* __restrict__ nbr_env,
                      const int32_t* __restrict__ orders, const int32_t* __restrict__ iAA2iA,
                      const double2* __restrict__ A, const double2* __restrict__ dphi,
                      double2* __restrict__ dAA) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnbr * nAA) return;
    const int64_t j = i / nAA;
    const int r = (int)(i % nAA);
    const int64_t env = nbr_env[j];
    const double2* Ae = A + env * nA;
    const int o = orders[r];
    double2 g[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    for (int a = 0; a < o; ++a) {
        double2 w = make_double2(1.0, 0.0);
        for (int t = 0; t < o; ++t)
            if (t != a) w = cmul(w, Ae[iAA2iA[r * maxord + t]]);
        const double2* dp = dphi + (j * nA + iAA2iA[r * maxord + a]) * 3;
        for (int d = 0; d < 3; ++d) {
            const double2 v = cmul(w, dp[d]);
            g[d].x += v.x;
            g[d].y += v.y;
        }
    }
    for (int d = 0; d < 3; ++d) dAA[(i * 3) + d] = g[d];
}

template <int PROP>
__global__ void k_dB(int64_t nnbr, int nB, int nAA, const int32_t* __restrict__ rowptr,
                     const int32_t* __restrict__ cols, const double* __restrict__ vals,
                     const double2* __restrict__ dAA, double* __restrict__ dB) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnbr * nB) return;
    const int64_t j = i / nB;
    const int b = (int)(i % nB);
    double s[NC][3];
#pragma unroll
    for (int p = 0; p < NC; ++p) s[p][0] = s[p][1] = s[p][2] = 0.0;
    for (int z = rowptr[b]; z < rowptr[b + 1]; ++z) {
        const double2* g = dAA + ((j * nAA) + cols[z]) * 3;
        for (int d = 0; d < 3; ++d) {
            if (PROP == 0) {
                s[0][d] += vals[z] * g[d].x;
            } else {
#pragma unroll
                for (int p = 0; p < NC; ++p)
                    s[p][d] += vals[6 * z + 2 * p] * g[d].x - vals[6 * z + 2 * p + 1] * g[d].y;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < NC; ++p)
        for (int d = 0; d < 3; ++d) dB[(i * NC + p) * 3 + d] = s[p][d];
}

__global__ void k_nbr_env(int64_t nenv, const int32_t* __restrict__ off, int32_t* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    for (int j = off[e]; j < off[e + 1]; ++j) out[j] = (int32_t)e;
}

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
                                             __bf16* __restrict__ Xn, double* __restrict__ E2, double e0) {
    const int64_t env = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (env >= nenv) return;
    double e = e0;
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

// ---------------------------------------------------------------------------
// host side

constexpr int MAX_PARTS = 4;  // workgroups per product-DAG tile: 1 to 4

struct WaveCfg {
    int W;
    std::vector<int32_t> ptr_all, list_all, ptr_e, list_e;
    DevBuf<int32_t> d_ptr_all, d_list_all, d_ptr_e, d_list_e;
    // task lists over parts * W virtual waves (Invariant), parts = 2 and 4
    DevBuf<int32_t> pp_all[MAX_PARTS + 1], pl_all[MAX_PARTS + 1], pp_e[MAX_PARTS + 1], pl_e[MAX_PARTS + 1];
};

constexpr int FORCE_NRMAX = 16;  // radial functions held in registers by the force kernels
constexpr int FF_QMAX = 256;  // (l, m >= 0, n) entries of the flat force kernel (npool <= 256)
// environments a 256-neighbour block stages in LDS: 32 when their G fits in
// FF_LDS (short environments), else 16
constexpr int FF_MAXE = 16;
constexpr int FF_MAXE_WIDE = 32;
constexpr size_t FF_LDS = 40 * 1024;

struct ForceFlatTab {
    int8_t nrl[ACE_LMAX + 1];                                  // n bound per l
    int16_t slot[(ACE_LMAX + 1) * (ACE_LMAX + 2) / 2][FORCE_NRMAX];  // G slot of (l, m >= 0, n)
    int16_t nq;  // number of (l, m, n) entries (their G slots in staging order: ace_ctx::ffq)
};

struct ace_ctx {
    int device = 0;
    int prop = 0, L = 0, nr = 0, nspecies = 0, tp = 2, pl = 0, pr = 0;
    double r0 = 1.0, tl = 0, tr = 0;
    int npool = 0, nA = 0, nAA = 0, maxorder = 0, nB = 0, nnz = 0;
    int nruns = 0, nleaves = 0, ntasks = 0, nterms = 0, ngslots = 0;
    int ncoef = 1;
    acehost::PackCopy host;
    // device
    DevBuf<double> rA, rB, rC;
    DevBuf<int32_t> pool, aspec_fid, aspec4, orders, iAA2iA, rowptr, cols;
    DevBuf<double> aspec_sgn, vals;
    DevBuf<BtRec> bt_rec;
    DevBuf<int32_t> ffq;  // G slots of the flat force kernel's staged (l, m, n) entries
    DevBuf<int16_t> kidx;
    int nrl[ACE_LMAX + 1] = {0};  // n bound per l over the pooled functions
    // flat force kernel (Invariant, one species, nr <= 16): slot table and the
    // split gradient slots folded into their primary slot before it runs
    bool flat_force = false;
    ForceFlatTab fflat;
    int nfold = 0;
    DevBuf<int32_t> gfold;
    DevBuf<int32_t> runs, leaf_fid, d_tasks, gslot_ptr, gslot_list;
    DevBuf<double> leaf_coef, task_const, e0v;
    DevBuf<double4> vleaf;  // EuclideanVector decoded leaves
    DevBuf<uint32_t> recs;
    // surrogate feature tables
    int nfeat = 0;
    std::vector<int32_t> feat_rows;
    DevBuf<int32_t> feat_ptr, feat_f, fgrad_ptr, fgrad_feat, fgrad_f;
    DevBuf<double> feat_coef, fgrad_coef, cfeat;
    int64_t surr_nenv = -1;
    double e0r = 0.0;
    std::vector<WaveCfg*> wcfg;
    int cur_w = 0;
    int ncu = 256;  // compute units of the device (tail sizing of the product-DAG grid)
    int parts = 0;  // workgroups per product-DAG tile: 0 = chosen per call, else 1 to 4
    // tiles at the end of the product-DAG grid that run one more part: -1 =
    // chosen per call (automatic parts only), else that many
    int tail = -1;
    std::map<std::pair<int64_t, int>, int64_t> tail_cache;  // automatic choice per (tiles, parts)
    // last eagerly evaluated batch: its neighbour count serves calls under stream capture
    const int32_t* batch_off = nullptr;
    int64_t batch_nenv = -1, batch_nnbr = 0;
    int last_parts = 1;
    // scratch
    DevBuf<double> Epart;
    DevBuf<double2> Aplus, Gbuf, Afull, AA, dphi, dAA, Wplus, Wfull;
    DevBuf<int32_t> nbr_env;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timing_valid = false;
    bool validate = true;
    ~ace_ctx() {
        for (auto* w : wcfg) delete w;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

static int upload_coeffs(ace_ctx* ctx, const double* c) {
    acehost::Coeffs cf = acehost::coefficients(ctx->host, c);
    if (ctx->leaf_coef.upload(cf.leaf.data(), cf.leaf.size())) return -1;
    if (ctx->prop != ACE_PROP_INVARIANT) {
        // decoded leaves: {factor id (low word), 0, c0, c1}, {c2, c3, c4, c5}
        const size_t nl = ctx->host.leaf_fid.size();
        std::vector<double4> vl(std::max<size_t>(2 * nl, 2));
        for (size_t i = 0; i < nl; ++i) {
            const double* c = cf.leaf.data() + 6 * i;
            uint64_t bits = (uint32_t)ctx->host.leaf_fid[i];
            double fid;
            std::memcpy(&fid, &bits, sizeof(fid));
            vl[2 * i] = make_double4(fid, 0.0, c[0], c[1]);
            vl[2 * i + 1] = make_double4(c[2], c[3], c[4], c[5]);
        }
        if (ctx->vleaf.upload(vl.data(), vl.size())) return -1;
    }
    if (ctx->prop == ACE_PROP_INVARIANT) {
        std::vector<uint32_t> rec;
        const std::string rerr = acehost::records(ctx->host, cf, rec);
        if (!rerr.empty()) return set_err(rerr);
        if (ctx->recs.upload(rec.data(), rec.size())) return -1;
    }
    if (ctx->task_const.upload(cf.task.data(), cf.task.size())) return -1;
    ctx->e0r = cf.e0[0];
    double ev[3] = {cf.e0[0], cf.e0[2], cf.e0[4]};
    if (ctx->nfeat > 0) {
        std::vector<double> cf_(ctx->nfeat);
        for (int f = 0; f < ctx->nfeat; ++f) cf_[f] = c[ctx->feat_rows[f]];
        if (ctx->cfeat.upload(cf_.data(), cf_.size())) return -1;
    }
    return ctx->e0v.upload(ev, 3);
}



extern "C" int ace_abi_version(void) { return ACEHIP_ABI_VERSION; }

extern "C" const char* ace_last_error(void) { return g_err.c_str(); }

static bool g_alp_ready[64] = {false};

static int ensure_alp(int device) {
    if (device < 0 || device >= 64) return set_err("device index out of range");
    if (g_alp_ready[device]) return 0;
    std::vector<double> A(ACE_NPMAX, 0.0), B(ACE_NPMAX, 0.0);
    for (int l = 2; l <= ACE_LMAX; ++l) {
        const double ls = (double)l * l, lm1s = (double)(l - 1) * (l - 1);
        for (int m = 0; m <= l - 2; ++m) {
            const double ms = (double)m * m;
            A[idxp(l, m)] = std::sqrt((4 * ls - 1.0) / (ls - ms));
            B[idxp(l, m)] = -std::sqrt((lm1s - ms) / (4 * lm1s - 1.0));
        }
    }
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_alpA), A.data(), sizeof(double) * ACE_NPMAX));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_alpB), B.data(), sizeof(double) * ACE_NPMAX));
    g_alp_ready[device] = true;
    return 0;
}

static int set_func_lds(const void* f, size_t bytes) {
    if (bytes > 64 * 1024) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return 0;
}

extern "C" int ace_ctx_create(const ace_pack* pk, int device, const double* c, int64_t nB,
                              ace_ctx** out) {
    if (!pk || !out || !c) return set_err("ace_ctx_create: null argument");
    *out = nullptr;
    const std::string verr = acehost::validate_pack(pk);
    if (!verr.empty()) return set_err("ace_ctx_create: " + verr);
    if (nB != pk->nB) return set_err("ace_ctx_create: len(c) != len(basis)");
    std::vector<acehost::WaveTable> wt;
    const std::string werr = acehost::wave_tables(pk, wt);
    if (!werr.empty()) return set_err("ace_ctx_create: " + werr);
    HIP_TRY(hipSetDevice(device));
    if (ensure_alp(device)) return -1;
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    ace_ctx* ctx = new ace_ctx();
    ctx->device = device;
    ctx->ncu = std::max(1, ncu);
    ctx->prop = pk->prop;
    ctx->L = pk->L;
    ctx->nr = pk->nr;
    ctx->nspecies = pk->nspecies;
    ctx->tp = pk->tp;
    ctx->pl = pk->pl;
    ctx->pr = pk->pr;
    ctx->r0 = pk->r0;
    ctx->tl = pk->tl;
    ctx->tr = pk->tr;
    ctx->npool = pk->npool;
    ctx->nA = pk->nA;
    ctx->nAA = pk->nAA;
    ctx->maxorder = pk->maxorder;
    ctx->nB = pk->nB;
    ctx->nnz = pk->a2b_nnz;
    ctx->nruns = pk->nruns;
    ctx->nleaves = pk->nleaves;
    ctx->ntasks = pk->ntasks;
    ctx->nterms = pk->nterms;
    ctx->ngslots = pk->ngslots;
    ctx->ncoef = pk->prop == ACE_PROP_INVARIANT ? 1 : 6;
    ctx->host = acehost::copy_pack(pk);
    const int nc = ctx->ncoef;
    auto fail = [&](const char* what) {
        std::string msg = g_err;
        delete ctx;
        return set_err(std::string("ace_ctx_create: ") + what + ": " + msg);
    };
    if (ctx->rA.upload(pk->rA, pk->nr) || ctx->rB.upload(pk->rB, pk->nr) ||
        ctx->rC.upload(pk->rC, pk->nr) || ctx->pool.upload(pk->pool, 4 * (size_t)pk->npool) ||
        ctx->aspec_fid.upload(pk->aspec_fid, pk->nA) ||
        ctx->aspec_sgn.upload(pk->aspec_sgn, pk->nA) ||
        ctx->orders.upload(pk->orders, pk->nAA) ||
        ctx->iAA2iA.upload(pk->iAA2iA, (size_t)pk->nAA * pk->maxorder) ||
        ctx->rowptr.upload(pk->a2b_rowptr, pk->nB + 1) ||
        ctx->cols.upload(pk->a2b_cols, pk->a2b_nnz) ||
        ctx->vals.upload(pk->a2b_vals, (size_t)pk->a2b_nnz * nc) ||
        ctx->runs.upload(pk->runs, 7 * (size_t)pk->nruns) ||
        ctx->leaf_fid.upload(pk->leaf_fid, pk->nleaves) ||
        ctx->d_tasks.upload(pk->tasks, 5 * (size_t)pk->ntasks) ||
        ctx->gslot_ptr.upload(pk->gslot_ptr, pk->npool + 1) ||
        ctx->gslot_list.upload(pk->gslot_list, pk->gslot_ptr[pk->npool]))
        return fail("upload");
    {
        // one record per A2B nonzero for the fused basis-value kernel: the AA
        // column's pooled factor ids and order, and (Invariant) the A2B value
        // times the product of the m < 0 factor signs (EuclideanVector: the sign)
        std::vector<BtRec> rec(std::max(1, pk->a2b_nnz));
        for (int z = 0; z < pk->a2b_nnz; ++z) {
            const int col = pk->a2b_cols[z];
            BtRec& r = rec[z];
            r.o = pk->orders[col];
            r.z = z;
            double sg = 1.0;
            for (int q = 0; q < 4; ++q) {
                r.f[q] = 0;
                if (q < r.o) {
                    const int iA = pk->iAA2iA[(size_t)col * pk->maxorder + q];
                    r.f[q] = pk->aspec_fid[iA];
                    sg *= pk->aspec_sgn[iA];
                }
            }
            r.coef = pk->prop == ACE_PROP_INVARIANT ? pk->a2b_vals[z] * sg : sg;
        }
        if (ctx->bt_rec.upload(rec.data(), rec.size())) return fail("upload");
        std::vector<int32_t> a4 = acehost::aspec_rows(pk);
        if (ctx->aspec4.upload(a4.data(), a4.size())) return fail("upload");
        std::vector<int16_t> kx = acehost::pool_lookup(pk);
        if (ctx->kidx.upload(kx.data(), kx.size())) return fail("upload");
        for (int k = 0; k < pk->npool; ++k)
            ctx->nrl[pk->pool[4 * k + 2]] = std::max(ctx->nrl[pk->pool[4 * k + 2]], (int)pk->pool[4 * k + 1]);
        ctx->flat_force = pk->prop == ACE_PROP_INVARIANT && pk->nspecies <= 1 && pk->nr <= FORCE_NRMAX &&
                          pk->ngslots > 0;
        if (ctx->flat_force) {
            std::memset(&ctx->fflat, 0, sizeof(ctx->fflat));
            for (int l = 0; l <= ACE_LMAX; ++l) ctx->fflat.nrl[l] = (int8_t)ctx->nrl[l];
            // every (l, m, n) inside the n bounds must be a pooled function (true
            // for the prefix-shaped pooled sets of RnYlm bases); otherwise the
            // per-environment kernel runs
            std::vector<int> seen((size_t)(ACE_LMAX + 1) * (ACE_LMAX + 2) / 2 * FORCE_NRMAX, 0);
            std::vector<int32_t> fold;
            for (int k = 0; k < pk->npool; ++k) {
                const int n = pk->pool[4 * k + 1] - 1, l = pk->pool[4 * k + 2], m = pk->pool[4 * k + 3];
                const int s0 = pk->gslot_ptr[k], s1 = pk->gslot_ptr[k + 1];
                if (s1 <= s0) {
                    ctx->flat_force = false;  // a pooled function without a gradient slot
                    break;
                }
                ctx->fflat.slot[idxp(l, m)][n] = (int16_t)pk->gslot_list[s0];
                seen[(size_t)idxp(l, m) * FORCE_NRMAX + n] = 1;
                for (int q = s0 + 1; q < s1; ++q) {
                    fold.push_back(pk->gslot_list[s0]);
                    fold.push_back(pk->gslot_list[q]);
                }
            }
            for (int l = 0; l <= pk->L && ctx->flat_force; ++l)
                for (int m = 0; m <= l; ++m)
                    for (int n = 0; n < ctx->nrl[l]; ++n)
                        if (!seen[(size_t)idxp(l, m) * FORCE_NRMAX + n]) ctx->flat_force = false;
            if (pk->ngslots > 32767) ctx->flat_force = false;
            // staging order of the flat force kernel: m, then l >= m, then n
            std::vector<int32_t> qslot;
            for (int m = 0; m <= pk->L && ctx->flat_force; ++m)
                for (int l = m; l <= pk->L; ++l)
                    for (int n = 0; n < ctx->nrl[l]; ++n) qslot.push_back(ctx->fflat.slot[idxp(l, m)][n]);
            if ((int)qslot.size() > FF_QMAX) ctx->flat_force = false;
            ctx->fflat.nq = (int16_t)std::min<size_t>(qslot.size(), FF_QMAX);
            if (ctx->flat_force && ctx->ffq.upload(qslot.data(), qslot.size())) return fail("upload");
            ctx->nfold = (int)fold.size() / 2;
            if (ctx->flat_force && ctx->nfold > 0 && ctx->gfold.upload(fold.data(), fold.size()))
                return fail("upload");
        }
    }
    for (auto& t : wt) {
        WaveCfg* wc = new WaveCfg();
        ctx->wcfg.push_back(wc);
        wc->W = t.W;
        wc->ptr_all = t.ptr_all;
        wc->list_all = t.list_all;
        wc->ptr_e = t.ptr_e;
        wc->list_e = t.list_e;
        if (wc->d_ptr_all.upload(wc->ptr_all.data(), wc->ptr_all.size()) ||
            wc->d_list_all.upload(wc->list_all.data(), wc->list_all.size()) ||
            wc->d_ptr_e.upload(wc->ptr_e.data(), wc->ptr_e.size()) ||
            wc->d_list_e.upload(wc->list_e.data(), wc->list_e.size()))
            return fail("upload");
        if (pk->prop == ACE_PROP_INVARIANT) {
            const std::vector<int> t_all = acehost::mode_tasks(ctx->host, true), t_e = acehost::mode_tasks(ctx->host, false);
            for (int P = 2; P <= MAX_PARTS; ++P) {
                std::vector<int32_t> ptr, list;
                acehost::lpt_lists(ctx->host, t_all, P * t.W, ptr, list);
                if (wc->pp_all[P].upload(ptr.data(), ptr.size()) || wc->pl_all[P].upload(list.data(), list.size()))
                    return fail("upload");
                acehost::lpt_lists(ctx->host, t_e, P * t.W, ptr, list);
                if (wc->pp_e[P].upload(ptr.data(), ptr.size()) || wc->pl_e[P].upload(list.data(), list.size()))
                    return fail("upload");
            }
        }
    }
    ctx->nfeat = pk->nfeat;
    if (pk->nfeat > 0) {
        const int nt = pk->feat_ptr[pk->nfeat], ng = pk->fgrad_ptr[pk->npool];
        ctx->feat_rows.assign(pk->feat_rows, pk->feat_rows + pk->nfeat);
        if (ctx->feat_ptr.upload(pk->feat_ptr, pk->nfeat + 1) || ctx->feat_f.upload(pk->feat_f, 2 * (size_t)nt) ||
            ctx->feat_coef.upload(pk->feat_coef, nt) || ctx->fgrad_ptr.upload(pk->fgrad_ptr, pk->npool + 1) ||
            ctx->fgrad_feat.upload(pk->fgrad_feat, ng) || ctx->fgrad_f.upload(pk->fgrad_f, ng) ||
            ctx->fgrad_coef.upload(pk->fgrad_coef, ng))
            return fail("upload");
    }
    if (upload_coeffs(ctx, c)) return fail("set_params");
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail("event");
    *out = ctx;
    return 0;
}

extern "C" int ace_set_params(ace_ctx* ctx, const double* c, int64_t nB) {
    if (!ctx || !c) return set_err("ace_set_params: null argument");
    if (nB != ctx->nB) return set_err("ace_set_params: len(c) != len(basis)");
    HIP_TRY(hipSetDevice(ctx->device));
    return upload_coeffs(ctx, c);
}

// Tail tiles for the automatic parts choice.  The grid runs in rounds of ncu
// workgroups of equal length (1/P of a tile's work).  Giving the last n tiles
// P + 1 parts fills the partial last round with shorter workgroups, each
// 1/(P+1) of the work plus 8 % (re-staging and thinner task lists: measured
// 0.200 ms per workgroup with 4 parts against 0.246 ms with 3 at 100k
// environments).  The makespan of a greedy dispatch of such workgroups onto
// ncu units is flat over a range of n; take the middle of that range.
static int64_t choose_tail(int64_t ntile, int P, int ncu) {
    const double d = 1.0 / P, dh = 1.08 / (P + 1);
    std::vector<double> h((size_t)ncu);
    auto makespan = [&](int64_t nhi) {
        // the P-part workgroups: full rounds, then r units busy one more
        const int64_t w = (ntile - nhi) * P, f = w / ncu, r = w % ncu;
        for (int i = 0; i < ncu; ++i) h[(size_t)i] = (i < r ? f + 1 : f) * d;
        std::make_heap(h.begin(), h.end(), std::greater<double>());
        for (int64_t k = 0; k < nhi * (P + 1); ++k) {
            std::pop_heap(h.begin(), h.end(), std::greater<double>());
            h.back() += dh;
            std::push_heap(h.begin(), h.end(), std::greater<double>());
        }
        return *std::max_element(h.begin(), h.end());
    };
    const int64_t nmax = std::min<int64_t>(ntile, ncu);
    std::vector<double> ms((size_t)nmax + 1);
    double best = 1e300;
    for (int64_t n = 0; n <= nmax; ++n) {
        ms[(size_t)n] = makespan(n);
        best = std::min(best, ms[(size_t)n]);
    }
    std::vector<int64_t> ties;
    for (int64_t n = 0; n <= nmax; ++n)
        if (ms[(size_t)n] <= best * (1.0 + 1e-9)) ties.push_back(n);
    return ties[ties.size() / 2];
}

extern "C" int ace_set_tail_parts(ace_ctx* ctx, int ntail) {
    if (!ctx) return set_err("ace_set_tail_parts: null context");
    if (ntail < -1) return set_err("ace_set_tail_parts: ntail must be -1 (automatic) or >= 0");
    ctx->tail = ntail;
    return 0;
}

extern "C" int ace_set_parts(ace_ctx* ctx, int parts) {
    if (!ctx) return set_err("ace_set_parts: null context");
    if (parts < 0 || parts > MAX_PARTS) return set_err("ace_set_parts: parts must be 0 (automatic) or 1-4");
    ctx->parts = parts;
    return 0;
}

extern "C" int ace_set_tile(ace_ctx* ctx, int W) {
    if (!ctx) return set_err("ace_set_tile: null context");
    for (size_t q = 0; q < ctx->wcfg.size(); ++q)
        if (ctx->wcfg[q]->W == W) {
            ctx->cur_w = (int)q;
            return 0;
        }
    return set_err("ace_set_tile: waves per tile not in the pack");
}

extern "C" int ace_set_validation(ace_ctx* ctx, int on) {
    if (!ctx) return set_err("ace_set_validation: null context");
    ctx->validate = on != 0;
    return 0;
}

extern "C" void ace_ctx_destroy(ace_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    delete ctx;
}

static RadialParams radial_params(ace_ctx* ctx) {
    RadialParams rp;
    rp.A = ctx->rA.p;
    rp.B = ctx->rB.p;
    rp.C = ctx->rC.p;
    rp.r0 = ctx->r0;
    rp.tl = ctx->tl;
    rp.tr = ctx->tr;
    rp.nr = ctx->nr;
    rp.tp = ctx->tp;
    rp.pl = ctx->pl;
    rp.pr = ctx->pr;
    return rp;
}

static int check_common(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                        const int8_t* species, int64_t* nnbr_out, void* stream) {
    if (!ctx) return set_err("null context");
    if (nenv < 0) return set_err("nenv < 0");
    if (nenv > 0 && (!off || !xyz)) return set_err("null env_offsets / xyz");
    if (ctx->nspecies > 0 && nenv > 0 && !species)
        return set_err("species array required for a categorical 1p basis");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing((hipStream_t)stream, &cap));
    if (cap != hipStreamCaptureStatusNone) {
        // stream capture (hipGraph): no host reads; the neighbour count comes
        // from the eager call on the same batch that must precede the capture
        // (it also sizes every scratch buffer)
        if (nenv > 0 && (off != ctx->batch_off || nenv != ctx->batch_nenv))
            return set_err("stream capture needs an eager call on the same batch (env_offsets, nenv) first");
        if (ctx->validate) return set_err("stream capture needs validation off (ace_set_validation(ctx, 0))");
        *nnbr_out = nenv > 0 ? ctx->batch_nnbr : 0;
        return 0;
    }
    int32_t last = 0;
    if (nenv > 0) {
        int32_t first = 0;
        HIP_TRY(hipMemcpy(&first, off, sizeof(int32_t), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&last, off + nenv, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (first != 0 || last < 0) return set_err("env_offsets must start at 0 and be non-negative");
        ctx->batch_off = off;
        ctx->batch_nenv = nenv;
        ctx->batch_nnbr = last;
    }
    *nnbr_out = last;
    return 0;
}

// host-side validation of offsets and species (bounds for every kernel)
static int validate_inputs(ace_ctx* ctx, int64_t nenv, const int32_t* off, int64_t nnbr,
                           const int8_t* species) {
    if (nenv == 0 || !ctx->validate) return 0;
    std::vector<int32_t> h(nenv + 1);
    HIP_TRY(hipMemcpy(h.data(), off, sizeof(int32_t) * (nenv + 1), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < nenv; ++e)
        if (h[e + 1] < h[e]) return set_err("env_offsets must be non-decreasing");
    if (species && ctx->nspecies > 0 && nnbr > 0) {
        std::vector<int8_t> s(nnbr);
        HIP_TRY(hipMemcpy(s.data(), species, nnbr, hipMemcpyDeviceToHost));
        for (int64_t j = 0; j < nnbr; ++j)
            if (s[j] < 0 || s[j] >= ctx->nspecies) return set_err("species index out of range (unknown category)");
    }
    return 0;
}

template <int L>
static void launch_pool(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off,
                        const double* xyz, int64_t nnbr, const int8_t* species, int64_t estride) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    // two environments per wave when a half-wave covers the pooled functions
    const int epw = ctx->npool <= 32 ? 2 : 1;
    const size_t lds = (size_t)epw * POOL_CH * (ctx->nr + 2 * NP + 1) * sizeof(double);
    int32_t* nbr_env = nullptr;
    if (ctx->flat_force && nnbr > 0) {
        if (ctx->nbr_env.ensure(nnbr)) return;  // error recorded; the force launch checks it again
        nbr_env = ctx->nbr_env.p;
    }
    const dim3 grid((unsigned)((nenv + epw - 1) / epw));
    const int8_t* sp = ctx->nspecies > 0 ? species : nullptr;
    if (epw == 2) {
        set_func_lds((const void*)k_pool<L, 2>, lds);
        hipLaunchKernelGGL((k_pool<L, 2>), grid, dim3(64), lds, st, radial_params(ctx), nenv, off, xyz, nnbr, sp,
                           ctx->pool.p, ctx->npool, ctx->Aplus.p, estride, nbr_env);
    } else {
        set_func_lds((const void*)k_pool<L, 1>, lds);
        hipLaunchKernelGGL((k_pool<L, 1>), grid, dim3(64), lds, st, radial_params(ctx), nenv, off, xyz, nnbr, sp,
                           ctx->pool.p, ctx->npool, ctx->Aplus.p, estride, nbr_env);
    }
}

// k_force: one wave per environment, lanes are neighbours.  The
// environment's adjoint is staged in LDS as a dense table
// Gd[mu][l, m >= 0][n] (zero where the pooled function is absent), so the
// contraction needs no index lookup.  R_n and dR_n stay in registers (nr <=
// 16); the ALP values are produced m-outer, l-inner, by the same operations
// as alp_eval_ed, so only the two previous degrees are live.
struct ForceTab {
    int8_t nrl[ACE_LMAX + 1];  // n bound per l (max pooled n over mu, m)
};

template <int L, int PROP>
__global__ __launch_bounds__(64) void k_force(RadialParams rp, ForceTab ft, int64_t nenv,
                                               const int32_t* __restrict__ off,
                                               const double* __restrict__ xyz, int64_t nnbr,
                                               const int8_t* __restrict__ species, int nsp,
                                               const int32_t* __restrict__ pool, int npool,
                                               const int32_t* __restrict__ gslot_ptr,
                                               const int32_t* __restrict__ gslot_list,
                                               const double2* __restrict__ Gbuf, int64_t estride,
                                               double* __restrict__ dE) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double2 Gd[];  // [nsp][NP][nr][NC]
    const int lane = threadIdx.x;
    const int64_t env = xcd_env(blockIdx.x, nenv);
    const int nr = rp.nr;
    if (env >= nenv) return;
    const int ntab = nsp * NP * nr * NC;
    for (int q = lane; q < ntab; q += 64) Gd[q] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int k = lane; k < npool; k += 64) {
        const int mu = max(0, pool[4 * k]), n = pool[4 * k + 1] - 1, l = pool[4 * k + 2], m = pool[4 * k + 3];
        const int d = ((mu * NP + idxp(l, m)) * nr + n) * NC;
#pragma unroll
        for (int p = 0; p < NC; ++p) {
            double2 g = make_double2(0.0, 0.0);
            for (int s = gslot_ptr[k]; s < gslot_ptr[k + 1]; ++s) {
                const double2 v = Gbuf[((int64_t)gslot_list[s] * NC + p) * estride + env];
                g.x += v.x;
                g.y += v.y;
            }
            Gd[d + p] = g;
        }
    }
    __syncthreads();
    const int j0 = off[env], j1 = off[env + 1];
    for (int j = j0 + lane; j < j1; j += 64) {
        const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
        const int mu = species ? (int)species[j] : 0;
        const double2* Gm = Gd + (int64_t)mu * NP * nr * NC;
        const Sph S = cart2spher(x, y, z);
        double t, dtdr;
        transform_ed(rp, S.r, t, dtdr);
        double R[FORCE_NRMAX], dR[FORCE_NRMAX];
        {
            double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
            R[0] = p0;
            dR[0] = d0;
            double p1 = 0.0, d1 = 0.0;
            if (nr > 1) {
                const double al = rp.A[1] * t + rp.B[1];
                p1 = al * p0;
                d1 = al * d0 + rp.A[1] * p0;
            }
            R[1] = p1;
            dR[1] = d1;
#pragma unroll
            for (int n = 2; n < FORCE_NRMAX; ++n) {
                if (n < nr) {
                    const double al = rp.A[n] * t + rp.B[n];
                    const double p2 = al * p1 + rp.C[n] * p0;
                    const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                    p0 = p1;
                    p1 = p2;
                    d0 = d1;
                    d1 = d2;
                }
                R[n] = p1;
                dR[n] = d1;
            }
        }
        const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
        const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
        const double st2 = S.sth * S.sth;
        double g[NC][3];
#pragma unroll
        for (int p = 0; p < NC; ++p) g[p][0] = g[p][1] = g[p][2] = 0.0;
        double epr = 0.7071067811865475, epi = 0.0;
        double Pmm = 0.3989422804014327, dPmm = 0.0;  // P_{m,m}, dP_{m,m} (P/sin(theta) for m >= 1)
        const double temp0 = 0.3989422804014327;
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            if (m == 1) {
                Pmm = -1.224744871391589 * temp0;
                dPmm = -1.224744871391589 * (S.cth * temp0);
            } else if (m >= 2) {
                const double f = sqrt(1.0 + 0.5 / m);
                const double nt1 = -f * S.sth * Pmm;
                const double ntd = -f * (S.cth * Pmm * S.sth + S.sth * dPmm);
                Pmm = nt1;
                dPmm = ntd;
            }
            if (m > 0) {
                const double nr_ = epr * S.cphi - epi * S.sphi;
                const double ni_ = epr * S.sphi + epi * S.cphi;
                epr = nr_;
                epi = ni_;
            }
            double Pa = 0.0, dPa = 0.0, Pb = Pmm, dPb = dPmm;  // degrees l-2, l-1 of order m
#pragma unroll
            for (int l = m; l <= L; ++l) {
                double Pl, dPl;
                if (l == m) {
                    Pl = Pmm;
                    dPl = dPmm;
                } else if (l == m + 1) {
                    if (m == 0) {
                        Pl = S.cth * 1.7320508075688772 * temp0;
                        dPl = -S.sth * 1.7320508075688772 * temp0;
                    } else {
                        const double c = sqrt(2.0 * m + 3.0);
                        Pl = c * S.cth * Pmm;
                        dPl = c * (-st2 * Pmm + S.cth * dPmm);
                    }
                } else {
                    const int i = idxp(l, m);
                    Pl = c_alpA[i] * (S.cth * Pb + c_alpB[i] * Pa);
                    dPl = m == 0 ? c_alpA[i] * (-S.sth * Pb + S.cth * dPb + c_alpB[i] * dPa)
                                 : c_alpA[i] * (-st2 * Pb + S.cth * dPb + c_alpB[i] * dPa);
                }
                if (l > m) {
                    Pa = Pb;
                    dPa = dPb;
                    Pb = Pl;
                    dPb = dPl;
                }
                // Y and its Cartesian gradient (cYlm_ed!, dspher_to_dcart)
                double Yr, Yi, fpr, fpi, ftr, fti;
                if (m == 0) {
                    Yr = Pl * 0.7071067811865475;
                    Yi = 0.0;
                    fpr = fpi = 0.0;
                    ftr = dPl * 0.7071067811865475;
                    fti = 0.0;
                } else {
                    Yr = epr * Pl * S.sth;
                    Yi = epi * Pl * S.sth;
                    fpr = -m * epi * Pl;
                    fpi = m * epr * Pl;
                    ftr = epr * dPl;
                    fti = epi * dPl;
                }
                const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                       (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                       (-S.sth * ftr) * rinv};
                const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                       (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                       (-S.sth * fti) * rinv};
                const double2* Gl = Gm + idxp(l, m) * nr * NC;
                const int nb = ft.nrl[l];
#pragma unroll
                for (int n = 0; n < FORCE_NRMAX; ++n) {
                    if (n < nb) {
                        const double Rn = R[n], dRn = dR[n] * dtdr;
#pragma unroll
                        for (int p = 0; p < NC; ++p) {
                            const double2 G = Gl[n * NC + p];
                            // Re(G * (dR rhat Y + R dY))
                            const double gy = G.x * Yr - G.y * Yi;
                            g[p][0] += dRn * gy * rx + Rn * (G.x * dYr[0] - G.y * dYi[0]);
                            g[p][1] += dRn * gy * ry + Rn * (G.x * dYr[1] - G.y * dYi[1]);
                            g[p][2] += dRn * gy * rz + Rn * (G.x * dYr[2] - G.y * dYi[2]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < NC; ++p)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) dE[(int64_t)(p * 3 + dd) * nnbr + j] = g[p][dd];
    }
}

// k_forcef: forces over the flat neighbour list (Invariant, one species,
// nr <= 16).  Every lane is a neighbour, whatever environment it belongs to,
// so no lane idles on short environments.  A lane reads its environment's
// adjoint G[slot][env] straight from memory (lanes of one environment read the
// same address; neighbouring environments are adjacent), with the slot of
// every (l, m, n) in the kernel arguments.  Radial values stay in registers
// and the ALP values are produced m-outer exactly as in k_force.

template <int L, int NRM, bool STAGED, int MAXE>
__device__ __forceinline__ void forcef_lane(const RadialParams& rp, const ForceFlatTab& ft, int64_t nnbr, int64_t j,
                                            const double* __restrict__ xyz, const double2* __restrict__ Ge,
                                            const double2* __restrict__ Gs, int64_t estride,
                                            double* __restrict__ dE) {
    const int nr = rp.nr;
    const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
    const Sph S = cart2spher(x, y, z);
    double t, dtdr;
    transform_ed(rp, S.r, t, dtdr);
    double R[NRM], dR[NRM];
    {
        double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
        R[0] = p0;
        dR[0] = d0;
        double p1 = 0.0, d1 = 0.0;
        if (nr > 1) {
            const double al = rp.A[1] * t + rp.B[1];
            p1 = al * p0;
            d1 = al * d0 + rp.A[1] * p0;
        }
        R[1] = p1;
        dR[1] = d1;
#pragma unroll
        for (int n = 2; n < NRM; ++n) {
            if (n < nr) {
                const double al = rp.A[n] * t + rp.B[n];
                const double p2 = al * p1 + rp.C[n] * p0;
                const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                p0 = p1;
                p1 = p2;
                d0 = d1;
                d1 = d2;
            }
            R[n] = p1;
            dR[n] = d1;
        }
    }
    const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
    const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
    const double st2 = S.sth * S.sth;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    int q = 0;  // running (l, m, n) index of the staged table
    double epr = 0.7071067811865475, epi = 0.0;
    double Pmm = 0.3989422804014327, dPmm = 0.0;
    const double temp0 = 0.3989422804014327;
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        if (m == 1) {
            Pmm = -1.224744871391589 * temp0;
            dPmm = -1.224744871391589 * (S.cth * temp0);
        } else if (m >= 2) {
            const double f = sqrt(1.0 + 0.5 / m);
            const double nt1 = -f * S.sth * Pmm;
            const double ntd = -f * (S.cth * Pmm * S.sth + S.sth * dPmm);
            Pmm = nt1;
            dPmm = ntd;
        }
        if (m > 0) {
            const double nr_ = epr * S.cphi - epi * S.sphi;
            const double ni_ = epr * S.sphi + epi * S.cphi;
            epr = nr_;
            epi = ni_;
        }
        double Pa = 0.0, dPa = 0.0, Pb = Pmm, dPb = dPmm;
#pragma unroll
        for (int l = m; l <= L; ++l) {
            double Pl, dPl;
            if (l == m) {
                Pl = Pmm;
                dPl = dPmm;
            } else if (l == m + 1) {
                if (m == 0) {
                    Pl = S.cth * 1.7320508075688772 * temp0;
                    dPl = -S.sth * 1.7320508075688772 * temp0;
                } else {
                    const double c = sqrt(2.0 * m + 3.0);
                    Pl = c * S.cth * Pmm;
                    dPl = c * (-st2 * Pmm + S.cth * dPmm);
                }
            } else {
                const int i = idxp(l, m);
                Pl = c_alpA[i] * (S.cth * Pb + c_alpB[i] * Pa);
                dPl = m == 0 ? c_alpA[i] * (-S.sth * Pb + S.cth * dPb + c_alpB[i] * dPa)
                             : c_alpA[i] * (-st2 * Pb + S.cth * dPb + c_alpB[i] * dPa);
            }
            if (l > m) {
                Pa = Pb;
                dPa = dPb;
                Pb = Pl;
                dPb = dPl;
            }
            double Yr, Yi, fpr, fpi, ftr, fti;
            if (m == 0) {
                Yr = Pl * 0.7071067811865475;
                Yi = 0.0;
                fpr = fpi = 0.0;
                ftr = dPl * 0.7071067811865475;
                fti = 0.0;
            } else {
                Yr = epr * Pl * S.sth;
                Yi = epi * Pl * S.sth;
                fpr = -m * epi * Pl;
                fpi = m * epr * Pl;
                ftr = epr * dPl;
                fti = epi * dPl;
            }
            const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                   (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                   (-S.sth * ftr) * rinv};
            const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                   (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                   (-S.sth * fti) * rinv};
            const int16_t* sl = ft.slot[idxp(l, m)];
            const int nb = ft.nrl[l];
#pragma unroll
            for (int n = 0; n < NRM; ++n) {
                if (n < nb) {
                    const double2 G = STAGED ? Gs[(q + n) * MAXE] : Ge[(int64_t)sl[n] * estride];
                    const double Rn = R[n], dRn = dR[n] * dtdr;
                    const double gy = G.x * Yr - G.y * Yi;
                    g0 += dRn * gy * rx + Rn * (G.x * dYr[0] - G.y * dYi[0]);
                    g1 += dRn * gy * ry + Rn * (G.x * dYr[1] - G.y * dYi[1]);
                    g2 += dRn * gy * rz + Rn * (G.x * dYr[2] - G.y * dYi[2]);
                }
            }
            q += nb;
        }
    }
    dE[j] = g0;
    dE[nnbr + j] = g1;
    dE[2 * nnbr + j] = g2;
}

template <int L, int NRM, int MAXE>
__global__ __launch_bounds__(256, 4) void k_forcef(RadialParams rp, ForceFlatTab ft, int64_t nnbr,
                                                 const double* __restrict__ xyz,
                                                 const int32_t* __restrict__ nbr_env,
                                                 const int32_t* __restrict__ ffq,
                                                 const double2* __restrict__ Gbuf, int64_t estride,
                                                 double* __restrict__ dE) {
    extern __shared__ double2 Gst[];  // [q][MAXE]
    const int64_t jb = (int64_t)blockIdx.x * 256;
    const int64_t j = jb + threadIdx.x;
    const int64_t jl = min(jb + 255, nnbr - 1);
    // the block's environments (neighbours are grouped by environment)
    const int64_t e0 = nbr_env[jb];
    const int ne = (int)(nbr_env[jl] - e0 + 1);
    if (ne <= MAXE) {
        for (int i = threadIdx.x; i < ft.nq * ne; i += 256) {
            const int qq = i / ne, el = i - qq * ne;
            Gst[qq * MAXE + el] = Gbuf[(int64_t)ffq[qq] * estride + e0 + el];
        }
        __syncthreads();
        if (j >= nnbr) return;  // no barriers below
        const int el = (int)(nbr_env[j] - e0);
        forcef_lane<L, NRM, true, MAXE>(rp, ft, nnbr, j, xyz, nullptr, Gst + el, estride, dE);
    } else {
        if (j >= nnbr) return;
        forcef_lane<L, NRM, false, MAXE>(rp, ft, nnbr, j, xyz, Gbuf + nbr_env[j], nullptr, estride, dE);
    }
}

// G[primary slot] += G[extra slot] for pooled functions whose gradient
// polynomial was split over several tasks (k_forcef reads one slot per k).
__global__ void k_gfold(int64_t nenv, int nfold, const int32_t* __restrict__ fold, double2* __restrict__ Gbuf,
                        int64_t estride) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    for (int q = 0; q < nfold; ++q) {
        double2* dst = Gbuf + (int64_t)fold[2 * q] * estride + e;
        const double2 v = Gbuf[(int64_t)fold[2 * q + 1] * estride + e];
        dst->x += v.x;
        dst->y += v.y;
    }
}

template <int L, int PROP>
static void launch_force(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off,
                         const double* xyz, int64_t nnbr, const int8_t* species, int64_t estride,
                         double* dE) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    constexpr int NP = (L + 1) * (L + 2) / 2;
    if (PROP == 0 && ctx->flat_force) {
        if (ctx->nfold > 0)
            hipLaunchKernelGGL(k_gfold, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, ctx->nfold,
                               ctx->gfold.p, ctx->Gbuf.p, estride);
        // radial functions held in registers: 12 when they suffice (fewer
        // registers, more waves per SIMD), else 16
        // blocks of short environments span more of them: stage 32 when that fits
        const bool wide = (size_t)ctx->fflat.nq * FF_MAXE_WIDE * sizeof(double2) <= FF_LDS;
        const size_t glds = (size_t)ctx->fflat.nq * (wide ? FF_MAXE_WIDE : FF_MAXE) * sizeof(double2);
        const dim3 grid((unsigned)((nnbr + 255) / 256));
        const int32_t* ne_ = ctx->nbr_env.p;
        if (ctx->nr <= 12 && wide)
            hipLaunchKernelGGL((k_forcef<L, 12, FF_MAXE_WIDE>), grid, dim3(256), glds, st, radial_params(ctx),
                               ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ctx->Gbuf.p, estride, dE);
        else if (ctx->nr <= 12)
            hipLaunchKernelGGL((k_forcef<L, 12, FF_MAXE>), grid, dim3(256), glds, st, radial_params(ctx),
                               ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ctx->Gbuf.p, estride, dE);
        else if (wide)
            hipLaunchKernelGGL((k_forcef<L, FORCE_NRMAX, FF_MAXE_WIDE>), grid, dim3(256), glds, st,
                               radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ctx->Gbuf.p, estride, dE);
        else
            hipLaunchKernelGGL((k_forcef<L, FORCE_NRMAX, FF_MAXE>), grid, dim3(256), glds, st,
                               radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ctx->Gbuf.p, estride, dE);
        return;
    }
    const int nsp = std::max(1, ctx->nspecies);
    const size_t dense = (size_t)nsp * NP * ctx->nr * NC * sizeof(double2);
    if (ctx->nr <= FORCE_NRMAX && dense <= 64 * 1024) {
        ForceTab ft;
        for (int l = 0; l <= ACE_LMAX; ++l) ft.nrl[l] = (int8_t)(l <= L ? ctx->nrl[l] : 0);
        set_func_lds((const void*)k_force<L, PROP>, dense);
        hipLaunchKernelGGL((k_force<L, PROP>), dim3((unsigned)nenv), dim3(64), dense, st, radial_params(ctx), ft,
                           nenv, off, xyz, nnbr, ctx->nspecies > 0 ? species : nullptr, nsp, ctx->pool.p,
                           ctx->npool, ctx->gslot_ptr.p, ctx->gslot_list.p, ctx->Gbuf.p, estride, dE);
        return;
    }
    const size_t lds = (size_t)ctx->npool * NC * sizeof(double2) + 64 * 2 * ctx->nr * sizeof(double);
    set_func_lds((const void*)k_force_lds<L, PROP>, lds);
    hipLaunchKernelGGL((k_force_lds<L, PROP>), dim3((unsigned)nenv), dim3(64), lds, st, radial_params(ctx),
                       nenv, off, xyz, nnbr, ctx->nspecies > 0 ? species : nullptr, ctx->kidx.p,
                       ctx->npool, ctx->gslot_ptr.p, ctx->gslot_list.p, ctx->Gbuf.p, estride, dE);
}

#define DISPATCH_L(L_, FN, ...)                          \
    switch (L_) {                                        \
        case 0: FN<0>(__VA_ARGS__); break;               \
        case 1: FN<1>(__VA_ARGS__); break;               \
        case 2: FN<2>(__VA_ARGS__); break;               \
        case 3: FN<3>(__VA_ARGS__); break;               \
        case 4: FN<4>(__VA_ARGS__); break;               \
        case 5: FN<5>(__VA_ARGS__); break;               \
        case 6: FN<6>(__VA_ARGS__); break;               \
        case 7: FN<7>(__VA_ARGS__); break;               \
        default: FN<8>(__VA_ARGS__); break;              \
    }

template <int L>
static void launch_force_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off,
                             const double* xyz, int64_t nnbr, const int8_t* species,
                             int64_t estride, double* dE) {
    if (ctx->prop == ACE_PROP_INVARIANT)
        launch_force<L, 0>(ctx, st, nenv, off, xyz, nnbr, species, estride, dE);
    else
        launch_force<L, 1>(ctx, st, nenv, off, xyz, nnbr, species, estride, dE);
}

static int run_path(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                    const int8_t* species, double* E_out, double* dE, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    if (!E_out) return set_err("null E_out");
    hipStream_t st = (hipStream_t)stream;
    const int NC = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    const int64_t ntile = (nenv + 63) / 64;
    const int64_t estride = ntile * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    const bool forces = dE != nullptr;
    if (forces && ctx->Gbuf.ensure((size_t)std::max(1, ctx->ngslots) * NC * estride)) return -1;
    WaveCfg* wc = ctx->wcfg[ctx->cur_w];
    // under stream capture (hipGraph) the timing events are left out of the graph
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    const bool timed = cap == hipStreamCaptureStatusNone;
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[0], st));
    DISPATCH_L(ctx->L, launch_pool, ctx, st, nenv, off, xyz, nnbr, species, estride);
    HIP_TRY(hipGetLastError());
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[1], st));
    DagArgs a;
    a.runs = ctx->runs.p;
    a.leaf_fid = ctx->leaf_fid.p;
    a.leaf_coef = ctx->leaf_coef.p;
    a.vleaf = ctx->vleaf.p;
    a.tasks = ctx->d_tasks.p;
    a.task_const = ctx->task_const.p;
    a.wtask_ptr = forces ? wc->d_ptr_all.p : wc->d_ptr_e.p;
    a.wtask_list = forces ? wc->d_list_all.p : wc->d_list_e.p;
    a.npool = ctx->npool;
    a.W = wc->W;
    size_t lds = (size_t)ctx->npool * 64 * sizeof(double2);
    lds = std::max(lds, (size_t)wc->W * NC * 64 * sizeof(double));
    if (ctx->prop == ACE_PROP_INVARIANT && lds > 160 * 1024)
        return set_err("pooled 1p basis too large for one LDS tile (npool > 160)");
    if (ctx->prop == ACE_PROP_INVARIANT) {
        // Parts per tile: the grid runs in rounds of ncu workgroups, and the last
        // round may be nearly empty; splitting each tile's tasks over P
        // workgroups shortens the rounds.  Pick P in 1..4 minimising
        // rounds(P) / P, with 2 % per extra part for re-staging the A+ tile.
        // A split is only considered while every virtual wave keeps at least 3
        // tasks (fewer leave the part workgroups unbalanced or idle).
        int P = ctx->parts;
        const int ntask_mode = forces ? wc->ptr_all[wc->W] : wc->ptr_e[wc->W];
        int pmax = 1;
        while (pmax < MAX_PARTS && ntask_mode >= 3 * (pmax + 1) * wc->W) ++pmax;
        if (P == 0) {
            double best = 1e300;
            for (int q = 1; q <= pmax; ++q) {
                const double rounds = (double)((ntile * q + ctx->ncu - 1) / ctx->ncu);
                const double cost = rounds / q * (1.0 + 0.02 * (q - 1));
                if (cost < best - 1e-12) {
                    best = cost;
                    P = q;
                }
            }
        }
        ctx->last_parts = P;
        // tail tiles with P + 1 parts (P >= 2, so every tile sums shares in k_esum)
        int64_t nhi = 0;
        if (ctx->tail >= 0) {
            nhi = std::min<int64_t>(ctx->tail, ntile);
            if (nhi > 0 && (P < 2 || P >= MAX_PARTS))
                return set_err("ace_set_tail_parts: tail tiles need 2 or 3 parts per tile (ace_set_parts)");
        } else if (ctx->parts == 0 && P >= 2 && P + 1 <= MAX_PARTS && ntask_mode >= 2 * (P + 1) * wc->W) {
            const auto key = std::make_pair(ntile, P);
            auto it = ctx->tail_cache.find(key);
            if (it == ctx->tail_cache.end()) {
                if (ctx->tail_cache.size() >= 64) ctx->tail_cache.clear();
                it = ctx->tail_cache.emplace(key, choose_tail(ntile, P, ctx->ncu)).first;
            }
            nhi = it->second;
        }
        DagRecArgs ra;
        ra.recs = ctx->recs.p;
        ra.tasks = ctx->d_tasks.p;
        ra.task_const = ctx->task_const.p;
        ra.wtask_ptr = P == 1 ? a.wtask_ptr : (forces ? wc->pp_all[P].p : wc->pp_e[P].p);
        ra.wtask_list = P == 1 ? a.wtask_list : (forces ? wc->pl_all[P].p : wc->pl_e[P].p);
        ra.npool = ctx->npool;
        ra.W = wc->W;
        ra.parts = P;
        ra.ntile_lo = ntile - nhi;
        ra.wtask_ptr_hi = nhi > 0 ? (forces ? wc->pp_all[P + 1].p : wc->pp_e[P + 1].p) : nullptr;
        ra.wtask_list_hi = nhi > 0 ? (forces ? wc->pl_all[P + 1].p : wc->pl_e[P + 1].p) : nullptr;
        const int nshare = nhi > 0 ? P + 1 : P;
        if (P > 1 && ctx->Epart.ensure((size_t)nshare * estride)) return -1;
        const int64_t nblk = (ntile - nhi) * P + nhi * (P + 1);
        set_func_lds((const void*)k_dagr, lds);
        hipLaunchKernelGGL(k_dagr, dim3((unsigned)nblk), dim3(64 * wc->W), lds, st, ra, nenv,
                           ctx->Aplus.p, estride, P == 1 ? E_out : ctx->Epart.p, ctx->Gbuf.p, ctx->e0r,
                           (const int32_t*)nullptr, (const int32_t*)nullptr);
        if (P > 1)
            hipLaunchKernelGGL(k_esum, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, P,
                               (ntile - nhi) * 64, ctx->Epart.p, estride, ctx->e0r, E_out);
    } else if ((size_t)ctx->npool * 64 * sizeof(double2) <= 160 * 1024) {
        set_func_lds((const void*)k_dag<1, 64>, lds);
        hipLaunchKernelGGL((k_dag<1, 64>), dim3((unsigned)ntile), dim3(64 * wc->W), lds, st, a, nenv,
                           ctx->Aplus.p, estride, E_out, ctx->Gbuf.p, ctx->e0r, ctx->e0v.p);
    } else {
        // 32-environment tiles, two task streams per wave: the W stream
        // assignment runs on W/2 waves
        if (wc->W < 2) return set_err("tile configuration needs an even number of streams");
        const int64_t nt32 = (nenv + 31) / 32;
        size_t lds32 = std::max((size_t)ctx->npool * 32 * sizeof(double2), (size_t)wc->W * NC * 32 * sizeof(double));
        if (lds32 > 160 * 1024) return set_err("pooled 1p basis too large for one LDS tile (npool > 320)");
        set_func_lds((const void*)k_dag<1, 32>, lds32);
        hipLaunchKernelGGL((k_dag<1, 32>), dim3((unsigned)nt32), dim3(64 * (wc->W / 2)), lds32, st, a, nenv,
                           ctx->Aplus.p, estride, E_out, ctx->Gbuf.p, ctx->e0r, ctx->e0v.p);
    }
    HIP_TRY(hipGetLastError());
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[2], st));
    if (forces && nnbr > 0) {
        DISPATCH_L(ctx->L, launch_force_any, ctx, st, nenv, off, xyz, nnbr, species, estride, dE);
        HIP_TRY(hipGetLastError());
    }
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[3], st));
    ctx->timing_valid = timed;
    return 0;
}

extern "C" int ace_energy(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                          const int8_t* species, double* E_out, void* stream) {
    return run_path(ctx, nenv, off, xyz, species, E_out, nullptr, stream);
}

extern "C" int ace_energy_forces(ace_ctx* ctx, int64_t nenv, const int32_t* off,
                                 const double* xyz, const int8_t* species, double* E_out,
                                 double* dE_dr, void* stream) {
    if (!dE_dr) return set_err("null dE_dr");
    return run_path(ctx, nenv, off, xyz, species, E_out, dE_dr, stream);
}

extern "C" int ace_last_timing(ace_ctx* ctx, double* dag_ms, double* pool_ms, double* force_ms) {
    if (!ctx) return set_err("null context");
    if (!ctx->timing_valid) return set_err("no timed call yet");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventSynchronize(ctx->ev[3]));
    float a = 0, b = 0, c = 0;
    HIP_TRY(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
    HIP_TRY(hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]));
    if (pool_ms) *pool_ms = a;
    if (dag_ms) *dag_ms = b;
    if (force_ms) *force_ms = c;
    return 0;
}

// ---- basis level

static int basis_A_internal(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, int64_t nnbr, hipStream_t st) {
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    if (ctx->Afull.ensure((size_t)nenv * ctx->nA)) return -1;
    DISPATCH_L(ctx->L, launch_pool, ctx, st, nenv, off, xyz, nnbr, species, estride);
    HIP_TRY(hipGetLastError());
    const int64_t n = nenv * ctx->nA;
    hipLaunchKernelGGL(k_expandA, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nenv, ctx->nA,
                       ctx->aspec_fid.p, ctx->aspec_sgn.p, ctx->Aplus.p, estride, ctx->Afull.p);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ace_basis_A(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                           const int8_t* species, double* A_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    HIP_TRY(hipMemcpyAsync(A_out, ctx->Afull.p, sizeof(double2) * nenv * ctx->nA,
                           hipMemcpyDeviceToDevice, st));
    return 0;
}

static int basis_AA_internal(ace_ctx* ctx, int64_t nenv, hipStream_t st) {
    if (ctx->AA.ensure((size_t)nenv * ctx->nAA)) return -1;
    const int64_t n = nenv * ctx->nAA;
    hipLaunchKernelGGL(k_AA, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nenv, ctx->nA, ctx->nAA,
                       ctx->maxorder, ctx->orders.p, ctx->iAA2iA.p, ctx->Afull.p, ctx->AA.p);
    HIP_TRY(hipGetLastError());
    return 0;
}

__global__ void k_realpart(int64_t n, const double2* __restrict__ in, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i].x;
}

extern "C" int ace_basis_AA(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, double* AA_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    if (basis_AA_internal(ctx, nenv, st)) return -1;
    const int64_t n = nenv * ctx->nAA;
    if (ctx->prop == ACE_PROP_INVARIANT) {
        hipLaunchKernelGGL(k_realpart, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ctx->AA.p, AA_out);
    } else {
        HIP_TRY(hipMemcpyAsync(AA_out, ctx->AA.p, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Fused basis values (k_Bt): whether the pooled tile fits LDS, and the launch.
static bool bt_fits(const ace_ctx* ctx, bool dir) {
    const int envs = dir ? 32 : 64;
    return (size_t)ctx->npool * envs * sizeof(double2) * (dir ? 2 : 1) <= 160 * 1024;
}

template <bool DIR>
static int launch_bt(ace_ctx* ctx, hipStream_t st, int64_t nenv, int64_t estride, double* out) {
    constexpr int ENVS = DIR ? 32 : 64;
    const size_t lds = (size_t)ctx->npool * ENVS * sizeof(double2) * (DIR ? 2 : 1);
    constexpr int NW = BT_NW;
    const int rpw = std::max(1, (ctx->nB + NW * 12 - 1) / (NW * 12));
    const int nchunk = (ctx->nB + NW * rpw - 1) / (NW * rpw);
    const int64_t nblk = (nenv + ENVS - 1) / ENVS * nchunk;
    if (nblk >= ((int64_t)1 << 31)) return set_err("batch too large for one launch");
    if (ctx->prop == ACE_PROP_INVARIANT) {
        set_func_lds((const void*)k_Bt<0, DIR, ENVS>, lds);
        hipLaunchKernelGGL((k_Bt<0, DIR, ENVS>), dim3((unsigned)nblk), dim3(64 * NW), lds, st, nenv, ctx->npool,
                           ctx->nB, nchunk, rpw, ctx->rowptr.p, ctx->bt_rec.p, ctx->vals.p, ctx->Aplus.p,
                           ctx->Wplus.p, estride, out);
    } else {
        if (DIR) return set_err("directional basis derivative: Invariant bases only");
        set_func_lds((const void*)k_Bt<1, false, 64>, lds);
        hipLaunchKernelGGL((k_Bt<1, false, 64>), dim3((unsigned)nblk), dim3(64 * NW), lds, st, nenv, ctx->npool,
                           ctx->nB, nchunk, rpw, ctx->rowptr.p, ctx->bt_rec.p, ctx->vals.p, ctx->Aplus.p,
                           ctx->Wplus.p, estride, out);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ace_basis_B(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                           const int8_t* species, double* B_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    {
        const int64_t estride = ((nenv + 63) / 64) * 64;
        const int fits = bt_fits(ctx, false);
        if (fits) {
            if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
            DISPATCH_L(ctx->L, launch_pool, ctx, st, nenv, off, xyz, nnbr, species, estride);
            HIP_TRY(hipGetLastError());
            return launch_bt<false>(ctx, st, nenv, estride, B_out);
        }
    }
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    if (basis_AA_internal(ctx, nenv, st)) return -1;
    const int64_t n = nenv * ctx->nB;
    if (ctx->prop == ACE_PROP_INVARIANT)
        hipLaunchKernelGGL(k_B<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nenv, ctx->nB, ctx->nAA,
                           ctx->rowptr.p, ctx->cols.p, ctx->vals.p, ctx->AA.p, B_out);
    else
        hipLaunchKernelGGL(k_B<1>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nenv, ctx->nB, ctx->nAA,
                           ctx->rowptr.p, ctx->cols.p, ctx->vals.p, ctx->AA.p, B_out);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int L>
static void launch_dphi(ace_ctx* ctx, hipStream_t st, int64_t nnbr, const double* xyz,
                        const int8_t* species) {
    hipLaunchKernelGGL(k_dphi<L>, dim3((unsigned)((nnbr + 63) / 64)), dim3(64), 0, st, radial_params(ctx),
                       nnbr, ctx->nbr_env.p, xyz, ctx->nspecies > 0 ? species : nullptr, ctx->aspec4.p,
                       ctx->nA, ctx->dphi.p);
}

extern "C" int ace_basis_dB(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, double* dB_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0 || nnbr == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    if (ctx->nbr_env.ensure(nnbr) || ctx->dphi.ensure((size_t)nnbr * ctx->nA * 3) ||
        ctx->dAA.ensure((size_t)nnbr * ctx->nAA * 3))
        return -1;
    hipLaunchKernelGGL(k_nbr_env, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, off,
                       ctx->nbr_env.p);
    DISPATCH_L(ctx->L, launch_dphi, ctx, st, nnbr, xyz, species);
    HIP_TRY(hipGetLastError());
    const int64_t n1 = nnbr * ctx->nAA;
    hipLaunchKernelGGL(k_dAA, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, nnbr, ctx->nA, ctx->nAA,
                       ctx->maxorder, ctx->nbr_env.p, ctx->orders.p, ctx->iAA2iA.p, ctx->Afull.p,
                       ctx->dphi.p, ctx->dAA.p);
    const int64_t n2 = nnbr * ctx->nB;
    if (ctx->prop == ACE_PROP_INVARIANT)
        hipLaunchKernelGGL(k_dB<0>, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, nnbr, ctx->nB,
                           ctx->nAA, ctx->rowptr.p, ctx->cols.p, ctx->vals.p, ctx->dAA.p, dB_out);
    else
        hipLaunchKernelGGL(k_dB<1>, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, nnbr, ctx->nB,
                           ctx->nAA, ctx->rowptr.p, ctx->cols.p, ctx->vals.p, ctx->dAA.p, dB_out);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int L>
static void launch_poolw(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                         const double* w, int64_t nnbr, const int8_t* species, int64_t estride) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    const size_t lds = (size_t)POOL_CH * (2 * ctx->nr + 4 * NP + 1) * sizeof(double);
    set_func_lds((const void*)k_poolw<L>, lds);
    hipLaunchKernelGGL(k_poolw<L>, dim3((unsigned)nenv), dim3(64), lds, st, radial_params(ctx), nenv, off, xyz, w,
                       nnbr, ctx->nspecies > 0 ? species : nullptr, ctx->pool.p, ctx->npool, ctx->Wplus.p,
                       estride);
}

extern "C" int ace_adjoint_EVAL_D(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                  const int8_t* species, const double* w, double* dB_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (ctx->prop != ACE_PROP_INVARIANT) return set_err("ace_adjoint_EVAL_D: Invariant bases only");
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    if (!dB_out || (nnbr > 0 && !w)) return set_err("ace_adjoint_EVAL_D: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (bt_fits(ctx, true)) {
        if (ctx->Aplus.ensure((size_t)ctx->npool * estride) || ctx->Wplus.ensure((size_t)ctx->npool * estride))
            return -1;
        DISPATCH_L(ctx->L, launch_pool, ctx, st, nenv, off, xyz, nnbr, species, estride);
        DISPATCH_L(ctx->L, launch_poolw, ctx, st, nenv, off, xyz, w, nnbr, species, estride);
        HIP_TRY(hipGetLastError());
        return launch_bt<true>(ctx, st, nenv, estride, dB_out);
    }
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    if (ctx->Wplus.ensure((size_t)ctx->npool * estride) || ctx->Wfull.ensure((size_t)nenv * ctx->nA) ||
        ctx->AA.ensure((size_t)nenv * ctx->nAA))
        return -1;
    DISPATCH_L(ctx->L, launch_poolw, ctx, st, nenv, off, xyz, w, nnbr, species, estride);
    HIP_TRY(hipGetLastError());
    const int64_t n = nenv * ctx->nA;
    hipLaunchKernelGGL(k_expandA, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nenv, ctx->nA,
                       ctx->aspec_fid.p, ctx->aspec_sgn.p, ctx->Wplus.p, estride, ctx->Wfull.p);
    const int64_t n1 = nenv * ctx->nAA;
    hipLaunchKernelGGL(k_AAdir, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, nenv, ctx->nA, ctx->nAA,
                       ctx->maxorder, ctx->orders.p, ctx->iAA2iA.p, ctx->Afull.p, ctx->Wfull.p, ctx->AA.p);
    const int64_t n2 = nenv * ctx->nB;
    hipLaunchKernelGGL(k_B<0>, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, nenv, ctx->nB, ctx->nAA,
                       ctx->rowptr.p, ctx->cols.p, ctx->vals.p, ctx->AA.p, dB_out);
    HIP_TRY(hipGetLastError());
    return 0;
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
    if (validate_inputs(ctx, nenv, off, nnbr, species)) return -1;
    if (nenv == 0) return 0;
    if (!xmean || !xinvstd || !E2) return set_err("ace_surr_features: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    DISPATCH_L(ctx->L, launch_pool, ctx, st, nenv, off, xyz, nnbr, species, estride);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_feat, dim3((unsigned)((nenv + 63) / 64)), dim3(64), 0, st, feat_args(ctx), nenv,
                       ctx->Aplus.p, estride, xmean, xinvstd, X, (__bf16*)Xn, E2, ctx->e0r);
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
    WaveCfg* wc = ctx->wcfg[ctx->cur_w];
    DagRecArgs ra;
    ra.recs = ctx->recs.p;
    ra.tasks = ctx->d_tasks.p;
    ra.task_const = ctx->task_const.p;
    ra.wtask_ptr = forces ? wc->d_ptr_all.p : wc->d_ptr_e.p;
    ra.wtask_list = forces ? wc->d_list_all.p : wc->d_list_e.p;
    ra.npool = ctx->npool;
    ra.W = wc->W;
    ra.parts = 1;
    ra.ntile_lo = ntile;
    ra.wtask_ptr_hi = nullptr;
    ra.wtask_list_hi = nullptr;
    size_t lds = std::max((size_t)ctx->npool * 64 * sizeof(double2), (size_t)wc->W * 64 * sizeof(double));
    if (lds > 160 * 1024) return set_err("pooled 1p basis too large for one LDS tile (npool > 160)");
    set_func_lds((const void*)k_dagr, lds);
    hipLaunchKernelGGL(k_dagr, dim3((unsigned)ntile), dim3(64 * wc->W), lds, st, ra, nenv, ctx->Aplus.p, estride,
                       E_out, ctx->Gbuf.p, ctx->e0r, (const int32_t*)fb_list, (const int32_t*)fb_count);
    HIP_TRY(hipGetLastError());
    if (forces && nnbr > 0) {
        DISPATCH_L(ctx->L, launch_force_any, ctx, st, nenv, off, xyz, nnbr, species, estride, dE);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}
