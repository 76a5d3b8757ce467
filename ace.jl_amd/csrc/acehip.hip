// libacehip.so — MI355X (gfx950) basis-evaluation path of the Atomic Cluster
// Expansion with ACE.jl v0.12.46 semantics.  FP64 throughout.
//
// Pipeline of an energy (+ forces) call:
//   k_pool   (ace_pool.hip): one wave per environment (two, one per
//              half-wave, for bases with at most 32 pooled values), lanes =
//              neighbours.  Radial basis and m >= 0 spherical harmonics per
//              neighbour, staged in LDS, then lanes = pooled 1p functions sum
//              over neighbours in order (product_1pbasis.jl:99-134) -> A+[k][env].
//   k_dagr / k_dag (this file): 64 (or 32) environments per workgroup,
//              lanes = environments, the A+ tile staged in LDS; every wave
//              interprets its share of the run programs (depth-first product
//              DAG: prefix products in registers, leaves streamed through
//              scalar loads).  Computes the site energy sum c~ Re AA
//              (evaluator.jl:121-147) and the adjoint polynomials G[k]
//              (evaluator.jl:161-185 restated on A+).
//   k_forcef / k_force (ace_force.hip): recomputes the radial/Ylm values and
//              gradients per neighbour and contracts them with G
//              (evaluator.jl:187-192): +dE/dr_j.
// This file also holds the context (creation, parameters, tile settings)
// and the energy entry points.

#include "ace_common.hpp"
#include "compile_host.hpp"

// ---------------------------------------------------------------------------
// k_dag: run-program interpreter, lanes = environments

struct DagArgs {
    const int32_t* runs;       // [nruns][ACE_RUN_W]
    const int32_t* leaf_fid;   // [nleaves]
    const double* leaf_coef;   // [nleaves][ncoef]
    const double4* vleaf;      // EuclideanVector: [nleaves][2] decoded leaves (factor id, then 3 complex coefs);
                               // class-split: [nleaves] (factor id, 0, one complex coef)
    const int32_t* tasks;      // [ntasks][3]
    const double* task_const;  // [ntasks][2 | 6]
    const int32_t* wtask_ptr;  // [W+1]
    const int32_t* wtask_list;
    int npool;
    int W;
};


// PROP 0: real coefficient, output Re(acc) (energy) or acc (G slot).
// PROP 1: 3 complex coefficients, output Re(acc_p) or acc_p.
// PROP 2: class-split EuclideanVector programs: one complex coefficient; an
//         energy task (slot -1 - c) adds acc to the class value Q_c, the
//         energy is E_p = e0_p + Re sum_c u[c][p] Q_c; a G slot stores acc.
// ENVS = 64: lanes = 64 environments, one task stream per wave.
// ENVS = 32 / 16: the lane groups of a wave are the same ENVS environments
//            running 64 / ENVS task streams (smaller LDS tiles for large
//            pooled bases).
// MAXD: deepest prefix of a run (3 up to correlation order 4, else 7); the
//            prefix products P[1..MAXD] stay in registers, P[0] = 1.
template <int PROP, int ENVS, int MAXD>
__global__ __launch_bounds__(1024) void k_dag(DagArgs a, int64_t nenv,
                                              const double2* __restrict__ Aplus,
                                              int64_t estride, double* __restrict__ E_out,
                                              double2* __restrict__ Gbuf, const double* __restrict__ e0v,
                                              const double2* __restrict__ cls_u) {
    constexpr int NC = PROP == 1 ? 3 : 1;             // complex accumulators per task
    constexpr int NE = PROP == 0 ? 1 : PROP == 1 ? 3 : 6;  // energy partials per lane
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

    double ep[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) ep[q] = 0.0;

    const int tb = a.wtask_ptr[stream], te = a.wtask_ptr[stream + 1];
    for (int ti = tb; ti < te; ++ti) {
        const int t = a.wtask_list[ti];
        const int rb = a.tasks[5 * t], re = a.tasks[5 * t + 1], slot = a.tasks[5 * t + 2];
        double2 acc[NC];
#pragma unroll
        for (int p = 0; p < NC; ++p)
            acc[p] = make_double2(a.task_const[(2 * NC) * t + 2 * p],
                                  a.task_const[(2 * NC) * t + 2 * p + 1]);
        double2 P[MAXD + 1];
#pragma unroll
        for (int lev = 0; lev <= MAXD; ++lev) P[lev] = make_double2(1.0, 0.0);
        for (int r = rb; r < re; ++r) {
            const int* run = a.runs + ACE_RUN_W * r;
            const int d = run[0], j = run[1];
            const int nl = run[2], lo = run[3];
            // levels j..d of the prefix change; level 1 is the factor itself
            if (j <= 1 && d >= 1) P[1] = lds_factor<ENVS>(As, run[4], el);
#pragma unroll
            for (int lev = 2; lev <= MAXD; ++lev)
                if (j <= lev && d >= lev) P[lev] = cmul(P[lev - 1], lds_factor<ENVS>(As, run[3 + lev], el));
            double2 Pd = P[0];
#pragma unroll
            for (int lev = 1; lev <= MAXD; ++lev)
                if (d == lev) Pd = P[lev];
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
            } else if (PROP == 2) {
                // one 32-byte decoded leaf: {factor id (low word), 0, c.re, c.im}
#pragma unroll 2
                for (int i = 0; i < nl; ++i) {
                    const double4 h = a.vleaf[lo + i];
                    const double2 v = lds_factor<ENVS>(As, (int)__double2loint(h.x), el);
                    S[0].x += h.z * v.x - h.w * v.y;
                    S[0].y += h.z * v.y + h.w * v.x;
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
            if (PROP == 2) {
                const int c = -1 - slot;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    if (c == q) {
                        ep[2 * q] += acc[0].x;
                        ep[2 * q + 1] += acc[0].y;
                    }
            } else {
#pragma unroll
                for (int p = 0; p < NC; ++p) ep[p] += acc[p].x;
            }
        } else if (valid) {
#pragma unroll
            for (int p = 0; p < NC; ++p) Gbuf[((int64_t)slot * NC + p) * estride + env] = acc[p];
        }
    }
    // reduce energy partials over streams (reuse the A tile region)
    __syncthreads();
    double* red = reinterpret_cast<double*>(As);
#pragma unroll
    for (int q = 0; q < NE; ++q) red[(stream * NE + q) * ENVS + el] = ep[q];
    __syncthreads();
    if (wave == 0 && lane < ENVS && valid && E_out) {
        if (PROP == 2) {
            double Q[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                Q[q] = 0.0;
                for (int w = 0; w < nwaves * SPW; ++w) Q[q] += red[(w * 6 + q) * ENVS + el];
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                double sum = e0v[p];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 u = cls_u[3 * c + p];
                    sum += u.x * Q[2 * c] - u.y * Q[2 * c + 1];
                }
                E_out[env * 3 + p] = sum;
            }
        } else {
#pragma unroll
            for (int p = 0; p < NC; ++p) {
                double sum = e0v[p];
                for (int w = 0; w < nwaves * SPW; ++w) sum += red[(w * NC + p) * ENVS + el];
                E_out[env * NC + p] = sum;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_dagr: the same programs as decoded records (Invariant, class-split
// EuclideanVector).  A record holds one run (or the continuation of a long
// run): up to four leaves with their coefficients and the prefix levels it
// recomputes, all as ready LDS byte offsets, so the loop spends few scalar
// instructions per record.
//   CPLX = false: real leaf coefficients, 64-byte records, one energy value.
//   CPLX = true:  one complex coefficient per leaf, 96-byte records; energy
//                 tasks (slot -1 - c) add to the class value Q_c and
//                 E_p = e0_p + Re sum_c u[c][p] Q_c.
//   HALF = false: 64 environments per tile, lane = environment.
//   HALF = true:  32 environments per tile (pooled bases of 161-320 values):
//                 lanes e and e + 32 hold environment e and run the same
//                 record stream, the low half on leaves 0 and 2, the high half
//                 on leaves 1 and 3; the halves' sums meet at the end of each
//                 task.

struct DagRecArgs {
    const uint32_t* recs;      // [nrec + 1][16 | 24]
    const int32_t* tasks;      // [ntasks][5]
    const double* task_const;  // [ntasks][2]
    const int32_t* wtask_ptr;  // [parts * W + 1]
    const int32_t* wtask_list;
    const double2* cls_u;      // CPLX: [3 classes][3 components]
    int npool;
    int W;
    int parts;                 // workgroups per tile, each running 1/parts of the tasks
    // tiles [ntile_lo, ntile) run parts + 1 workgroups each, with these task
    // lists: shorter workgroups that fill the grid's last round
    int64_t ntile_lo;
    const int32_t* wtask_ptr_hi;  // [(parts + 1) * W + 1]
    const int32_t* wtask_list_hi;
    // segment kernel (null: the per-record kernel)
    const uint32_t* segw;
    const uint4* segt;
    const int32_t* seg_ptr;
    bool deep;  // prefixes of more than 3 factors (correlation order > 4)
};

// Per-lane choice between two record words (the half-wave mode), taken by
// value: a conditional between the fields themselves is an lvalue, a choice
// of address, which can put the record registers in scratch memory.
__device__ __forceinline__ uint32_t pick(bool hi, uint32_t lo_w, uint32_t hi_w) { return hi ? hi_w : lo_w; }

__device__ __forceinline__ double rec_coef(uint32_t lo, uint32_t hi) {
    return __hiloint2double((int)hi, (int)lo);
}

// S (+)= c * v for a complex coefficient held as a uint4 (re lo, re hi, im
// lo, im hi), as one chain of fused multiply-adds per component (FIRST: S
// starts from this product)
template <bool FIRST>
__device__ __forceinline__ void rec_cfma(double2& S, const uint4& c, const double2& v) {
    const double cr = rec_coef(c.x, c.y), ci = rec_coef(c.z, c.w);
    if (FIRST) {
        S.x = cr * v.x;
        S.y = cr * v.y;
    } else {
        S.x = fma(cr, v.x, S.x);
        S.y = fma(cr, v.y, S.y);
    }
    S.x = fma(-ci, v.y, S.x);
    S.y = fma(ci, v.x, S.y);
}

// One record of a given type (MORE: leaves 2 and 3 present; U1 / U2 / U3:
// prefix levels recomputed): its LDS reads first, then its arithmetic.
// h2..h5: the coefficients (real: h2 = c0 c1, h3 = c2 c3; complex: one each).
template <bool MORE, bool U1, bool U2, bool U3, bool CPLX, bool HALF>
__device__ __forceinline__ void dagr_record(const char* lds, uint32_t ctl, const uint4& h0, const uint4& h1,
                                            const uint4& h2, const uint4& h3, const uint4& h4, const uint4& h5,
                                            bool hi, double2& acc, double2& P1, double2& P2, double2& Pc) {
    auto at = [&](uint32_t off) { return *reinterpret_cast<const double2*>(lds + off); };
    double2 v0, v1, v2, v3, a1, a2, a3;
    if (HALF) {
        // low half: leaves 0 and 2, high half: leaves 1 and 3 (records of at
        // most two leaves take one per half); absent leaves have offset 0 and
        // coefficient 0
        v0 = at(pick(hi, h0.y, h0.z));
        if (MORE) v1 = at(pick(hi, h0.w, h1.x));
    } else {
        v0 = at(h0.y);
        v1 = at(h0.z);
        if (MORE) {
            v2 = at(h0.w);
            v3 = at(h1.x);
        }
    }
    if (U1) a1 = at(h1.y);
    if (U2) a2 = at(h1.z);
    if (U3) a3 = at(h1.w);
    // S = sum of the leaves' coefficient times value, accumulated in one
    // fused multiply-add chain per component
    double2 S;
    if (CPLX) {
        if (HALF) {
            // per-lane choice of the coefficient words (field by field: a
            // whole-record select would go through memory)
            const uint4 c0 = make_uint4(pick(hi, h2.x, h3.x), pick(hi, h2.y, h3.y), pick(hi, h2.z, h3.z),
                                        pick(hi, h2.w, h3.w));
            rec_cfma<true>(S, c0, v0);
            if (MORE) {
                const uint4 c1 = make_uint4(pick(hi, h4.x, h5.x), pick(hi, h4.y, h5.y), pick(hi, h4.z, h5.z),
                                            pick(hi, h4.w, h5.w));
                rec_cfma<false>(S, c1, v1);
            }
        } else {
            rec_cfma<true>(S, h2, v0);
            rec_cfma<false>(S, h3, v1);
            if (MORE) {
                rec_cfma<false>(S, h4, v2);
                rec_cfma<false>(S, h5, v3);
            }
        }
    } else if (HALF) {
        const double c0 = rec_coef(pick(hi, h2.x, h2.z), pick(hi, h2.y, h2.w));
        S.x = c0 * v0.x;
        S.y = c0 * v0.y;
        if (MORE) {
            const double c1 = rec_coef(pick(hi, h3.x, h3.z), pick(hi, h3.y, h3.w));
            S.x = fma(c1, v1.x, S.x);
            S.y = fma(c1, v1.y, S.y);
        }
    } else {
        const double c0 = rec_coef(h2.x, h2.y), c1 = rec_coef(h2.z, h2.w);
        S.x = fma(c1, v1.x, c0 * v0.x);
        S.y = fma(c1, v1.y, c0 * v0.y);
        if (MORE) {
            const double c2 = rec_coef(h3.x, h3.y), c3 = rec_coef(h3.z, h3.w);
            S.x = fma(c3, v3.x, fma(c2, v2.x, S.x));
            S.y = fma(c3, v3.y, fma(c2, v2.y, S.y));
        }
    }
    S.y = flip_sign_if(S.y, ctl >> 16);  // bit 31, conjugated leaves: conj(sum c A) = sum c conj(A)
    if (U1) {
        a1.y = flip_sign_if(a1.y, ctl >> 13);  // bit 28
        P1 = a1;
        Pc = a1;
    }
    if (U2) {
        a2.y = flip_sign_if(a2.y, ctl >> 14);  // bit 29
        P2 = cmul(P1, a2);
        Pc = P2;
    }
    if (U3) {
        a3.y = flip_sign_if(a3.y, ctl >> 15);  // bit 30
        Pc = cmul(P2, a3);
    }
    acc.x = fma(-Pc.y, S.y, fma(Pc.x, S.x, acc.x));
    acc.y = fma(Pc.y, S.x, fma(Pc.x, S.y, acc.y));
}

// Record words through the constant address space: the address is
// wave-uniform, so every fetch is a scalar load and the coefficients and LDS
// offsets are scalar operands without lane-to-scalar copies.
typedef unsigned int RecWords __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 rec_load(const uint4* p) {
    const RecWords v = *(const __attribute__((address_space(4))) RecWords*)(uintptr_t)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

template <bool CPLX, bool HALF>
__global__ __launch_bounds__(1024) void k_dagr(DagRecArgs a, const uint4* __restrict__ recs, int64_t nenv,
                                               const double2* __restrict__ Aplus,
                                               int64_t estride, double* __restrict__ E_out,
                                               double2* __restrict__ Gbuf, const double* __restrict__ e0p,
                                               const int32_t* __restrict__ env_list,
                                               const int32_t* __restrict__ env_count) {
    constexpr int ET = HALF ? 32 : 64;     // environments per tile
    constexpr int RQ = CPLX ? 6 : 4;       // uint4 per record
    constexpr int NE = CPLX ? 6 : 1;       // energy partials per lane
    extern __shared__ double2 As[];  // [npool][ET]
    const int lane = threadIdx.x & 63;
    const int el = HALF ? (lane & 31) : lane;
    const bool hi = HALF && lane >= 32;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // workgroup = (tile, part): the parts of a tile run disjoint task lists
    const int64_t blo = a.ntile_lo * a.parts;
    const bool tail = (int64_t)blockIdx.x >= blo;
    const int P = tail ? a.parts + 1 : a.parts;
    const int64_t bt = tail ? (int64_t)blockIdx.x - blo : (int64_t)blockIdx.x;
    const int64_t tile = (tail ? a.ntile_lo : 0) + bt / P;
    const int part = (int)(bt % P);
    const int32_t* const wptr = tail ? a.wtask_ptr_hi : a.wtask_ptr;
    const int32_t* const wlist = tail ? a.wtask_list_hi : a.wtask_list;
    int64_t env = tile * ET + el;
    bool valid = env < nenv;
    if (env_list) {
        // environments listed by the caller (fallback of the learned surrogate)
        const int64_t cnt = *env_count;
        if (tile * ET >= cnt) return;  // whole workgroup idle: before any barrier
        valid = env < cnt;
        env = valid ? env_list[env] : 0;
    }
    {
        const int k0 = HALF ? 2 * wave + (hi ? 1 : 0) : wave, kstep = HALF ? 2 * a.W : a.W;
        for (int k = k0; k < a.npool; k += kstep)
            As[k * ET + el] = valid ? nt_load2(Aplus + (int64_t)k * estride + env) : make_double2(0.0, 0.0);
    }
    __syncthreads();

    const char* lds = reinterpret_cast<const char*>(As) + 16 * el;
    double ep[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) ep[q] = 0.0;
    const int tb = wptr[part * a.W + wave], te = wptr[part * a.W + wave + 1];
    for (int ti = tb; ti < te; ++ti) {
        const int t = wlist[ti];
        const int slot = a.tasks[5 * t + 2], rb = a.tasks[5 * t + 3], re = a.tasks[5 * t + 4];
        double2 acc = (HALF && hi) ? make_double2(0.0, 0.0) : make_double2(a.task_const[2 * t], a.task_const[2 * t + 1]);
        // P1, P2: prefix products of levels 1 and 2; Pc: the product of the
        // current record's whole prefix.  Runs are sorted, so every new run
        // recomputes its deepest level and Pc is always the last level computed
        // (1 for the prefix-free runs, which come first).
        double2 P1 = make_double2(1.0, 0.0), P2 = P1, Pc = P1;
        const uint4* rp = recs + RQ * (int64_t)rb;
        uint4 h0 = rec_load(rp), h1 = rec_load(rp + 1), h2 = rec_load(rp + 2), h3 = rec_load(rp + 3), h4, h5;
        if (CPLX) {
            h4 = rec_load(rp + 4);
            h5 = rec_load(rp + 5);
        }
        // unrolled by 2: the prefetched and current record headers alternate
        // between two register sets instead of being copied every record
        // (measured 18% faster than the rolled loop)
#pragma unroll 2
        for (int r = rb; r < re; ++r) {
            // next record (record nrec exists and is all zeros)
            const uint4* np = recs + RQ * (int64_t)(r + 1);
            const uint4 n0 = rec_load(np), n1 = rec_load(np + 1), n2 = rec_load(np + 2), n3 = rec_load(np + 3);
            uint4 n4, n5;
            if (CPLX) {
                n4 = rec_load(np + 4);
                n5 = rec_load(np + 5);
            }
            const uint32_t ctl = h0.x;
#define ACE_REC(T)                                                                                          \
    case T:                                                                                                 \
        dagr_record<(T & 1) != 0, (T & 2) != 0, (T & 4) != 0, (T & 8) != 0, CPLX, HALF>(                   \
            lds, ctl, h0, h1, h2, h3, h4, h5, hi, acc, P1, P2, Pc);                                         \
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
            if (CPLX) {
                h4 = n4;
                h5 = n5;
            }
        }
        if (HALF) {
            // the two halves' partial sums of the same environment
            acc.x += __shfl_xor(acc.x, 32);
            acc.y += __shfl_xor(acc.y, 32);
        }
        if (slot < 0) {
            if (CPLX) {
                const int c = -1 - slot;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    if (c == q) {
                        ep[2 * q] += acc.x;
                        ep[2 * q + 1] += acc.y;
                    }
            } else {
                ep[0] += acc.x;
            }
        } else if (valid && !hi) {
            nt_store2(Gbuf + (int64_t)slot * estride + env, acc);
        }
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(As);  // [W][NE][ET]
    if (!hi) {
#pragma unroll
        for (int q = 0; q < NE; ++q) red[(wave * NE + q) * ET + el] = ep[q];
    }
    __syncthreads();
    if (wave == 0 && !hi && valid && E_out) {
        double Q[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            Q[q] = (!CPLX && a.parts == 1) ? *e0p : 0.0;
            for (int w = 0; w < a.W; ++w) Q[q] += red[(w * NE + q) * ET + el];
        }
        if (a.parts > 1) {
            // this part's share, summed by k_esum
#pragma unroll
            for (int q = 0; q < NE; ++q) E_out[((int64_t)part * NE + q) * estride + env] = Q[q];
        } else if (CPLX) {
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                double sum = e0p[p];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 u = a.cls_u[3 * c + p];
                    sum += u.x * Q[2 * c] - u.y * Q[2 * c + 1];
                }
                E_out[env * 3 + p] = sum;
            }
        } else {
            E_out[env] = Q[0];  // e0 + the waves' sums, in that order
        }
    }
}

// ---------------------------------------------------------------------------
// k_dags: the record programs regrouped into kind-sorted segments of
// self-contained records.  Every record recomputes its whole prefix product,
// so the records of a task can be reordered: each task is a short list of
// segments, each a run of records of one kind (prefix depth D = 0..7, NL =
// 1..8 leaves), and the kernel switches once per segment, not once per
// record.  The loop over a segment has no branches, so the LDS reads of
// consecutive records overlap.  Record words (flags, prefix offsets, leaf
// offsets, padding to an even count, coefficients):
//   w[0]              bit 31: the leaves are conjugated; bit 31 - t: prefix
//                     factor t (1..7) is conjugated
//   w[1 .. D]         LDS byte offsets of the prefix factors
//   w[1 + D ..]       LDS byte offsets of the NL leaves
//   then NL FP64 coefficients (CPLX: NL complex).
// Narrow tiles (more than 160 pooled values): ET < 64 environments, as
// many as fit the LDS; the lanes past ET repeat lane 0's work and store
// nothing.  Running whole records per lane costs fewer instructions than
// splitting a record's leaves between lanes that hold the same environment,
// which needs a per-lane choice of every record word (two vector instructions
// per 32-bit word, as an instruction reads one scalar register).
template <int D, int NL, bool CPLX>
struct SegShape {
    static constexpr int HW = (2 + D + NL) & ~1;           // header words, even
    static constexpr int RW = HW + NL * (CPLX ? 4 : 2);     // record words
};

typedef const __attribute__((address_space(4))) uint32_t SegWord;

__device__ __forceinline__ double xor_sign(double c, uint32_t m) {
    // negate c when bit 31 of m is set
    return __hiloint2double(__double2hiint(c) ^ (int)(m & 0x80000000u), __double2loint(c));
}

template <int D, int NL, bool CPLX>
__device__ __forceinline__ void dags_segment(const char* lds, const uint32_t* wp, int count, double2& acc) {
    using SH = SegShape<D, NL, CPLX>;
    constexpr int HW = SH::HW, RW = SH::RW;
    auto at = [&](uint32_t off) { return *reinterpret_cast<const double2*>(lds + off); };
    const SegWord* w = (const SegWord*)(uintptr_t)wp;
    // unrolled by 3 (real coefficients) or 2 (complex), by one less for
    // records of more than 4 leaves (register use): measured fastest
#pragma unroll(CPLX ? (NL > 4 ? 1 : 2) : (NL > 4 ? 2 : 3))
    for (int i = 0; i < count; ++i, w += RW) {
        uint32_t h[RW];
#pragma unroll
        for (int q = 0; q < RW; ++q) h[q] = w[q];
        double2 v[NL], f[D > 0 ? D : 1];
#pragma unroll
        for (int q = 0; q < NL; ++q) v[q] = at(h[1 + D + q]);
#pragma unroll
        for (int t = 0; t < D; ++t) f[t] = at(h[1 + t]);
        // S = sum of coefficient times leaf value: one fused multiply-add
        // chain per component
        double2 S;
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            if (CPLX) {
                const uint4 c = make_uint4(h[HW + 4 * q], h[HW + 4 * q + 1], h[HW + 4 * q + 2], h[HW + 4 * q + 3]);
                if (q == 0)
                    rec_cfma<true>(S, c, v[q]);
                else
                    rec_cfma<false>(S, c, v[q]);
            } else {
                const double c = rec_coef(h[HW + 2 * q], h[HW + 2 * q + 1]);
                if (q == 0) {
                    S.x = c * v[q].x;
                    S.y = c * v[q].y;
                } else {
                    S.x = fma(c, v[q].x, S.x);
                    S.y = fma(c, v[q].y, S.y);
                }
            }
        }
        S.y = xor_sign(S.y, h[0]);
        if (D == 0) {
            acc.x += S.x;
            acc.y += S.y;
        } else {
            double2 P = f[0];
            P.y = xor_sign(P.y, h[0] << 1);
#pragma unroll
            for (int t = 1; t < D; ++t) {
                double2 g = f[t];
                g.y = xor_sign(g.y, h[0] << (t + 1));
                P = cmul(P, g);
            }
            acc.x = fma(-P.y, S.y, fma(P.x, S.x, acc.x));
            acc.y = fma(P.y, S.x, fma(P.x, S.y, acc.y));
        }
    }
}

// ET environments per tile (ET <= 64, as many as the pooled values of which
// fit the 160 KB of LDS): lane e < ET holds environment e; the lanes past ET
// repeat lane 0's work and store nothing.
template <bool CPLX, bool DEEP>
__global__ __launch_bounds__(1024) void k_dags(DagRecArgs a, const uint32_t* __restrict__ segw,
                                               const uint4* __restrict__ segt, const int32_t* __restrict__ seg_ptr,
                                               int ET, int64_t nenv, const double2* __restrict__ Aplus,
                                               int64_t estride, double* __restrict__ E_out,
                                               double2* __restrict__ Gbuf, const double* __restrict__ e0p,
                                               const int32_t* __restrict__ env_list,
                                               const int32_t* __restrict__ env_count) {
    constexpr int NE = CPLX ? 6 : 1;       // energy partials per lane
    extern __shared__ double2 As[];  // [npool][ET]
    const int lane = threadIdx.x & 63;
    const bool act = lane < ET;
    const int el = act ? lane : 0;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t blo = a.ntile_lo * a.parts;
    const bool tail = (int64_t)blockIdx.x >= blo;
    const int P = tail ? a.parts + 1 : a.parts;
    // the parts of a tile on one XCD (the first tiles' blocks XCD-contiguous),
    // so that the A+ tile they each load meets in one L2; not for a listed
    // subset, whose working tiles are a prefix (XCD-contiguous blocks would
    // leave the last XCDs idle)
    const int64_t bt = tail ? (int64_t)blockIdx.x - blo : env_list ? (int64_t)blockIdx.x : xcd_env(blockIdx.x, blo);
    const int64_t tile = (tail ? a.ntile_lo : 0) + bt / P;
    const int part = (int)(bt % P);
    const int32_t* const wptr = tail ? a.wtask_ptr_hi : a.wtask_ptr;
    const int32_t* const wlist = tail ? a.wtask_list_hi : a.wtask_list;
    int64_t env = tile * ET + el;
    bool valid = act && env < nenv;
    if (env_list) {
        const int64_t cnt = *env_count;
        if (tile * ET >= cnt) return;  // whole workgroup idle: before any barrier
        valid = act && env < cnt;
        env = valid ? env_list[env] : 0;
    }
    if (act) {
        for (int k = wave; k < a.npool; k += a.W)
            As[k * ET + el] = valid ? nt_load2(Aplus + (int64_t)k * estride + env) : make_double2(0.0, 0.0);
    }
    __syncthreads();

    const char* lds = reinterpret_cast<const char*>(As) + 16 * el;
    double ep[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) ep[q] = 0.0;
    const SegWord* sp = (const SegWord*)(uintptr_t)seg_ptr;
    const int tb = wptr[part * a.W + wave], te = wptr[part * a.W + wave + 1];
    for (int ti = tb; ti < te; ++ti) {
        const int t = wlist[ti];
        const int slot = a.tasks[5 * t + 2];
        double2 acc = make_double2(a.task_const[2 * t], a.task_const[2 * t + 1]);
        const int sb = (int)sp[t], se = (int)sp[t + 1];
        for (int sg = sb; sg < se; ++sg) {
            const SegWord* st = (const SegWord*)(uintptr_t)(segt + sg);
            const uint32_t kind = st[0], count = st[1], off = st[2];
            const uint32_t* wp = segw + off;
#define ACE_SEG(DD, NN)                                                                  \
    case DD * 8 + NN - 1:                                                                \
        dags_segment<DD, NN, CPLX>(lds, wp, (int)count, acc);                            \
        break;
#define ACE_SEG4(DD) ACE_SEG(DD, 1) ACE_SEG(DD, 2) ACE_SEG(DD, 3) ACE_SEG(DD, 4)
#define ACE_SEG8(DD) ACE_SEG4(DD) ACE_SEG(DD, 5) ACE_SEG(DD, 6) ACE_SEG(DD, 7) ACE_SEG(DD, 8)
            // kind = depth * 8 + leaves - 1: records of up to 8 leaves; DEEP:
            // prefixes of 4 to 7 factors (correlation orders 5 to 8), records
            // of up to 4 leaves
            if constexpr (DEEP) {
                switch (kind) {
                    ACE_SEG4(0) ACE_SEG4(1) ACE_SEG4(2) ACE_SEG4(3) ACE_SEG4(4) ACE_SEG4(5) ACE_SEG4(6) ACE_SEG4(7)
                }
            } else {
                switch (kind) { ACE_SEG8(0) ACE_SEG8(1) ACE_SEG8(2) ACE_SEG8(3) }
            }
#undef ACE_SEG8
#undef ACE_SEG4
#undef ACE_SEG
        }
        if (slot < 0) {
            if (CPLX) {
                const int c = -1 - slot;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    if (c == q) {
                        ep[2 * q] += acc.x;
                        ep[2 * q + 1] += acc.y;
                    }
            } else {
                ep[0] += acc.x;
            }
        } else if (valid) {
            nt_store2(Gbuf + (int64_t)slot * estride + env, acc);
        }
    }
    __syncthreads();
    double* red = reinterpret_cast<double*>(As);  // [W][NE][ET]
    if (act) {
#pragma unroll
        for (int q = 0; q < NE; ++q) red[(wave * NE + q) * ET + el] = ep[q];
    }
    __syncthreads();
    if (wave == 0 && valid && E_out) {
        double Q[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            Q[q] = (!CPLX && a.parts == 1) ? *e0p : 0.0;
            for (int w = 0; w < a.W; ++w) Q[q] += red[(w * NE + q) * ET + el];
        }
        if (a.parts > 1) {
#pragma unroll
            for (int q = 0; q < NE; ++q) E_out[((int64_t)part * NE + q) * estride + env] = Q[q];
        } else if (CPLX) {
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                double sum = e0p[p];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 u = a.cls_u[3 * c + p];
                    sum += u.x * Q[2 * c] - u.y * Q[2 * c + 1];
                }
                E_out[env * 3 + p] = sum;
            }
        } else {
            E_out[env] = Q[0];
        }
    }
}

// E = e0 + sum of the parts' energy shares (fixed order: deterministic).
// Environments from env_hi on belong to tiles with parts + 1 shares.  CPLX:
// the shares are the class values Q_c, E_p = e0_p + Re sum_c u[c][p] Q_c.
template <bool CPLX>
__global__ void k_esum(int64_t nenv, int parts, int64_t env_hi, const double* __restrict__ Ep, int64_t estride,
                       const double* __restrict__ e0p, const double2* __restrict__ cls_u, double* __restrict__ E_out) {
    constexpr int NE = CPLX ? 6 : 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    const int np = e >= env_hi ? parts + 1 : parts;
    double Q[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        Q[q] = CPLX ? 0.0 : *e0p;
        for (int p = 0; p < np; ++p) Q[q] += Ep[((int64_t)p * NE + q) * estride + e];
    }
    if (CPLX) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            double sum = e0p[p];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double2 u = cls_u[3 * c + p];
                sum += u.x * Q[2 * c] - u.y * Q[2 * c + 1];
            }
            E_out[e * 3 + p] = sum;
        }
    } else {
        E_out[e] = Q[0];
    }
}

// ---------------------------------------------------------------------------
// The scratch of ace_energy_forces_noalloc: a fixed device array, one copy
// per device, held for the life of the process.  Host code reaches it only
// through the address hipGetSymbolAddress returns (ace_ctx::scratch).
constexpr size_t ACE_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[ACE_SCRATCH_BYTES];

// ---------------------------------------------------------------------------
// host side

static int upload_coeffs(ace_ctx* ctx, const double* c) {
    acehost::Coeffs cf = acehost::coefficients(ctx->host, c);
    if (ctx->leaf_coef.upload(cf.leaf.data(), cf.leaf.size())) return -1;
    if (ctx->classed) {
        // decoded leaves: {factor id (low word), 0, c.re, c.im}
        const size_t nl = ctx->host.leaf_fid.size();
        std::vector<double4> vl(std::max<size_t>(nl, 1));
        for (size_t i = 0; i < nl; ++i) {
            uint64_t bits = (uint32_t)ctx->host.leaf_fid[i];
            double fid;
            std::memcpy(&fid, &bits, sizeof(fid));
            vl[i] = make_double4(fid, 0.0, cf.leaf[2 * i], cf.leaf[2 * i + 1]);
        }
        if (ctx->vleaf.upload(vl.data(), vl.size())) return -1;
    } else if (ctx->prop != ACE_PROP_INVARIANT) {
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
    if (ctx->rec_et > 0) {
        std::vector<uint32_t> rec;
        const std::string rerr = acehost::records(ctx->host, cf, rec, ctx->rec_et == 64 ? 10 : 9);
        if (!rerr.empty()) return set_err(rerr);
        if (ctx->recs.upload(rec.data(), rec.size())) return -1;
    }
    if (ctx->seg_et > 0) {
        acehost::Segments sg;
        const std::string serr = acehost::segments(ctx->host, cf, 16 * ctx->seg_et, acehost::seg_leaves(ctx->maxorder), sg);
        if (!serr.empty()) return set_err(serr);
        if (ctx->segw.upload(sg.words.data(), sg.words.size()) || ctx->segt.upload(sg.segt.data(), sg.segt.size()) ||
            ctx->seg_ptr.upload(sg.seg_ptr.data(), sg.seg_ptr.size()))
            return -1;
    }
    if (ctx->task_const.upload(cf.task.data(), cf.task.size())) return -1;
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
    ctx->nrec = pk->nrec;
    ctx->ncoef = pk->prop == ACE_PROP_INVARIANT ? 1 : 6;
    ctx->classed = pk->classed;
    // the record kernels' tiles (Invariant and class-split programs): the
    // segment kernel takes as many environments as fit the 160 KB of LDS (16 B
    // per pooled value), at most 64 (npool <= 160), at least 20 (npool <= 512);
    // deeper prefixes than 3 factors only on 64-environment tiles (else the
    // interpreter).  The per-record kernel: 64 or 32 environments, for bases
    // that carry records (correlation order <= 4, npool <= 320).
    if (pk->prop == ACE_PROP_INVARIANT || pk->classed) {
        const int et = (int)std::min<size_t>(64, (160 * 1024) / ((size_t)pk->npool * sizeof(double2)));
        ctx->seg_et = pk->maxorder > 4 && et < 64 ? 0 : et;
        if (pk->nrec > 0)
            ctx->rec_et = (size_t)pk->npool * 64 * sizeof(double2) <= 160 * 1024   ? 64
                          : (size_t)pk->npool * 32 * sizeof(double2) <= 160 * 1024 ? 32
                                                                                   : 0;
    }
    ctx->host = acehost::copy_pack(pk);
    const int nc = ctx->ncoef;
    auto fail = [&](const char* what) {
        std::string msg = g_err;
        delete ctx;
        return set_err(std::string("ace_ctx_create: ") + what + ": " + msg);
    };
    // this device's copy of the allocation-free entry point's scratch array
    if (hipGetSymbolAddress((void**)&ctx->scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
        delete ctx;
        return set_err("ace_ctx_create: hipGetSymbolAddress of the scratch array failed");
    }
    if (hipHostMalloc((void**)&ctx->h_words, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        ctx->h_words = nullptr;
        delete ctx;
        return set_err("ace_ctx_create: pinned host allocation failed");
    }
    if (ctx->rA.upload(pk->rA, pk->nr) || ctx->rB.upload(pk->rB, pk->nr) ||
        ctx->rC.upload(pk->rC, pk->nr) || ctx->pool.upload(pk->pool, 4 * (size_t)pk->npool) ||
        ctx->aspec_fid.upload(pk->aspec_fid, pk->nA) ||
        ctx->aspec_sgn.upload(pk->aspec_sgn, pk->nA) ||
        ctx->orders.upload(pk->orders, pk->nAA) ||
        ctx->iAA2iA.upload(pk->iAA2iA, (size_t)pk->nAA * pk->maxorder) ||
        ctx->rowptr.upload(pk->a2b_rowptr, pk->nB + 1) ||
        ctx->cols.upload(pk->a2b_cols, pk->a2b_nnz) ||
        ctx->vals.upload(pk->a2b_vals, (size_t)pk->a2b_nnz * nc) ||
        ctx->runs.upload(pk->runs, ACE_RUN_W * (size_t)pk->nruns) ||
        ctx->leaf_fid.upload(pk->leaf_fid, pk->nleaves) ||
        ctx->d_tasks.upload(pk->tasks, 5 * (size_t)pk->ntasks) ||
        ctx->gslot_ptr.upload(pk->gslot_ptr, pk->npool + 1) ||
        ctx->gslot_list.upload(pk->gslot_list, pk->gslot_ptr[pk->npool]))
        return fail("upload");
    if (ctx->classed) {
        double2 u[9];
        for (int c = 0; c < 3; ++c)
            for (int p = 0; p < 3; ++p) u[3 * c + p] = make_double2(ctx->host.cv.u[c][2 * p], ctx->host.cv.u[c][2 * p + 1]);
        if (ctx->cls_u.upload(u, 9) || ctx->gslot_cls.upload(pk->gslot_class, std::max(1, pk->ngslots)))
            return fail("upload");
    }
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
        // every row's nonzeros ordered by the order of their AA column, and
        // where each order's records start, so that the kernel runs one
        // straight-line loop per order (orders 0 to 4; larger bases do not
        // use the fused kernel)
        std::vector<int32_t> rowo((size_t)std::max(1, pk->nB) * BT_ROWO, 0);
        for (int b = 0; b < pk->nB; ++b) {
            const int z0 = pk->a2b_rowptr[b], z1 = pk->a2b_rowptr[b + 1];
            std::stable_sort(rec.begin() + z0, rec.begin() + z1, [](const BtRec& x, const BtRec& y) { return x.o < y.o; });
            int32_t* ro = rowo.data() + (size_t)b * BT_ROWO;
            ro[0] = z0;
            for (int o = 0; o + 1 < BT_ROWO; ++o) {
                int n = 0;
                for (int z = z0; z < z1; ++z) n += rec[z].o == o;
                ro[o + 1] = ro[o] + n;
            }
        }
        if (ctx->bt_rec.upload(rec.data(), rec.size()) || ctx->bt_rowo.upload(rowo.data(), rowo.size()))
            return fail("upload");
        std::vector<int32_t> a4 = acehost::aspec_rows(pk);
        if (ctx->aspec4.upload(a4.data(), a4.size())) return fail("upload");
        std::vector<int16_t> kx = acehost::pool_lookup(pk);
        if (ctx->kidx.upload(kx.data(), kx.size())) return fail("upload");
        // the pooling kernel's shared-Y lane sets at its pass count
        std::vector<int32_t> pslot;
        if (acehost::pool_lane_sets(pk->pool, pk->npool, pool_passes(pk->npool), pslot)) {
            if (ctx->pool_slot.upload(pslot.data(), pslot.size())) return fail("upload");
            ctx->pool_shy_kp = pool_passes(pk->npool);
        }
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
            std::vector<int32_t> fslot(seen.size(), 0);  // G slot of (l, m >= 0, n)
            std::vector<int32_t> fold;
            for (int k = 0; k < pk->npool; ++k) {
                const int n = pk->pool[4 * k + 1] - 1, l = pk->pool[4 * k + 2], m = pk->pool[4 * k + 3];
                const int s0 = pk->gslot_ptr[k], s1 = pk->gslot_ptr[k + 1];
                if (s1 <= s0) {
                    ctx->flat_force = false;  // a pooled function without a gradient slot
                    break;
                }
                fslot[(size_t)idxp(l, m) * FORCE_NRMAX + n] = pk->gslot_list[s0];
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
                    for (int n = 0; n < ctx->nrl[l]; ++n) qslot.push_back(fslot[(size_t)idxp(l, m) * FORCE_NRMAX + n]);
            if ((int)qslot.size() > FF_QMAX) ctx->flat_force = false;
            ctx->fflat.nq = (int16_t)std::min<size_t>(qslot.size(), FF_QMAX);
            if (ctx->flat_force && ctx->ffq.upload(qslot.data(), qslot.size())) return fail("upload");
            ctx->nfold = (int)fold.size() / 2;
            if (ctx->flat_force && ctx->nfold > 0 && ctx->gfold.upload(fold.data(), fold.size()))
                return fail("upload");
        }
    }
    {
        // tables of the basis-gradient weight kernel (ace_basis_dB)
        const int NC = pk->prop == ACE_PROP_INVARIANT ? 1 : 3;
        const size_t fixed = db_fixed_lds(pk->npool);
        const size_t per_slot = (size_t)NC * sizeof(double2) + sizeof(int32_t);
        if (fixed + 64 * per_slot <= DB_LDS_BUDGET) {
            const int max_slots = (int)((DB_LDS_BUDGET - fixed) / per_slot);
            ctx->db_host = acehost::db_tables(pk, 256, max_slots);
            ctx->db_usable = ctx->db_host.max_chunk_slots <= max_slots;
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
        if (ctx->seg_et > 0 || ctx->rec_et > 0) {
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

extern "C" int ace_ctx_create_jl(const ace_jl_basis* basis, int device, const double* c, int64_t nB,
                                 ace_ctx** out) {
    if (!basis || !out || !c) return set_err("ace_ctx_create_jl: null argument");
    *out = nullptr;
    acehost::PackStore ps;
    const std::string err = acehost::compile_jl(basis, ps);
    if (!err.empty()) return set_err("ace_ctx_create_jl: " + err);
    const ace_pack pk = ps.view();
    return ace_ctx_create(&pk, device, c, nB, out);
}

extern "C" int ace_compile_check(const ace_jl_basis* basis, int64_t* sizes) {
    if (!basis || !sizes) return set_err("ace_compile_check: null argument");
    acehost::PackStore ps;
    std::string err = acehost::compile_jl(basis, ps);
    if (err.empty()) {
        const ace_pack pk = ps.view();
        err = acehost::validate_pack(&pk);
        if (err.empty()) {
            sizes[0] = pk.npool;
            sizes[1] = pk.nruns;
            sizes[2] = pk.nleaves;
            sizes[3] = pk.ntasks;
            sizes[4] = pk.nrec;
            sizes[5] = pk.ngslots;
            sizes[6] = pk.nfeat;
        }
    }
    return err.empty() ? 0 : set_err("ace_compile_check: " + err);
}

extern "C" int ace_last_dag_shape(ace_ctx* ctx, int* parts, int64_t* tail_tiles, int* kernel) {
    if (!ctx) return set_err("ace_last_dag_shape: null context");
    if (parts) *parts = ctx->last_parts;
    if (tail_tiles) *tail_tiles = ctx->last_tail;
    if (kernel) *kernel = ctx->last_kernel;
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

extern "C" int ace_set_pool_shape(ace_ctx* ctx, int chunk, int envs_per_wave) {
    if (!ctx) return set_err("ace_set_pool_shape: null context");
    if (chunk != 0 && chunk != 8 && chunk != 16 && chunk != 32)
        return set_err("ace_set_pool_shape: chunk must be 0 (default 16), 8, 16 or 32");
    if (envs_per_wave != 0 && envs_per_wave != 1 && envs_per_wave != 2)
        return set_err("ace_set_pool_shape: envs_per_wave must be 0 (automatic), 1 or 2");
    ctx->pool_ch = chunk;
    ctx->pool_epw = envs_per_wave;
    return 0;
}

extern "C" int ace_set_kernel(ace_ctx* ctx, int mode) {
    if (!ctx) return set_err("ace_set_kernel: null context");
    if (mode < 0 || mode > 2)
        return set_err("ace_set_kernel: mode must be 0 (automatic), 1 (run interpreter) or 2 (per-record dispatch)");
    ctx->force_interp = mode == 1;
    ctx->rec_dispatch = mode == 2;
    return 0;
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

int check_common(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
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
        // read on the caller's stream, after whatever that stream has queued
        // that writes env_offsets; the call waits for these two words only
        hipStream_t st = (hipStream_t)stream;
        HIP_TRY(hipMemcpyAsync(ctx->h_words, off, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(ctx->h_words + 1, off + nenv, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const int32_t first = ctx->h_words[0];
        last = ctx->h_words[1];
        if (first != 0 || last < 0) return set_err("env_offsets must start at 0 and be non-negative");
        ctx->batch_off = off;
        ctx->batch_nenv = nenv;
        ctx->batch_nnbr = last;
    }
    *nnbr_out = last;
    return 0;
}

// host-side validation of offsets and species (bounds for every kernel)
int validate_inputs(ace_ctx* ctx, int64_t nenv, const int32_t* off, int64_t nnbr, const int8_t* species,
                    hipStream_t st) {
    if (nenv == 0 || !ctx->validate) return 0;
    std::vector<int32_t> h(nenv + 1);
    HIP_TRY(hipMemcpyAsync(h.data(), off, sizeof(int32_t) * (nenv + 1), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int64_t e = 0; e < nenv; ++e)
        if (h[e + 1] < h[e]) return set_err("env_offsets must be non-decreasing");
    if (species && ctx->nspecies > 0 && nnbr > 0) {
        std::vector<int8_t> s(nnbr);
        HIP_TRY(hipMemcpyAsync(s.data(), species, nnbr, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int64_t j = 0; j < nnbr; ++j)
            if (s[j] < 0 || s[j] >= ctx->nspecies) return set_err("species index out of range (unknown category)");
    }
    return 0;
}

// The run interpreter for bases the record kernel does not take: a tile of
// 64, 32 or 16 environments (the largest whose pooled values fit LDS) and
// prefix products up to depth 3 or 7.
template <int PROP, int ENVS, int MAXD>
static int launch_dag_tile(ace_ctx* ctx, hipStream_t st, const DagArgs& a, int64_t nenv, int64_t estride,
                           const Workspace& ws, double* E_out) {
    constexpr int NE = PROP == 0 ? 1 : PROP == 1 ? 3 : 6;  // energy partials per lane
    constexpr int SPW = 64 / ENVS;  // task streams per wave
    if (a.W % SPW) return set_err("tile configuration needs a multiple of 64 / ENVS task streams");
    const size_t lds = std::max((size_t)ctx->npool * ENVS * sizeof(double2), (size_t)a.W * NE * ENVS * sizeof(double));
    const int64_t nt = (nenv + ENVS - 1) / ENVS;
    set_func_lds((const void*)k_dag<PROP, ENVS, MAXD>, lds);
    hipLaunchKernelGGL((k_dag<PROP, ENVS, MAXD>), dim3((unsigned)nt), dim3(64 * (a.W / SPW)), lds, st, a, nenv,
                       ws.Aplus, estride, E_out, ws.Gbuf, ctx->e0v.p, ctx->cls_u.p);
    return 0;
}

template <int PROP, int MAXD>
static int launch_dag_prop(ace_ctx* ctx, hipStream_t st, const DagArgs& a, int64_t nenv, int64_t estride,
                           const Workspace& ws, double* E_out) {
    constexpr int NE = PROP == 0 ? 1 : PROP == 1 ? 3 : 6;
    auto fits = [&](int envs) {
        return std::max((size_t)ctx->npool * envs * sizeof(double2), (size_t)a.W * NE * envs * sizeof(double)) <=
               160 * 1024;
    };
    ctx->last_kernel = fits(64) ? 64 : fits(32) ? 32 : 16;
    if (fits(64)) return launch_dag_tile<PROP, 64, MAXD>(ctx, st, a, nenv, estride, ws, E_out);
    if (fits(32)) return launch_dag_tile<PROP, 32, MAXD>(ctx, st, a, nenv, estride, ws, E_out);
    if (fits(16)) return launch_dag_tile<PROP, 16, MAXD>(ctx, st, a, nenv, estride, ws, E_out);
    return set_err("pooled 1p basis too large for the LDS tile of 16 environments");
}

static int launch_dag_interp(ace_ctx* ctx, hipStream_t st, const DagArgs& a, int64_t nenv, int64_t estride,
                             const Workspace& ws, double* E_out) {
    const bool deep = ctx->maxorder > 4;
    if (ctx->prop == ACE_PROP_INVARIANT)
        return deep ? launch_dag_prop<0, 7>(ctx, st, a, nenv, estride, ws, E_out)
                    : launch_dag_prop<0, 3>(ctx, st, a, nenv, estride, ws, E_out);
    if (ctx->classed)
        return deep ? launch_dag_prop<2, 7>(ctx, st, a, nenv, estride, ws, E_out)
                    : launch_dag_prop<2, 3>(ctx, st, a, nenv, estride, ws, E_out);
    return deep ? launch_dag_prop<1, 7>(ctx, st, a, nenv, estride, ws, E_out)
                : launch_dag_prop<1, 3>(ctx, st, a, nenv, estride, ws, E_out);
}

// The record kernel for a tile of ET environments: the segment kernel k_dags
// when ra carries segments (any ET <= 64), else k_dagr (ET 64 or 32).
template <bool CPLX>
static void launch_dagr(int ET, dim3 grid, dim3 block, size_t lds, hipStream_t st, const DagRecArgs& ra,
                        int64_t nenv, const double2* Aplus, int64_t estride, double* E_out, double2* Gbuf,
                        const double* e0p, const int32_t* env_list, const int32_t* env_count) {
    if (ra.segw) {
        if (!ra.deep) {
            set_func_lds((const void*)k_dags<CPLX, false>, lds);
            hipLaunchKernelGGL((k_dags<CPLX, false>), grid, block, lds, st, ra, ra.segw, ra.segt, ra.seg_ptr, ET, nenv,
                               Aplus, estride, E_out, Gbuf, e0p, env_list, env_count);
        } else {
            set_func_lds((const void*)k_dags<CPLX, true>, lds);
            hipLaunchKernelGGL((k_dags<CPLX, true>), grid, block, lds, st, ra, ra.segw, ra.segt, ra.seg_ptr, ET, nenv,
                               Aplus, estride, E_out, Gbuf, e0p, env_list, env_count);
        }
    } else if (ET == 64) {
        set_func_lds((const void*)k_dagr<CPLX, false>, lds);
        hipLaunchKernelGGL((k_dagr<CPLX, false>), grid, block, lds, st, ra, reinterpret_cast<const uint4*>(ra.recs),
                           nenv, Aplus, estride, E_out, Gbuf, e0p, env_list, env_count);
    } else {
        set_func_lds((const void*)k_dagr<CPLX, true>, lds);
        hipLaunchKernelGGL((k_dagr<CPLX, true>), grid, block, lds, st, ra, reinterpret_cast<const uint4*>(ra.recs),
                           nenv, Aplus, estride, E_out, Gbuf, e0p, env_list, env_count);
    }
}

// The launches of an energy (+ forces) evaluation, into a workspace that
// holds A+ [npool][estride], G [ngslots * NC][estride] (forces), MAX_PARTS
// energy shares [estride] and, for the flat force kernel, nnbr map entries.
static int eval_path(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                     int64_t nnbr, const int8_t* species, double* E_out, double* dE, const Workspace& ws) {
    const int64_t estride = ((nenv + 63) / 64) * 64;
    const bool forces = dE != nullptr;
    WaveCfg* wc = ctx->wcfg[ctx->cur_w];
    // under stream capture (hipGraph) the timing events are left out of the graph
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    const bool timed = cap == hipStreamCaptureStatusNone;
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[0], st));
    if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, forces && ctx->flat_force)) return -1;
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
    // the record kernel: Invariant and class-split EuclideanVector bases of
    // order <= 4 with at most 320 pooled values; everything else runs the run
    // interpreter
    const int rec_et = ctx->rec_dispatch ? ctx->rec_et : ctx->seg_et;
    const bool use_records = rec_et > 0 && !ctx->force_interp;
    if (ctx->rec_dispatch && ctx->rec_et == 0 && !ctx->force_interp)
        return set_err("ace_set_kernel mode 2: per-record dispatch needs records (correlation order <= 4, npool <= 320)");
    ctx->last_tail = 0;
    if (use_records) {
        const int ET = rec_et;
        const int NE = ctx->classed ? 6 : 1;
        const int64_t ntile = (nenv + ET - 1) / ET;
        const size_t lds = std::max((size_t)ctx->npool * ET * sizeof(double2), (size_t)wc->W * NE * ET * sizeof(double));
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
        ctx->last_kernel = ET == 64 ? 0 : 1;  // 1: narrower tiles
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
        ctx->last_tail = nhi;
        DagRecArgs ra;
        ra.recs = ctx->recs.p;
        ra.segw = ctx->rec_dispatch ? nullptr : ctx->segw.p;
        ra.segt = reinterpret_cast<const uint4*>(ctx->segt.p);
        ra.seg_ptr = ctx->seg_ptr.p;
        ra.deep = ctx->maxorder > 4;
        ra.tasks = ctx->d_tasks.p;
        ra.task_const = ctx->task_const.p;
        ra.wtask_ptr = P == 1 ? a.wtask_ptr : (forces ? wc->pp_all[P].p : wc->pp_e[P].p);
        ra.wtask_list = P == 1 ? a.wtask_list : (forces ? wc->pl_all[P].p : wc->pl_e[P].p);
        ra.cls_u = ctx->cls_u.p;
        ra.npool = ctx->npool;
        ra.W = wc->W;
        ra.parts = P;
        ra.ntile_lo = ntile - nhi;
        ra.wtask_ptr_hi = nhi > 0 ? (forces ? wc->pp_all[P + 1].p : wc->pp_e[P + 1].p) : nullptr;
        ra.wtask_list_hi = nhi > 0 ? (forces ? wc->pl_all[P + 1].p : wc->pl_e[P + 1].p) : nullptr;
        const int64_t nblk = (ntile - nhi) * P + nhi * (P + 1);
        double* Eo = P == 1 ? E_out : ws.Epart;
        if (ctx->classed)
            launch_dagr<true>(ET, dim3((unsigned)nblk), dim3(64 * wc->W), lds, st, ra, nenv, ws.Aplus, estride, Eo,
                              ws.Gbuf, ctx->e0v.p, nullptr, nullptr);
        else
            launch_dagr<false>(ET, dim3((unsigned)nblk), dim3(64 * wc->W), lds, st, ra, nenv, ws.Aplus, estride, Eo,
                               ws.Gbuf, ctx->e0v.p, nullptr, nullptr);
        if (P > 1) {
            const dim3 g((unsigned)((nenv + 255) / 256)), b(256);
            if (ctx->classed)
                hipLaunchKernelGGL(k_esum<true>, g, b, 0, st, nenv, P, (ntile - nhi) * ET, ws.Epart, estride,
                                   ctx->e0v.p, ctx->cls_u.p, E_out);
            else
                hipLaunchKernelGGL(k_esum<false>, g, b, 0, st, nenv, P, (ntile - nhi) * ET, ws.Epart, estride,
                                   ctx->e0v.p, ctx->cls_u.p, E_out);
        }
    } else {
        ctx->last_parts = 1;
        if (launch_dag_interp(ctx, st, a, nenv, estride, ws, E_out)) return -1;
    }
    HIP_TRY(hipGetLastError());
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[2], st));
    if (forces && nnbr > 0 && launch_force_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE)) return -1;
    if (timed) HIP_TRY(hipEventRecord(ctx->ev[3], st));
    ctx->timing_valid = timed;
    return 0;
}

static int run_path(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                    const int8_t* species, double* E_out, double* dE, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
    if (nenv == 0) return 0;
    if (!E_out) return set_err("null E_out");
    const int64_t estride = ((nenv + 63) / 64) * 64;
    const bool forces = dE != nullptr;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    if (forces && ctx->Gbuf.ensure((size_t)std::max(1, ctx->ngslots) * gslot_width(ctx) * estride)) return -1;
    if ((ctx->seg_et > 0 || ctx->rec_et > 0) && ctx->Epart.ensure((size_t)MAX_PARTS * (ctx->classed ? 6 : 1) * estride))
        return -1;
    if (forces && ctx->flat_force && nnbr > 0 && ctx->nbr_env.ensure(nnbr)) return -1;
    return eval_path(ctx, (hipStream_t)stream, nenv, off, xyz, nnbr, species, E_out, dE, ctx_buffers(ctx));
}

extern "C" int ace_energy_forces_noalloc(ace_ctx* ctx, int64_t nenv, const int32_t* off, int64_t nnbr,
                                         const double* xyz, const int8_t* species, double* E_out, double* dE_dr,
                                         void* stream) {
    if (!ctx) return set_err("ace_energy_forces_noalloc: null context");
    if (nenv < 0 || nnbr < 0) return set_err("ace_energy_forces_noalloc: negative size");
    if (nenv == 0) return 0;
    if (!off || !E_out || !dE_dr || (nnbr > 0 && !xyz)) return set_err("ace_energy_forces_noalloc: null argument");
    if (ctx->nspecies > 0 && nnbr > 0 && !species)
        return set_err("species array required for a categorical 1p basis");
    if (!ctx->scratch) return set_err("ace_energy_forces_noalloc: no scratch array on this device");
    HIP_TRY(hipSetDevice(ctx->device));
    // carve the fixed array: A+, G, the energy shares, the neighbour map
    const int64_t estride = ((nenv + 63) / 64) * 64;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bA = up((size_t)ctx->npool * estride * sizeof(double2));
    const size_t bG = up((size_t)std::max(1, ctx->ngslots) * gslot_width(ctx) * estride * sizeof(double2));
    const size_t bE = ctx->seg_et > 0 || ctx->rec_et > 0
                          ? up((size_t)MAX_PARTS * (ctx->classed ? 6 : 1) * estride * sizeof(double))
                          : 0;
    const size_t bN = ctx->flat_force ? up((size_t)nnbr * sizeof(int32_t)) : 0;
    const size_t need = bA + bG + bE + bN;
    if (need > ACE_SCRATCH_BYTES)
        return set_err("ace_energy_forces_noalloc: the batch needs " + std::to_string(need) + " bytes of scratch, " +
                       "more than the fixed " + std::to_string(ACE_SCRATCH_BYTES) + "; split it into smaller batches");
    Workspace ws;
    ws.Aplus = reinterpret_cast<double2*>(ctx->scratch);
    ws.Gbuf = reinterpret_cast<double2*>(ctx->scratch + bA);
    ws.Epart = bE ? reinterpret_cast<double*>(ctx->scratch + bA + bG) : nullptr;
    ws.nbr_env = bN ? reinterpret_cast<int32_t*>(ctx->scratch + bA + bG + bE) : nullptr;
    ws.nbr_cap = bN ? (size_t)nnbr : 0;
    return eval_path(ctx, (hipStream_t)stream, nenv, off, xyz, nnbr, species, E_out, dE_dr, ws);
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

int launch_dag_listed(ace_ctx* ctx, hipStream_t st, int64_t nenv, int64_t estride, bool forces, double* E_out,
                      const int32_t* env_list, const int32_t* env_count) {
    if (ctx->prop != ACE_PROP_INVARIANT) return set_err("listed product-DAG evaluation: Invariant bases only");
    if (ctx->rec_dispatch && ctx->nrec == 0)
        return set_err("ace_set_kernel mode 2: per-record dispatch needs correlation order <= 4");
    if ((ctx->rec_dispatch ? ctx->rec_et : ctx->seg_et) != 64)
        return set_err("pooled 1p basis too large for one LDS tile (npool > 160)");
    WaveCfg* wc = ctx->wcfg[ctx->cur_w];
    const int64_t ntile = (nenv + 63) / 64;
    DagRecArgs ra;
    ra.recs = ctx->recs.p;
    ra.segw = ctx->rec_dispatch ? nullptr : ctx->segw.p;
    ra.segt = reinterpret_cast<const uint4*>(ctx->segt.p);
    ra.seg_ptr = ctx->seg_ptr.p;
    ra.deep = ctx->maxorder > 4;
    ra.tasks = ctx->d_tasks.p;
    ra.task_const = ctx->task_const.p;
    ra.wtask_ptr = forces ? wc->d_ptr_all.p : wc->d_ptr_e.p;
    ra.wtask_list = forces ? wc->d_list_all.p : wc->d_list_e.p;
    ra.cls_u = nullptr;
    ra.npool = ctx->npool;
    ra.W = wc->W;
    ra.parts = 1;
    ra.ntile_lo = ntile;
    ra.wtask_ptr_hi = nullptr;
    ra.wtask_list_hi = nullptr;
    const size_t lds = std::max((size_t)ctx->npool * 64 * sizeof(double2), (size_t)wc->W * 64 * sizeof(double));
    launch_dagr<false>(64, dim3((unsigned)ntile), dim3(64 * wc->W), lds, st, ra, nenv, ctx->Aplus.p, estride, E_out,
                       forces ? ctx->Gbuf.p : nullptr, ctx->e0v.p, env_list, env_count);
    HIP_TRY(hipGetLastError());
    return 0;
}
