// Shared definitions of the translation units of libacehip.so: the
// per-neighbour device helpers (radial basis, spherical coordinates, ALP
// recursion), the device-side table formats, the context struct and the
// launchers one unit calls in another.
//
// Units: acehip.hip (context, product DAG, energy entry points),
// ace_pool.hip (1p basis and pooling), ace_force.hip (adjoint force
// kernels), ace_basis.hip (basis-level evaluate / evaluate_d,
// adjoint_EVAL_D), ace_surr.hip (learned-surrogate support).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/acehip.h"
#include "host_util.hpp"
#include "pack_host.hpp"

#define ACE_LMAX 10
#define ACE_NPMAX ((ACE_LMAX + 1) * (ACE_LMAX + 2) / 2)
#define ACE_NRMAX 32
#define ACE_CONJ 0x8000

__host__ __device__ constexpr int idxp(int l, int m) { return m + l * (l + 1) / 2; }
// k_pool's neighbour slot (doubles): R_n from 0, Y_lm (NP complex values) from
// the even offset pool_yoff, then the species; an even count in all, so that
// every Y_lm is a 16-byte aligned LDS read
__host__ __device__ constexpr int pool_yoff(int nr) { return (nr + 1) & ~1; }
__host__ __device__ constexpr int pool_rstride(int nr, int np) { return (pool_yoff(nr) + 2 * np + 2) & ~1; }
// k_pool's pooled functions per lane with one environment per wave: the
// fewest passes of 64 lanes that cover npool among 2, 3, 4 and 8
inline int pool_passes(int npool) { return npool <= 128 ? 2 : npool <= 192 ? 3 : npool <= 256 ? 4 : 8; }

// Per-neighbour parameters of the 1p basis: the radial recursion
// coefficients and envelope (orthpolys.jl:82-99, distancetransforms.jl:73).
struct RadialParams {
    const double* A;
    const double* B;
    const double* C;
    double r0, tl, tr;
    int nr, tp, pl, pr;
};

// ALP recursion coefficients a_lm and b_lm (sphericalharmonics.jl:146-159).
// Every caller's loops over l and m are unrolled, so these fold to
// compile-time constants (IEEE square roots, the same bits as the reference's
// table): no table, no loads.
__device__ __forceinline__ double alp_a(int l, int m) {
    const double ls = (double)l * l, ms = (double)m * m;
    return sqrt((4 * ls - 1.0) / (ls - ms));
}
__device__ __forceinline__ double alp_b(int l, int m) {
    const double ls = (double)(l - 1) * (l - 1), ms = (double)m * m;
    return -sqrt((ls - ms) / (4 * ls - 1.0));
}

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

// ALP values (P/sin(theta) for m >= 1) and d/dtheta, sphericalharmonics.jl:212-267
template <int L>
__device__ __forceinline__ void alp_eval_ed(const RadialParams& rp, const Sph& S, double* P, double* dP) {
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
            const double a = alp_a(l, 0), b = alp_b(l, 0);
            P[i] = a * (S.cth * P[i1] + b * P[i2]);
            dP[i] = a * (-S.sth * P[i1] + S.cth * dP[i1] + b * dP[i2]);
        }
#pragma unroll
        for (int m = 1; m <= l - 2; ++m) {
            const int i = idxp(l, m), i1 = idxp(l - 1, m), i2 = idxp(l - 2, m);
            const double a = alp_a(l, m), b = alp_b(l, m);
            P[i] = a * (S.cth * P[i1] + b * P[i2]);
            dP[i] = a * (-st2 * P[i1] + S.cth * dP[i1] + b * dP[i2]);
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

constexpr int POOL_CH = 16;
// LDS of the basis-gradient weight kernel (k_dBw), and neighbours staged per
// barrier: measured at config 2, one neighbour per stage within 48 KB (three
// workgroups per CU) beat 2 or 4 neighbours per stage (fewer functions per
// chunk, fewer workgroups per CU) and a 32 KB budget
constexpr int DB_NJ_MAX = 1;
constexpr size_t DB_LDS_BUDGET = 48 * 1024;
// k_dBw's LDS besides the weights: A+ [npool] and a stage of up to DB_NJ_MAX
// neighbours' dphi+ [npool][3] within 2048 double2 (at least one neighbour)
inline size_t db_fixed_lds(int npool) {
    const size_t nq = 3 * (size_t)npool;
    const size_t njs = std::max<size_t>(1, std::min<size_t>(DB_NJ_MAX, 2048 / nq));
    return ((size_t)npool + njs * nq) * sizeof(double2);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int ENVS>
__device__ __forceinline__ double2 lds_factor(const double2* As, int f, int lane) {
    double2 v = As[(f & 0x7fff) * ENVS + lane];
    if (f & ACE_CONJ) v.y = -v.y;
    return v;
}

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
// per basis row of the fused kernel: where its records of order 0, 1, ..., 4
// start, and the row's end
constexpr int BT_ROWO = 6;

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
// environments a 256-neighbour block stages in LDS per round: 32 when their G
// fits in FF_LDS (short environments), else 16
constexpr int FF_MAXE = 16;
constexpr int FF_MAXE_WIDE = 32;
constexpr size_t FF_LDS = 40 * 1024;

struct ForceFlatTab {
    int8_t nrl[ACE_LMAX + 1];  // n bound per l
    int16_t nq;  // number of (l, m, n) entries (their G slots in staging order: ace_ctx::ffq)
};

struct ace_ctx {
    int device = 0;
    int prop = 0, L = 0, nr = 0, nspecies = 0, tp = 2, pl = 0, pr = 0;
    double r0 = 1.0, tl = 0, tr = 0;
    int npool = 0, nA = 0, nAA = 0, maxorder = 0, nB = 0, nnz = 0;
    int nruns = 0, nleaves = 0, ntasks = 0, nterms = 0, ngslots = 0;
    int ncoef = 1;
    // class-split EuclideanVector programs: tasks accumulate one complex
    // polynomial of class c = M + 1, G slots hold one complex value, and the
    // Cartesian components are sum_c u[c][p] (class value)
    int classed = 0;
    // per-record kernel tile: 64 or 32 environments (0: no records);
    // segment kernel tile: at most 64 environments, as many as fit the LDS
    // (0: the run interpreter only); force_interp: run the interpreter even so
    // (ace_set_kernel)
    int rec_et = 0;
    int seg_et = 0;
    bool force_interp = false;
    // rec_dispatch: the record kernel that switches per record (ace_set_kernel
    // mode 2) instead of the segment kernel
    bool rec_dispatch = false;
    // pooling launch shape (ace_set_pool_shape): neighbours per chunk (8, 16,
    // 32; 0 = 16) and environments per wave (0 = automatic, 1 = one)
    int pool_ch = 0, pool_epw = 0;
    // shared-Y lane sets of k_pool: pool_slot[hl + 64 p] is the pooled function
    // of lane hl in pass p (-1: none); the KP functions of a lane share
    // (species, l, m).  pool_shy_kp: the pass count the table is laid out for
    // (pool_passes(npool)), 0 when the sets need more than 64 lanes
    int pool_shy_kp = 0;
    DevBuf<int32_t> pool_slot;
    // basis-gradient weight tables: built on the host at context creation,
    // uploaded at the first ace_basis_dB call; db_usable: they fit the
    // weight kernel's LDS budget (DB_LDS_BUDGET)
    acehost::DbTables db_host;
    bool db_usable = false, db_ready = false;
    int db_nchunk = 0, db_max_slots = 0;
    DevBuf<int32_t> db_chunk_b, db_chunk_s, db_b_slot, db_slot_k;
    DevBuf<uint32_t> db_ztab;
    DevBuf<double2> cls_u;        // [3 classes][3 components]
    DevBuf<int32_t> gslot_cls;    // [ngslots] class of every G slot
    acehost::PackCopy host;
    // device
    DevBuf<double> rA, rB, rC;
    DevBuf<int32_t> pool, aspec_fid, aspec4, orders, iAA2iA, rowptr, cols;
    DevBuf<double> aspec_sgn, vals;
    DevBuf<BtRec> bt_rec;
    DevBuf<int32_t> bt_rowo;  // [nB][BT_ROWO] record ranges per order
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
    DevBuf<double4> vleaf;  // EuclideanVector decoded leaves (two per leaf; one when class-split)
    DevBuf<uint32_t> recs;
    DevBuf<uint32_t> segw, segt;  // segment kernel: record words, segments [kind, count, offset, 0]
    DevBuf<int32_t> seg_ptr;      // [ntasks + 1] segments of every task
    // surrogate feature tables
    int nfeat = 0;
    std::vector<int32_t> feat_rows;
    DevBuf<int32_t> feat_ptr, feat_f, fgrad_ptr, fgrad_feat, fgrad_f;
    DevBuf<double> feat_coef, fgrad_coef, cfeat;
    int64_t surr_nenv = -1;
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
    int64_t last_tail = 0;  // tiles of the last call that ran one more part
    int last_kernel = 0;    // product-DAG kernel of the last call: 0 records, else the interpreter's tile width
    int nrec = 0;           // program records (0: the run interpreter evaluates every call)
    unsigned char* scratch = nullptr;  // device address of this device's allocation-free scratch array
    // scratch
    DevBuf<double> Epart;
    DevBuf<double2> Aplus, Gbuf, Afull, AA, dphi, dAA, Wplus, Wfull;
    DevBuf<int32_t> nbr_env;
    // host-memory calls (ace_*_host): their stream and device staging buffers
    hipStream_t io_stream = nullptr;
    DevBuf<int32_t> io_off;
    DevBuf<double> io_xyz, io_w, io_out1, io_out2;
    DevBuf<int8_t> io_sp;
    // pinned host words for the stream-ordered reads of env_offsets[0] and
    // env_offsets[nenv] (hipHostMalloc)
    int32_t* h_words = nullptr;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timing_valid = false;
    bool validate = true;
    ~ace_ctx() {
        if (h_words) (void)hipHostFree(h_words);
        if (io_stream) (void)hipStreamDestroy(io_stream);
        for (auto* w : wcfg) delete w;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Complex values per G slot: the Cartesian components (3) of a
// EuclideanVector basis, or one (Invariant, class-split programs).
inline int gslot_width(const ace_ctx* ctx) { return ctx->prop == ACE_PROP_INVARIANT || ctx->classed ? 1 : 3; }

// Device scratch of one evaluation call: the pooled values A+
// [npool][estride], the adjoint G [ngslots * gslot_width][estride], the per-part
// energy shares [parts][estride] and the neighbour -> environment map of the
// flat force kernel [nnbr].  Either the context's own buffers or carved from
// the allocation-free entry point's fixed scratch array.
struct Workspace {
    double2* Aplus = nullptr;
    double2* Gbuf = nullptr;
    double* Epart = nullptr;
    int32_t* nbr_env = nullptr;
    size_t nbr_cap = 0;  // entries nbr_env holds
};

// The context's own buffers (as sized by their last ensure).
inline Workspace ctx_buffers(const ace_ctx* ctx) {
    Workspace w;
    w.Aplus = ctx->Aplus.p;
    w.Gbuf = ctx->Gbuf.p;
    w.Epart = ctx->Epart.p;
    w.nbr_env = ctx->nbr_env.p;
    w.nbr_cap = ctx->nbr_env.n;
    return w;
}

inline int set_func_lds(const void* f, size_t bytes) {
    if (bytes > 64 * 1024) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return 0;
}

inline RadialParams radial_params(const ace_ctx* ctx) {
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

// Calls FN<L>(...) for the context's Lmax (compile-time L of the
// per-neighbour kernels).
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
        case 8: FN<8>(__VA_ARGS__); break;               \
        case 9: FN<9>(__VA_ARGS__); break;               \
        default: FN<10>(__VA_ARGS__); break;             \
    }

// ---- launchers shared between units (all stream-ordered; 0 or -1 with the
// error recorded)

// acehip.hip: argument checks of every evaluation call.  check_common
// returns the neighbour count env_offsets[nenv] (read in stream order);
// validate_inputs checks offsets and species on the host.
int check_common(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz, const int8_t* species,
                 int64_t* nnbr_out, void* stream);
int validate_inputs(ace_ctx* ctx, int64_t nenv, const int32_t* off, int64_t nnbr, const int8_t* species,
                    hipStream_t st);
// acehip.hip: the product DAG over the environments listed in env_list
// (device count env_count): energies, and the adjoint G with forces.
int launch_dag_listed(ace_ctx* ctx, hipStream_t st, int64_t nenv, int64_t estride, bool forces, double* E_out,
                      const int32_t* env_list, const int32_t* env_count);
// ace_pool.hip: A+[k][env] into ws.Aplus (stride estride); with nbr_map
// also the neighbour -> environment map ws.nbr_env (flat force kernel).
int launch_pool_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                    int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws, bool nbr_map);
// ace_pool.hip: dA+w[k][env] = sum_j w_j . grad_j A+[k] into ctx->Wplus.
int launch_poolw_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                     const double* w, int64_t nnbr, const int8_t* species, int64_t estride);
// ace_force.hip: +dE/dr_j from the adjoint ws.Gbuf (and ws.nbr_env).
int launch_force_any(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                     int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws, double* dE);
