// libacehip.so, basis level: Product1pBasis A (product_1pbasis.jl:123-134),
// PIBasis AA (pibasis.jl:265-275), SymmetricBasis B = A2B AA
// (symmbasis.jl:248-315) and their derivatives dB (pibasis.jl:338-432,
// symmbasis.jl:322-336), and adjoint_EVAL_D (evaluator.jl:204-244): the calls
// behind grad_params and the least-squares design matrices.
#include "ace_common.hpp"

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
    alp_eval_ed<L>(rp, S, P, dP);
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



// Factor f of the staged tile, conjugated when f carries ACE_CONJ (the sign
// bit of the imaginary part flipped without a branch).
template <int ENVS>
__device__ __forceinline__ double2 bt_factor(const double2* As, int f, int e) {
    const double2 v = As[(f & 0x7fff) * ENVS + e];
    return make_double2(v.x, __hiloint2double(__double2hiint(v.y) ^ ((f & ACE_CONJ) << 16), __double2loint(v.y)));
}

// The records [za, zb) of one row whose AA columns have order O: each
// recomputes its product (or dual-number derivative, DIR) left to right
// from the staged tiles, straight-line code, four records per iteration so
// that their scalar loads and LDS reads overlap (two for the dual numbers of
// DIR: registers).
template <int O, int PROP, bool DIR, int ENVS>
__device__ __forceinline__ void bt_order(int za, int zb, const BtRec* __restrict__ recs,
                                         const double* __restrict__ vals, const double2* As, const double2* Ws,
                                         int e, double* s) {
#pragma unroll(DIR ? 2 : 4)
    for (int z = za; z < zb; ++z) {
        const BtRec r = recs[z];
        double2 p = make_double2(1.0, 0.0), dp = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < O; ++q) {
            const double2 a = bt_factor<ENVS>(As, r.f[q], e);
            if (DIR) {
                const double2 da = bt_factor<ENVS>(Ws, r.f[q], e);
                const double2 x = cmul(dp, a), y = cmul(p, da);
                dp = make_double2(x.x + y.x, x.y + y.y);
            }
            p = cmul(p, a);
        }
        const double2 v = DIR ? dp : p;
        if (PROP == 0) {
            s[0] += r.coef * v.x;
        } else {
            const double sg = r.coef;
#pragma unroll
            for (int q = 0; q < 3; ++q) s[q] += sg * (vals[6 * r.z + 2 * q] * v.x - vals[6 * r.z + 2 * q + 1] * v.y);
        }
    }
}

// k_Bt: basis values B = A2B . AA (DIR = false; pibasis.jl:265-275,
// symmbasis.jl:248-264) or their directional derivative A2B . dAA along
// per-neighbour weights (DIR = true; evaluator.jl:225-240) for a tile of
// ENVS environments, without materialising AA.  The pooled A+ tile (and
// dA+w for DIR) is staged in LDS as in k_dagr; every A2B nonzero (one BtRec,
// read with scalar loads) recomputes its AA product from it.  A row's
// nonzeros are ordered by the order of their AA column (rowo: where each
// order starts), one straight-line loop per order.  A workgroup is (tile,
// chunk of rows); each wave takes rpw consecutive rows (with ENVS = 32 the two
// half-waves take alternate rows).  Workgroups of one tile are consecutive on
// one XCD so the staged tile is read from that XCD's L2.
constexpr int BT_NW = 16;  // waves per k_Bt workgroup

template <int PROP, bool DIR, int ENVS>
__global__ __launch_bounds__(64 * BT_NW) void k_Bt(int64_t nenv, int npool, int nB, int nchunk, int rpw,
                                            const int32_t* __restrict__ rowo, const BtRec* __restrict__ recs,
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
        const int32_t* ro = rowo + (int64_t)b * BT_ROWO;
        bt_order<0, PROP, DIR, ENVS>(ro[0], ro[1], recs, vals, As, Ws, e, s);
        bt_order<1, PROP, DIR, ENVS>(ro[1], ro[2], recs, vals, As, Ws, e, s);
        bt_order<2, PROP, DIR, ENVS>(ro[2], ro[3], recs, vals, As, Ws, e, s);
        bt_order<3, PROP, DIR, ENVS>(ro[3], ro[4], recs, vals, As, Ws, e, s);
        bt_order<4, PROP, DIR, ENVS>(ro[4], ro[5], recs, vals, As, Ws, e, s);
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

// k_dBw: per-neighbour basis gradients through per-environment adjoint
// weights (symmbasis.jl:322-336 with pibasis.jl:402-432 reorganised).
// dB[j][b] = Re sum_k W+[b][k] dphi+[j][k], where W+ collects, for every A2B
// entry z of row b and factor t of its AA column, the A2B value times the
// product of the other factors (prefix and suffix products) times the
// factor's sign, conjugated for conjugated (m < 0) factors, in the slot of
// the factor's pooled function.  One workgroup per (environment, chunk of up
// to 256 basis functions); thread = basis function.  W+ of the chunk lives in
// LDS; the neighbours' pooled gradients dphi+ [npool][3] are staged DB_NJ at a
// time (as many as DB_STAGE double2 hold), the next stage loaded into
// registers while the current one is used.  The weights are built once per
// workgroup, so a (environment, chunk) pair costs O(A2B entries x order) for
// the weights and O(neighbours x slots) for the gradients, instead of the
// dAA[neighbour][AA function] the materialising path writes.
struct DbArgs {
    const int32_t* chunk_b;   // [nchunk + 1]
    const int32_t* chunk_s;   // [nchunk + 1]
    const int32_t* b_slot;    // [nB + 1]
    const int32_t* slot_k;    // [nslots]
    const uint4* ztab;        // [nnz][MO / 4]: factor and slot codes per A2B entry
    const int32_t* rowptr;
    const double* vals;       // [nnz] or [nnz][6]
    int npool, nB;
};

constexpr int DB_NJ = DB_NJ_MAX;  // neighbours per stage (at most)
constexpr int DB_STAGE = 2048;    // double2 per stage (8 per thread)

template <int PROP, int MO>
__global__ __launch_bounds__(256) void k_dBw(DbArgs a, int64_t nenv, const int32_t* __restrict__ off,
                                             const double2* __restrict__ Aplus, int64_t estride,
                                             const double2* __restrict__ dphi, double* __restrict__ dB) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double2 sm[];
    const int64_t env = blockIdx.x;
    const int c = blockIdx.y;
    const int tid = threadIdx.x;
    const int b0 = a.chunk_b[c], b1 = a.chunk_b[c + 1], s0 = a.chunk_s[c], s1 = a.chunk_s[c + 1];
    const int ns = s1 - s0;
    const int nq = 3 * a.npool;                           // double2 per neighbour
    const int njs = max(1, min(DB_NJ, DB_STAGE / nq));     // neighbours per stage
    double2* As = sm;                       // [npool]
    double2* dp = As + a.npool;             // [njs][npool][3]
    double2* W = dp + (size_t)njs * nq;     // [ns][NC]
    int* sk = reinterpret_cast<int*>(W + (size_t)ns * NC);  // [ns]
    for (int k = tid; k < a.npool; k += 256) As[k] = Aplus[(int64_t)k * estride + env];
    for (int q = tid; q < ns; q += 256) {
        sk[q] = a.slot_k[s0 + q];
#pragma unroll
        for (int p = 0; p < NC; ++p) W[q * NC + p] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int b = b0 + tid;
    const bool active = b < b1;
    const int sb = active ? a.b_slot[b] - s0 : 0;
    const int nsb = active ? a.b_slot[b + 1] - a.b_slot[b] : 0;
    if (active) {
        for (int z = a.rowptr[b]; z < a.rowptr[b + 1]; ++z) {
            uint32_t code[MO];
#pragma unroll
            for (int h = 0; h < MO / 4; ++h) {
                const uint4 w4 = a.ztab[(size_t)z * (MO / 4) + h];
                code[4 * h] = w4.x;
                code[4 * h + 1] = w4.y;
                code[4 * h + 2] = w4.z;
                code[4 * h + 3] = w4.w;
            }
            if (!(code[0] & (1u << 13))) continue;  // order 0: no gradient
            double2 f[MO];
            double sg[MO];
#pragma unroll
            for (int q = 0; q < MO; ++q) {
                f[q] = make_double2(1.0, 0.0);
                sg[q] = 1.0;
                if (code[q] & (1u << 13)) {
                    double2 v = As[code[q] & 0x3ffu];
                    if (code[q] & 0x8000u) v.y = -v.y;
                    sg[q] = (code[q] & 0x4000u) ? -1.0 : 1.0;
                    f[q] = make_double2(sg[q] * v.x, sg[q] * v.y);
                }
            }
            double2 pre[MO];
            pre[0] = make_double2(1.0, 0.0);
#pragma unroll
            for (int q = 1; q < MO; ++q) pre[q] = cmul(pre[q - 1], f[q - 1]);
            double2 suf = make_double2(1.0, 0.0);
#pragma unroll
            for (int t = MO - 1; t >= 0; --t) {
                if (code[t] & (1u << 13)) {
                    const double2 P = cmul(pre[t], suf);
                    const uint32_t sc = code[t] >> 16;
                    const bool cj = (sc & 0x8000u) != 0;
                    double2* w = W + (size_t)(sb + (int)(sc & 0x7fffu)) * NC;
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        double2 v;
                        if (PROP == 0) {
                            const double cv = a.vals[z] * sg[t];
                            v = make_double2(cv * P.x, cv * P.y);
                        } else {
                            const double2 cv = make_double2(a.vals[6 * z + 2 * p] * sg[t], a.vals[6 * z + 2 * p + 1] * sg[t]);
                            v = cmul(cv, P);
                        }
                        w[p].x += v.x;
                        w[p].y += cj ? -v.y : v.y;
                    }
                    suf = cmul(suf, f[t]);
                }
            }
        }
    }
    __syncthreads();
    const int j0 = off[env], j1 = off[env + 1];
    // stages of njs neighbours: their dphi+ rows are consecutive in memory
    double2 pf[DB_STAGE / 256];
    auto fetch = [&](int jb) {
        const int n = min(njs, j1 - jb) * nq;
        const double2* src = dphi + (int64_t)jb * nq;
#pragma unroll
        for (int r = 0; r < DB_STAGE / 256; ++r) {
            const int q = tid + 256 * r;
            if (q < n) pf[r] = src[q];
        }
    };
    if (j0 < j1) fetch(j0);
    for (int jb = j0; jb < j1; jb += njs) {
        const int nj = min(njs, j1 - jb);
#pragma unroll
        for (int r = 0; r < DB_STAGE / 256; ++r) {
            const int q = tid + 256 * r;
            if (q < nj * nq) dp[q] = pf[r];
        }
        __syncthreads();
        if (jb + njs < j1) fetch(jb + njs);  // in flight while this stage is used
        if (active) {
            for (int jj = 0; jj < nj; ++jj) {
                const double2* d = dp + (size_t)jj * nq;
                double acc[NC][3];
#pragma unroll
                for (int p = 0; p < NC; ++p) acc[p][0] = acc[p][1] = acc[p][2] = 0.0;
                for (int q = 0; q < nsb; ++q) {
                    const int k = sk[sb + q];
                    const double2 x0 = d[3 * k], x1 = d[3 * k + 1], x2 = d[3 * k + 2];
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        const double2 w = W[(size_t)(sb + q) * NC + p];
                        acc[p][0] += w.x * x0.x - w.y * x0.y;
                        acc[p][1] += w.x * x1.x - w.y * x1.y;
                        acc[p][2] += w.x * x2.x - w.y * x2.y;
                    }
                }
                double* out = dB + ((int64_t)(jb + jj) * a.nB + b) * NC * 3;
#pragma unroll
                for (int p = 0; p < NC; ++p) {
                    out[3 * p] = acc[p][0];
                    out[3 * p + 1] = acc[p][1];
                    out[3 * p + 2] = acc[p][2];
                }
            }
        }
        __syncthreads();
    }
}

__global__ void k_nbr_env(int64_t nenv, const int32_t* __restrict__ off, int32_t* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    for (int j = off[e]; j < off[e + 1]; ++j) out[j] = (int32_t)e;
}

// ---- basis level

static int basis_A_internal(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, int64_t nnbr, hipStream_t st) {
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
    if (ctx->Afull.ensure((size_t)nenv * ctx->nA)) return -1;
    if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), false)) return -1;
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
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
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
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
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

// Fused basis values (k_Bt): the tile it runs on (environments; 0 when the
// pooled tile does not fit LDS), and the launch.  Basis values take 64
// environments when their A+ tile fits, else 32 (npool <= 320); the
// directional derivative stages A+ and dA+w, 32 environments.
static int bt_envs(const ace_ctx* ctx, bool dir) {
    if (ctx->maxorder > 4) return 0;  // BtRec holds four factors
    const size_t row = (size_t)ctx->npool * sizeof(double2) * (dir ? 2 : 1);
    if (!dir && row * 64 <= 160 * 1024) return 64;
    return row * 32 <= 160 * 1024 ? 32 : 0;
}

template <int PROP, bool DIR, int ENVS>
static int launch_bt_shape(ace_ctx* ctx, hipStream_t st, int64_t nenv, int64_t estride, double* out) {
    const size_t lds = (size_t)ctx->npool * ENVS * sizeof(double2) * (DIR ? 2 : 1);
    constexpr int NW = BT_NW;
    const int rpw = std::max(1, (ctx->nB + NW * 12 - 1) / (NW * 12));
    const int nchunk = (ctx->nB + NW * rpw - 1) / (NW * rpw);
    const int64_t nblk = (nenv + ENVS - 1) / ENVS * nchunk;
    if (nblk >= ((int64_t)1 << 31)) return set_err("batch too large for one launch");
    set_func_lds((const void*)k_Bt<PROP, DIR, ENVS>, lds);
    hipLaunchKernelGGL((k_Bt<PROP, DIR, ENVS>), dim3((unsigned)nblk), dim3(64 * NW), lds, st, nenv, ctx->npool,
                       ctx->nB, nchunk, rpw, ctx->bt_rowo.p, ctx->bt_rec.p, ctx->vals.p, ctx->Aplus.p, ctx->Wplus.p,
                       estride, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <bool DIR>
static int launch_bt(ace_ctx* ctx, hipStream_t st, int64_t nenv, int64_t estride, double* out) {
    const int envs = bt_envs(ctx, DIR);
    if (envs == 0) return set_err("pooled tile does not fit LDS");
    if (DIR) {
        if (ctx->prop != ACE_PROP_INVARIANT) return set_err("directional basis derivative: Invariant bases only");
        return launch_bt_shape<0, true, 32>(ctx, st, nenv, estride, out);
    }
    if (ctx->prop == ACE_PROP_INVARIANT)
        return envs == 64 ? launch_bt_shape<0, false, 64>(ctx, st, nenv, estride, out)
                          : launch_bt_shape<0, false, 32>(ctx, st, nenv, estride, out);
    return envs == 64 ? launch_bt_shape<1, false, 64>(ctx, st, nenv, estride, out)
                      : launch_bt_shape<1, false, 32>(ctx, st, nenv, estride, out);
}

extern "C" int ace_basis_B(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                           const int8_t* species, double* B_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
    if (nenv == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    {
        const int64_t estride = ((nenv + 63) / 64) * 64;
        if (bt_envs(ctx, false) > 0) {
            if (ctx->Aplus.ensure((size_t)ctx->npool * estride)) return -1;
            if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), false)) return -1;
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

template <int L>
static void launch_dphi_pool(ace_ctx* ctx, hipStream_t st, int64_t nnbr, const double* xyz,
                             const int8_t* species) {
    // the pooled (m >= 0) functions: pool rows are (mu, n, l, m) like aspec rows
    hipLaunchKernelGGL(k_dphi<L>, dim3((unsigned)((nnbr + 63) / 64)), dim3(64), 0, st, radial_params(ctx),
                       nnbr, ctx->nbr_env.p, xyz, ctx->nspecies > 0 ? species : nullptr, ctx->pool.p,
                       ctx->npool, ctx->dphi.p);
}

// Basis gradients through per-environment adjoint weights (k_dBw): 0, or 1
// when the tables do not fit the kernel's LDS (the caller then runs the
// materialising path).
static int basis_dB_weights(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, int64_t nnbr, double* dB_out) {
    if (!ctx->db_usable) return 1;
    const int NC = ctx->prop == ACE_PROP_INVARIANT ? 1 : 3;
    const size_t fixed = db_fixed_lds(ctx->npool);
    const size_t per_slot = (size_t)NC * sizeof(double2) + sizeof(int32_t);
    if (!ctx->db_ready) {
        acehost::DbTables& t = ctx->db_host;
        // no empty uploads (a basis without A2B entries)
        if (t.slot_k.empty()) t.slot_k.push_back(0);
        if (t.ztab.size() < 4) t.ztab.resize(4, 0u);
        if (ctx->db_chunk_b.upload(t.chunk_b.data(), t.chunk_b.size()) ||
            ctx->db_chunk_s.upload(t.chunk_s.data(), t.chunk_s.size()) ||
            ctx->db_b_slot.upload(t.b_slot.data(), t.b_slot.size()) ||
            ctx->db_slot_k.upload(t.slot_k.data(), t.slot_k.size()) ||
            ctx->db_ztab.upload(t.ztab.data(), t.ztab.size()))
            return -1;
        ctx->db_nchunk = (int)t.chunk_b.size() - 1;
        ctx->db_max_slots = t.max_chunk_slots;
        ctx->db_ready = true;
    }
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (ctx->Aplus.ensure((size_t)ctx->npool * estride) || ctx->nbr_env.ensure(nnbr) ||
        ctx->dphi.ensure((size_t)nnbr * ctx->npool * 3))
        return -1;
    if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), false)) return -1;
    hipLaunchKernelGGL(k_nbr_env, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, off,
                       ctx->nbr_env.p);
    DISPATCH_L(ctx->L, launch_dphi_pool, ctx, st, nnbr, xyz, species);
    HIP_TRY(hipGetLastError());
    DbArgs a;
    a.chunk_b = ctx->db_chunk_b.p;
    a.chunk_s = ctx->db_chunk_s.p;
    a.b_slot = ctx->db_b_slot.p;
    a.slot_k = ctx->db_slot_k.p;
    a.ztab = reinterpret_cast<const uint4*>(ctx->db_ztab.p);
    a.rowptr = ctx->rowptr.p;
    a.vals = ctx->vals.p;
    a.npool = ctx->npool;
    a.nB = ctx->nB;
    if (ctx->db_nchunk == 0) return 0;  // no basis functions
    const size_t lds = fixed + (size_t)ctx->db_max_slots * per_slot;
    const dim3 grid((unsigned)nenv, (unsigned)ctx->db_nchunk);
    auto go = [&](auto kern) {
        set_func_lds((const void*)kern, lds);
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, a, nenv, off, ctx->Aplus.p, estride, ctx->dphi.p, dB_out);
    };
    const bool wide = ctx->db_host.zw == 8;
    if (ctx->prop == ACE_PROP_INVARIANT)
        wide ? go(k_dBw<0, 8>) : go(k_dBw<0, 4>);
    else
        wide ? go(k_dBw<1, 8>) : go(k_dBw<1, 4>);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ace_basis_dB(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                            const int8_t* species, double* dB_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
    if (nenv == 0 || nnbr == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int rc = basis_dB_weights(ctx, st, nenv, off, xyz, species, nnbr, dB_out);
    if (rc <= 0) return rc;
    // bases whose tables exceed the weight kernel's LDS: dAA per neighbour
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

extern "C" int ace_adjoint_EVAL_D(ace_ctx* ctx, int64_t nenv, const int32_t* off, const double* xyz,
                                  const int8_t* species, const double* w, double* dB_out, void* stream) {
    int64_t nnbr = 0;
    if (check_common(ctx, nenv, off, xyz, species, &nnbr, stream)) return -1;
    if (ctx->prop != ACE_PROP_INVARIANT) return set_err("ace_adjoint_EVAL_D: Invariant bases only");
    if (validate_inputs(ctx, nenv, off, nnbr, species, (hipStream_t)stream)) return -1;
    if (nenv == 0) return 0;
    if (!dB_out || (nnbr > 0 && !w)) return set_err("ace_adjoint_EVAL_D: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t estride = ((nenv + 63) / 64) * 64;
    if (bt_envs(ctx, true) > 0) {
        if (ctx->Aplus.ensure((size_t)ctx->npool * estride) || ctx->Wplus.ensure((size_t)ctx->npool * estride))
            return -1;
        if (launch_pool_any(ctx, st, nenv, off, xyz, nnbr, species, estride, ctx_buffers(ctx), false) ||
            launch_poolw_any(ctx, st, nenv, off, xyz, w, nnbr, species, estride))
            return -1;
        return launch_bt<true>(ctx, st, nenv, estride, dB_out);
    }
    if (basis_A_internal(ctx, nenv, off, xyz, species, nnbr, st)) return -1;
    if (ctx->Wplus.ensure((size_t)ctx->npool * estride) || ctx->Wfull.ensure((size_t)nenv * ctx->nA) ||
        ctx->AA.ensure((size_t)nenv * ctx->nAA))
        return -1;
    if (launch_poolw_any(ctx, st, nenv, off, xyz, w, nnbr, species, estride)) return -1;
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
