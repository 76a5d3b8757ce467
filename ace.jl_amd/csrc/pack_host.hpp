// Host-side (CPU) logic of libacehip.so on the flat basis pack: validation of
// every table before any device use, the effective coefficients
// c~ = A2B^T c (evaluator.jl:59-67, symmbasis.jl:267-285) and the program
// leaf / task coefficients derived from them, and the lookup tables of the
// kernels.  Plain C++ (no HIP) so it is also built and tested on the CPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/acehip.h"

namespace acehost {

struct WaveTable {
    int W;
    std::vector<int32_t> ptr_all, list_all, ptr_e, list_e;
};

// Class c = M + 1 of every AA row of a pack, M = the total m of its factors
// (pooled m, negated for conjugated factors); -1 for the order-0 row.
inline std::vector<int8_t> pack_row_classes(const ace_pack* pk) {
    std::vector<int8_t> out((size_t)pk->nAA, (int8_t)-1);
    for (int i = 0; i < pk->nAA; ++i) {
        const int o = pk->orders[i];
        if (!o) continue;
        int M = 0;
        for (int t = 0; t < o; ++t) {
            const int f = pk->aspec_fid[pk->iAA2iA[(size_t)i * pk->maxorder + t]];
            const int m = pk->pool[4 * (f & 0x7fff) + 3];
            M += (f & 0x8000) ? -m : m;
        }
        out[(size_t)i] = (int8_t)std::max(-2, std::min(M + 1, 3));  // outside [0, 2]: not a vector class
    }
    return out;
}

// The class structure of a EuclideanVector A2B map (symmbasis.jl:267-285 with
// the l = 1 coupling of the vector property): every value of a column of
// class c should be one complex number times a unit vector u_c, and u_c, u_{2-c}
// should be conjugate up to a phase (Y_1^{-M} = (-1)^M conj(Y_1^M)).
//   u[c]    the value of largest norm among the class's entries (the first
//           such), normalised; a class without entries takes conj(u) of its
//           mirror (or stays zero)
//   sig[c]  <conj(u_c), u_{2-c}> = sum_p conj(u_c,p) conj(u_{2-c},p), so that
//           conj(u_c) = sig[c] u_{2-c}
//   ok      every entry lies within tol * (largest norm of the map) of its
//           class line, and |sig[c]| >= 1 - tol wherever both classes occur
struct ClassVectors {
    double u[3][6] = {};
    double sig[3][2] = {};
    bool ok = false;
};

inline ClassVectors class_vectors(const int32_t* cols, const double* vals, size_t nnz, const int8_t* row_class,
                                  double tol = 1e-12) {
    ClassVectors cv;
    auto norm = [](const double* v) {
        double n2 = 0.0;
        for (int q = 0; q < 6; ++q) n2 += v[q] * v[q];
        return std::sqrt(n2);
    };
    double best[3] = {0.0, 0.0, 0.0}, maxn = 0.0;
    long ref[3] = {-1, -1, -1};
    for (size_t z = 0; z < nnz; ++z) {
        const int c = row_class[cols[z]];
        if (c < 0 || c > 2) return cv;
        const double n = norm(vals + 6 * z);
        maxn = std::max(maxn, n);
        if (n > best[c]) {
            best[c] = n;
            ref[c] = (long)z;
        }
    }
    for (int c = 0; c < 3; ++c)
        if (ref[c] >= 0)
            for (int q = 0; q < 6; ++q) cv.u[c][q] = vals[6 * ref[c] + q] / best[c];
    for (int c = 0; c < 3; ++c)
        if (ref[c] < 0 && ref[2 - c] >= 0)
            for (int p = 0; p < 3; ++p) {
                cv.u[c][2 * p] = cv.u[2 - c][2 * p];
                cv.u[c][2 * p + 1] = -cv.u[2 - c][2 * p + 1];
            }
    for (size_t z = 0; z < nnz; ++z) {
        const int c = row_class[cols[z]];
        const double* v = vals + 6 * z;
        const double* u = cv.u[c];
        double dr = 0.0, di = 0.0;  // <v, u> = sum_p v_p conj(u_p)
        for (int p = 0; p < 3; ++p) {
            dr += v[2 * p] * u[2 * p] + v[2 * p + 1] * u[2 * p + 1];
            di += v[2 * p + 1] * u[2 * p] - v[2 * p] * u[2 * p + 1];
        }
        double r[6];
        for (int p = 0; p < 3; ++p) {
            r[2 * p] = v[2 * p] - (dr * u[2 * p] - di * u[2 * p + 1]);
            r[2 * p + 1] = v[2 * p + 1] - (dr * u[2 * p + 1] + di * u[2 * p]);
        }
        if (norm(r) > tol * maxn) return cv;
    }
    for (int c = 0; c < 3; ++c) {
        const double *a = cv.u[c], *b = cv.u[2 - c];
        double sr = 0.0, si = 0.0;  // sum_p conj(a_p) conj(b_p) = conj(sum_p a_p b_p)
        for (int p = 0; p < 3; ++p) {
            sr += a[2 * p] * b[2 * p] - a[2 * p + 1] * b[2 * p + 1];
            si += a[2 * p] * b[2 * p + 1] + a[2 * p + 1] * b[2 * p];
        }
        cv.sig[c][0] = sr;
        cv.sig[c][1] = -si;
        if (ref[c] >= 0 && ref[2 - c] >= 0 && std::sqrt(sr * sr + si * si) < 1.0 - tol) return cv;
    }
    cv.ok = true;
    return cv;
}

// Returns "" when the pack is consistent, else a message.
inline std::string validate_pack(const ace_pack* pk) {
    if (!pk) return "null pack";
    if (pk->abi_version != ACEHIP_ABI_VERSION) return "ABI version mismatch";
    if (pk->prop != ACE_PROP_INVARIANT && pk->prop != ACE_PROP_EUCLIDEAN_VECTOR) return "unknown property";
    if (pk->L < 0 || pk->L > 10) return "L must be in [0, 10]";
    if (pk->nr < 1 || pk->nr > 32) return "nr must be in [1, 32]";
    if (pk->npool < 1 || pk->npool > 512) return "npool out of range (at most 512 pooled functions)";
    if (pk->nspecies < 0 || pk->nspecies > 127) return "nspecies out of range";
    if (pk->tp < 1 || pk->tp > 16) return "transform power out of range";
    if (pk->pl < 0 || pk->pr < 0 || pk->pl > 16 || pk->pr > 16) return "envelope powers out of range";
    if (pk->nA < 0 || pk->nAA < 0 || pk->nB < 0 || pk->a2b_nnz < 0) return "negative table size";
    if (pk->maxorder < 1 || pk->maxorder > ACE_MAX_ORDER) return "maxorder must be in [1, 8]";
    if (!pk->rA || !pk->rB || !pk->rC || !pk->pool) return "null radial/pool table";
    for (int k = 0; k < pk->npool; ++k) {
        const int mu = pk->pool[4 * k], n = pk->pool[4 * k + 1], l = pk->pool[4 * k + 2],
                  m = pk->pool[4 * k + 3];
        if (n < 1 || n > pk->nr || l < 0 || l > pk->L || m < 0 || m > l) return "pool entry out of range";
        if (pk->nspecies > 0 ? (mu < 0 || mu >= pk->nspecies) : (mu != -1)) return "pool species out of range";
    }
    for (int i = 0; i < pk->nA; ++i)
        if (pk->aspec_fid[i] < 0 || (pk->aspec_fid[i] & 0x7fff) >= pk->npool) return "aspec_fid out of range";
    for (int r = 0; r < pk->nAA; ++r) {
        if (pk->orders[r] < 0 || pk->orders[r] > pk->maxorder) return "orders out of range";
        for (int t = 0; t < pk->orders[r]; ++t) {
            const int v = pk->iAA2iA[r * pk->maxorder + t];
            if (v < 0 || v >= pk->nA) return "iAA2iA out of range";
        }
    }
    if (!pk->a2b_rowptr || pk->a2b_rowptr[0] != 0 || pk->a2b_rowptr[pk->nB] != pk->a2b_nnz) return "bad A2B rowptr";
    for (int b = 0; b < pk->nB; ++b)
        if (pk->a2b_rowptr[b + 1] < pk->a2b_rowptr[b]) return "bad A2B rowptr";
    for (int z = 0; z < pk->a2b_nnz; ++z)
        if (pk->a2b_cols[z] < 0 || pk->a2b_cols[z] >= pk->nAA) return "A2B column out of range";
    if (pk->nruns < 0 || pk->nleaves < 0 || pk->ntasks < 0 || pk->nterms < 0 || pk->ngslots < 0)
        return "negative program size";
    for (int r = 0; r < pk->nruns; ++r) {
        const int32_t* run = pk->runs + ACE_RUN_W * (size_t)r;
        if (run[0] < 0 || run[0] > ACE_MAX_ORDER - 1 || run[0] > pk->maxorder - 1 || run[1] < 1 ||
            run[1] > run[0] + 1 || run[2] < 1 || run[3] < 0 || (int64_t)run[3] + run[2] > pk->nleaves)
            return "run out of range";
        for (int q = 0; q < run[0]; ++q)
            if (run[4 + q] < 0 || (run[4 + q] & 0x7fff) >= pk->npool) return "run factor out of range";
    }
    for (int i = 0; i < pk->nleaves; ++i)
        if (pk->leaf_fid[i] < 0 || (pk->leaf_fid[i] & 0x7fff) >= pk->npool) return "leaf factor out of range";
    if (pk->leaf_term_ptr[0] != 0) return "bad leaf term table";
    for (int i = 0; i < pk->nleaves; ++i)
        if (pk->leaf_term_ptr[i + 1] < pk->leaf_term_ptr[i] || pk->leaf_term_ptr[i + 1] > pk->nterms)
            return "bad leaf term table";
    if (pk->nrec < 0 || (pk->nrec > 0 && (!pk->rec_hdr || !pk->rec_leaf))) return "bad record table";
    for (int r = 0; r < pk->nrec; ++r) {
        const uint32_t* h = pk->rec_hdr + 8 * (size_t)r;
        const int d = h[0] & 3, j = (h[0] >> 2) & 7, nl = (h[0] >> 5) & 7;
        if (h[0] >> 8) return "record header out of range";
        if (j < 1 || j > d + 1 || nl < 1 || nl > 4) return "record header out of range";
        const uint32_t f[3] = {h[1] & 0xffff, h[1] >> 16, h[2] & 0xffff};
        for (int q = 0; q < 3; ++q)
            if ((f[q] & 0x7fff) >= (uint32_t)pk->npool) return "record factor out of range";
        for (int i = 0; i < 4; ++i) {
            const uint32_t w = h[3 + i / 2];
            const uint32_t fid = (i & 1) ? (w >> 16) : (w & 0xffff);
            if ((fid & 0x7fff) >= (uint32_t)pk->npool) return "record factor out of range";
            const int lf = pk->rec_leaf[4 * (size_t)r + i];
            if ((i < nl) != (lf >= 0) || lf >= pk->nleaves) return "record leaf out of range";
        }
    }
    for (int t = 0; t < pk->ntasks; ++t) {
        const int32_t* tk = pk->tasks + 5 * (size_t)t;
        if (tk[0] < 0 || tk[1] < tk[0] || tk[1] > pk->nruns || tk[2] < (pk->classed ? -3 : -1) ||
            tk[2] >= pk->ngslots)
            return "task out of range";
        if (tk[3] < 0 || tk[4] < tk[3] || tk[4] > pk->nrec) return "task out of range";
        if (pk->task_term_ptr[t] < 0 || pk->task_term_ptr[t + 1] < pk->task_term_ptr[t] ||
            pk->task_term_ptr[t + 1] > pk->nterms)
            return "bad task term table";
    }
    for (int tt = 0; tt < pk->nterms; ++tt)
        if (pk->term_row[tt] < 0 || pk->term_row[tt] >= pk->nAA) return "term row out of range";
    if (!pk->gslot_ptr || pk->gslot_ptr[0] != 0) return "bad gslot table";
    for (int k = 0; k < pk->npool; ++k)
        if (pk->gslot_ptr[k + 1] < pk->gslot_ptr[k]) return "bad gslot table";
    if (pk->gslot_ptr[pk->npool] > pk->ngslots) return "bad gslot table";
    for (int s = 0; s < pk->gslot_ptr[pk->npool]; ++s)
        if (pk->gslot_list[s] < 0 || pk->gslot_list[s] >= pk->ngslots) return "gslot out of range";
    if (pk->classed != 0 && pk->classed != 1) return "classed must be 0 or 1";
    if (pk->classed) {
        if (pk->prop != ACE_PROP_EUCLIDEAN_VECTOR) return "class-split programs need a EuclideanVector basis";
        if (pk->ngslots > 0 && !pk->gslot_class) return "null gslot_class";
        for (int s = 0; s < pk->ngslots; ++s)
            if (pk->gslot_class[s] < 0 || pk->gslot_class[s] > 2) return "gslot class out of range";
        const std::vector<int8_t> rc = pack_row_classes(pk);
        if (!class_vectors(pk->a2b_cols, pk->a2b_vals, (size_t)pk->a2b_nnz, rc.data()).ok)
            return "classed pack, but the A2B map is not one complex number times a class vector per column";
    }
    if (pk->nfeat < 0 || pk->nfeat > 160) return "nfeat must be in [0, 160]";
    if (pk->nfeat > 0) {
        if (!pk->feat_rows || !pk->feat_ptr || !pk->feat_f || !pk->feat_coef || !pk->fgrad_ptr)
            return "null feature table";
        if (pk->feat_ptr[0] != 0) return "bad feature table";
        for (int f = 0; f < pk->nfeat; ++f) {
            if (pk->feat_rows[f] < 0 || pk->feat_rows[f] >= pk->nB) return "feature row out of range";
            if (pk->feat_ptr[f + 1] < pk->feat_ptr[f]) return "bad feature table";
        }
        const int nt = pk->feat_ptr[pk->nfeat];
        for (int z = 0; z < 2 * nt; ++z)
            if (pk->feat_f[z] < -1 || (pk->feat_f[z] >= 0 && (pk->feat_f[z] & 0x7fff) >= pk->npool))
                return "feature factor out of range";
        if (pk->fgrad_ptr[0] != 0) return "bad feature gradient table";
        for (int k = 0; k < pk->npool; ++k)
            if (pk->fgrad_ptr[k + 1] < pk->fgrad_ptr[k]) return "bad feature gradient table";
        const int ng = pk->fgrad_ptr[pk->npool];
        for (int z = 0; z < ng; ++z) {
            if (pk->fgrad_feat[z] < 0 || pk->fgrad_feat[z] >= pk->nfeat) return "feature gradient entry out of range";
            if (pk->fgrad_f[z] < -1 || (pk->fgrad_f[z] >= 0 && (pk->fgrad_f[z] & 0x7fff) >= pk->npool))
                return "feature gradient entry out of range";
        }
    }
    if (pk->nwave_cfg < 1) return "no wave configuration";
    return "";
}

// Parse and check the per-tile-shape task assignments.
inline std::string wave_tables(const ace_pack* pk, std::vector<WaveTable>& out) {
    out.clear();
    size_t pa = 0, la = 0, pe = 0, le = 0;
    for (int q = 0; q < pk->nwave_cfg; ++q) {
        const int W = pk->wave_cfg_W[q];
        if (W < 1 || W > 16) return "waves per tile must be in [1, 16]";
        WaveTable wt;
        wt.W = W;
        wt.ptr_all.assign(pk->wave_task_ptr + pa, pk->wave_task_ptr + pa + W + 1);
        wt.ptr_e.assign(pk->wave_etask_ptr + pe, pk->wave_etask_ptr + pe + W + 1);
        if (wt.ptr_all[0] != 0 || wt.ptr_e[0] != 0) return "bad wave task table";
        for (int w = 0; w < W; ++w)
            if (wt.ptr_all[w + 1] < wt.ptr_all[w] || wt.ptr_e[w + 1] < wt.ptr_e[w]) return "bad wave task table";
        const int na = wt.ptr_all[W], ne = wt.ptr_e[W];
        wt.list_all.assign(pk->wave_task_list + la, pk->wave_task_list + la + na);
        wt.list_e.assign(pk->wave_etask_list + le, pk->wave_etask_list + le + ne);
        for (int x : wt.list_all)
            if (x < 0 || x >= pk->ntasks) return "bad wave task table";
        for (int x : wt.list_e)
            if (x < 0 || x >= pk->ntasks) return "bad wave task table";
        pa += W + 1;
        pe += W + 1;
        la += na;
        le += ne;
        out.push_back(std::move(wt));
    }
    return "";
}

// Coefficient state derived from c (all host vectors).
struct Coeffs {
    std::vector<double> leaf;   // [nleaves][ncoef]
    std::vector<double> task;   // [ntasks][2 | 6]
    double e0[6] = {0, 0, 0, 0, 0, 0};  // constant (order-0 AA rows), complex components
};

// The parts of the pack kept after ace_ctx_create for ace_set_params.
struct PackCopy {
    int prop = 0, nAA = 0, nB = 0, nleaves = 0, ntasks = 0, nrec = 0;
    std::vector<uint32_t> rec_hdr;
    std::vector<int32_t> rec_leaf;
    std::vector<int32_t> leaf_fid;  // [nleaves] factor id (k | 0x8000 conjugated)
    std::vector<int32_t> tasks;  // [ntasks][5] (run_b, run_e, slot, rec_b, rec_e)
    std::vector<int32_t> runs;   // [nruns][ACE_RUN_W] depth, jstart, nleaf, leaf_off, prefix factors
    std::vector<int32_t> rowptr, cols, leaf_term_ptr, task_term_ptr, term_row, term_conj, orders;
    std::vector<double> vals, term_scale;
    // class-split EuclideanVector programs: row classes and class vectors
    int classed = 0;
    std::vector<int8_t> row_class;
    ClassVectors cv;
    int ncoef() const { return prop == ACE_PROP_INVARIANT ? 1 : 6; }
    // doubles per leaf / task coefficient: 1 real, 2 (one complex, class-split), 6
    int lcoef() const { return prop == ACE_PROP_INVARIANT ? 1 : classed ? 2 : 6; }
};

inline PackCopy copy_pack(const ace_pack* pk) {
    PackCopy p;
    p.prop = pk->prop;
    p.nAA = pk->nAA;
    p.nB = pk->nB;
    p.nleaves = pk->nleaves;
    p.ntasks = pk->ntasks;
    const int nc = p.ncoef();
    p.rowptr.assign(pk->a2b_rowptr, pk->a2b_rowptr + pk->nB + 1);
    p.cols.assign(pk->a2b_cols, pk->a2b_cols + pk->a2b_nnz);
    p.vals.assign(pk->a2b_vals, pk->a2b_vals + (size_t)pk->a2b_nnz * nc);
    p.leaf_term_ptr.assign(pk->leaf_term_ptr, pk->leaf_term_ptr + pk->nleaves + 1);
    p.leaf_fid.assign(pk->leaf_fid, pk->leaf_fid + pk->nleaves);
    p.task_term_ptr.assign(pk->task_term_ptr, pk->task_term_ptr + pk->ntasks + 1);
    p.term_row.assign(pk->term_row, pk->term_row + pk->nterms);
    p.term_scale.assign(pk->term_scale, pk->term_scale + pk->nterms);
    p.term_conj.assign(pk->term_conj, pk->term_conj + pk->nterms);
    p.orders.assign(pk->orders, pk->orders + pk->nAA);
    p.nrec = pk->nrec;
    p.rec_hdr.assign(pk->rec_hdr, pk->rec_hdr + 8 * (size_t)pk->nrec);
    p.rec_leaf.assign(pk->rec_leaf, pk->rec_leaf + 4 * (size_t)pk->nrec);
    p.tasks.assign(pk->tasks, pk->tasks + 5 * (size_t)pk->ntasks);
    p.runs.assign(pk->runs, pk->runs + (size_t)ACE_RUN_W * pk->nruns);
    p.classed = pk->classed;
    if (p.classed) {
        p.row_class = pack_row_classes(pk);
        p.cv = class_vectors(pk->a2b_cols, pk->a2b_vals, (size_t)pk->a2b_nnz, p.row_class.data());
    }
    return p;
}

// c~ = A2B^T c, then every leaf / task coefficient as a sum of its terms.
inline Coeffs coefficients(const PackCopy& p, const double* c) {
    const int nc = p.ncoef();
    std::vector<double> ct((size_t)p.nAA * nc, 0.0);
    for (int b = 0; b < p.nB; ++b)
        for (int z = p.rowptr[b]; z < p.rowptr[b + 1]; ++z)
            for (int q = 0; q < nc; ++q) ct[(size_t)p.cols[z] * nc + q] += p.vals[(size_t)z * nc + q] * c[b];
    auto add_term = [&](int tt, double* out) {
        const int row = p.term_row[tt];
        const double s = p.term_scale[tt];
        const bool cj = p.term_conj[tt] != 0;
        for (int q = 0; q < nc; ++q) {
            double v = ct[(size_t)row * nc + q] * s;
            if (cj && (q & 1)) v = -v;
            out[q] += v;
        }
    };
    Coeffs out;
    for (int r = 0; r < p.nAA; ++r)
        if (p.orders[r] == 0)
            for (int q = 0; q < nc; ++q) out.e0[q] += ct[(size_t)r * nc + q];
    if (p.classed) {
        // c~[row] = s[row] u_M with s[row] = <c~[row], u_M>; a term's scalar is
        // scale s[row], or scale sig_M conj(s[row]) when conjugated (class -M)
        std::vector<double> sc((size_t)p.nAA * 2, 0.0);
        for (int r = 0; r < p.nAA; ++r) {
            const int c = p.row_class[(size_t)r];
            if (c < 0) continue;
            const double* v = &ct[(size_t)r * 6];
            const double* u = p.cv.u[c];
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < 3; ++q) {
                sr += v[2 * q] * u[2 * q] + v[2 * q + 1] * u[2 * q + 1];
                si += v[2 * q + 1] * u[2 * q] - v[2 * q] * u[2 * q + 1];
            }
            sc[2 * (size_t)r] = sr;
            sc[2 * (size_t)r + 1] = si;
        }
        auto add_scalar = [&](int tt, double* o) {
            const int row = p.term_row[tt];
            const double s = p.term_scale[tt], sr = sc[2 * (size_t)row], si = sc[2 * (size_t)row + 1];
            if (p.term_conj[tt]) {
                const double* g = p.cv.sig[p.row_class[(size_t)row]];
                o[0] += s * (g[0] * sr + g[1] * si);  // sig conj(s)
                o[1] += s * (g[1] * sr - g[0] * si);
            } else {
                o[0] += s * sr;
                o[1] += s * si;
            }
        };
        out.leaf.assign((size_t)p.nleaves * 2, 0.0);
        for (int i = 0; i < p.nleaves; ++i)
            for (int tt = p.leaf_term_ptr[i]; tt < p.leaf_term_ptr[i + 1]; ++tt) add_scalar(tt, &out.leaf[2 * (size_t)i]);
        out.task.assign((size_t)p.ntasks * 2, 0.0);
        for (int t = 0; t < p.ntasks; ++t)
            for (int tt = p.task_term_ptr[t]; tt < p.task_term_ptr[t + 1]; ++tt) add_scalar(tt, &out.task[2 * (size_t)t]);
        return out;
    }
    out.leaf.assign((size_t)p.nleaves * nc, 0.0);
    for (int i = 0; i < p.nleaves; ++i)
        for (int tt = p.leaf_term_ptr[i]; tt < p.leaf_term_ptr[i + 1]; ++tt) add_term(tt, &out.leaf[(size_t)i * nc]);
    const int tcn = p.prop == ACE_PROP_INVARIANT ? 2 : 6;
    out.task.assign((size_t)p.ntasks * tcn, 0.0);
    for (int t = 0; t < p.ntasks; ++t) {
        double tmp[6] = {0, 0, 0, 0, 0, 0};
        for (int tt = p.task_term_ptr[t]; tt < p.task_term_ptr[t + 1]; ++tt) add_term(tt, tmp);
        for (int q = 0; q < nc; ++q) out.task[(size_t)t * tcn + q] = tmp[q];
    }
    return out;
}

// Device records, decoded for the kernel: 64 bytes each (real leaf
// coefficients, Invariant), 96 bytes (one complex coefficient per leaf,
// class-split EuclideanVector programs).
//   dw0    control: REC_MORE (leaves 2 and 3 present), REC_U1 / REC_U2 /
//          REC_U3 (prefix level 1 / 2 / 3 is recomputed) -- together the
//          record type in bits 2-5 --, REC_CONJ (the leaves are conjugated:
//          one flag per record, since the leaf's m is fixed by the prefix),
//          REC_CONJ1 / 2 / 3 (prefix factor t is conjugated)
//   dw1-4  leaf i: LDS byte offset of its pooled value (k << lds_shift: 64
//          lanes x 16 B at shift 10, 32 lanes at shift 9)
//   dw5-7  prefix factor t: LDS byte offset
//   dw8-15 leaf coefficients (FP64; zero for absent leaves), or
//   dw8-23 complex leaf coefficients (re, im), stored conjugated when the
//          leaves are: sum c conj(A) = conj(sum conj(c) A)
// One zero record follows the last so the kernel can always load record r + 1.
constexpr uint32_t REC_MORE = 1u << 2, REC_U1 = 1u << 3, REC_U2 = 1u << 4, REC_U3 = 1u << 5;
constexpr uint32_t REC_CONJ1 = 1u << 28, REC_CONJ2 = 1u << 29, REC_CONJ3 = 1u << 30, REC_CONJ = 1u << 31;

inline int rec_words(const PackCopy& p) { return p.classed ? 24 : 16; }

// Returns "" or a message (a record whose leaves mix conjugation).
inline std::string records(const PackCopy& p, const Coeffs& cf, std::vector<uint32_t>& out, int lds_shift = 10) {
    const int RW = rec_words(p);
    out.assign((size_t)RW * ((size_t)p.nrec + 1), 0u);
    auto off = [&](uint32_t f) { return (f & 0x7fffu) << lds_shift; };
    for (int r = 0; r < p.nrec; ++r) {
        const uint32_t* h = p.rec_hdr.data() + 8 * (size_t)r;
        uint32_t* o = out.data() + (size_t)RW * r;
        const int d = h[0] & 3, j = (h[0] >> 2) & 7, nl = (h[0] >> 5) & 7;
        uint32_t ctl = 0;
        if (nl > 2) ctl |= REC_MORE;
        if (j <= d && j <= 1) ctl |= REC_U1;
        if (j <= d && j <= 2 && d >= 2) ctl |= REC_U2;
        if (j <= d && d >= 3) ctl |= REC_U3;
        const uint32_t lid[4] = {h[3] & 0xffffu, h[3] >> 16, h[4] & 0xffffu, h[4] >> 16};
        const uint32_t fid[3] = {h[1] & 0xffffu, h[1] >> 16, h[2] & 0xffffu};
        const bool cj = (lid[0] & 0x8000u) != 0;
        for (int i = 0; i < nl; ++i)
            if (((lid[i] & 0x8000u) != 0) != cj) return "record leaves mix conjugation";
        if (cj) ctl |= REC_CONJ;
        for (int t = 0; t < 3; ++t)
            if (fid[t] & 0x8000u) ctl |= REC_CONJ1 << t;
        o[0] = ctl;
        for (int i = 0; i < 4; ++i) o[1 + i] = i < nl ? off(lid[i]) : 0u;
        for (int t = 0; t < 3; ++t) o[5 + t] = off(fid[t]);
        if (p.classed) {
            double c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 4; ++i) {
                const int lf = p.rec_leaf[4 * (size_t)r + i];
                if (lf < 0) continue;
                c[2 * i] = cf.leaf[2 * (size_t)lf];
                c[2 * i + 1] = cj ? -cf.leaf[2 * (size_t)lf + 1] : cf.leaf[2 * (size_t)lf + 1];
            }
            std::memcpy(o + 8, c, sizeof(c));
        } else {
            double c[4] = {0, 0, 0, 0};
            for (int i = 0; i < 4; ++i) {
                const int lf = p.rec_leaf[4 * (size_t)r + i];
                if (lf >= 0) c[i] = cf.leaf[lf];
            }
            std::memcpy(o + 8, c, sizeof(c));
        }
    }
    return "";
}

// Leaves per record of the per-record kernel (and the record count that
// stands for a task's cost in the task lists): runs longer than this continue
// in further records with the same prefix.
constexpr int SEG_LEAVES = 4;
// Deepest prefix of a segment record (correlation order 8).
constexpr int SEG_MAXD = 7;
// Leaves per segment record: 8 for prefixes of at most 3 factors, 4 for
// deeper ones.
inline int seg_leaves(int maxorder) { return maxorder > 4 ? 4 : 8; }

// Records of task t when its runs are cut into records of at most SEG_LEAVES
// leaves (the task's cost in the task lists).
inline int task_records(const PackCopy& p, int t) {
    int n = 0;
    for (int r = p.tasks[5 * (size_t)t]; r < p.tasks[5 * (size_t)t + 1]; ++r)
        n += (std::max(p.runs[(size_t)ACE_RUN_W * r + 2], 1) + SEG_LEAVES - 1) / SEG_LEAVES;
    return n;
}

// The programs as kind-sorted segments of self-contained records (the
// segment kernel).  Every run is cut into records of at most maxleaves
// leaves; a record carries its whole prefix, so the records of a task can be
// grouped by kind = depth * 8 + leaves - 1 (stable within a kind), one
// segment per kind present.  Record words: flags (bit 31 conjugated leaves,
// bit 31 - t conjugated prefix factor t), the depth prefix offsets, the leaf
// offsets, padding to an even count, then the leaf coefficients (FP64, or
// complex for class-split programs).
//   segt[4 * s]     kind, record count, word offset of the first record, 0
//   seg_ptr[t]      the segments of task t: [seg_ptr[t], seg_ptr[t + 1])
struct Segments {
    std::vector<uint32_t> words, segt;
    std::vector<int32_t> seg_ptr;
};

// row_bytes: LDS bytes per pooled value (16 times the tile's environments).
inline std::string segments(const PackCopy& p, const Coeffs& cf, int row_bytes, int maxleaves, Segments& out) {
    if (maxleaves < 1 || maxleaves > 8) return "segment records hold 1 to 8 leaves";
    out.words.clear();
    out.segt.clear();
    out.seg_ptr.assign(1, 0);
    auto off = [&](uint32_t f) { return (f & 0x7fffu) * (uint32_t)row_bytes; };
    const int CW = p.classed ? 4 : 2;  // coefficient words per leaf
    constexpr int NK = 8 * (SEG_MAXD + 1);
    struct Rec {
        int run, q0, n;
    };
    for (int t = 0; t < p.ntasks; ++t) {
        std::vector<std::vector<Rec>> by_kind(NK);
        for (int r = p.tasks[5 * (size_t)t]; r < p.tasks[5 * (size_t)t + 1]; ++r) {
            const int32_t* run = p.runs.data() + (size_t)ACE_RUN_W * r;
            const int d = run[0], nl = run[2];
            if (d < 0 || d > SEG_MAXD) return "run prefix deeper than 7 factors";
            if (nl < 1) return "run without leaves";
            for (int q0 = 0; q0 < nl; q0 += maxleaves) {
                const int n = std::min(maxleaves, nl - q0);
                by_kind[d * 8 + n - 1].push_back({r, q0, n});
            }
        }
        for (int kind = 0; kind < NK; ++kind) {
            const std::vector<Rec>& rs = by_kind[kind];
            if (rs.empty()) continue;
            const int d = kind / 8, ns = kind % 8 + 1;
            const int hw = (2 + d + ns) & ~1, rw = hw + ns * CW;
            out.segt.insert(out.segt.end(), {(uint32_t)kind, (uint32_t)rs.size(), (uint32_t)out.words.size(), 0u});
            for (const Rec& rc : rs) {
                const int32_t* run = p.runs.data() + (size_t)ACE_RUN_W * rc.run;
                const int lo = run[3] + rc.q0;
                const bool cj = (p.leaf_fid[lo] & 0x8000) != 0;
                for (int i = 0; i < rc.n; ++i)
                    if (((p.leaf_fid[lo + i] & 0x8000) != 0) != cj) return "record leaves mix conjugation";
                std::vector<uint32_t> w((size_t)rw, 0u);
                uint32_t fl = cj ? 0x80000000u : 0u;
                for (int q = 0; q < d; ++q) {
                    const uint32_t f = (uint32_t)run[4 + q];
                    if (f & 0x8000u) fl |= 0x80000000u >> (q + 1);
                    w[1 + q] = off(f);
                }
                w[0] = fl;
                for (int i = 0; i < ns; ++i) w[1 + d + i] = off((uint32_t)p.leaf_fid[lo + i]);
                for (int i = 0; i < rc.n; ++i) {
                    const size_t lf = (size_t)lo + i;
                    double c[2];
                    if (p.classed) {
                        c[0] = cf.leaf[2 * lf];
                        c[1] = cj ? -cf.leaf[2 * lf + 1] : cf.leaf[2 * lf + 1];
                    } else {
                        c[0] = cf.leaf[lf];
                    }
                    std::memcpy(w.data() + hw + (size_t)CW * i, c, sizeof(double) * (CW / 2));
                }
                out.words.insert(out.words.end(), w.begin(), w.end());
            }
        }
        out.seg_ptr.push_back((int32_t)(out.segt.size() / 4));
    }
    // slack after the last record (the kernel never reads past a segment)
    out.words.resize(out.words.size() + 64, 0u);
    if (out.segt.empty()) out.segt.assign(4, 0u);
    return "";
}

// Task lists of the two kernel modes: every task (energy + forces), energy only.
inline std::vector<int> mode_tasks(const PackCopy& p, bool forces) {
    std::vector<int> out;
    for (int t = 0; t < p.ntasks; ++t)
        if (forces || p.tasks[5 * (size_t)t + 2] < 0) out.push_back(t);
    return out;
}

// Longest-first assignment of tasks to V virtual waves by record count (each
// to the least loaded wave), every list ascending: ptr[V + 1], list.
inline void lpt_lists(const PackCopy& p, const std::vector<int>& tasks, int V, std::vector<int32_t>& ptr,
                      std::vector<int32_t>& list) {
    auto cost = [&](int t) { return std::max(1, task_records(p, t)); };
    std::vector<int> order(tasks);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
    std::vector<long> load(V, 0);
    std::vector<std::vector<int>> lists(V);
    for (int t : order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[w].push_back(t);
        load[w] += cost(t);
    }
    ptr.assign(1, 0);
    list.clear();
    for (auto& l : lists) {
        std::sort(l.begin(), l.end());
        list.insert(list.end(), l.begin(), l.end());
        ptr.push_back((int32_t)list.size());
    }
}

// Tables of the per-environment basis-gradient kernel: basis function b's
// gradient is Re sum_k W+[b][k] dphi+[j][k] over the pooled (m >= 0)
// functions k it depends on, with W+ built per environment from the A2B
// entries of row b.  For entry z (AA column col, factors iA_t) and factor t,
// the product of the other factors times the A2B value goes to the slot of
// pooled function k(iA_t) in b's slot list, conjugated when iA_t is a
// conjugated (m < 0) function.
//   b_slot[nB + 1]      CSR of slots per basis function
//   slot_k[nslots]      pooled function of a slot
//   ztab[nnz][zw]       per A2B entry, one word per factor t < zw (zw = 4
//                       for correlation order <= 4, else 8): low half the
//                       factor (pooled k in bits 0-9, bit 13 present, bit 14
//                       negative sign, bit 15 conjugated), high half the slot
//                       index within b's list (bits 0-14) | conjugated << 15
//   chunk_b[nchunk + 1], chunk_s[nchunk + 1]: basis-function chunks of at
//   most max_b functions and max_slots slots (one workgroup each)
struct DbTables {
    std::vector<int32_t> b_slot, slot_k, chunk_b, chunk_s;
    std::vector<uint32_t> ztab;
    int zw = 4;
    int max_chunk_slots = 0;
};

inline DbTables db_tables(const ace_pack* pk, int max_b, int max_slots) {
    DbTables t;
    const int mo = pk->maxorder;
    t.zw = mo <= 4 ? 4 : 8;
    t.b_slot.assign((size_t)pk->nB + 1, 0);
    t.ztab.assign((size_t)pk->a2b_nnz * t.zw, 0u);
    std::vector<int> where((size_t)pk->npool, -1);
    for (int b = 0; b < pk->nB; ++b) {
        const int s0 = (int)t.slot_k.size();
        for (int z = pk->a2b_rowptr[b]; z < pk->a2b_rowptr[b + 1]; ++z) {
            const int col = pk->a2b_cols[z];
            for (int q = 0; q < pk->orders[col]; ++q) {
                const int iA = pk->iAA2iA[(size_t)col * mo + q];
                const int f = pk->aspec_fid[iA];
                const int k = f & 0x7fff;
                if (where[(size_t)k] < 0) {
                    where[(size_t)k] = (int)t.slot_k.size() - s0;
                    t.slot_k.push_back(k);
                }
                const uint32_t cj = (f & 0x8000) ? 1u : 0u;
                const uint32_t fc = (uint32_t)k | (1u << 13) | (pk->aspec_sgn[iA] < 0 ? 1u << 14 : 0u) | (cj << 15);
                const uint32_t sc = (uint32_t)where[(size_t)k] | (cj << 15);
                t.ztab[(size_t)z * t.zw + q] = fc | (sc << 16);
            }
        }
        for (size_t i = (size_t)s0; i < t.slot_k.size(); ++i) where[(size_t)t.slot_k[i]] = -1;
        t.b_slot[(size_t)b + 1] = (int32_t)t.slot_k.size();
    }
    t.chunk_b.push_back(0);
    t.chunk_s.push_back(0);
    for (int b = 0; b < pk->nB;) {
        int e = b;
        while (e < pk->nB && e - b < max_b && t.b_slot[(size_t)e + 1] - t.b_slot[(size_t)b] <= max_slots) ++e;
        if (e == b) e = b + 1;  // one function above the slot budget: a chunk of its own
        t.max_chunk_slots = std::max(t.max_chunk_slots, t.b_slot[(size_t)e] - t.b_slot[(size_t)b]);
        t.chunk_b.push_back(e);
        t.chunk_s.push_back(t.b_slot[(size_t)e]);
        b = e;
    }
    return t;
}

// (mu, l, m, n) -> pooled index, -1 when absent (force kernel lookup).
inline std::vector<int16_t> pool_lookup(const ace_pack* pk) {
    const int ns = std::max(1, (int)pk->nspecies);
    std::vector<int16_t> kx((size_t)ns * (pk->L + 1) * (pk->L + 1) * pk->nr, (int16_t)-1);
    for (int k = 0; k < pk->npool; ++k) {
        const int mu = std::max(0, (int)pk->pool[4 * k]), n = pk->pool[4 * k + 1] - 1, l = pk->pool[4 * k + 2],
                  m = pk->pool[4 * k + 3];
        kx[(((size_t)mu * (pk->L + 1) + l) * (pk->L + 1) + m) * pk->nr + n] = (int16_t)k;
    }
    return kx;
}

// 1p table as (mu, n, l, m) rows with signed m (basis-level gradient kernel).
inline std::vector<int32_t> aspec_rows(const ace_pack* pk) {
    std::vector<int32_t> a4((size_t)pk->nA * 4);
    for (int i = 0; i < pk->nA; ++i) {
        const int f = pk->aspec_fid[i], k = f & 0x7fff;
        a4[4 * i] = pk->pool[4 * k];
        a4[4 * i + 1] = pk->pool[4 * k + 1];
        a4[4 * i + 2] = pk->pool[4 * k + 2];
        a4[4 * i + 3] = (f & 0x8000) ? -pk->pool[4 * k + 3] : pk->pool[4 * k + 3];
    }
    return a4;
}

// Lane sets of the pooling kernel for kp pooled functions per lane: the
// pooled functions grouped by (species, l, m), each group cut into sets of at
// most kp, set g on lane g; slot[g + 64 p] is the set's p-th member (-1:
// none).  The kp functions of a lane then share one Y_lm.  Returns false (and
// leaves slot empty) when the sets need more than 64 lanes.
inline bool pool_lane_sets(const int32_t* pool, int npool, int kp, std::vector<int32_t>& slot) {
    slot.clear();
    if (kp < 1 || npool < 1) return false;
    std::vector<int> order(npool);
    for (int k = 0; k < npool; ++k) order[k] = k;
    auto key = [&](int k) { return std::make_tuple(pool[4 * k], pool[4 * k + 2], pool[4 * k + 3]); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
    std::vector<int32_t> out(64 * (size_t)kp, -1);
    int g = -1, fill = kp;
    for (int i = 0; i < npool; ++i) {
        const int k = order[i];
        if (fill == kp || key(k) != key(order[i - 1])) {
            if (++g >= 64) return false;
            fill = 0;
        }
        out[g + 64 * (size_t)fill++] = k;
    }
    slot.swap(out);
    return true;
}

}  // namespace acehost
