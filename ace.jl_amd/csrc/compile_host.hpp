// Host-side compiler of the device tables (plain C++, no HIP): from a basis
// as ACE.jl holds it (ace_jl_basis: Product1pBasis.indices, PIBasisSpec,
// A2Bmap in CSC, the radial OrthPolyBasis) to the flat ace_pack that
// ace_ctx_create uploads.
//
// What is compiled.  The site energy of a LinearACEModel is a polynomial in
// the 1p values, P(A) = sum_iAA c~[iAA] prod_t A[iAA2iA[iAA, t]]
// (evaluator.jl:121-147, c~ = A2B^T c).  With R_n real and
// Y_l^-m = (-1)^m conj(Y_l^m) (sphericalharmonics.jl:394-398) every 1p value
// is a signed, possibly conjugated, pooled m >= 0 value A+[k], so P is a
// polynomial in A+ and conj(A+).  The compiler
//   1. builds the pooled table A+ (sorted by degree n - 1 + l, then species,
//      l, m, n) and the map A -> (k, conj, sign);
//   2. collects the monomials of P with symbolic coefficients (AA row, scale,
//      conj) so that new parameters only need c~ and a gather;
//   3. merges every monomial with its mirror image (only Re P matters, and
//      Re(c x) = Re(conj(c) conj(x)));
//   4. differentiates: G[k] = dP/dA+[k] + conj(dP/dconj(A+[k])), the adjoint
//      that the force kernels contract with the 1p gradients
//      (evaluator.jl:150-200 restated on A+);
//   5. writes each polynomial as runs (a shared prefix product plus its leaf
//      factors), sorted so that consecutive runs share leading factors (the
//      depth-first order of the product DAG), splits them into tasks and
//      assigns the tasks to the waves of a tile (longest first);
//   6. decodes the runs into 64-byte records for the record kernel
//      (Invariant, order <= 4), and extracts the order <= 2 invariants.
// Every step is deterministic: iteration follows first insertion, sorting is
// stable with total keys, and ties in the task assignment go to the lowest
// wave, so the tables equal those of the Python host (acehip/pack.py,
// acehip/programs.py) entry for entry.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "pack_host.hpp"

namespace acehost {

constexpr int32_t DEV_CONJ = 0x8000;
constexpr int ENERGY_PARTS = 16;        // the energy polynomial is split into this many tasks
constexpr int REC_LEAVES = 4;           // leaves per 64-byte record
constexpr double MAX_TASK_COST = 1200;  // gradient polynomials with more monomials are split
constexpr int WAVE_CFGS[3] = {4, 8, 16};  // waves per workgroup of the product-DAG kernels
constexpr int MAX_FEATURES = 160;       // order <= 2 invariants the surrogate kernels take

struct Term {
    int32_t row;
    double scale;
    int32_t conj;
};
using Mono = std::vector<int32_t>;  // internal factor ids 2k + conj, ascending

// Polynomial with insertion-ordered monomials.
struct Poly {
    std::vector<Mono> keys;
    std::vector<std::vector<Term>> terms;
    std::map<Mono, size_t> index;
    size_t slot(const Mono& m) {
        auto it = index.find(m);
        if (it != index.end()) return it->second;
        index.emplace(m, keys.size());
        keys.push_back(m);
        terms.emplace_back();
        return keys.size() - 1;
    }
    void add(const Mono& m, const Term& t) { terms[slot(m)].push_back(t); }
    size_t size() const { return keys.size(); }
    bool empty() const { return keys.empty(); }
};

// Owner of every array an ace_pack points to.
struct PackStore {
    int prop = 0, L = 0, nspecies = 0, tp = 2, pl = 0, pr = 0, maxorder = 1, nB = 0, ngslots = 0;
    double r0 = 1.0, tl = 0.0, tr = 0.0;
    std::vector<double> rA, rB, rC;
    std::vector<int32_t> pool;  // [npool][4] (mu, n, l, m >= 0)
    std::vector<int32_t> aspec_fid;
    std::vector<double> aspec_sgn;
    std::vector<int32_t> orders, iAA2iA;  // iAA2iA [nAA][maxorder], 0-based, -1 padding
    std::vector<int32_t> a2b_rowptr, a2b_cols;
    std::vector<double> a2b_vals;
    std::vector<int32_t> runs, leaf_fid, leaf_term_ptr, tasks, task_term_ptr, term_row, term_conj;
    std::vector<double> term_scale;
    std::vector<int32_t> gslot_ptr, gslot_list, gslot_class;
    int classed = 0;
    std::vector<uint32_t> rec_hdr;
    std::vector<int32_t> rec_leaf;
    std::vector<int32_t> feat_rows, feat_ptr, feat_f, fgrad_ptr, fgrad_feat, fgrad_f;
    std::vector<double> feat_coef, fgrad_coef;
    std::vector<int32_t> wave_cfg_W, wave_task_ptr, wave_task_list, wave_etask_ptr, wave_etask_list;
    int32_t ntasks_energy = 0;

    ace_pack view() const {
        ace_pack p;
        std::memset(&p, 0, sizeof(p));
        p.abi_version = ACEHIP_ABI_VERSION;
        p.prop = prop;
        p.L = L;
        p.nr = (int32_t)rA.size();
        p.nspecies = nspecies;
        p.tp = tp;
        p.r0 = r0;
        p.tl = tl;
        p.tr = tr;
        p.pl = pl;
        p.pr = pr;
        p.rA = rA.data();
        p.rB = rB.data();
        p.rC = rC.data();
        p.npool = (int32_t)(pool.size() / 4);
        p.pool = pool.data();
        p.nA = (int32_t)aspec_fid.size();
        p.aspec_fid = aspec_fid.data();
        p.aspec_sgn = aspec_sgn.data();
        p.nAA = (int32_t)orders.size();
        p.maxorder = maxorder;
        p.orders = orders.data();
        p.iAA2iA = iAA2iA.data();
        p.nB = nB;
        p.a2b_nnz = (int32_t)a2b_cols.size();
        p.a2b_rowptr = a2b_rowptr.data();
        p.a2b_cols = a2b_cols.data();
        p.a2b_vals = a2b_vals.data();
        p.nruns = (int32_t)(runs.size() / ACE_RUN_W);
        p.runs = runs.data();
        p.nleaves = (int32_t)leaf_fid.size();
        p.leaf_fid = leaf_fid.data();
        p.leaf_term_ptr = leaf_term_ptr.data();
        p.ntasks = (int32_t)(tasks.size() / 5);
        p.tasks = tasks.data();
        p.task_term_ptr = task_term_ptr.data();
        p.nterms = (int32_t)term_row.size();
        p.term_row = term_row.data();
        p.term_scale = term_scale.data();
        p.term_conj = term_conj.data();
        p.ngslots = ngslots;
        p.gslot_ptr = gslot_ptr.data();
        p.gslot_list = gslot_list.data();
        p.classed = classed;
        p.gslot_class = gslot_class.data();
        p.nrec = (int32_t)(rec_hdr.size() / 8);
        p.rec_hdr = rec_hdr.data();
        p.rec_leaf = rec_leaf.data();
        p.nfeat = (int32_t)feat_rows.size();
        p.feat_rows = feat_rows.data();
        p.feat_ptr = feat_ptr.data();
        p.feat_f = feat_f.data();
        p.feat_coef = feat_coef.data();
        p.fgrad_ptr = fgrad_ptr.data();
        p.fgrad_feat = fgrad_feat.data();
        p.fgrad_f = fgrad_f.data();
        p.fgrad_coef = fgrad_coef.data();
        p.nwave_cfg = (int32_t)wave_cfg_W.size();
        p.wave_cfg_W = wave_cfg_W.data();
        p.wave_task_ptr = wave_task_ptr.data();
        p.wave_task_list = wave_task_list.data();
        p.wave_etask_ptr = wave_etask_ptr.data();
        p.wave_etask_list = wave_etask_list.data();
        return p;
    }
};

// ---- input checks of the ACE.jl tables (nothing is read out of bounds)

inline std::string validate_jl(const ace_jl_basis* b) {
    if (!b) return "null basis";
    if (b->abi_version != ACEHIP_ABI_VERSION) return "ABI version mismatch";
    if (b->prop != ACE_PROP_INVARIANT && b->prop != ACE_PROP_EUCLIDEAN_VECTOR) return "unknown property";
    if (b->nr < 1 || b->nr > 32 || !b->A || !b->B || !b->C) return "radial basis: need 1 to 32 functions";
    if (b->maxL < 0 || b->maxL > 10) return "maxL must be in [0, 10]";
    if (b->ncategories < 0 || b->ncategories > 127) return "ncategories out of range";
    if (b->tp < 1 || b->tp > 16 || b->pl < 0 || b->pr < 0 || b->pl > 16 || b->pr > 16)
        return "transform or envelope powers out of range";
    if (b->ncomp < 2 || b->ncomp > 3 || !b->comp_kind) return "Product1pBasis: need Rn, Ylm and optionally a Categorical1pBasis";
    int seen[3] = {0, 0, 0};
    for (int c = 0; c < b->ncomp; ++c) {
        if (b->comp_kind[c] < 0 || b->comp_kind[c] > 2) return "unknown 1p component kind";
        seen[b->comp_kind[c]]++;
    }
    if (seen[ACE_COMP_RN] != 1 || seen[ACE_COMP_YLM] != 1 || seen[ACE_COMP_CATEGORICAL] > 1)
        return "Product1pBasis: need exactly one Rn and one Ylm component";
    if ((seen[ACE_COMP_CATEGORICAL] == 1) != (b->ncategories > 0))
        return "Categorical1pBasis component and ncategories disagree";
    if (b->nA < 1 || b->nA > (1 << 20) || !b->indices) return "1p basis size out of range";
    if (b->nAA < 1 || b->nAA > (1 << 24) || !b->orders || !b->iAA2iA) return "AA basis size out of range";
    if (b->maxorder < 1 || b->maxorder > ACE_MAX_ORDER) return "correlation order must be in [1, 8]";
    if (b->nB < 0 || b->nB > (1 << 24) || !b->colptr || (b->nB > 0 && (!b->rowval || !b->nzval)))
        return "A2B map size out of range";
    for (int64_t i = 0; i < b->nA; ++i)
        for (int c = 0; c < b->ncomp; ++c) {
            const int64_t v = b->indices[i * b->ncomp + c];
            switch (b->comp_kind[c]) {
                case ACE_COMP_CATEGORICAL:
                    if (v < 1 || v > b->ncategories) return "Categorical1pBasis index out of range";
                    break;
                case ACE_COMP_RN:
                    if (v < 1 || v > b->nr) return "Rn index out of range";
                    break;
                default:
                    if (v < 1 || v > (b->maxL + 1) * (b->maxL + 1)) return "Ylm index out of range";
            }
        }
    for (int64_t i = 0; i < b->nAA; ++i) {
        const int64_t o = b->orders[i];
        if (o < 0 || o > b->maxorder) return "orders out of range";
        for (int64_t t = 0; t < b->maxorder; ++t) {
            const int64_t v = b->iAA2iA[t * b->nAA + i];
            if (t < o ? (v < 1 || v > b->nA) : (v != 0)) return "iAA2iA entry out of range";
        }
    }
    if (b->colptr[0] != 1) return "A2B colptr must start at 1";
    for (int64_t j = 0; j < b->nAA; ++j)
        if (b->colptr[j + 1] < b->colptr[j]) return "A2B colptr must be non-decreasing";
    const int64_t nnz = b->colptr[b->nAA] - 1;
    if (nnz > (int64_t)1 << 28) return "A2B map too large";
    for (int64_t j = 0; j < b->nAA; ++j)
        for (int64_t z = b->colptr[j] - 1; z < b->colptr[j + 1] - 1; ++z) {
            if (b->rowval[z] < 1 || b->rowval[z] > b->nB) return "A2B row index out of range";
            if (z > b->colptr[j] - 1 && b->rowval[z] <= b->rowval[z - 1]) return "A2B rows must ascend within a column";
        }
    return "";
}

// ---- the compiler

namespace detail {

inline int32_t conj_id(int32_t x, const std::vector<char>& is_m0) {
    const int32_t k = x >> 1;
    return is_m0[k] ? 2 * k : (x ^ 1);
}

inline int32_t dev_id(int32_t f) { return (f >> 1) | ((f & 1) ? DEV_CONJ : 0); }

// Split into nparts with balanced cost, keeping prefix groups together:
// groups by decreasing size (then prefix), each to the least loaded part.
inline std::vector<Poly> split_poly(const Poly& poly, int nparts) {
    std::map<Mono, std::vector<size_t>> groups;
    std::vector<Mono> order;  // prefixes in first-insertion order
    for (size_t i = 0; i < poly.size(); ++i) {
        const Mono& m = poly.keys[i];
        const Mono pre = m.empty() ? Mono() : Mono(m.begin(), m.end() - 1);  // a constant joins the empty prefix
        auto it = groups.find(pre);
        if (it == groups.end()) {
            groups.emplace(pre, std::vector<size_t>{i});
            order.push_back(pre);
        } else {
            it->second.push_back(i);
        }
    }
    std::vector<const Mono*> keys;
    for (auto& kv : groups) keys.push_back(&kv.first);
    std::stable_sort(keys.begin(), keys.end(), [&](const Mono* a, const Mono* b) {
        const size_t la = groups[*a].size(), lb = groups[*b].size();
        if (la != lb) return la > lb;
        return *a < *b;
    });
    std::vector<Poly> parts((size_t)nparts);
    std::vector<double> load((size_t)nparts, 0.0);
    for (const Mono* p : keys) {
        const size_t i = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        const auto& g = groups[*p];
        for (size_t mi : g) {
            const size_t s = parts[i].slot(poly.keys[mi]);
            parts[i].terms[s] = poly.terms[mi];
        }
        load[i] += (double)g.size() + 2.0;
    }
    std::vector<Poly> out;
    for (auto& p : parts)
        if (!p.empty()) out.push_back(std::move(p));
    return out;
}

struct Program {
    std::vector<int32_t> runs;  // [nruns][ACE_RUN_W]
    std::vector<int32_t> leaf_fid;
    std::vector<std::vector<Term>> leaf_terms;
    std::vector<int32_t> tasks;  // [ntasks][3] run_begin, run_end, out_slot
    std::vector<std::vector<Term>> task_terms;

    // One task: the runs of a polynomial (monomials grouped by their prefix,
    // prefixes in lexicographic order, leaves by factor id).
    std::string add_task(const Poly& poly, int out_slot, int& task_id) {
        std::vector<Term> consts;
        std::map<Mono, std::vector<std::pair<int32_t, size_t>>> groups;
        for (size_t i = 0; i < poly.size(); ++i) {
            const Mono& m = poly.keys[i];
            if (m.empty()) {
                consts.insert(consts.end(), poly.terms[i].begin(), poly.terms[i].end());
                continue;
            }
            if ((int)m.size() > ACE_MAX_ORDER) return "correlation order above 8";
            groups[Mono(m.begin(), m.end() - 1)].emplace_back(m.back(), i);
        }
        const int32_t r0 = (int32_t)(runs.size() / ACE_RUN_W);
        const Mono* prev = nullptr;
        for (auto& kv : groups) {
            const Mono& pre = kv.first;
            auto leaves = kv.second;
            std::stable_sort(leaves.begin(), leaves.end(),
                             [](const std::pair<int32_t, size_t>& a, const std::pair<int32_t, size_t>& b) {
                                 return a.first < b.first;
                             });
            const int d = (int)pre.size();
            int j = 1;
            while (prev && j <= d && j <= (int)prev->size() && pre[j - 1] == (*prev)[j - 1]) ++j;
            int32_t run[ACE_RUN_W] = {0};
            run[0] = d;
            run[1] = j;
            run[2] = (int32_t)leaves.size();
            run[3] = (int32_t)leaf_fid.size();
            for (int q = 0; q < d; ++q) run[4 + q] = dev_id(pre[q]);
            runs.insert(runs.end(), run, run + ACE_RUN_W);
            for (auto& lf : leaves) {
                leaf_fid.push_back(dev_id(lf.first));
                leaf_terms.push_back(poly.terms[lf.second]);
            }
            prev = &pre;
        }
        tasks.push_back(r0);
        tasks.push_back((int32_t)(runs.size() / ACE_RUN_W));
        tasks.push_back(out_slot);
        task_terms.push_back(consts);
        task_id = (int)task_terms.size() - 1;
        return "";
    }

    double cost(int t) const {
        double leaves = 0.0;
        for (int r = tasks[3 * t]; r < tasks[3 * t + 1]; ++r) leaves += runs[(size_t)r * ACE_RUN_W + 2];
        return leaves + 2.0 * (tasks[3 * t + 1] - tasks[3 * t]);
    }
};

inline int term_class(const Term& t, const std::vector<int8_t>& row_class) {
    const int c = row_class[(size_t)t.row];
    return t.conj ? 2 - c : c;
}

// The polynomial as three polynomials, one per term class.
inline std::vector<Poly> split_by_class(const Poly& poly, const std::vector<int8_t>& row_class) {
    std::vector<Poly> out(3);
    for (size_t i = 0; i < poly.size(); ++i)
        for (const Term& t : poly.terms[i]) out[(size_t)term_class(t, row_class)].add(poly.keys[i], t);
    return out;
}

// Longest-first assignment of tasks to W waves (ties to the lowest wave);
// every list ascending.
inline std::vector<std::vector<int>> lpt_assign(const std::vector<double>& costs, int W) {
    std::vector<int> order(costs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return -costs[a] < -costs[b]; });
    std::vector<double> load((size_t)W, 0.0);
    std::vector<std::vector<int>> lists((size_t)W);
    for (int t : order) {
        const size_t w = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[w].push_back(t);
        load[w] += costs[t];
    }
    for (auto& l : lists) std::sort(l.begin(), l.end());
    return lists;
}

}  // namespace detail

inline std::string compile_jl(const ace_jl_basis* b, PackStore& ps) {
    using namespace detail;
    std::string err = validate_jl(b);
    if (!err.empty()) return err;
    ps = PackStore();
    ps.prop = b->prop;
    ps.L = (int)b->maxL;
    ps.nspecies = (int)b->ncategories;
    ps.tp = (int)b->tp;
    ps.pl = (int)b->pl;
    ps.pr = (int)b->pr;
    ps.r0 = b->r0;
    ps.tl = b->tl;
    ps.tr = b->tr;
    ps.rA.assign(b->A, b->A + b->nr);
    ps.rB.assign(b->B, b->B + b->nr);
    ps.rC.assign(b->C, b->C + b->nr);
    const int nA = (int)b->nA, nAA = (int)b->nAA, mo = (int)b->maxorder;
    // 1. the 1p spec (mu, n, l, m) and the pooled m >= 0 table
    std::vector<std::array<int, 4>> aspec((size_t)nA);
    for (int i = 0; i < nA; ++i) {
        int mu = -1, n = 0, l = 0, m = 0;
        for (int c = 0; c < b->ncomp; ++c) {
            const int64_t v = b->indices[(int64_t)i * b->ncomp + c];
            if (b->comp_kind[c] == ACE_COMP_CATEGORICAL) {
                mu = (int)v - 1;
            } else if (b->comp_kind[c] == ACE_COMP_RN) {
                n = (int)v;
            } else {
                const int iy = (int)v - 1;  // index_y(l, m) - 1 = l^2 + l + m
                l = (int)std::floor(std::sqrt((double)iy));
                while (l * l > iy) --l;
                while ((l + 1) * (l + 1) <= iy) ++l;
                m = iy - l * l - l;
            }
        }
        aspec[(size_t)i] = {mu, n, l, m};
    }
    std::map<std::array<int, 5>, int> keyset;  // (degree, mu, l, |m|, n) -> 0
    for (auto& a : aspec) keyset[{a[1] - 1 + a[2], a[0], a[2], std::abs(a[3]), a[1]}] = 0;
    if (keyset.size() > 512) return "pooled 1p basis too large (more than 512 functions with m >= 0)";
    std::map<std::array<int, 4>, int> kindex;
    for (auto& kv : keyset) {
        const int k = (int)kindex.size();
        const auto& q = kv.first;
        kindex[{q[1], q[4], q[2], q[3]}] = k;
        ps.pool.insert(ps.pool.end(), {q[1], q[4], q[2], q[3]});
    }
    const int npool = (int)kindex.size();
    std::vector<int> kk((size_t)nA), cj((size_t)nA);
    std::vector<double> sgn((size_t)nA, 1.0);
    for (int i = 0; i < nA; ++i) {
        const auto& a = aspec[(size_t)i];
        kk[(size_t)i] = kindex[{a[0], a[1], a[2], std::abs(a[3])}];
        if (a[3] < 0) {
            cj[(size_t)i] = 1;
            sgn[(size_t)i] = (a[3] % 2) ? -1.0 : 1.0;
        }
        ps.aspec_fid.push_back(kk[(size_t)i] | (cj[(size_t)i] ? DEV_CONJ : 0));
        ps.aspec_sgn.push_back(sgn[(size_t)i]);
    }
    std::vector<char> is_m0((size_t)npool);
    for (int k = 0; k < npool; ++k) is_m0[(size_t)k] = ps.pool[4 * (size_t)k + 3] == 0;
    // AA table (0-based, -1 padding) and A2B in CSR (rows ascending, columns ascending)
    ps.maxorder = mo;
    ps.orders.resize((size_t)nAA);
    ps.iAA2iA.assign((size_t)nAA * mo, -1);
    for (int i = 0; i < nAA; ++i) {
        ps.orders[(size_t)i] = (int32_t)b->orders[i];
        for (int t = 0; t < b->orders[i]; ++t)
            ps.iAA2iA[(size_t)i * mo + t] = (int32_t)(b->iAA2iA[(int64_t)t * nAA + i] - 1);
    }
    const int nc = b->prop == ACE_PROP_INVARIANT ? 1 : 6;
    const int64_t nnz = b->colptr[nAA] - 1;
    ps.nB = (int)b->nB;
    ps.a2b_rowptr.assign((size_t)ps.nB + 1, 0);
    for (int64_t z = 0; z < nnz; ++z) ps.a2b_rowptr[(size_t)b->rowval[z]]++;
    for (int r = 0; r < ps.nB; ++r) ps.a2b_rowptr[(size_t)r + 1] += ps.a2b_rowptr[(size_t)r];
    ps.a2b_cols.assign((size_t)nnz, 0);
    ps.a2b_vals.assign((size_t)nnz * nc, 0.0);
    {
        std::vector<int32_t> fill(ps.a2b_rowptr.begin(), ps.a2b_rowptr.end() - 1);
        for (int j = 0; j < nAA; ++j)
            for (int64_t z = b->colptr[j] - 1; z < b->colptr[j + 1] - 1; ++z) {
                const size_t dst = (size_t)fill[(size_t)(b->rowval[z] - 1)]++;
                ps.a2b_cols[dst] = j;
                for (int q = 0; q < nc; ++q) ps.a2b_vals[dst * nc + q] = b->nzval[z * nc + q];
            }
    }
    // 2. monomials of P with symbolic coefficients (row, scale, conj)
    Poly poly;
    for (int i = 0; i < nAA; ++i) {
        const int o = ps.orders[(size_t)i];
        if (o == 0) continue;  // the constant: e0 of the coefficient upload
        Mono mono;
        double s = 1.0;
        for (int t = 0; t < o; ++t) {
            const int a = ps.iAA2iA[(size_t)i * mo + t];
            s *= sgn[(size_t)a];
            mono.push_back(2 * kk[(size_t)a] + cj[(size_t)a]);
        }
        std::sort(mono.begin(), mono.end());
        poly.add(mono, Term{i, s, 0});
    }
    // 3. mirror merging
    Poly merged;
    for (size_t i = 0; i < poly.size(); ++i) {
        const Mono& m = poly.keys[i];
        Mono c;
        for (int32_t f : m) c.push_back(conj_id(f, is_m0));
        std::sort(c.begin(), c.end());
        if (c < m) {
            const size_t s = merged.slot(c);
            for (const Term& t : poly.terms[i]) merged.terms[s].push_back(Term{t.row, t.scale, 1 - t.conj});
        } else {
            const size_t s = merged.slot(m);
            merged.terms[s].insert(merged.terms[s].end(), poly.terms[i].begin(), poly.terms[i].end());
        }
    }
    // 4. adjoint polynomials G[k]
    std::vector<Poly> G((size_t)npool);
    for (size_t i = 0; i < merged.size(); ++i) {
        const Mono& m = merged.keys[i];
        for (size_t p = 0; p < m.size(); ++p) {
            if (p > 0 && m[p] == m[p - 1]) continue;  // each distinct factor once
            const int32_t f = m[p];
            const double mult = (double)std::count(m.begin(), m.end(), f);
            Mono rest(m);
            rest.erase(std::find(rest.begin(), rest.end(), f));
            const int32_t k = f >> 1;
            if ((f & 1) == 0) {
                const size_t s = G[(size_t)k].slot(rest);
                for (const Term& t : merged.terms[i]) G[(size_t)k].terms[s].push_back(Term{t.row, t.scale * mult, t.conj});
            } else {
                Mono key;
                for (int32_t x : rest) key.push_back(conj_id(x, is_m0));
                std::sort(key.begin(), key.end());
                const size_t s = G[(size_t)k].slot(key);
                for (const Term& t : merged.terms[i])
                    G[(size_t)k].terms[s].push_back(Term{t.row, t.scale * mult, 1 - t.conj});
            }
        }
    }
    // 5. tasks: the energy in ENERGY_PARTS parts, then each G[k] (split when
    // long).  EuclideanVector maps that allow it are split by class: the
    // energy into one task group per class polynomial Q_M (output slot
    // -1 - class), each G[k] into its class parts (slot class in gslot_class).
    const ace_pack v0 = ps.view();
    const std::vector<int8_t> row_class = pack_row_classes(&v0);
    ps.classed = (b->prop == ACE_PROP_EUCLIDEAN_VECTOR &&
                  class_vectors(ps.a2b_cols.data(), ps.a2b_vals.data(), ps.a2b_cols.size(), row_class.data()).ok)
                     ? 1
                     : 0;
    Program prog;
    std::vector<int> energy_tasks;
    {
        std::vector<std::pair<int, Poly>> groups;
        if (ps.classed) {
            std::vector<Poly> q = split_by_class(merged, row_class);
            for (int c = 0; c < 3; ++c)
                if (!q[(size_t)c].empty()) groups.emplace_back(-1 - c, std::move(q[(size_t)c]));
        } else {
            groups.emplace_back(-1, merged);
        }
        for (auto& gr : groups)
            for (const Poly& part : split_poly(gr.second, ENERGY_PARTS)) {
                int t = 0;
                err = prog.add_task(part, gr.first, t);
                if (!err.empty()) return err;
                energy_tasks.push_back(t);
            }
    }
    std::vector<std::vector<int32_t>> gslots((size_t)npool);
    int nslot = 0;
    for (int k = 0; k < npool; ++k) {
        const Poly& g = G[(size_t)k];
        if (g.empty()) continue;
        std::vector<Poly> byc;
        if (ps.classed)
            byc = split_by_class(g, row_class);
        else
            byc.push_back(g);
        for (size_t c = 0; c < byc.size(); ++c) {
            const Poly& gc = byc[c];
            if (gc.empty()) continue;
            const int nparts = std::max(1, (int)std::ceil((double)gc.size() / MAX_TASK_COST));
            std::vector<Poly> parts;
            if (nparts > 1)
                parts = split_poly(gc, nparts);
            else
                parts.push_back(gc);
            for (const Poly& part : parts) {
                int t = 0;
                err = prog.add_task(part, nslot, t);
                if (!err.empty()) return err;
                gslots[(size_t)k].push_back(nslot);
                if (ps.classed) ps.gslot_class.push_back((int32_t)c);
                ++nslot;
            }
        }
    }
    ps.ngslots = nslot;
    ps.ntasks_energy = (int32_t)energy_tasks.size();
    ps.gslot_ptr.assign(1, 0);
    for (auto& s : gslots) {
        ps.gslot_list.insert(ps.gslot_list.end(), s.begin(), s.end());
        ps.gslot_ptr.push_back((int32_t)ps.gslot_list.size());
    }
    ps.runs = prog.runs;
    ps.leaf_fid = prog.leaf_fid;
    // 6. records (Invariant and class-split EuclideanVector bases of order <= 4
    // run the record kernel)
    const int ntasks = (int)prog.task_terms.size();
    const bool records = mo <= 4 && (b->prop == ACE_PROP_INVARIANT || ps.classed);
    for (int t = 0; t < ntasks; ++t) {
        const int32_t rb = (int32_t)(ps.rec_hdr.size() / 8);
        for (int r = prog.tasks[3 * (size_t)t]; records && r < prog.tasks[3 * (size_t)t + 1]; ++r) {
            const int32_t* run = prog.runs.data() + (size_t)r * ACE_RUN_W;
            const int d = run[0], j = run[1], nl = run[2], lo = run[3];
            for (int q0 = 0; q0 < std::max(nl, 1); q0 += REC_LEAVES) {
                const int n = std::min(REC_LEAVES, nl - q0);
                const int jj = q0 == 0 ? j : d + 1;
                uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                h[0] = (uint32_t)(d | (jj << 2) | (n << 5));
                h[1] = (uint32_t)run[4] | ((uint32_t)run[5] << 16);
                h[2] = (uint32_t)run[6];
                uint32_t ids[REC_LEAVES];
                for (int i = 0; i < REC_LEAVES; ++i) ids[i] = i < n ? (uint32_t)prog.leaf_fid[(size_t)(lo + q0 + i)] : 0u;
                for (int i = 0; i < REC_LEAVES; i += 2) h[3 + i / 2] = ids[i] | (ids[i + 1] << 16);
                ps.rec_hdr.insert(ps.rec_hdr.end(), h, h + 8);
                for (int i = 0; i < REC_LEAVES; ++i) ps.rec_leaf.push_back(i < n ? lo + q0 + i : -1);
            }
        }
        ps.tasks.insert(ps.tasks.end(), {prog.tasks[3 * (size_t)t], prog.tasks[3 * (size_t)t + 1],
                                         prog.tasks[3 * (size_t)t + 2], rb, (int32_t)(ps.rec_hdr.size() / 8)});
    }
    // coefficient terms: the leaves', then the tasks' constants, in the same arrays
    ps.leaf_term_ptr.assign(1, 0);
    for (auto& ts : prog.leaf_terms) {
        for (const Term& t : ts) {
            ps.term_row.push_back(t.row);
            ps.term_scale.push_back(t.scale);
            ps.term_conj.push_back(t.conj);
        }
        ps.leaf_term_ptr.push_back((int32_t)ps.term_row.size());
    }
    ps.task_term_ptr.assign(1, (int32_t)ps.term_row.size());
    for (auto& ts : prog.task_terms) {
        for (const Term& t : ts) {
            ps.term_row.push_back(t.row);
            ps.term_scale.push_back(t.scale);
            ps.term_conj.push_back(t.conj);
        }
        ps.task_term_ptr.push_back((int32_t)ps.term_row.size());
    }
    // task -> wave lists per tile shape: all tasks, and the energy tasks alone
    std::vector<double> costs((size_t)ntasks);
    for (int t = 0; t < ntasks; ++t) costs[(size_t)t] = prog.cost(t);
    std::vector<double> ecost;
    for (int t : energy_tasks) ecost.push_back(costs[(size_t)t]);
    for (int W : WAVE_CFGS) {
        ps.wave_cfg_W.push_back(W);
        auto lists = lpt_assign(costs, W);
        ps.wave_task_ptr.push_back(0);
        int acc = 0;
        for (auto& l : lists) {
            acc += (int)l.size();
            ps.wave_task_ptr.push_back(acc);
            ps.wave_task_list.insert(ps.wave_task_list.end(), l.begin(), l.end());
        }
        auto elists = lpt_assign(ecost, W);
        ps.wave_etask_ptr.push_back(0);
        acc = 0;
        for (auto& l : elists) {
            acc += (int)l.size();
            ps.wave_etask_ptr.push_back(acc);
            for (int i : l) ps.wave_etask_list.push_back(energy_tasks[(size_t)i]);
        }
    }
    // 7. order 1 and 2 invariants (surrogate inputs): the B rows whose AA
    // columns have order 1 or 2; skipped when there are more than the
    // surrogate kernels take
    ps.fgrad_ptr.assign((size_t)npool + 1, 0);
    ps.feat_ptr.assign(1, 0);
    if (b->prop == ACE_PROP_INVARIANT) {
        std::vector<int32_t> rows;
        for (int r = 0; r < ps.nB; ++r) {
            int maxo = 0;
            for (int z = ps.a2b_rowptr[(size_t)r]; z < ps.a2b_rowptr[(size_t)r + 1]; ++z)
                maxo = std::max(maxo, (int)ps.orders[(size_t)ps.a2b_cols[(size_t)z]]);
            if (maxo >= 1 && maxo <= 2) rows.push_back(r);
        }
        if ((int)rows.size() <= MAX_FEATURES) {
            std::vector<std::vector<std::array<double, 3>>> grad((size_t)npool);  // (feature, other, coef)
            for (size_t fi = 0; fi < rows.size(); ++fi) {
                const int r = rows[fi];
                ps.feat_rows.push_back(r);
                for (int z = ps.a2b_rowptr[(size_t)r]; z < ps.a2b_rowptr[(size_t)r + 1]; ++z) {
                    const int col = ps.a2b_cols[(size_t)z];
                    const int od = ps.orders[(size_t)col];
                    double s = 1.0;
                    int32_t f[2] = {-1, -1};
                    for (int t = 0; t < od && t < 2; ++t) {
                        const int a = ps.iAA2iA[(size_t)col * mo + t];
                        s *= sgn[(size_t)a];
                        f[t] = ps.aspec_fid[(size_t)a];
                    }
                    const double c = ps.a2b_vals[(size_t)z] * s;
                    ps.feat_f.push_back(f[0]);
                    ps.feat_f.push_back(f[1]);
                    ps.feat_coef.push_back(c);
                    if (od == 1) {
                        grad[(size_t)(f[0] & 0x7fff)].push_back({(double)fi, -1.0, c});
                    } else if (od == 2) {
                        for (int a_ = 0; a_ < 2; ++a_) {
                            const int32_t fa = f[a_], fb = f[1 - a_];
                            const int k = fa & 0x7fff;
                            int32_t other = fb;
                            if ((fa & DEV_CONJ) && !is_m0[(size_t)(fb & 0x7fff)]) other = fb ^ DEV_CONJ;
                            grad[(size_t)k].push_back({(double)fi, (double)other, c});
                        }
                    }
                }
                ps.feat_ptr.push_back((int32_t)ps.feat_coef.size());
            }
            for (int k = 0; k < npool; ++k) {
                for (auto& e : grad[(size_t)k]) {
                    ps.fgrad_feat.push_back((int32_t)e[0]);
                    ps.fgrad_f.push_back((int32_t)e[1]);
                    ps.fgrad_coef.push_back(e[2]);
                }
                ps.fgrad_ptr[(size_t)k + 1] = (int32_t)ps.fgrad_feat.size();
            }
        }
    }
    return "";
}

}  // namespace acehost
