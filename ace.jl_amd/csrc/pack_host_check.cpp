// CPU test shim over pack_host.hpp and compile_host.hpp: the host-side pack
// logic and the program compiler of libacehip.so built without HIP (tests
// build it with the address sanitizer, and plain for the larger bases).
#include <cstring>

#include "compile_host.hpp"
#include "nbrlist_host.hpp"
#include "pack_host.hpp"

extern "C" int acehost_validate(const ace_pack* pk, char* msg, int len) {
    std::string e = acehost::validate_pack(pk);
    if (e.empty()) {
        std::vector<acehost::WaveTable> wt;
        e = acehost::wave_tables(pk, wt);
    }
    if (msg && len > 0) {
        std::strncpy(msg, e.c_str(), (size_t)len - 1);
        msg[len - 1] = 0;
    }
    return e.empty() ? 0 : -1;
}

extern "C" int acehost_coeffs(const ace_pack* pk, const double* c, double* leaf, double* task, double* e0) {
    if (!acehost::validate_pack(pk).empty()) return -1;
    acehost::PackCopy p = acehost::copy_pack(pk);
    acehost::Coeffs cf = acehost::coefficients(p, c);
    std::memcpy(leaf, cf.leaf.data(), cf.leaf.size() * sizeof(double));
    std::memcpy(task, cf.task.data(), cf.task.size() * sizeof(double));
    std::memcpy(e0, cf.e0, 6 * sizeof(double));
    return 0;
}

extern "C" int acehost_lookup(const ace_pack* pk, short* kidx, int* a4) {
    if (!acehost::validate_pack(pk).empty()) return -1;
    std::vector<int16_t> kx = acehost::pool_lookup(pk);
    std::memcpy(kidx, kx.data(), kx.size() * sizeof(int16_t));
    std::vector<int32_t> a = acehost::aspec_rows(pk);
    std::memcpy(a4, a.data(), a.size() * sizeof(int32_t));
    return 0;
}

// Lane sets of the pooling kernel: writes 64 kp slots, returns 1 when the
// sets fit 64 lanes, else 0.
extern "C" int acehost_pool_sets(const int32_t* pool, int npool, int kp, int32_t* slot) {
    std::vector<int32_t> v;
    if (!acehost::pool_lane_sets(pool, npool, kp, v)) return 0;
    std::memcpy(slot, v.data(), v.size() * sizeof(int32_t));
    return 1;
}

// Neighbour-list binning (nbrlist_host.hpp).  bin_start must hold
// 2 * natoms + 66 entries (the bin count never exceeds 2 natoms + 64).
extern "C" int acehost_nbr_grid(int64_t natoms, const double* pos, const double* cell, const int32_t* pbc,
                                double rcut, int* nb, int* R, int32_t* bin, int32_t* bin_start, int32_t* bin_atoms,
                                int32_t* wrap, double* pw, char* msg, int len) {
    acehost::NbrGrid g;
    std::string e = acehost::nbr_grid(natoms, pos, cell, pbc, rcut, g);
    if (msg && len > 0) {
        std::strncpy(msg, e.c_str(), (size_t)len - 1);
        msg[len - 1] = 0;
    }
    if (!e.empty()) return -1;
    for (int d = 0; d < 3; ++d) {
        nb[d] = g.nb[d];
        R[d] = g.R[d];
    }
    std::memcpy(bin, g.bin.data(), g.bin.size() * sizeof(int32_t));
    std::memcpy(bin_start, g.bin_start.data(), g.bin_start.size() * sizeof(int32_t));
    std::memcpy(bin_atoms, g.bin_atoms.data(), g.bin_atoms.size() * sizeof(int32_t));
    std::memcpy(wrap, g.wrap.data(), g.wrap.size() * sizeof(int32_t));
    std::memcpy(pw, g.p.data(), g.p.size() * sizeof(double));
    return 0;
}

// The program compiler (compile_host.hpp): compiles an ace_jl_basis, keeps the
// tables, and hands out any of them by name for comparison with the Python
// host's pack.  acehost_compiled_array returns the element count (copying
// into out when it is not null) or -1 for an unknown name.
extern "C" void* acehost_compile(const ace_jl_basis* b, char* msg, int len) {
    auto* ps = new acehost::PackStore();
    std::string e = acehost::compile_jl(b, *ps);
    if (e.empty()) {
        const ace_pack pk = ps->view();
        e = acehost::validate_pack(&pk);
    }
    if (msg && len > 0) {
        std::strncpy(msg, e.c_str(), (size_t)len - 1);
        msg[len - 1] = 0;
    }
    if (!e.empty()) {
        delete ps;
        return nullptr;
    }
    return ps;
}

extern "C" void acehost_compiled_free(void* p) { delete static_cast<acehost::PackStore*>(p); }

extern "C" int64_t acehost_compiled_array(void* p, const char* name, void* out) {
    const auto* ps = static_cast<const acehost::PackStore*>(p);
    auto put = [&](const auto& v) -> int64_t {
        if (out && !v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
        return (int64_t)v.size();
    };
    const std::string n(name);
#define ACE_ARR(F) \
    if (n == #F) return put(ps->F);
    ACE_ARR(rA) ACE_ARR(rB) ACE_ARR(rC) ACE_ARR(pool) ACE_ARR(aspec_fid) ACE_ARR(aspec_sgn) ACE_ARR(orders)
    ACE_ARR(iAA2iA) ACE_ARR(a2b_rowptr) ACE_ARR(a2b_cols) ACE_ARR(a2b_vals) ACE_ARR(runs) ACE_ARR(leaf_fid)
    ACE_ARR(leaf_term_ptr) ACE_ARR(tasks) ACE_ARR(task_term_ptr) ACE_ARR(term_row) ACE_ARR(term_scale)
    ACE_ARR(term_conj) ACE_ARR(gslot_ptr) ACE_ARR(gslot_list) ACE_ARR(gslot_class) ACE_ARR(rec_hdr) ACE_ARR(rec_leaf) ACE_ARR(feat_rows)
    ACE_ARR(feat_ptr) ACE_ARR(feat_f) ACE_ARR(feat_coef) ACE_ARR(fgrad_ptr) ACE_ARR(fgrad_feat) ACE_ARR(fgrad_f)
    ACE_ARR(fgrad_coef) ACE_ARR(wave_cfg_W) ACE_ARR(wave_task_ptr) ACE_ARR(wave_task_list) ACE_ARR(wave_etask_ptr)
    ACE_ARR(wave_etask_list)
#undef ACE_ARR
    return -1;
}
