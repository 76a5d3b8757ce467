// CPU test shim over pack_host.hpp: the host-side pack logic of libacehip.so
// built without HIP (tests build it with the address sanitizer).
#include <cstring>

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
