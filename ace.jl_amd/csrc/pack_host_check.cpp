// CPU test shim over pack_host.hpp: the host-side pack logic of libacehip.so
// built without HIP (tests build it with the address sanitizer).
#include <cstring>

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
