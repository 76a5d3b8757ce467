// Host side of the neighbour-list builder (no HIP): cell geometry, wrapping
// of periodic coordinates, binning and the bin search ranges.  Shared by
// libacehip.so and the CPU test shim.
//
// Conventions.  cell[9] holds the lattice vectors a, b, c as rows; a
// position is x = s . cell with fractional coordinates s.  Along a periodic
// axis s is wrapped into [0, 1) (wrapped position p = x - w . cell with the
// integer wrap count w).  The search keeps the pair (i -> j) with integer
// image shift S when |(p_j - p_i) + (S0 a + S1 b + S2 c)| < rcut, evaluated in
// that order, so that (j -> i, -S) gives exactly the negated vector and the
// same decision: the list is exactly symmetric.  The pair i -> i with S = 0
// is excluded.  Shifts are reported relative to the caller's (unwrapped)
// positions: S' = S - w_j + w_i, so that r_ij = (x_j - x_i) + S' . cell.
// Bins are at least rcut (1 + 1e-9) wide in Cartesian distance along each
// axis, so rounding in the fractional coordinates cannot hide a pair.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace acehost {

struct NbrGrid {
    double cell[9];
    int pbc[3];
    double rcut;
    int nb[3];        // bins per axis
    int R[3];         // bin search range per axis (bins on each side)
    double s0[3];     // fractional origin of the bins
    double w[3];      // fractional bin width
    std::vector<double> p;           // wrapped positions, SoA [3][natoms]
    std::vector<int32_t> wrap;       // wrap counts, SoA [3][natoms]
    std::vector<int32_t> bin;        // [natoms] linear bin index
    std::vector<int32_t> bin_start;  // [nbins + 1]
    std::vector<int32_t> bin_atoms;  // atoms by bin, ascending index within a bin
    int nbins() const { return nb[0] * nb[1] * nb[2]; }
};

inline bool invert3(const double* m, double* inv) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
    return true;
}

// Cell set-up: validated inputs, the inverse cell (fractional s = x . inv,
// column d of inv gives axis d) and the plane spacings h[d] = volume / |e x f|.
struct NbrGeom {
    double cell[9], inv[9], h[3], rcut;
    int pbc[3];
};

inline std::string nbr_setup(const double* cell, const int32_t* pbc, double rcut, NbrGeom& geo) {
    if (!(rcut > 0.0) || !std::isfinite(rcut)) return "rcut must be positive";
    for (int q = 0; q < 9; ++q) {
        geo.cell[q] = cell[q];
        if (!std::isfinite(cell[q])) return "cell must be finite";
    }
    for (int d = 0; d < 3; ++d) {
        if (pbc[d] != 0 && pbc[d] != 1) return "pbc entries must be 0 or 1";
        geo.pbc[d] = pbc[d];
    }
    geo.rcut = rcut;
    if (!invert3(geo.cell, geo.inv)) return "cell is singular";
    const double* a = geo.cell;
    const double* b = geo.cell + 3;
    const double* c = geo.cell + 6;
    auto cross_norm = [](const double* u, const double* v) {
        const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
        return std::sqrt(x * x + y * y + z * z);
    };
    const double vol = std::fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                                 a[2] * (b[0] * c[1] - b[1] * c[0]));
    geo.h[0] = vol / cross_norm(b, c);
    geo.h[1] = vol / cross_norm(c, a);
    geo.h[2] = vol / cross_norm(a, b);
    return "";
}

// Fractional coordinate of atom i along axis d, wrapped into [0, 1) along
// periodic axes (wrap count in *w).  The device binning kernel evaluates the
// same operations in the same order.
inline double nbr_frac(const NbrGeom& geo, double x, double y, double z, int d, int64_t* w) {
    double s = x * geo.inv[d] + y * geo.inv[3 + d] + z * geo.inv[6 + d];
    *w = 0;
    if (geo.pbc[d]) {
        const double f = std::floor(s);
        // wraps beyond +-1e6 cells are rejected by the callers; saturate first
        *w = std::fabs(f) > 1e6 ? (f > 0 ? 2000000 : -2000000) : (int64_t)f;
        s -= f;
        if (s >= 1.0) s = 0.0;  // rounding at the upper face
    }
    return s;
}

// Bin counts nb, search ranges R, origin s0 and width w from the extents
// smin / smax of the fractional coordinates along open axes (periodic axes
// span [0, 1)).  Bins are at least rcut (1 + 1e-9) wide in Cartesian distance
// along each axis, at most 1024 per axis and 2 natoms + 64 in total.
inline std::string nbr_params(const NbrGeom& geo, int64_t natoms, const double* smin_in, const double* smax_in,
                              NbrGrid& g) {
    for (int q = 0; q < 9; ++q) g.cell[q] = geo.cell[q];
    for (int d = 0; d < 3; ++d) g.pbc[d] = geo.pbc[d];
    g.rcut = geo.rcut;
    const double rmarg = geo.rcut * (1.0 + 1e-9);
    double span[3], len[3];
    for (int d = 0; d < 3; ++d) {
        double smin = geo.pbc[d] ? 0.0 : smin_in[d], smax = geo.pbc[d] ? 1.0 : smax_in[d];
        if (natoms == 0) {
            smin = 0.0;
            smax = 1.0;
        }
        span[d] = geo.pbc[d] ? 1.0 : std::max(smax - smin, 1e-12);
        len[d] = span[d] * geo.h[d];  // Cartesian extent normal to the planes
        g.nb[d] = (int)std::max(1.0, std::min(std::floor(len[d] / rmarg), 1024.0));
        g.s0[d] = geo.pbc[d] ? 0.0 : smin;
    }
    // sparse systems: coarsen the widest axis until bins <= 2 natoms + 64
    while ((int64_t)g.nb[0] * g.nb[1] * g.nb[2] > 2 * natoms + 64) {
        const int d = g.nb[0] >= g.nb[1] ? (g.nb[0] >= g.nb[2] ? 0 : 2) : (g.nb[1] >= g.nb[2] ? 1 : 2);
        g.nb[d] = std::max(1, g.nb[d] / 2);
    }
    for (int d = 0; d < 3; ++d) {
        g.w[d] = span[d] / g.nb[d];
        // bins on each side that can hold a neighbour
        const double r = std::ceil(rmarg / (len[d] / g.nb[d]));
        if (!g.pbc[d]) {
            // no images along open axes: at most the other bins of the axis
            g.R[d] = (int)std::min(std::max(r, 1.0), (double)(g.nb[d] - 1));
            continue;
        }
        if (r > 64) return "cell too small for rcut (more than 64 bins or images per side)";
        g.R[d] = std::max(1, (int)r);
    }
    return "";
}

// Bin of fractional coordinate s along axis d (clamped into the grid).
inline int nbr_bin1(const NbrGrid& g, double s, int d) {
    const int v = (int)std::floor((s - g.s0[d]) / g.w[d]);
    return std::min(std::max(v, 0), g.nb[d] - 1);
}

// Host binning of positions pos[3][natoms] (SoA): wrapped positions, wrap
// counts, bins and the atoms sorted by bin (stable).  Returns "" or a message.
inline std::string nbr_grid(int64_t natoms, const double* pos, const double* cell, const int32_t* pbc, double rcut,
                            NbrGrid& g) {
    if (natoms < 0) return "natoms < 0";
    if (natoms >= (int64_t)1 << 31) return "too many atoms";
    NbrGeom geo;
    const std::string e = nbr_setup(cell, pbc, rcut, geo);
    if (!e.empty()) return e;
    const double* a = geo.cell;
    const double* b = geo.cell + 3;
    const double* c = geo.cell + 6;
    std::vector<double> s(3 * (size_t)natoms);
    g.p.assign(3 * (size_t)natoms, 0.0);
    g.wrap.assign(3 * (size_t)natoms, 0);
    double smin[3] = {1e300, 1e300, 1e300}, smax[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < natoms; ++i) {
        const double x = pos[i], y = pos[natoms + i], z = pos[2 * natoms + i];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) return "positions must be finite";
        int64_t wrap[3];
        for (int d = 0; d < 3; ++d) {
            const double si = nbr_frac(geo, x, y, z, d, &wrap[d]);
            if (std::llabs(wrap[d]) > 1000000) return "positions too far outside the cell";
            s[3 * i + d] = si;
            g.wrap[d * (size_t)natoms + i] = (int32_t)wrap[d];
            smin[d] = std::min(smin[d], si);
            smax[d] = std::max(smax[d], si);
        }
        // wrapped Cartesian position: x - wrap . cell
        for (int q = 0; q < 3; ++q)
            g.p[q * (size_t)natoms + i] = pos[q * natoms + i] - ((double)wrap[0] * a[q] + (double)wrap[1] * b[q] +
                                                                 (double)wrap[2] * c[q]);
    }
    const std::string pe = nbr_params(geo, natoms, smin, smax, g);
    if (!pe.empty()) return pe;
    // stable counting sort of the atoms by bin
    const int nbins = g.nbins();
    g.bin.assign((size_t)natoms, 0);
    g.bin_start.assign((size_t)nbins + 1, 0);
    for (int64_t i = 0; i < natoms; ++i) {
        g.bin[i] = (nbr_bin1(g, s[3 * i + 2], 2) * g.nb[1] + nbr_bin1(g, s[3 * i + 1], 1)) * g.nb[0] +
                   nbr_bin1(g, s[3 * i], 0);
        ++g.bin_start[g.bin[i] + 1];
    }
    for (int q = 0; q < nbins; ++q) g.bin_start[q + 1] += g.bin_start[q];
    g.bin_atoms.assign((size_t)natoms, 0);
    std::vector<int32_t> fill(g.bin_start.begin(), g.bin_start.end() - 1);
    for (int64_t i = 0; i < natoms; ++i) g.bin_atoms[fill[g.bin[i]]++] = (int32_t)i;
    return "";
}

}  // namespace acehost
