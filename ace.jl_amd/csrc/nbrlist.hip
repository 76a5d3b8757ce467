// Environment extraction and force assembly on either side of the basis
// evaluation (the atoms-level callers of ACE.jl: src/testing/testlsq.jl:
// 27-35,42-69 call energy(V, at) / forces(V, at) on whole configurations).
//
//   ace_nbrlist_build   cell-list neighbour search with periodic images:
//                       host bins the atoms (nbrlist_host.hpp), the device
//                       counts and fills every atom's environment (relative
//                       positions, neighbour index, image shift, species)
//                       and the reverse index of each pair.
//   ace_nbrlist_update  recomputes the relative positions of an existing
//                       list from new positions (a list built with a skin
//                       stays valid for several MD steps); stream-ordered,
//                       no host synchronization.
//   ace_nbrlist_forces_multi  the same for many fields at once, e.g. the
//                       per-basis-function forces of least-squares fitting.
//   ace_nbrlist_forces  F_k = -dE/dx_k from the per-neighbour gradients
//                       g_m = +dE_i/dr_m (the ACE.jl grad_config sign) and
//                       the virial W = -sum_m g_m r_m^T, deterministically
//                       (a gather over each atom's own entries, no atomics).
//
// Entry m of atom i is the pair (i -> j, S') with r_m = (x_j - x_i) + S' . cell.
// E = sum_i E_i(r_m : m in N(i)), so
//   dE/dx_k = sum_{m : j(m) = k} g_m - sum_{m in N(k)} g_m
//           = sum_{m in N(k)} (g_rev(m) - g_m),
// where rev(m) is the entry (j -> i, -S') of the symmetric list.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "../../include/acehip.h"
#include "host_util.hpp"
#include "nbrlist_host.hpp"

// No multiply-add contraction in this file: the search evaluates every pair
// in both orientations and in two passes, and the decisions must agree
// exactly (the negated operands of the reverse pair round identically only
// when the same separately rounded operations are used).
#pragma clang fp contract(off)

struct NbrArgs {
    double cell[9];
    double rc2;
    int nb[3], R[3], pbc[3];
    int natoms;
    const double* pw;          // wrapped positions [3][natoms]
    const int32_t* wrap;       // wrap counts [3][natoms]
    const int32_t* bin;        // [natoms]
    const int32_t* bin_start;  // [nbins + 1]
    const int32_t* bin_atoms;  // [natoms]
};

__device__ __forceinline__ int floordiv(int z, int n) { return z >= 0 ? z / n : -((n - 1 - z) / n); }

// Shift vector S0 a + S1 b + S2 c and the squared distance.
__device__ __forceinline__ void shift_vec(const double* cell, int s0, int s1, int s2, double& vx, double& vy,
                                          double& vz) {
    const double a = s0, b = s1, c = s2;
    vx = (a * cell[0] + b * cell[3]) + c * cell[6];
    vy = (a * cell[1] + b * cell[4]) + c * cell[7];
    vz = (a * cell[2] + b * cell[5]) + c * cell[8];
}

__device__ __forceinline__ double norm2(double x, double y, double z) { return (x * x + y * y) + z * z; }

// One thread per atom, atoms taken in bin order so that a wave scans the
// same few bins.  FILL = false counts, FILL = true writes the entries in the
// same order: bins (z, y, x ascending, images included), atoms of a bin
// ascending.
template <bool FILL>
__global__ __launch_bounds__(256) void k_nbr_search(NbrArgs a, int32_t* __restrict__ cnt,
                                                    const int32_t* __restrict__ off, const double* __restrict__ pos,
                                                    int64_t nnbr, int32_t* __restrict__ nbr_j,
                                                    int32_t* __restrict__ shift, double* __restrict__ xyz,
                                                    const int8_t* __restrict__ zat, int8_t* __restrict__ znbr,
                                                    int32_t* __restrict__ flag) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.natoms) return;
    const int n = a.natoms;
    const int i = a.bin_atoms[t];
    const int b = a.bin[i];
    const int bx = b % a.nb[0], by = (b / a.nb[0]) % a.nb[1], bz = b / (a.nb[0] * a.nb[1]);
    const double pix = a.pw[i], piy = a.pw[n + i], piz = a.pw[2 * n + i];
    int64_t m = 0, mend = 0;
    int wi0 = 0, wi1 = 0, wi2 = 0;
    double xix = 0, xiy = 0, xiz = 0;
    if (FILL) {
        m = off[i];
        mend = off[i + 1];
        wi0 = a.wrap[i];
        wi1 = a.wrap[n + i];
        wi2 = a.wrap[2 * n + i];
        xix = pos[i];
        xiy = pos[n + i];
        xiz = pos[2 * n + i];
    }
    int c = 0;
    for (int dz = -a.R[2]; dz <= a.R[2]; ++dz) {
        int z = bz + dz, S2 = 0;
        if (a.pbc[2]) {
            S2 = floordiv(z, a.nb[2]);
            z -= S2 * a.nb[2];
        } else if (z < 0 || z >= a.nb[2]) {
            continue;
        }
        for (int dy = -a.R[1]; dy <= a.R[1]; ++dy) {
            int y = by + dy, S1 = 0;
            if (a.pbc[1]) {
                S1 = floordiv(y, a.nb[1]);
                y -= S1 * a.nb[1];
            } else if (y < 0 || y >= a.nb[1]) {
                continue;
            }
            for (int dx = -a.R[0]; dx <= a.R[0]; ++dx) {
                int x = bx + dx, S0 = 0;
                if (a.pbc[0]) {
                    S0 = floordiv(x, a.nb[0]);
                    x -= S0 * a.nb[0];
                } else if (x < 0 || x >= a.nb[0]) {
                    continue;
                }
                double vx, vy, vz;
                shift_vec(a.cell, S0, S1, S2, vx, vy, vz);
                const bool zero = (S0 | S1 | S2) == 0;
                const int bb = (z * a.nb[1] + y) * a.nb[0] + x;
                const int q1 = a.bin_start[bb + 1];
                for (int q = a.bin_start[bb]; q < q1; ++q) {
                    const int j = a.bin_atoms[q];
                    if (zero && j == i) continue;
                    const double ex = (a.pw[j] - pix) + vx;
                    const double ey = (a.pw[n + j] - piy) + vy;
                    const double ez = (a.pw[2 * n + j] - piz) + vz;
                    if (!(norm2(ex, ey, ez) < a.rc2)) continue;
                    if (FILL) {
                        if (m < mend) {
                            const int s0 = S0 - a.wrap[j] + wi0;
                            const int s1 = S1 - a.wrap[n + j] + wi1;
                            const int s2 = S2 - a.wrap[2 * n + j] + wi2;
                            double ux, uy, uz;
                            shift_vec(a.cell, s0, s1, s2, ux, uy, uz);
                            nbr_j[m] = j;
                            shift[m] = s0;
                            shift[nnbr + m] = s1;
                            shift[2 * nnbr + m] = s2;
                            xyz[m] = (pos[j] - xix) + ux;
                            xyz[nnbr + m] = (pos[n + j] - xiy) + uy;
                            xyz[2 * nnbr + m] = (pos[2 * n + j] - xiz) + uz;
                            if (znbr) znbr[m] = zat[j];
                        }
                        ++m;
                    } else {
                        ++c;
                    }
                }
            }
        }
    }
    if (!FILL) cnt[i] = c;
    if (FILL && m != mend) atomicOr(flag, 1);
}

// Centre of entry m: the atom i with off[i] <= m < off[i+1] (off ascending).
__device__ __forceinline__ int owner(const int32_t* off, int natoms, int m) {
    int lo = 0, hi = natoms;  // off[lo] <= m < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= m) lo = mid;
        else hi = mid;
    }
    return lo;
}

// rev[m] = index of (j -> i, -S') in the list of j, one thread per entry;
// flag bit 2 if absent.
__global__ __launch_bounds__(256) void k_nbr_rev(int natoms, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ nbr_j, const int32_t* __restrict__ shift,
                                                 int64_t nnbr, int32_t* __restrict__ rev, int32_t* __restrict__ flag) {
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= nnbr) return;
    const int i = owner(off, natoms, (int)m);
    const int j = nbr_j[m];
    const int s0 = -shift[m], s1 = -shift[nnbr + m], s2 = -shift[2 * nnbr + m];
    int found = -1;
    const int q1 = off[j + 1];
    for (int q = off[j]; q < q1; ++q)
        if (nbr_j[q] == i && shift[q] == s0 && shift[nnbr + q] == s1 && shift[2 * nnbr + q] == s2) {
            found = q;
            break;
        }
    rev[m] = found;
    if (found < 0) atomicOr(flag, 2);
}

__global__ __launch_bounds__(256) void k_nbr_update(int natoms, NbrArgs a, const double* __restrict__ pos,
                                                    const int32_t* __restrict__ off, const int32_t* __restrict__ nbr_j,
                                                    const int32_t* __restrict__ shift, int64_t nnbr,
                                                    double* __restrict__ xyz) {
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= nnbr) return;
    const int i = owner(off, natoms, (int)m);
    const int j = nbr_j[m];
    double ux, uy, uz;
    shift_vec(a.cell, shift[m], shift[nnbr + m], shift[2 * nnbr + m], ux, uy, uz);
    const int n = natoms;
    xyz[m] = (pos[j] - pos[i]) + ux;
    xyz[nnbr + m] = (pos[n + j] - pos[n + i]) + uy;
    xyz[2 * nnbr + m] = (pos[2 * n + j] - pos[2 * n + i]) + uz;
}

// F_k = sum_{m in N(k)} (g_m - g_rev(m)) for FA_ATOMS atoms per block.  The
// block's entries [off[a0], off[a0 + FA_ATOMS]) are read coalesced by all
// threads, FA_CH at a time, as d_m = g_m - g_rev(m) into LDS; then one thread
// per atom sums its own entries in order.  Virial partials -g_m r_m^T are
// summed per thread over its strided entries and reduced in a fixed order.
constexpr int FA_ATOMS = 64;
constexpr int FA_CH = 1024;

__global__ __launch_bounds__(256) void k_nbr_forces(int natoms, const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ rev, int64_t nnbr,
                                                    const double* __restrict__ g, const double* __restrict__ xyz,
                                                    double* __restrict__ F, double* __restrict__ vpart) {
    static_assert(9 * 256 <= 3 * FA_CH, "the virial reduction reuses the staging array");
    __shared__ double dl[3][FA_CH];
    double (*red)[256] = reinterpret_cast<double (*)[256]>(&dl[0][0]);
    const int a0 = blockIdx.x * FA_ATOMS;
    const int a1 = min(a0 + FA_ATOMS, natoms);
    const int tid = threadIdx.x;
    const int m0 = off[a0], m1 = off[a1];
    // this thread's atom (tid < FA_ATOMS) and its entry range
    const int k = a0 + tid;
    const bool own = tid < FA_ATOMS && k < a1;
    const int mk0 = own ? off[k] : 0, mk1 = own ? off[k + 1] : 0;
    double fx = 0, fy = 0, fz = 0;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c0 = m0; c0 < m1; c0 += FA_CH) {
        const int c1 = min(c0 + FA_CH, m1);
        for (int m = c0 + tid; m < c1; m += 256) {
            const int r = rev[m];
            const double gx = g[m], gy = g[nnbr + m], gz = g[2 * nnbr + m];
            dl[0][m - c0] = gx - g[r];
            dl[1][m - c0] = gy - g[nnbr + r];
            dl[2][m - c0] = gz - g[2 * nnbr + r];
            if (vpart) {
                const double rx = xyz[m], ry = xyz[nnbr + m], rz = xyz[2 * nnbr + m];
                v[0] -= gx * rx;
                v[1] -= gx * ry;
                v[2] -= gx * rz;
                v[3] -= gy * rx;
                v[4] -= gy * ry;
                v[5] -= gy * rz;
                v[6] -= gz * rx;
                v[7] -= gz * ry;
                v[8] -= gz * rz;
            }
        }
        __syncthreads();
        if (own) {
            const int s0 = max(mk0, c0), s1 = min(mk1, c1);
            for (int m = s0; m < s1; ++m) {
                fx += dl[0][m - c0];
                fy += dl[1][m - c0];
                fz += dl[2][m - c0];
            }
        }
        __syncthreads();
    }
    if (own) {
        F[k] = fx;
        F[natoms + k] = fy;
        F[2 * natoms + k] = fz;
    }
    if (!vpart) return;  // uniform over the block
    for (int c = 0; c < 9; ++c) red[c][tid] = v[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
            for (int c = 0; c < 9; ++c) red[c][tid] += red[c][tid + s];
        __syncthreads();
    }
    if (tid < 9) vpart[(int64_t)blockIdx.x * 9 + tid] = red[tid][0];
}

// F[k][v][d] = sum_{m in N(k)} (g[m][v][d] - g[rev(m)][v][d]) for nvec
// fields at once (per-basis-function forces): thread per (atom, field),
// consecutive threads read consecutive fields of one entry.
__global__ __launch_bounds__(256) void k_nbr_forces_multi(int natoms, const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ rev, int64_t nvec,
                                                          const double* __restrict__ g, double* __restrict__ F) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)natoms * nvec) return;
    const int k = (int)(t / nvec);
    const int64_t v = t - (int64_t)k * nvec;
    double fx = 0, fy = 0, fz = 0;
    const int m1 = off[k + 1];
    for (int m = off[k]; m < m1; ++m) {
        const double* a = g + ((int64_t)m * nvec + v) * 3;
        const double* b = g + ((int64_t)rev[m] * nvec + v) * 3;
        fx += a[0] - b[0];
        fy += a[1] - b[1];
        fz += a[2] - b[2];
    }
    double* o = F + t * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
}

// W[c] = sum over blocks of vpart[b][c] in a fixed order: 64 lanes per
// component take every 64th block, then a tree over the lanes.
__global__ __launch_bounds__(576) void k_nbr_vsum(int nblk, const double* __restrict__ vpart,
                                                  double* __restrict__ W) {
    __shared__ double part[9][64];
    const int c = threadIdx.x >> 6, l = threadIdx.x & 63;
    double s = 0.0;
    for (int b = l; b < nblk; b += 64) s += vpart[(int64_t)b * 9 + c];
    part[c][l] = s;
    __syncthreads();
    for (int h = 32; h > 0; h >>= 1) {
        if (l < h) part[c][l] += part[c][l + h];
        __syncthreads();
    }
    if (l == 0) W[c] = part[c][0];
}

__global__ void k_nbr_zero(int32_t* p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// ---------------------------------------------------------------------------
// binning on the device (the operations of nbrlist_host.hpp: nbr_frac,
// nbr_bin1, wrapped positions), so that the bins match the host's exactly

struct NbrFracArgs {
    double cell[9], inv[9];
    int pbc[3];
    int natoms;
};

// Wrapped fractional coordinates sf[3][n], wrap counts, wrapped positions
// pw[3][n], and per-block minima / maxima of sf (bmm[block][6]).  flag bit 4:
// a non-finite position; bit 8: a wrap beyond 1e6 cells.
__global__ __launch_bounds__(256) void k_nbr_frac(NbrFracArgs a, const double* __restrict__ pos,
                                                  double* __restrict__ sf, int32_t* __restrict__ wrap,
                                                  double* __restrict__ pw, double* __restrict__ bmm,
                                                  int32_t* __restrict__ flag) {
    __shared__ double red[6][256];
    const int n = a.natoms;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (i < n) {
        const double x = pos[i], y = pos[n + i], z = pos[2 * n + i];
        int w[3] = {0, 0, 0};
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) atomicOr(flag, 4);
        for (int d = 0; d < 3; ++d) {
            double sv = x * a.inv[d] + y * a.inv[3 + d] + z * a.inv[6 + d];
            if (a.pbc[d]) {
                const double f = floor(sv);
                if (!(fabs(f) <= 1e6)) {
                    atomicOr(flag, 8);
                } else {
                    w[d] = (int)f;
                }
                sv -= f;
                if (sv >= 1.0) sv = 0.0;
            }
            sf[d * n + i] = sv;
            wrap[d * n + i] = w[d];
            lo[d] = sv;
            hi[d] = sv;
        }
        for (int q = 0; q < 3; ++q)
            pw[q * n + i] = pos[q * n + i] - ((double)w[0] * a.cell[q] + (double)w[1] * a.cell[3 + q] +
                                              (double)w[2] * a.cell[6 + q]);
    }
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + h]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + h]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) bmm[(int64_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// mm[0..2] = minima, mm[3..5] = maxima over the blocks (one block).
__global__ __launch_bounds__(256) void k_nbr_minmax(int nblk, const double* __restrict__ bmm, double* __restrict__ mm) {
    __shared__ double red[6][256];
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int b = threadIdx.x; b < nblk; b += 256)
        for (int c = 0; c < 6; ++c) v[c] = c < 3 ? fmin(v[c], bmm[(int64_t)b * 6 + c]) : fmax(v[c], bmm[(int64_t)b * 6 + c]);
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int c = 0; c < 6; ++c)
                red[c][threadIdx.x] = c < 3 ? fmin(red[c][threadIdx.x], red[c][threadIdx.x + h])
                                            : fmax(red[c][threadIdx.x], red[c][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 6) mm[threadIdx.x] = red[threadIdx.x][0];
}

struct NbrBinArgs {
    int nb[3];
    double s0[3], w[3];
    int natoms;
};

__device__ __forceinline__ int bin1(const NbrBinArgs& g, double s, int d) {
    const int v = (int)floor((s - g.s0[d]) / g.w[d]);
    return min(max(v, 0), g.nb[d] - 1);
}

// bin[i] and the sort values (atom index)
__global__ __launch_bounds__(256) void k_nbr_bin(NbrBinArgs g, const double* __restrict__ sf,
                                                 int32_t* __restrict__ bin, int32_t* __restrict__ idx) {
    const int n = g.natoms;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    bin[i] = (bin1(g, sf[2 * n + i], 2) * g.nb[1] + bin1(g, sf[n + i], 1)) * g.nb[0] + bin1(g, sf[i], 0);
    idx[i] = i;
}

// bin_start[b] = first position of bin b in the sorted keys (b <= nbins)
__global__ __launch_bounds__(256) void k_nbr_binstart(int nbins, int n, const int32_t* __restrict__ keys,
                                                      int32_t* __restrict__ bin_start) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b > nbins) return;
    int lo = 0, hi = n;  // first q with keys[q] >= b
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < b) lo = mid + 1;
        else hi = mid;
    }
    bin_start[b] = lo;
}

// ---------------------------------------------------------------------------
// host side

struct ace_nbrlist {
    int device = 0;
    acehost::NbrGrid g;  // grid parameters (nb, R, s0, w) of the last build
    std::vector<int32_t> hcnt, hoff;
    int32_t hflag = 0;
    double hmm[6] = {0, 0, 0, 0, 0, 0};
    DevBuf<double> pw, vpart, sf, bmm, mm;
    DevBuf<int32_t> wrap, bin, bin_start, bin_atoms, cnt, flag, keys, idx;
    DevBuf<char> sort_tmp;
};

static NbrArgs grid_args(ace_nbrlist* nl, const double* cell) {
    NbrArgs a{};
    for (int q = 0; q < 9; ++q) a.cell[q] = cell[q];
    const acehost::NbrGrid& g = nl->g;
    a.rc2 = g.rcut * g.rcut;
    for (int d = 0; d < 3; ++d) {
        a.nb[d] = g.nb[d];
        a.R[d] = g.R[d];
        a.pbc[d] = g.pbc[d];
    }
    a.natoms = (int)nl->hcnt.size();
    a.pw = nl->pw.p;
    a.wrap = nl->wrap.p;
    a.bin = nl->bin.p;
    a.bin_start = nl->bin_start.p;
    a.bin_atoms = nl->bin_atoms.p;
    return a;
}

extern "C" int ace_nbrlist_create(int device, ace_nbrlist** out) {
    if (!out) return set_err("ace_nbrlist_create: null argument");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_err("ace_nbrlist_create: device index out of range");
    ace_nbrlist* nl = new ace_nbrlist();
    nl->device = device;
    *out = nl;
    return 0;
}

extern "C" void ace_nbrlist_destroy(ace_nbrlist* nl) {
    if (!nl) return;
    (void)hipSetDevice(nl->device);
    delete nl;
}

extern "C" int ace_nbrlist_build(ace_nbrlist* nl, int64_t natoms, const double* pos, const double* cell,
                                 const int32_t* pbc, double rcut, const int8_t* atom_species, int64_t capacity,
                                 int32_t* env_offsets, int32_t* nbr_j, int32_t* shift, double* xyz, int32_t* rev,
                                 int8_t* nbr_species, int64_t* nnbr_out, void* stream) {
    if (!nl || !cell || !pbc || !env_offsets || !nnbr_out) return set_err("ace_nbrlist_build: null argument");
    if (natoms < 0 || natoms >= ((int64_t)1 << 31) - 1) return set_err("ace_nbrlist_build: natoms out of range");
    if (natoms > 0 && !pos) return set_err("ace_nbrlist_build: null positions");
    if (capacity < 0) return set_err("ace_nbrlist_build: capacity < 0");
    if (capacity > 0 && (!nbr_j || !shift || !xyz)) return set_err("ace_nbrlist_build: null neighbour arrays");
    if (nbr_species && !atom_species) return set_err("ace_nbrlist_build: nbr_species needs atom_species");
    *nnbr_out = 0;
    HIP_TRY(hipSetDevice(nl->device));
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)natoms;
    acehost::NbrGeom geo;
    const std::string gerr = acehost::nbr_setup(cell, pbc, rcut, geo);
    if (!gerr.empty()) return set_err("ace_nbrlist_build: " + gerr);
    nl->hoff.assign((size_t)n + 1, 0);
    if (n == 0) {
        HIP_TRY(hipMemcpyAsync(env_offsets, nl->hoff.data(), sizeof(int32_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
        return 0;
    }
    const unsigned nblk = (unsigned)((n + 255) / 256);
    if (nl->pw.ensure(3 * (size_t)n) || nl->wrap.ensure(3 * (size_t)n) || nl->sf.ensure(3 * (size_t)n) ||
        nl->bin.ensure(n) || nl->idx.ensure(n) || nl->keys.ensure(n) || nl->bin_atoms.ensure(n) ||
        nl->cnt.ensure(n) || nl->flag.ensure(1) || nl->bmm.ensure((size_t)nblk * 6) || nl->mm.ensure(6))
        return -1;
    // wrap and bin on the device: fractional coordinates, extents (for open axes)
    NbrFracArgs fa{};
    for (int q = 0; q < 9; ++q) {
        fa.cell[q] = geo.cell[q];
        fa.inv[q] = geo.inv[q];
    }
    for (int d = 0; d < 3; ++d) fa.pbc[d] = geo.pbc[d];
    fa.natoms = n;
    hipLaunchKernelGGL(k_nbr_zero, dim3(1), dim3(64), 0, st, nl->flag.p, 1);
    hipLaunchKernelGGL(k_nbr_frac, dim3(nblk), dim3(256), 0, st, fa, pos, nl->sf.p, nl->wrap.p, nl->pw.p, nl->bmm.p,
                       nl->flag.p);
    hipLaunchKernelGGL(k_nbr_minmax, dim3(1), dim3(256), 0, st, (int)nblk, (const double*)nl->bmm.p, nl->mm.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(nl->hmm, nl->mm.p, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&nl->hflag, nl->flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nl->hflag & 4) return set_err("ace_nbrlist_build: positions must be finite");
    if (nl->hflag & 8) return set_err("ace_nbrlist_build: positions too far outside the cell");
    acehost::NbrGrid& g = nl->g;
    const std::string perr = acehost::nbr_params(geo, natoms, nl->hmm, nl->hmm + 3, g);
    if (!perr.empty()) return set_err("ace_nbrlist_build: " + perr);
    const int nbins = g.nbins();
    if (nl->bin_start.ensure((size_t)nbins + 1)) return -1;
    NbrBinArgs ba{};
    for (int d = 0; d < 3; ++d) {
        ba.nb[d] = g.nb[d];
        ba.s0[d] = g.s0[d];
        ba.w[d] = g.w[d];
    }
    ba.natoms = n;
    hipLaunchKernelGGL(k_nbr_bin, dim3(nblk), dim3(256), 0, st, ba, (const double*)nl->sf.p, nl->bin.p, nl->idx.p);
    HIP_TRY(hipGetLastError());
    // stable sort of the atoms by bin (ascending atom index within a bin)
    int end_bit = 1;
    while (end_bit < 31 && (1 << end_bit) < nbins) ++end_bit;
    size_t tmp_bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, nl->bin.p, nl->keys.p, nl->idx.p,
                                               nl->bin_atoms.p, n, 0, end_bit, st));
    if (nl->sort_tmp.ensure(std::max<size_t>(tmp_bytes, 1))) return -1;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nl->sort_tmp.p, tmp_bytes, nl->bin.p, nl->keys.p, nl->idx.p,
                                               nl->bin_atoms.p, n, 0, end_bit, st));
    hipLaunchKernelGGL(k_nbr_binstart, dim3((unsigned)((nbins + 1 + 255) / 256)), dim3(256), 0, st, nbins, n,
                       (const int32_t*)nl->keys.p, nl->bin_start.p);
    HIP_TRY(hipGetLastError());
    nl->hcnt.resize(n);
    const NbrArgs a = grid_args(nl, cell);
    hipLaunchKernelGGL(k_nbr_search<false>, dim3(nblk), dim3(256), 0, st, a, nl->cnt.p, nullptr, nullptr,
                       (int64_t)0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(nl->hcnt.data(), nl->cnt.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        total += nl->hcnt[i];
        if (total >= ((int64_t)1 << 31)) return set_err("ace_nbrlist_build: more than 2^31 - 1 neighbour entries");
        nl->hoff[i + 1] = (int32_t)total;
    }
    *nnbr_out = total;
    if (total > capacity) {
        set_err("ace_nbrlist_build: capacity " + std::to_string(capacity) + " < " + std::to_string(total) +
                " neighbour entries (*nnbr_out holds the size needed)");
        return ACE_ERR_CAPACITY;
    }
    HIP_TRY(hipMemcpyAsync(env_offsets, nl->hoff.data(), ((size_t)n + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                           st));
    hipLaunchKernelGGL(k_nbr_zero, dim3(1), dim3(64), 0, st, nl->flag.p, 1);
    hipLaunchKernelGGL(k_nbr_search<true>, dim3(nblk), dim3(256), 0, st, a, nl->cnt.p, (const int32_t*)env_offsets,
                       pos, total, nbr_j, shift, xyz, atom_species, nbr_species, nl->flag.p);
    HIP_TRY(hipGetLastError());
    if (rev && total > 0) {
        hipLaunchKernelGGL(k_nbr_rev, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, n,
                           (const int32_t*)env_offsets, (const int32_t*)nbr_j, (const int32_t*)shift, total, rev,
                           nl->flag.p);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(&nl->hflag, nl->flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nl->hflag & 1) return set_err("ace_nbrlist_build: fill pass disagrees with the count pass");
    if (nl->hflag & 2) return set_err("ace_nbrlist_build: neighbour list not symmetric (reverse entry missing)");
    return 0;
}

extern "C" int ace_nbrlist_update(ace_nbrlist* nl, int64_t natoms, const double* pos, const double* cell,
                                  int64_t nnbr, const int32_t* env_offsets, const int32_t* nbr_j,
                                  const int32_t* shift, double* xyz, void* stream) {
    if (!nl || !cell) return set_err("ace_nbrlist_update: null argument");
    if (natoms < 0 || natoms >= ((int64_t)1 << 31) - 1 || nnbr < 0 || nnbr >= ((int64_t)1 << 31))
        return set_err("ace_nbrlist_update: size out of range");
    if (nnbr == 0) return 0;
    if (!pos || !env_offsets || !nbr_j || !shift || !xyz || natoms == 0)
        return set_err("ace_nbrlist_update: null argument");
    HIP_TRY(hipSetDevice(nl->device));
    NbrArgs a{};
    for (int q = 0; q < 9; ++q) a.cell[q] = cell[q];
    hipLaunchKernelGGL(k_nbr_update, dim3((unsigned)((nnbr + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (int)natoms, a, pos, env_offsets, nbr_j, shift, nnbr, xyz);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ace_nbrlist_forces(ace_nbrlist* nl, int64_t natoms, const int32_t* env_offsets, const int32_t* rev,
                                  int64_t nnbr, const double* xyz, const double* dE_dr, double* F, double* virial,
                                  void* stream) {
    if (!nl) return set_err("ace_nbrlist_forces: null argument");
    if (natoms < 0 || natoms >= ((int64_t)1 << 31) - 1 || nnbr < 0 || nnbr >= ((int64_t)1 << 31))
        return set_err("ace_nbrlist_forces: size out of range");
    if (natoms > 0 && (!env_offsets || !F)) return set_err("ace_nbrlist_forces: null argument");
    if (nnbr > 0 && (!rev || !dE_dr)) return set_err("ace_nbrlist_forces: null argument");
    if (virial && nnbr > 0 && !xyz) return set_err("ace_nbrlist_forces: the virial needs xyz");
    HIP_TRY(hipSetDevice(nl->device));
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)natoms;
    if (n == 0) {
        if (virial) hipLaunchKernelGGL(k_nbr_vsum, dim3(1), dim3(576), 0, st, 0, (const double*)nullptr, virial);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    const int nblk = (n + FA_ATOMS - 1) / FA_ATOMS;
    if (virial && nl->vpart.ensure((size_t)nblk * 9)) return -1;
    hipLaunchKernelGGL(k_nbr_forces, dim3((unsigned)nblk), dim3(256), 0, st, n, env_offsets, rev, nnbr, dE_dr, xyz, F,
                       virial ? nl->vpart.p : nullptr);
    HIP_TRY(hipGetLastError());
    if (virial) {
        hipLaunchKernelGGL(k_nbr_vsum, dim3(1), dim3(576), 0, st, nblk, (const double*)nl->vpart.p, virial);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

extern "C" int ace_nbrlist_forces_multi(ace_nbrlist* nl, int64_t natoms, const int32_t* env_offsets,
                                        const int32_t* rev, int64_t nnbr, int64_t nvec, const double* g, double* F,
                                        void* stream) {
    if (!nl) return set_err("ace_nbrlist_forces_multi: null argument");
    if (natoms < 0 || natoms >= ((int64_t)1 << 31) - 1 || nnbr < 0 || nnbr >= ((int64_t)1 << 31) || nvec < 0)
        return set_err("ace_nbrlist_forces_multi: size out of range");
    const int64_t n = natoms * nvec;
    if (n == 0) return 0;
    if (!env_offsets || !F || (nnbr > 0 && (!rev || !g))) return set_err("ace_nbrlist_forces_multi: null argument");
    HIP_TRY(hipSetDevice(nl->device));
    hipLaunchKernelGGL(k_nbr_forces_multi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (int)natoms, env_offsets, rev, nvec, g, F);
    HIP_TRY(hipGetLastError());
    return 0;
}
