// Force kernels of libacehip.so (included by the ace_force*.hip units, each
// of which instantiates the launcher for some values of Lmax: the kernels
// are fully unrolled over (l, m), so splitting by Lmax keeps the units'
// compile times short).
//
// Forces: +dE/dr_j = Re sum_k G[k] grad(R_n Y_lm)(r_j), the
// adjoint of the site energy (evaluator.jl:150-200, restated on the pooled
// values) contracted with the per-neighbour 1p gradients, which are
// recomputed (orthpolys.jl:287-306, sphericalharmonics.jl:212-267,410-443,
// 60-65) rather than stored.
#pragma once

#include "ace_common.hpp"

namespace {

// ---------------------------------------------------------------------------
// k_force: +dE/dr_j = Re sum_k G[k] grad(R_n Y_lm)(r_j)

// The NC components of G[k] for one environment: sums over k's slots.  With
// class-split programs (gcls set) a slot holds one complex value of class
// gcls[slot] (read once), and component p is sum_slots cls_u[class][p] * value.
template <int NC>
__device__ __forceinline__ void gather_g(const int32_t* __restrict__ gslot_ptr,
                                         const int32_t* __restrict__ gslot_list,
                                         const int32_t* __restrict__ gcls, const double2* __restrict__ cls_u,
                                         const double2* __restrict__ Gbuf, int64_t estride, int64_t env, int k,
                                         double2 (&g)[NC]) {
#pragma unroll
    for (int p = 0; p < NC; ++p) g[p] = make_double2(0.0, 0.0);
    for (int s = gslot_ptr[k]; s < gslot_ptr[k + 1]; ++s) {
        const int slot = gslot_list[s];
        if (NC > 1 && gcls) {
            const double2 v = Gbuf[(int64_t)slot * estride + env];
            const double2* u = cls_u + 3 * gcls[slot];
#pragma unroll
            for (int p = 0; p < NC; ++p) {
                g[p].x += u[p].x * v.x - u[p].y * v.y;
                g[p].y += u[p].x * v.y + u[p].y * v.x;
            }
        } else {
#pragma unroll
            for (int p = 0; p < NC; ++p) {
                const double2 v = Gbuf[((int64_t)slot * NC + p) * estride + env];
                g[p].x += v.x;
                g[p].y += v.y;
            }
        }
    }
}

// The force contraction Re sum_{l,m,n} G[l,m,n] grad(R_n Y_lm)(r_j),
// factored over the radial index: with a = sum_n G dR_n/dt and b = sum_n G
// R_n (complex; 4 fused multiply-adds per n) the (l, m) term is
// Re(a Y) rhat dt/dr + Re(b grad Y), whose angular part needs only the
// azimuthal and polar derivative factors of Y (cYlm_ed! and dspher_to_dcart,
// sphericalharmonics.jl:60-65,410-443; Pl = P_lm / sin(theta) for m >= 1).
// force_lm_add adds one (l, m) term to the gradient.
__device__ __forceinline__ void force_lm_add(double Pl, double dPl, int m, double epr, double epi, const Sph& S,
                                             double dtdr, double rinv, double rx, double ry, double rz, double ax,
                                             double ay, double bx, double by, double& g0, double& g1, double& g2) {
    double sr, su, sv;  // Re(a Y), Re(b fp), Re(b ft): fp = dY/dphi / sin(theta), ft = dY/dtheta
    if (m == 0) {
        constexpr double isq2 = 0.7071067811865475;
        sr = isq2 * Pl * ax;
        su = 0.0;
        sv = isq2 * dPl * bx;
    } else {
        sr = S.sth * Pl * (epr * ax - epi * ay);
        su = -(double)m * Pl * (epi * bx + epr * by);
        sv = dPl * (epr * bx - epi * by);
    }
    const double a = dtdr * sr, ct = S.cth * sv;
    g0 += fma(a, rx, rinv * (S.cphi * ct - S.sphi * su));
    g1 += fma(a, ry, rinv * (S.sphi * ct + S.cphi * su));
    g2 += fma(a, rz, rinv * (-S.sth * sv));
}

// The running gradient sums materialised after every (l, m) term: without
// this the compiler defers some of the sums to the end of the kernel and
// spills every term's inputs (measured: hundreds of spilled registers).
template <int NC>
__device__ __forceinline__ void force_keep(double (&g)[NC][3]) {
#pragma unroll
    for (int p = 0; p < NC; ++p) asm volatile("" : "+v"(g[p][0]), "+v"(g[p][1]), "+v"(g[p][2]));
}

template <int L, int PROP>
__global__ __launch_bounds__(64) void k_force_lds(RadialParams rp, int64_t nenv,
                                               const int32_t* __restrict__ off,
                                               const double* __restrict__ xyz, int64_t nnbr,
                                               const int8_t* __restrict__ species,
                                               const int16_t* __restrict__ kidx, int npool,
                                               const int32_t* __restrict__ gslot_ptr,
                                               const int32_t* __restrict__ gslot_list,
                                               const int32_t* __restrict__ gcls,
                                               const double2* __restrict__ cls_u,
                                               const double2* __restrict__ Gbuf, int64_t estride,
                                               double* __restrict__ dE) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const int64_t env = xcd_env(blockIdx.x, nenv);
    const int nr = rp.nr;
    double2* Gs = reinterpret_cast<double2*>(smem);
    double* Rs = reinterpret_cast<double*>(Gs + npool * NC);  // [64][2*nr]
    if (env >= nenv) return;
    for (int k = lane; k < npool; k += 64) {
        double2 g[NC];
        gather_g<NC>(gslot_ptr, gslot_list, gcls, cls_u, Gbuf, estride, env, k, g);
#pragma unroll
        for (int p = 0; p < NC; ++p) Gs[k * NC + p] = g[p];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int j0 = off[env], j1 = off[env + 1];
    double* myR = Rs + lane * 2 * nr;
    for (int j = j0 + lane; j < j1; j += 64) {
        const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
        const int mu = species ? (int)species[j] : 0;
        Sph S = cart2spher(x, y, z);
        double t, dtdr;
        transform_ed(rp, S.r, t, dtdr);
        {
            double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
            myR[0] = p0;
            myR[nr] = d0;
            if (nr > 1) {
                double al = rp.A[1] * t + rp.B[1];
                double p1 = al * p0, d1 = al * d0 + rp.A[1] * p0;
                myR[1] = p1;
                myR[nr + 1] = d1;
                for (int n = 2; n < nr; ++n) {
                    al = rp.A[n] * t + rp.B[n];
                    const double p2 = al * p1 + rp.C[n] * p0;
                    const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                    p0 = p1;
                    p1 = p2;
                    d0 = d1;
                    d1 = d2;
                    myR[n] = p2;
                    myR[nr + n] = d2;
                }
            }
        }
        double P[NP], dP[NP];
        alp_eval_ed<L>(rp, S, P, dP);
        const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
        const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
        double g[NC][3];
#pragma unroll
        for (int p = 0; p < NC; ++p) g[p][0] = g[p][1] = g[p][2] = 0.0;
        double epr = 0.7071067811865475, epi = 0.0;
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            if (m > 0) {
                const double nr_ = epr * S.cphi - epi * S.sphi;
                const double ni_ = epr * S.sphi + epi * S.cphi;
                epr = nr_;
                epi = ni_;
            }
#pragma unroll
            for (int l = m; l <= L; ++l) {
                const int ip = idxp(l, m);
                const int16_t* kr = kidx + ((int64_t)(mu * (L + 1) + l) * (L + 1) + m) * nr;
                double ax[NC], ay[NC], bx[NC], by[NC];
#pragma unroll
                for (int p = 0; p < NC; ++p) ax[p] = ay[p] = bx[p] = by[p] = 0.0;
                for (int n = 0; n < nr; ++n) {
                    const int k = kr[n];
                    if (k < 0) continue;
                    const double R = myR[n], dR = myR[nr + n];
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        const double2 G = Gs[k * NC + p];
                        ax[p] = fma(G.x, dR, ax[p]);
                        ay[p] = fma(G.y, dR, ay[p]);
                        bx[p] = fma(G.x, R, bx[p]);
                        by[p] = fma(G.y, R, by[p]);
                    }
                }
#pragma unroll
                for (int p = 0; p < NC; ++p)
                    force_lm_add(P[ip], dP[ip], m, epr, epi, S, dtdr, rinv, rx, ry, rz, ax[p], ay[p], bx[p], by[p],
                                 g[p][0], g[p][1], g[p][2]);
                force_keep<NC>(g);
            }
        }
#pragma unroll
        for (int p = 0; p < NC; ++p)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) dE[(int64_t)(p * 3 + dd) * nnbr + j] = g[p][dd];
    }
}

// k_force: one wave per environment, lanes are neighbours.  The
// environment's adjoint is staged in LDS as a table Gd[mu][(l, m >= 0), n <
// nrl[l]] (zero where the pooled function is absent; each (l, m) holds only
// the n its degree admits, half the dense [l, m][nr] table at Lmax 5-6), so
// the contraction needs no index lookup.  R_n and dR_n stay in registers (NRM
// of them: 12 when nr <= 12, else 16); the ALP values are produced m-outer,
// l-inner, by the same operations as alp_eval_ed, so only the two previous
// degrees are live.
struct ForceTab {
    int8_t nrl[ACE_LMAX + 1];                            // n bound per l (max pooled n over mu, m)
    int16_t off[(ACE_LMAX + 1) * (ACE_LMAX + 2) / 2];    // first entry of (l, m) within a species' table
    int32_t spstride;                                    // entries per species: NC sum_l (l + 1) nrl[l]
};

template <int L, int PROP, int NRM>
__global__ __launch_bounds__(64) void k_force(RadialParams rp, ForceTab ft, int64_t nenv,
                                               const int32_t* __restrict__ off,
                                               const double* __restrict__ xyz, int64_t nnbr,
                                               const int8_t* __restrict__ species, int nsp,
                                               const int32_t* __restrict__ pool, int npool,
                                               const int32_t* __restrict__ gslot_ptr,
                                               const int32_t* __restrict__ gslot_list,
                                               const int32_t* __restrict__ gcls,
                                               const double2* __restrict__ cls_u,
                                               const double2* __restrict__ Gbuf, int64_t estride,
                                               double* __restrict__ dE) {
    constexpr int NP = (L + 1) * (L + 2) / 2;
    constexpr int NC = PROP == 0 ? 1 : 3;
    extern __shared__ double2 Gd[];  // [nsp][spstride]
    const int lane = threadIdx.x;
    const int64_t env = xcd_env(blockIdx.x, nenv);
    const int nr = rp.nr;
    if (env >= nenv) return;
    const int ntab = nsp * ft.spstride;
    for (int q = lane; q < ntab; q += 64) Gd[q] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int k = lane; k < npool; k += 64) {
        const int mu = max(0, pool[4 * k]), n = pool[4 * k + 1] - 1, l = pool[4 * k + 2], m = pool[4 * k + 3];
        const int d = mu * ft.spstride + ft.off[idxp(l, m)] + n * NC;
        double2 g[NC];
        gather_g<NC>(gslot_ptr, gslot_list, gcls, cls_u, Gbuf, estride, env, k, g);
#pragma unroll
        for (int p = 0; p < NC; ++p) Gd[d + p] = g[p];
    }
    __syncthreads();
    const int j0 = off[env], j1 = off[env + 1];
    for (int j = j0 + lane; j < j1; j += 64) {
        const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
        const int mu = species ? (int)species[j] : 0;
        const double2* Gm = Gd + (int64_t)mu * ft.spstride;
        const Sph S = cart2spher(x, y, z);
        double t, dtdr;
        transform_ed(rp, S.r, t, dtdr);
        double R[NRM], dR[NRM];
        {
            double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
            R[0] = p0;
            dR[0] = d0;
            double p1 = 0.0, d1 = 0.0;
            if (nr > 1) {
                const double al = rp.A[1] * t + rp.B[1];
                p1 = al * p0;
                d1 = al * d0 + rp.A[1] * p0;
            }
            R[1] = p1;
            dR[1] = d1;
#pragma unroll
            for (int n = 2; n < NRM; ++n) {
                if (n < nr) {
                    const double al = rp.A[n] * t + rp.B[n];
                    const double p2 = al * p1 + rp.C[n] * p0;
                    const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                    p0 = p1;
                    p1 = p2;
                    d0 = d1;
                    d1 = d2;
                }
                R[n] = p1;
                dR[n] = d1;
            }
        }
        const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
        const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
        const double st2 = S.sth * S.sth;
        double g[NC][3];
#pragma unroll
        for (int p = 0; p < NC; ++p) g[p][0] = g[p][1] = g[p][2] = 0.0;
        double epr = 0.7071067811865475, epi = 0.0;
        double Pmm = 0.3989422804014327, dPmm = 0.0;  // P_{m,m}, dP_{m,m} (P/sin(theta) for m >= 1)
        const double temp0 = 0.3989422804014327;
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            if (m == 1) {
                Pmm = -1.224744871391589 * temp0;
                dPmm = -1.224744871391589 * (S.cth * temp0);
            } else if (m >= 2) {
                const double f = sqrt(1.0 + 0.5 / m);
                const double nt1 = -f * S.sth * Pmm;
                const double ntd = -f * (S.cth * Pmm * S.sth + S.sth * dPmm);
                Pmm = nt1;
                dPmm = ntd;
            }
            if (m > 0) {
                const double nr_ = epr * S.cphi - epi * S.sphi;
                const double ni_ = epr * S.sphi + epi * S.cphi;
                epr = nr_;
                epi = ni_;
            }
            double Pa = 0.0, dPa = 0.0, Pb = Pmm, dPb = dPmm;  // degrees l-2, l-1 of order m
#pragma unroll
            for (int l = m; l <= L; ++l) {
                double Pl, dPl;
                if (l == m) {
                    Pl = Pmm;
                    dPl = dPmm;
                } else if (l == m + 1) {
                    if (m == 0) {
                        Pl = S.cth * 1.7320508075688772 * temp0;
                        dPl = -S.sth * 1.7320508075688772 * temp0;
                    } else {
                        const double c = sqrt(2.0 * m + 3.0);
                        Pl = c * S.cth * Pmm;
                        dPl = c * (-st2 * Pmm + S.cth * dPmm);
                    }
                } else {
                    const double al = alp_a(l, m), bl = alp_b(l, m);
                    Pl = al * (S.cth * Pb + bl * Pa);
                    dPl = m == 0 ? al * (-S.sth * Pb + S.cth * dPb + bl * dPa)
                                 : al * (-st2 * Pb + S.cth * dPb + bl * dPa);
                }
                if (l > m) {
                    Pa = Pb;
                    dPa = dPb;
                    Pb = Pl;
                    dPb = dPl;
                }
                const double2* Gl = Gm + ft.off[idxp(l, m)];
                const int nb = ft.nrl[l];
                double ax[NC], ay[NC], bx[NC], by[NC];
#pragma unroll
                for (int p = 0; p < NC; ++p) ax[p] = ay[p] = bx[p] = by[p] = 0.0;
#pragma unroll
                for (int n = 0; n < NRM; ++n) {
                    if (n < nb) {
#pragma unroll
                        for (int p = 0; p < NC; ++p) {
                            const double2 G = Gl[n * NC + p];
                            ax[p] = fma(G.x, dR[n], ax[p]);
                            ay[p] = fma(G.y, dR[n], ay[p]);
                            bx[p] = fma(G.x, R[n], bx[p]);
                            by[p] = fma(G.y, R[n], by[p]);
                        }
                    }
                }
#pragma unroll
                for (int p = 0; p < NC; ++p)
                    force_lm_add(Pl, dPl, m, epr, epi, S, dtdr, rinv, rx, ry, rz, ax[p], ay[p], bx[p], by[p], g[p][0],
                                 g[p][1], g[p][2]);
                force_keep<NC>(g);
            }
        }
#pragma unroll
        for (int p = 0; p < NC; ++p)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) dE[(int64_t)(p * 3 + dd) * nnbr + j] = g[p][dd];
    }
}

// k_forcef: forces over the flat neighbour list (Invariant, one species,
// nr <= 16).  Every lane is a neighbour, whatever environment it belongs to,
// so no lane idles on short environments.  A block stages the adjoint G of
// its environments in LDS, MAXE environments per round, as [(l, m, n)][env]
// in the kernel's own (m, l, n) order; each lane computes in the round that
// holds its environment.  Radial values stay in registers and the ALP values
// are produced m-outer exactly as in k_force.

template <int L, int NRM, int MAXE>
__device__ __forceinline__ void forcef_lane(const RadialParams& rp, const ForceFlatTab& ft, int64_t nnbr, int64_t j,
                                            const double* __restrict__ xyz, const double2* __restrict__ Gs,
                                            double* __restrict__ dE) {
    const int nr = rp.nr;
    const double x = xyz[j], y = xyz[nnbr + j], z = xyz[2 * nnbr + j];
    const Sph S = cart2spher(x, y, z);
    double t, dtdr;
    transform_ed(rp, S.r, t, dtdr);
    double R[NRM], dR[NRM];
    {
        double p0 = rp.A[0] * fcut(rp, t), d0 = rp.A[0] * fcut_d(rp, t);
        R[0] = p0;
        dR[0] = d0;
        double p1 = 0.0, d1 = 0.0;
        if (nr > 1) {
            const double al = rp.A[1] * t + rp.B[1];
            p1 = al * p0;
            d1 = al * d0 + rp.A[1] * p0;
        }
        R[1] = p1;
        dR[1] = d1;
#pragma unroll
        for (int n = 2; n < NRM; ++n) {
            if (n < nr) {
                const double al = rp.A[n] * t + rp.B[n];
                const double p2 = al * p1 + rp.C[n] * p0;
                const double d2 = al * d1 + rp.C[n] * d0 + rp.A[n] * p1;
                p0 = p1;
                p1 = p2;
                d0 = d1;
                d1 = d2;
            }
            R[n] = p1;
            dR[n] = d1;
        }
    }
    const double rinv = 1.0 / (S.r + 2.220446049250313e-16);
    const double rx = x / S.r, ry = y / S.r, rz = z / S.r;
    const double st2 = S.sth * S.sth;
    double g[1][3] = {{0.0, 0.0, 0.0}};
    int q = 0;  // running (l, m, n) index of the staged table
    double epr = 0.7071067811865475, epi = 0.0;
    double Pmm = 0.3989422804014327, dPmm = 0.0;
    const double temp0 = 0.3989422804014327;
#pragma unroll
    for (int m = 0; m <= L; ++m) {
        if (m == 1) {
            Pmm = -1.224744871391589 * temp0;
            dPmm = -1.224744871391589 * (S.cth * temp0);
        } else if (m >= 2) {
            const double f = sqrt(1.0 + 0.5 / m);
            const double nt1 = -f * S.sth * Pmm;
            const double ntd = -f * (S.cth * Pmm * S.sth + S.sth * dPmm);
            Pmm = nt1;
            dPmm = ntd;
        }
        if (m > 0) {
            const double nr_ = epr * S.cphi - epi * S.sphi;
            const double ni_ = epr * S.sphi + epi * S.cphi;
            epr = nr_;
            epi = ni_;
        }
        double Pa = 0.0, dPa = 0.0, Pb = Pmm, dPb = dPmm;
#pragma unroll
        for (int l = m; l <= L; ++l) {
            double Pl, dPl;
            if (l == m) {
                Pl = Pmm;
                dPl = dPmm;
            } else if (l == m + 1) {
                if (m == 0) {
                    Pl = S.cth * 1.7320508075688772 * temp0;
                    dPl = -S.sth * 1.7320508075688772 * temp0;
                } else {
                    const double c = sqrt(2.0 * m + 3.0);
                    Pl = c * S.cth * Pmm;
                    dPl = c * (-st2 * Pmm + S.cth * dPmm);
                }
            } else {
                const double al = alp_a(l, m), bl = alp_b(l, m);
                Pl = al * (S.cth * Pb + bl * Pa);
                dPl = m == 0 ? al * (-S.sth * Pb + S.cth * dPb + bl * dPa)
                             : al * (-st2 * Pb + S.cth * dPb + bl * dPa);
            }
            if (l > m) {
                Pa = Pb;
                dPa = dPb;
                Pb = Pl;
                dPb = dPl;
            }
            // a = sum_n G dR_n/dt, b = sum_n G R_n
            const int nb = ft.nrl[l];
            double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
#pragma unroll
            for (int n = 0; n < NRM; ++n) {
                if (n < nb) {
                    const double2 G = Gs[(q + n) * MAXE];
                    ax = fma(G.x, dR[n], ax);
                    ay = fma(G.y, dR[n], ay);
                    bx = fma(G.x, R[n], bx);
                    by = fma(G.y, R[n], by);
                }
            }
            q += nb;
            force_lm_add(Pl, dPl, m, epr, epi, S, dtdr, rinv, rx, ry, rz, ax, ay, bx, by, g[0][0], g[0][1], g[0][2]);
            force_keep<1>(g);
        }
    }
    dE[j] = g[0][0];
    dE[nnbr + j] = g[0][1];
    dE[2 * nnbr + j] = g[0][2];
}

// WIDE = false: the blocks whose environments fit one round (the others
// exit); WIDE = true: the blocks that need several rounds (the others exit).
// Two kernels, so that the common one keeps the registers of a single round.
// BS neighbours per block: 256, or 128 for short environments whose G table
// only fits 16 environments (fewer environments per block, one round).
template <int L, int NRM, int MAXE, bool WIDE, int BS>
__global__ __launch_bounds__(BS, 4) void k_forcef(RadialParams rp, ForceFlatTab ft, int64_t nnbr,
                                                 const double* __restrict__ xyz,
                                                 const int32_t* __restrict__ nbr_env,
                                                 const int32_t* __restrict__ ffq,
                                                 const double2* __restrict__ Gbuf, int64_t estride,
                                                 double* __restrict__ dE) {
    extern __shared__ double2 Gst[];  // [q][MAXE]
    // blocks XCD-contiguous: neighbouring blocks, which stage G rows of the
    // same cache lines (adjacent environments), share one L2
    const int64_t jb = xcd_env(blockIdx.x, gridDim.x) * BS;
    const int64_t j = jb + threadIdx.x;
    const int64_t jl = min(jb + BS - 1, nnbr - 1);
    // the block's environments (neighbours are grouped by environment)
    const int64_t e0 = nbr_env[jb];
    const int ne = (int)(nbr_env[jl] - e0 + 1);
    if ((ne > MAXE) != WIDE) return;  // the other kernel's block: before any barrier
    const int el = j < nnbr ? (int)(nbr_env[j] - e0) : -1;
    for (int r0 = 0; r0 < ne; r0 += MAXE) {
        const int nre = min(MAXE, ne - r0);
        if (r0 > 0) __syncthreads();  // the previous round's reads are done
        for (int i = threadIdx.x; i < ft.nq * nre; i += BS) {
            const int qq = i / nre, ee = i - qq * nre;
            Gst[qq * MAXE + ee] = Gbuf[(int64_t)ffq[qq] * estride + e0 + r0 + ee];
        }
        __syncthreads();
        if (!WIDE) {
            if (el < 0) return;  // no barriers below
            forcef_lane<L, NRM, MAXE>(rp, ft, nnbr, j, xyz, Gst + el, dE);
            return;
        }
        if (el >= r0 && el < r0 + nre) forcef_lane<L, NRM, MAXE>(rp, ft, nnbr, j, xyz, Gst + (el - r0), dE);
    }
}

// G[primary slot] += G[extra slot] for pooled functions whose gradient
// polynomial was split over several tasks (k_forcef reads one slot per k).
__global__ void k_gfold(int64_t nenv, int nfold, const int32_t* __restrict__ fold, double2* __restrict__ Gbuf,
                        int64_t estride) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nenv) return;
    for (int q = 0; q < nfold; ++q) {
        double2* dst = Gbuf + (int64_t)fold[2 * q] * estride + e;
        const double2 v = Gbuf[(int64_t)fold[2 * q + 1] * estride + e];
        dst->x += v.x;
        dst->y += v.y;
    }
}

template <int L, int PROP>
static void launch_force(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off,
                         const double* xyz, int64_t nnbr, const int8_t* species, int64_t estride,
                         const Workspace& ws, double* dE) {
    constexpr int NC = PROP == 0 ? 1 : 3;
    constexpr int NP = (L + 1) * (L + 2) / 2;
    if (PROP == 0 && ctx->flat_force) {
        if (ctx->nfold > 0)
            hipLaunchKernelGGL(k_gfold, dim3((unsigned)((nenv + 255) / 256)), dim3(256), 0, st, nenv, ctx->nfold,
                               ctx->gfold.p, ws.Gbuf, estride);
        // radial functions held in registers: 12 when they suffice (fewer
        // registers, more waves per SIMD), else 16
        // blocks of short environments span more of them: stage 32 when that fits
        const bool wide = (size_t)ctx->fflat.nq * FF_MAXE_WIDE * sizeof(double2) <= FF_LDS;
        const size_t glds = (size_t)ctx->fflat.nq * (wide ? FF_MAXE_WIDE : FF_MAXE) * sizeof(double2);
        // short environments (fewer than 18 neighbours on average: a 256-neighbour
        // block would span more than ~14) with a 16-environment stage:
        // 128-neighbour blocks, so that a block's environments fit one round
        const bool half = !wide && nnbr < (int64_t)18 * nenv;
        const int32_t* ne_ = ws.nbr_env;
#define ACE_FORCEF(NRM_, MAXE_, BS_)                                                                           \
    do {                                                                                                        \
        const dim3 grid((unsigned)((nnbr + BS_ - 1) / BS_));                                                    \
        hipLaunchKernelGGL((k_forcef<L, NRM_, MAXE_, false, BS_>), grid, dim3(BS_), glds, st,                   \
                           radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);   \
        hipLaunchKernelGGL((k_forcef<L, NRM_, MAXE_, true, BS_>), grid, dim3(BS_), glds, st,                    \
                           radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);   \
    } while (0)
        if (ctx->nr <= 12 && wide)
            ACE_FORCEF(12, FF_MAXE_WIDE, 256);
        else if (ctx->nr <= 12 && half)
            ACE_FORCEF(12, FF_MAXE, 128);
        else if (ctx->nr <= 12)
            ACE_FORCEF(12, FF_MAXE, 256);
        else if (wide)
            ACE_FORCEF(FORCE_NRMAX, FF_MAXE_WIDE, 256);
        else if (half)
            ACE_FORCEF(FORCE_NRMAX, FF_MAXE, 128);
        else
            ACE_FORCEF(FORCE_NRMAX, FF_MAXE, 256);
#undef ACE_FORCEF
        return;
    }
    const int nsp = std::max(1, ctx->nspecies);
    const int32_t* gcls = ctx->classed ? ctx->gslot_cls.p : nullptr;
    // the per-environment table: each (l, m >= 0) holds the n < nrl[l]
    ForceTab ft;
    std::memset(&ft, 0, sizeof(ft));
    int ent = 0;
    for (int l = 0; l <= ACE_LMAX; ++l) {
        ft.nrl[l] = (int8_t)(l <= L ? ctx->nrl[l] : 0);
        for (int m = 0; m <= l && l <= L; ++m) {
            ft.off[idxp(l, m)] = (int16_t)ent;
            ent += ctx->nrl[l] * NC;
        }
    }
    ft.spstride = ent;
    const size_t table = (size_t)nsp * ent * sizeof(double2);
    if (ctx->nr <= FORCE_NRMAX && table <= 64 * 1024) {
#define ACE_FORCE(NRM_)                                                                                          \
    do {                                                                                                          \
        set_func_lds((const void*)k_force<L, PROP, NRM_>, table);                                                 \
        hipLaunchKernelGGL((k_force<L, PROP, NRM_>), dim3((unsigned)nenv), dim3(64), table, st, radial_params(ctx), \
                           ft, nenv, off, xyz, nnbr, ctx->nspecies > 0 ? species : nullptr, nsp, ctx->pool.p,     \
                           ctx->npool, ctx->gslot_ptr.p, ctx->gslot_list.p, gcls, ctx->cls_u.p, ws.Gbuf, estride, \
                           dE);                                                                                   \
    } while (0)
        // radial functions held in registers: 12 when they suffice, else 16
        if (ctx->nr <= 12)
            ACE_FORCE(12);
        else
            ACE_FORCE(FORCE_NRMAX);
#undef ACE_FORCE
        return;
    }
    const size_t lds = (size_t)ctx->npool * NC * sizeof(double2) + 64 * 2 * ctx->nr * sizeof(double);
    set_func_lds((const void*)k_force_lds<L, PROP>, lds);
    hipLaunchKernelGGL((k_force_lds<L, PROP>), dim3((unsigned)nenv), dim3(64), lds, st, radial_params(ctx),
                       nenv, off, xyz, nnbr, ctx->nspecies > 0 ? species : nullptr, ctx->kidx.p,
                       ctx->npool, ctx->gslot_ptr.p, ctx->gslot_list.p, gcls, ctx->cls_u.p, ws.Gbuf, estride, dE);
}

}  // namespace

template <int L>
int launch_force_l(ace_ctx* ctx, hipStream_t st, int64_t nenv, const int32_t* off, const double* xyz,
                   int64_t nnbr, const int8_t* species, int64_t estride, const Workspace& ws, double* dE) {
    if (ctx->prop == ACE_PROP_INVARIANT)
        launch_force<L, 0>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
    else
        launch_force<L, 1>(ctx, st, nenv, off, xyz, nnbr, species, estride, ws, dE);
    HIP_TRY(hipGetLastError());
    return 0;
}
