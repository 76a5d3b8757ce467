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
                // Y and its Cartesian gradient (cYlm_ed!, dspher_to_dcart)
                double Yr, Yi, fpr, fpi, ftr, fti;
                if (m == 0) {
                    Yr = P[ip] * 0.7071067811865475;
                    Yi = 0.0;
                    fpr = fpi = 0.0;
                    ftr = dP[ip] * 0.7071067811865475;
                    fti = 0.0;
                } else {
                    const double pds = P[ip];
                    Yr = epr * pds * S.sth;
                    Yi = epi * pds * S.sth;
                    // dep/dphi = i m ep
                    fpr = -m * epi * pds;
                    fpi = m * epr * pds;
                    ftr = epr * dP[ip];
                    fti = epi * dP[ip];
                }
                const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                       (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                       (-S.sth * ftr) * rinv};
                const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                       (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                       (-S.sth * fti) * rinv};
                const int16_t* kr = kidx + ((int64_t)(mu * (L + 1) + l) * (L + 1) + m) * nr;
                for (int n = 0; n < nr; ++n) {
                    const int k = kr[n];
                    if (k < 0) continue;
                    const double R = myR[n], dR = myR[nr + n] * dtdr;
#pragma unroll
                    for (int p = 0; p < NC; ++p) {
                        const double2 G = Gs[k * NC + p];
                        // Re(G * (dR rhat Y + R dY))
                        const double gy = G.x * Yr - G.y * Yi;
                        g[p][0] += dR * gy * rx + R * (G.x * dYr[0] - G.y * dYi[0]);
                        g[p][1] += dR * gy * ry + R * (G.x * dYr[1] - G.y * dYi[1]);
                        g[p][2] += dR * gy * rz + R * (G.x * dYr[2] - G.y * dYi[2]);
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < NC; ++p)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) dE[(int64_t)(p * 3 + dd) * nnbr + j] = g[p][dd];
    }
}

// k_force: one wave per environment, lanes are neighbours.  The
// environment's adjoint is staged in LDS as a dense table
// Gd[mu][l, m >= 0][n] (zero where the pooled function is absent), so the
// contraction needs no index lookup.  R_n and dR_n stay in registers (nr <=
// 16); the ALP values are produced m-outer, l-inner, by the same operations
// as alp_eval_ed, so only the two previous degrees are live.
struct ForceTab {
    int8_t nrl[ACE_LMAX + 1];  // n bound per l (max pooled n over mu, m)
};

template <int L, int PROP>
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
    extern __shared__ double2 Gd[];  // [nsp][NP][nr][NC]
    const int lane = threadIdx.x;
    const int64_t env = xcd_env(blockIdx.x, nenv);
    const int nr = rp.nr;
    if (env >= nenv) return;
    const int ntab = nsp * NP * nr * NC;
    for (int q = lane; q < ntab; q += 64) Gd[q] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int k = lane; k < npool; k += 64) {
        const int mu = max(0, pool[4 * k]), n = pool[4 * k + 1] - 1, l = pool[4 * k + 2], m = pool[4 * k + 3];
        const int d = ((mu * NP + idxp(l, m)) * nr + n) * NC;
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
        const double2* Gm = Gd + (int64_t)mu * NP * nr * NC;
        const Sph S = cart2spher(x, y, z);
        double t, dtdr;
        transform_ed(rp, S.r, t, dtdr);
        double R[FORCE_NRMAX], dR[FORCE_NRMAX];
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
            for (int n = 2; n < FORCE_NRMAX; ++n) {
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
                    const int i = idxp(l, m);
                    Pl = rp.alpA[i] * (S.cth * Pb + rp.alpB[i] * Pa);
                    dPl = m == 0 ? rp.alpA[i] * (-S.sth * Pb + S.cth * dPb + rp.alpB[i] * dPa)
                                 : rp.alpA[i] * (-st2 * Pb + S.cth * dPb + rp.alpB[i] * dPa);
                }
                if (l > m) {
                    Pa = Pb;
                    dPa = dPb;
                    Pb = Pl;
                    dPb = dPl;
                }
                // Y and its Cartesian gradient (cYlm_ed!, dspher_to_dcart)
                double Yr, Yi, fpr, fpi, ftr, fti;
                if (m == 0) {
                    Yr = Pl * 0.7071067811865475;
                    Yi = 0.0;
                    fpr = fpi = 0.0;
                    ftr = dPl * 0.7071067811865475;
                    fti = 0.0;
                } else {
                    Yr = epr * Pl * S.sth;
                    Yi = epi * Pl * S.sth;
                    fpr = -m * epi * Pl;
                    fpi = m * epr * Pl;
                    ftr = epr * dPl;
                    fti = epi * dPl;
                }
                const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                       (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                       (-S.sth * ftr) * rinv};
                const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                       (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                       (-S.sth * fti) * rinv};
                const double2* Gl = Gm + idxp(l, m) * nr * NC;
                const int nb = ft.nrl[l];
#pragma unroll
                for (int n = 0; n < FORCE_NRMAX; ++n) {
                    if (n < nb) {
                        const double Rn = R[n], dRn = dR[n] * dtdr;
#pragma unroll
                        for (int p = 0; p < NC; ++p) {
                            const double2 G = Gl[n * NC + p];
                            // Re(G * (dR rhat Y + R dY))
                            const double gy = G.x * Yr - G.y * Yi;
                            g[p][0] += dRn * gy * rx + Rn * (G.x * dYr[0] - G.y * dYi[0]);
                            g[p][1] += dRn * gy * ry + Rn * (G.x * dYr[1] - G.y * dYi[1]);
                            g[p][2] += dRn * gy * rz + Rn * (G.x * dYr[2] - G.y * dYi[2]);
                        }
                    }
                }
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
// so no lane idles on short environments.  A lane reads its environment's
// adjoint G[slot][env] straight from memory (lanes of one environment read the
// same address; neighbouring environments are adjacent), with the slot of
// every (l, m, n) in the kernel arguments.  Radial values stay in registers
// and the ALP values are produced m-outer exactly as in k_force.

template <int L, int NRM, bool STAGED, int MAXE>
__device__ __forceinline__ void forcef_lane(const RadialParams& rp, const ForceFlatTab& ft, int64_t nnbr, int64_t j,
                                            const double* __restrict__ xyz, const double2* __restrict__ Ge,
                                            const double2* __restrict__ Gs, int64_t estride,
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
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
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
                const int i = idxp(l, m);
                Pl = rp.alpA[i] * (S.cth * Pb + rp.alpB[i] * Pa);
                dPl = m == 0 ? rp.alpA[i] * (-S.sth * Pb + S.cth * dPb + rp.alpB[i] * dPa)
                             : rp.alpA[i] * (-st2 * Pb + S.cth * dPb + rp.alpB[i] * dPa);
            }
            if (l > m) {
                Pa = Pb;
                dPa = dPb;
                Pb = Pl;
                dPb = dPl;
            }
            double Yr, Yi, fpr, fpi, ftr, fti;
            if (m == 0) {
                Yr = Pl * 0.7071067811865475;
                Yi = 0.0;
                fpr = fpi = 0.0;
                ftr = dPl * 0.7071067811865475;
                fti = 0.0;
            } else {
                Yr = epr * Pl * S.sth;
                Yi = epi * Pl * S.sth;
                fpr = -m * epi * Pl;
                fpi = m * epr * Pl;
                ftr = epr * dPl;
                fti = epi * dPl;
            }
            const double dYr[3] = {(-S.sphi * fpr + S.cphi * S.cth * ftr) * rinv,
                                   (S.cphi * fpr + S.sphi * S.cth * ftr) * rinv,
                                   (-S.sth * ftr) * rinv};
            const double dYi[3] = {(-S.sphi * fpi + S.cphi * S.cth * fti) * rinv,
                                   (S.cphi * fpi + S.sphi * S.cth * fti) * rinv,
                                   (-S.sth * fti) * rinv};
            const int16_t* sl = ft.slot[idxp(l, m)];
            const int nb = ft.nrl[l];
#pragma unroll
            for (int n = 0; n < NRM; ++n) {
                if (n < nb) {
                    const double2 G = STAGED ? Gs[(q + n) * MAXE] : Ge[(int64_t)sl[n] * estride];
                    const double Rn = R[n], dRn = dR[n] * dtdr;
                    const double gy = G.x * Yr - G.y * Yi;
                    g0 += dRn * gy * rx + Rn * (G.x * dYr[0] - G.y * dYi[0]);
                    g1 += dRn * gy * ry + Rn * (G.x * dYr[1] - G.y * dYi[1]);
                    g2 += dRn * gy * rz + Rn * (G.x * dYr[2] - G.y * dYi[2]);
                }
            }
            q += nb;
        }
    }
    dE[j] = g0;
    dE[nnbr + j] = g1;
    dE[2 * nnbr + j] = g2;
}

template <int L, int NRM, int MAXE>
__global__ __launch_bounds__(256, 4) void k_forcef(RadialParams rp, ForceFlatTab ft, int64_t nnbr,
                                                 const double* __restrict__ xyz,
                                                 const int32_t* __restrict__ nbr_env,
                                                 const int32_t* __restrict__ ffq,
                                                 const double2* __restrict__ Gbuf, int64_t estride,
                                                 double* __restrict__ dE) {
    extern __shared__ double2 Gst[];  // [q][MAXE]
    const int64_t jb = (int64_t)blockIdx.x * 256;
    const int64_t j = jb + threadIdx.x;
    const int64_t jl = min(jb + 255, nnbr - 1);
    // the block's environments (neighbours are grouped by environment)
    const int64_t e0 = nbr_env[jb];
    const int ne = (int)(nbr_env[jl] - e0 + 1);
    if (ne <= MAXE) {
        for (int i = threadIdx.x; i < ft.nq * ne; i += 256) {
            const int qq = i / ne, el = i - qq * ne;
            Gst[qq * MAXE + el] = Gbuf[(int64_t)ffq[qq] * estride + e0 + el];
        }
        __syncthreads();
        if (j >= nnbr) return;  // no barriers below
        const int el = (int)(nbr_env[j] - e0);
        forcef_lane<L, NRM, true, MAXE>(rp, ft, nnbr, j, xyz, nullptr, Gst + el, estride, dE);
    } else {
        if (j >= nnbr) return;
        forcef_lane<L, NRM, false, MAXE>(rp, ft, nnbr, j, xyz, Gbuf + nbr_env[j], nullptr, estride, dE);
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
        const dim3 grid((unsigned)((nnbr + 255) / 256));
        const int32_t* ne_ = ws.nbr_env;
        if (ctx->nr <= 12 && wide)
            hipLaunchKernelGGL((k_forcef<L, 12, FF_MAXE_WIDE>), grid, dim3(256), glds, st, radial_params(ctx),
                               ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);
        else if (ctx->nr <= 12)
            hipLaunchKernelGGL((k_forcef<L, 12, FF_MAXE>), grid, dim3(256), glds, st, radial_params(ctx),
                               ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);
        else if (wide)
            hipLaunchKernelGGL((k_forcef<L, FORCE_NRMAX, FF_MAXE_WIDE>), grid, dim3(256), glds, st,
                               radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);
        else
            hipLaunchKernelGGL((k_forcef<L, FORCE_NRMAX, FF_MAXE>), grid, dim3(256), glds, st,
                               radial_params(ctx), ctx->fflat, nnbr, xyz, ne_, ctx->ffq.p, ws.Gbuf, estride, dE);
        return;
    }
    const int nsp = std::max(1, ctx->nspecies);
    const int32_t* gcls = ctx->classed ? ctx->gslot_cls.p : nullptr;
    const size_t dense = (size_t)nsp * NP * ctx->nr * NC * sizeof(double2);
    if (ctx->nr <= FORCE_NRMAX && dense <= 64 * 1024) {
        ForceTab ft;
        for (int l = 0; l <= ACE_LMAX; ++l) ft.nrl[l] = (int8_t)(l <= L ? ctx->nrl[l] : 0);
        set_func_lds((const void*)k_force<L, PROP>, dense);
        hipLaunchKernelGGL((k_force<L, PROP>), dim3((unsigned)nenv), dim3(64), dense, st, radial_params(ctx), ft,
                           nenv, off, xyz, nnbr, ctx->nspecies > 0 ? species : nullptr, nsp, ctx->pool.p,
                           ctx->npool, ctx->gslot_ptr.p, ctx->gslot_list.p, gcls, ctx->cls_u.p, ws.Gbuf, estride,
                           dE);
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
