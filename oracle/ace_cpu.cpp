// CPU restatement of ACE.jl's ProductEvaluator — TEST INFRASTRUCTURE ONLY.
//
// A C++/OpenMP port of the reference algorithm for the Invariant and
// EuclideanVector LinearACEModel, used as the second oracle and as the timed
// CPU baseline
// (cpu_baseline.kind = "port").  It follows ACE.jl v0.12.46 loop by loop:
//   * Rn: polytransform + OrthPolyBasis evaluate! / evaluate_ed!
//     (src/transforms/distancetransforms.jl:73; src/polynomials/orthpolys.jl:
//     250-306), gradient dP * t'(r) * r/|r| (src/chain.jl:40-52)
//   * Ylm: cart2spher with atan + sincos, ALP evaluate! / _evaluate_ed!, cYlm! /
//     cYlm_ed!, dspher_to_dcart (src/polynomials/sphericalharmonics.jl:43-65,
//     167-267, 379-443)
//   * A and dA per neighbour: _add_into_A_dA! (src/product_1pbasis.jl:169-244)
//   * energy: evaluator.jl:121-147 (naive left-to-right products per AA row);
//     EuclideanVector: the three components Re(AA c~_p), c~ complex
//     (properties.jl:165-248)
//   * forces: _rrule_evaluate with _AA_local_adjoints! (evaluator.jl:161-200,
//     pibasis.jl:338-390): dAco, then g_j = sum_iA Re(dAco[iA] dA[iA, j])
//     (per component p for EuclideanVector: g[j][p][d]).
// One environment per OpenMP iteration, schedule(dynamic).
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

using cplx = std::complex<double>;

extern "C" {

struct acecpu_tables {
    int32_t nr, tp, pl, pr, L;
    double r0, tl, tr;
    const double* rA;
    const double* rB;
    const double* rC;
    int32_t nA;
    const int32_t* aspec;  // [nA][4] (mu, n 1-based, l, m)
    int32_t nAA, maxorder;
    const int32_t* orders;
    const int32_t* iAA2iA;  // [nAA][maxorder], 0-based
    const double* ctilde;   // [nAA] (ncomp 1) or [nAA][3][re, im] (ncomp 3)
    int32_t ncomp;          // 1 Invariant, 3 EuclideanVector
};

}  // extern "C"

namespace {

inline int idxp(int l, int m) { return m + l * (l + 1) / 2; }
inline int idxy(int l, int m) { return m + l + l * l; }

struct Work {
    std::vector<double> P, dP, R, dR;
    std::vector<cplx> Y;
    std::vector<cplx> dY;  // [nY][3]
    std::vector<cplx> A, dA, dAco, dAAdA;  // dAco [nA][ncomp]
    std::vector<double> alpA, alpB;
};

void alp_coeffs(int L, std::vector<double>& A, std::vector<double>& B) {
    const int np = (L + 1) * (L + 2) / 2;
    A.assign(np, 0.0);
    B.assign(np, 0.0);
    for (int l = 2; l <= L; ++l) {
        const double ls = (double)l * l, lm1s = (double)(l - 1) * (l - 1);
        for (int m = 0; m <= l - 2; ++m) {
            const double ms = (double)m * m;
            A[idxp(l, m)] = std::sqrt((4 * ls - 1.0) / (ls - ms));
            B[idxp(l, m)] = -std::sqrt((lm1s - ms) / (4 * lm1s - 1.0));
        }
    }
}

inline double ipow(double x, int p) {
    double y = 1.0;
    for (int i = 0; i < p; ++i) y *= x;
    return y;
}

struct Sph {
    double r, cphi, sphi, cth, sth;
};

inline Sph cart2spher(const double* R) {
    Sph s;
    s.r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const double phi = std::atan2(R[1], R[0]);
    s.sphi = std::sin(phi);
    s.cphi = std::cos(phi);
    s.cth = R[2] / s.r;
    s.sth = std::sqrt(R[0] * R[0] + R[1] * R[1]) / s.r;
    return s;
}

void radial(const acecpu_tables& T, double r, bool grad, Work& w) {
    const double u = (1.0 + T.r0) / (1.0 + r);
    const double t = ipow(u, T.tp);
    double fc = 0.0, fcd = 0.0;
    if (!((T.pl > 0 && t < T.tl) || (T.pr > 0 && t > T.tr))) {
        fc = ipow(t - T.tl, T.pl) * ipow(t - T.tr, T.pr);
        if (T.pl > 0) fcd += T.pl * ipow(t - T.tl, T.pl - 1) * ipow(t - T.tr, T.pr);
        if (T.pr > 0) fcd += ipow(t - T.tl, T.pl) * T.pr * ipow(t - T.tr, T.pr - 1);
    }
    const int N = T.nr;
    w.R[0] = T.rA[0] * fc;
    if (!grad) {
        if (N > 1) w.R[1] = (T.rA[1] * t + T.rB[1]) * w.R[0];
        for (int n = 2; n < N; ++n) w.R[n] = (T.rA[n] * t + T.rB[n]) * w.R[n - 1] + T.rC[n] * w.R[n - 2];
        return;
    }
    w.dR[0] = T.rA[0] * fcd;
    if (N > 1) {
        const double al = T.rA[1] * t + T.rB[1];
        w.R[1] = al * w.R[0];
        w.dR[1] = al * w.dR[0] + T.rA[1] * w.R[0];
    }
    for (int n = 2; n < N; ++n) {
        const double al = T.rA[n] * t + T.rB[n];
        w.R[n] = al * w.R[n - 1] + T.rC[n] * w.R[n - 2];
        w.dR[n] = al * w.dR[n - 1] + T.rC[n] * w.dR[n - 2] + T.rA[n] * w.R[n - 1];
    }
    if (grad) {
        const double du = -(1.0 + T.r0) / ((1.0 + r) * (1.0 + r));
        const double dt = T.tp * ipow(u, T.tp - 1) * du;
        for (int n = 0; n < N; ++n) w.dR[n] *= dt;  // dR/dr; times rhat later
    }
}

void ylm(int L, const Sph& S, Work& w) {
    double* P = w.P.data();
    double temp = std::sqrt(0.5 / M_PI);
    P[0] = temp;
    if (L > 0) {
        P[idxp(1, 0)] = S.cth * std::sqrt(3.0) * temp;
        temp = -std::sqrt(1.5) * S.sth * temp;
        P[idxp(1, 1)] = temp;
        for (int l = 2; l <= L; ++l) {
            for (int m = 0; m <= l - 2; ++m)
                P[idxp(l, m)] = w.alpA[idxp(l, m)] * (S.cth * P[idxp(l - 1, m)] + w.alpB[idxp(l, m)] * P[idxp(l - 2, m)]);
            P[idxp(l, l - 1)] = S.cth * std::sqrt(2.0 * (l - 1) + 3) * temp;
            temp = -std::sqrt(1.0 + 0.5 / l) * S.sth * temp;
            P[idxp(l, l)] = temp;
        }
    }
    cplx ep(1.0 / std::sqrt(2.0), 0.0);
    for (int l = 0; l <= L; ++l) w.Y[idxy(l, 0)] = P[idxp(l, 0)] * ep;
    double sig = 1.0;
    const cplx fac(S.cphi, S.sphi);
    for (int m = 1; m <= L; ++m) {
        sig = -sig;
        ep *= fac;
        const cplx em = sig * std::conj(ep);
        for (int l = m; l <= L; ++l) {
            const double p = P[idxp(l, m)];
            w.Y[idxy(l, -m)] = em * p;
            w.Y[idxy(l, m)] = ep * p;
        }
    }
}

inline void dspher_to_dcart(const Sph& S, cplx fphi, cplx fth, cplx* out) {
    const double r = S.r + 2.220446049250313e-16;
    out[0] = (-(S.sphi * fphi) + (S.cphi * S.cth * fth)) / r;
    out[1] = ((S.cphi * fphi) + (S.sphi * S.cth * fth)) / r;
    out[2] = (-(S.sth * fth)) / r;
}

void ylm_ed(int L, const Sph& S, Work& w) {
    double* P = w.P.data();
    double* dP = w.dP.data();
    double temp = std::sqrt(0.5 / M_PI);
    P[0] = temp;
    dP[0] = 0.0;
    double temp_d = 0.0;
    if (L > 0) {
        const double s3 = std::sqrt(3.0);
        P[idxp(1, 0)] = S.cth * s3 * temp;
        dP[idxp(1, 0)] = -S.sth * s3 * temp + S.cth * s3 * temp_d;
        double temp1 = -std::sqrt(1.5) * temp;
        temp_d = -std::sqrt(1.5) * (S.cth * temp + S.sth * temp_d);
        P[idxp(1, 1)] = temp1;
        dP[idxp(1, 1)] = temp_d;
        const double st2 = S.sth * S.sth;
        for (int l = 2; l <= L; ++l) {
            int i = idxp(l, 0), i1 = idxp(l - 1, 0), i2 = idxp(l - 2, 0);
            P[i] = w.alpA[i] * (S.cth * P[i1] + w.alpB[i] * P[i2]);
            dP[i] = w.alpA[i] * (-S.sth * P[i1] + S.cth * dP[i1] + w.alpB[i] * dP[i2]);
            for (int m = 1; m <= l - 2; ++m) {
                i = idxp(l, m), i1 = idxp(l - 1, m), i2 = idxp(l - 2, m);
                P[i] = w.alpA[i] * (S.cth * P[i1] + w.alpB[i] * P[i2]);
                dP[i] = w.alpA[i] * (-st2 * P[i1] + S.cth * dP[i1] + w.alpB[i] * dP[i2]);
            }
            const double c = std::sqrt(2.0 * (l - 1) + 3);
            P[idxp(l, l - 1)] = c * S.cth * temp1;
            dP[idxp(l, l - 1)] = c * (-st2 * temp1 + S.cth * temp_d);
            const double f = std::sqrt(1.0 + 0.5 / l);
            const double nt1 = -f * S.sth * temp1;
            const double ntd = -f * (S.cth * temp1 * S.sth + S.sth * temp_d);
            temp1 = nt1;
            temp_d = ntd;
            P[idxp(l, l)] = temp1;
            dP[idxp(l, l)] = temp_d;
        }
    }
    const double isq2 = 1.0 / std::sqrt(2.0);
    for (int l = 0; l <= L; ++l) {
        w.Y[idxy(l, 0)] = P[idxp(l, 0)] * isq2;
        dspher_to_dcart(S, 0.0, dP[idxp(l, 0)] * isq2, &w.dY[3 * idxy(l, 0)]);
    }
    double sig = 1.0;
    cplx ep(isq2, 0.0);
    const cplx fac(S.cphi, S.sphi);
    for (int m = 1; m <= L; ++m) {
        sig = -sig;
        ep *= fac;
        const cplx em = sig * std::conj(ep);
        const cplx dep = cplx(0, m) * ep, dem = cplx(0, -m) * em;
        for (int l = m; l <= L; ++l) {
            const double pds = P[idxp(l, m)];
            w.Y[idxy(l, -m)] = em * pds * S.sth;
            w.Y[idxy(l, m)] = ep * pds * S.sth;
            const double dpt = dP[idxp(l, m)];
            dspher_to_dcart(S, dem * pds, em * dpt, &w.dY[3 * idxy(l, -m)]);
            dspher_to_dcart(S, dep * pds, ep * dpt, &w.dY[3 * idxy(l, m)]);
        }
    }
}

void env_eval(const acecpu_tables& T, int nX, const double* R, const int8_t* sp, bool forces, Work& w,
              double* E, double* g) {
    const int nA = T.nA;
    std::fill(w.A.begin(), w.A.end(), cplx(0, 0));
    if (forces) w.dA.assign((size_t)nA * nX * 3, cplx(0, 0));
    for (int j = 0; j < nX; ++j) {
        const double* rr = R + 3 * j;
        const Sph S = cart2spher(rr);
        radial(T, S.r, forces, w);
        if (forces)
            ylm_ed(T.L, S, w);
        else
            ylm(T.L, S, w);
        const double rh[3] = {rr[0] / S.r, rr[1] / S.r, rr[2] / S.r};
        for (int iA = 0; iA < nA; ++iA) {
            const int mu = T.aspec[4 * iA], n = T.aspec[4 * iA + 1] - 1, l = T.aspec[4 * iA + 2],
                      m = T.aspec[4 * iA + 3];
            if (mu >= 0 && sp && sp[j] != mu) continue;
            const int iy = idxy(l, m);
            w.A[iA] += w.R[n] * w.Y[iy];
            if (forces) {
                cplx* d = &w.dA[((size_t)iA * nX + j) * 3];
                for (int q = 0; q < 3; ++q) d[q] = (w.dR[n] * rh[q]) * w.Y[iy] + w.dY[3 * iy + q] * w.R[n];
            }
        }
    }
    // energy (evaluator.jl:121-147)
    const int NC = T.ncomp;
    auto coef = [&](int r, int p) {
        return NC == 1 ? cplx(T.ctilde[r], 0.0) : cplx(T.ctilde[6 * (size_t)r + 2 * p], T.ctilde[6 * (size_t)r + 2 * p + 1]);
    };
    double val[3] = {0.0, 0.0, 0.0};
    int i0 = 0;
    if (T.nAA > 0 && T.orders[0] == 0) {
        for (int p = 0; p < NC; ++p) val[p] += coef(0, p).real();
        i0 = 1;
    }
    for (int r = i0; r < T.nAA; ++r) {
        const int* ix = T.iAA2iA + (size_t)r * T.maxorder;
        cplx aa = w.A[ix[0]];
        for (int t = 1; t < T.orders[r]; ++t) aa *= w.A[ix[t]];
        if (NC == 1)
            val[0] += aa.real() * T.ctilde[r];
        else
            for (int p = 0; p < 3; ++p) val[p] += (aa * coef(r, p)).real();
    }
    for (int p = 0; p < NC; ++p) E[p] = val[p];
    if (!forces) return;
    // dAco (evaluator.jl:174-185, pibasis.jl:338-390)
    std::fill(w.dAco.begin(), w.dAco.end(), cplx(0, 0));
    cplx* d = w.dAAdA.data();
    for (int r = i0; r < T.nAA; ++r) {
        const int ord = T.orders[r];
        const int* ix = T.iAA2iA + (size_t)r * T.maxorder;
        if (ord == 1) {
            d[0] = 1.0;
        } else if (ord == 2) {
            d[0] = w.A[ix[1]];
            d[1] = w.A[ix[0]];
        } else {
            const cplx A1 = w.A[ix[0]], A2 = w.A[ix[1]];
            d[0] = 1.0;
            d[1] = A1;
            cplx fwd = A1 * A2;
            for (int a = 2; a < ord - 1; ++a) {
                d[a] = fwd;
                fwd *= w.A[ix[a]];
            }
            d[ord - 1] = fwd;
            cplx bwd = w.A[ix[ord - 1]];
            for (int a = ord - 2; a >= 2; --a) {
                d[a] *= bwd;
                bwd *= w.A[ix[a]];
            }
            d[1] *= bwd;
            bwd *= A2;
            d[0] *= bwd;
        }
        if (NC == 1)
            for (int t = 0; t < ord; ++t) w.dAco[ix[t]] += d[t] * T.ctilde[r];
        else
            for (int t = 0; t < ord; ++t)
                for (int p = 0; p < 3; ++p) w.dAco[3 * (size_t)ix[t] + p] += d[t] * coef(r, p);
    }
    for (int j = 0; j < nX; ++j)
        for (int p = 0; p < NC; ++p) {
            double gx = 0, gy = 0, gz = 0;
            for (int iA = 0; iA < nA; ++iA) {
                const cplx* da = &w.dA[((size_t)iA * nX + j) * 3];
                const cplx dc = w.dAco[(size_t)NC * iA + p];
                gx += (dc * da[0]).real();
                gy += (dc * da[1]).real();
                gz += (dc * da[2]).real();
            }
            double* gj = g + 3 * ((size_t)NC * j + p);
            gj[0] = gx;
            gj[1] = gy;
            gj[2] = gz;
        }
}

}  // namespace

extern "C" int acecpu_eval(const acecpu_tables* T, int64_t nenv, const int32_t* off, const double* R,
                           const int8_t* species, double* E, double* g, int forces, int nthreads) {
    if (!T || nenv < 0 || (T->ncomp != 1 && T->ncomp != 3)) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        Work w;
        const int np = (T->L + 1) * (T->L + 2) / 2;
        w.P.assign(np, 0.0);
        w.dP.assign(np, 0.0);
        w.R.assign(T->nr, 0.0);
        w.dR.assign(T->nr, 0.0);
        w.Y.assign((T->L + 1) * (T->L + 1), cplx(0, 0));
        w.dY.assign(3 * (T->L + 1) * (T->L + 1), cplx(0, 0));
        w.A.assign(T->nA, cplx(0, 0));
        w.dAco.assign((size_t)T->nA * T->ncomp, cplx(0, 0));
        w.dAAdA.assign(T->maxorder + 1, cplx(0, 0));
        alp_coeffs(T->L, w.alpA, w.alpB);
#pragma omp for schedule(dynamic, 16)
        for (int64_t e = 0; e < nenv; ++e) {
            const int j0 = off[e], nX = off[e + 1] - off[e];
            env_eval(*T, nX, R + 3 * (size_t)j0, species ? species + j0 : nullptr, forces != 0, w,
                     E + (size_t)T->ncomp * e, forces ? g + 3 * (size_t)T->ncomp * j0 : nullptr);
        }
    }
    return 0;
}

extern "C" int acecpu_max_threads(void) { return omp_get_max_threads(); }
