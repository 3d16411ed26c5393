// chem_device.hpp -- per-reaction rate evaluation shared by the wave-per-cell integrator (chem.hip)
// and the lane-per-cell explicit screen (chem_lane.hip).
#pragma once

#include "chem.hpp"
#include "fcfd_device.hpp"

namespace fcfd {

// x**n by binary powering (the product order of the CPU restatement); negative n as 1 / x**(-n).
// Written without recursion so it inlines: a call from a kernel holding hundreds of live registers
// costs far more than the multiplications.  The common exponents 0..3 are straight-line.
__device__ __forceinline__ double ipow(double x, int n) {
    const bool neg = n < 0;
    int m = neg ? -n : n;
    double r;
    if (m == 1) r = x;
    else if (m == 2) r = x * x;
    else if (m == 0) r = 1.0;
    else {
        r = 1.0;
        bool first = true;
        while (m) {
            if (m & 1) { r = first ? x : r * x; first = false; }
            m >>= 1;
            if (m) x = x * x;
        }
    }
    return neg ? 1.0 / r : r;
}

// CKSMH (cklib.f:7540-7560): S/R - H/RT of one species.
__device__ __forceinline__ double smh_species(const Nasa &a, double T) {
    const double t2 = T * T, t3 = T * t2, t4 = t2 * t2;
    double s = (fm_log(T) - 1.0) * a.a0;
    s = s + T / 2.0 * a.a1;
    s = s + t2 / 6.0 * a.a2;
    s = s + t3 / 12.0 * a.a3;
    s = s + t4 / 20.0 * a.a4;
    return s + a.a6 - a.a5 / T;
}

// Effective rate constants (after fall-off and third-body factors) and rates of progress of
// reaction i: CKRATT + CKRATX (cklib.f:6662-6718, 6851-6948).
struct Rop {
    double kf, kr, qf, qr;
};

// c and smh hold species k at index k * STRIDE (STRIDE 1: one cell's arrays; 64: one lane's column of
// a species-major [k][64] block).
template <int STRIDE>
__device__ __forceinline__ Rop reaction(const DevKin *__restrict__ K, int i, double T, double alogt, double pfac,
                                        double ctot, const double *__restrict__ c, const double *__restrict__ smh) {
    // exp(0) = 1 exactly: rows without temperature dependence (B = E = 0) skip the exponential
    double kf = K->A[i];
    if (!K->arr0[i]) kf = kf * fm_exp(K->B[i] * alogt - K->E[i] / T);
    double kr = 0.0;
    if (K->rev[i]) {   // irreversible rows need no equilibrium constant
        double s = 0.0;
        for (int n = 0; n < 6; ++n) {
            const int k = K->nk[i][n];
            if (k >= 0) s = s + (double)K->nu[i][n] * smh[k * STRIDE];
        }
        double eqk = s != 0.0 ? fm_exp(fmin(s, K->exparg)) : 1.0;
        eqk = eqk * ipow(pfac, K->nus[i]);
        kr = kf / fmax(eqk, K->small);
    }
    double ctb = 1.0;
    const int t = K->tb[i];
    if (t >= 0) {
        // the row's efficiencies are loaded eight at a time, so their memory latency is paid once per
        // eight entries rather than once per entry; the sum keeps the entry order
        ctb = ctot;
        const int nl = K->ntbs[t];
        constexpr int TB = 8;
        for (int l0 = 0; l0 < nl; l0 += TB) {
            double ta[TB];
            int tk[TB];
#pragma unroll
            for (int u = 0; u < TB; ++u) {
                const int l = min(l0 + u, MAXTB - 1);
                ta[u] = K->tba[t][l];
                tk[u] = K->tbk[t][l];
            }
#pragma unroll
            for (int u = 0; u < TB; ++u)
                if (l0 + u < nl) ctb = ctb + (ta[u] - 1.0) * c[(tk[u] >= 0 ? tk[u] : 0) * STRIDE];
        }
    }
    const int f = K->fo[i];
    if (f >= 0) {
        const double *fp = K->fpar[f];
        const double rklow = fp[0] * fm_exp(fp[1] * alogt - fp[2] / T);
        double pr;
        if (K->kfal[f] < 0) { pr = rklow * ctb / kf; ctb = 1.0; }
        else pr = rklow * c[K->kfal[f] * STRIDE] / kf;
        double pcor = pr / (1.0 + pr);
        if (K->ifop[f] > 1) {
            const double prlog = fm_log10(fmax(pr, K->small));
            double fc;
            if (K->ifop[f] == 2) {
                const double xp = 1.0 / (1.0 + prlog * prlog);
                fc = fm_pow(fp[3] * fm_exp(-fp[4] / T) + fm_exp(-T / fp[5]), xp) * fp[6] * fm_pow(T, fp[7]);
            } else {
                double fcent = (1.0 - fp[3]) * fm_exp(-T / fp[4]) + fp[3] * fm_exp(-T / fp[5]);
                if (K->ifop[f] == 4) fcent = fcent + fm_exp(-fp[6] / T);
                const double fclog = fm_log10(fmax(fcent, K->small));
                // single-precision literals of cklib.f:6915-6917
                const double xn = 0.75 - (double)1.27f * fclog;
                const double cprlog = prlog - ((double)0.4f + (double)0.67f * fclog);
                const double q = cprlog / (xn - (double)0.14f * cprlog);
                fc = fm_exp10(fclog / (1.0 + q * q));
            }
            pcor = fc * pcor;
        }
        kf = kf * pcor;
        kr = kr * pcor;
    }
    kf = kf * ctb * 1.0;
    kr = kr * ctb * 1.0;
    Rop r{kf, kr, 0.0, 0.0};
    const int *nu = K->nu[i], *nk = K->nk[i];
    if (nu[0] != 0) {
        r.qf = kf * ipow(c[nk[0] * STRIDE], abs(nu[0]));
        r.qr = kr * ipow(c[nk[3] * STRIDE], nu[3]);
        if (nk[1] >= 0) {
            r.qf = r.qf * ipow(c[nk[1] * STRIDE], abs(nu[1]));
            if (nk[2] >= 0) r.qf = r.qf * ipow(c[nk[2] * STRIDE], abs(nu[2]));
        }
        if (nk[4] >= 0) {
            r.qr = r.qr * ipow(c[nk[4] * STRIDE], nu[4]);
            if (nk[5] >= 0) r.qr = r.qr * ipow(c[nk[5] * STRIDE], nu[5]);
        }
    }
    return r;
}

// reaction() for kernels whose lanes evaluate different rows (the integrator, the rates kernel): the
// same operations with the same operands, but the third-body sum and every exponential a row can
// need are evaluated before the branches on the row's kind -- the Arrhenius factor, the equilibrium
// constant, and for fall-off rows the low-pressure factor and the Troe terms -- so their dependent
// chains run side by side instead of one after another down the lanes' divergent paths.  Values a row does not use are
// discarded; a row with B = E = 0 multiplies A by exp(0) = 1 (exact).  Results are bitwise those of
// reaction().
template <int STRIDE>
__device__ __forceinline__ Rop reaction_ilp(const DevKin *__restrict__ K, int i, double T, double alogt, double pfac,
                                            double ctot, const double *__restrict__ c, const double *__restrict__ smh) {
    // the third-body concentration first: its loads and chain of additions do not depend on the rate
    // constants, so as straight-line code they overlap the exponentials below.  The first eight
    // efficiencies are added under selects (past the row's count the previous sum is kept), any
    // others in a loop: the additions of reaction(), in its order.
    const int t = K->tb[i];
    const int tt = t >= 0 ? t : 0;
    const int nl = t >= 0 ? K->ntbs[tt] : 0;
    double ctb = ctot;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int tk = K->tbk[tt][u];
        const double v = ctb + (K->tba[tt][u] - 1.0) * c[(tk >= 0 ? tk : 0) * STRIDE];
        ctb = u < nl ? v : ctb;
    }
    for (int l = 8; l < nl; ++l) {
        const int tk = K->tbk[tt][l];
        ctb = ctb + (K->tba[tt][l] - 1.0) * c[(tk >= 0 ? tk : 0) * STRIDE];
    }
    if (t < 0) ctb = 1.0;
    const double ekf = fm_exp(K->B[i] * alogt - K->E[i] / T);
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {       // empty slots keep the sum (selects, no branches)
        const int k = K->nk[i][n];
        const double v = s + (double)K->nu[i][n] * smh[(k >= 0 ? k : 0) * STRIDE];
        s = k >= 0 ? v : s;
    }
    const double eks = fm_exp(fmin(s, K->exparg));
    const int f = K->fo[i];
    const double *fp = K->fpar[f >= 0 ? f : 0];
    // the fall-off terms in every lane (a row without them reads row 0's parameters and discards the
    // values), outside any branch, so they interleave with the two exponentials above
    const double erk = fm_exp(fp[1] * alogt - fp[2] / T);
    const double e4 = fm_exp(-T / fp[4]);
    const double e5 = fm_exp(-T / fp[5]);
    const double e6 = fm_exp(-fp[6] / T);
    double kf = K->A[i];
    if (!K->arr0[i]) kf = kf * ekf;
    double kr = 0.0;
    if (K->rev[i]) {
        double eqk = s != 0.0 ? eks : 1.0;
        eqk = eqk * ipow(pfac, K->nus[i]);
        kr = kf / fmax(eqk, K->small);
    }
    if (f >= 0) {
        const double rklow = fp[0] * erk;
        double pr;
        if (K->kfal[f] < 0) { pr = rklow * ctb / kf; ctb = 1.0; }
        else pr = rklow * c[K->kfal[f] * STRIDE] / kf;
        double pcor = pr / (1.0 + pr);
        if (K->ifop[f] > 1) {
            const double prlog = fm_log10(fmax(pr, K->small));
            double fc;
            if (K->ifop[f] == 2) {
                const double xp = 1.0 / (1.0 + prlog * prlog);
                fc = fm_pow(fp[3] * fm_exp(-fp[4] / T) + fm_exp(-T / fp[5]), xp) * fp[6] * fm_pow(T, fp[7]);
            } else {
                double fcent = (1.0 - fp[3]) * e4 + fp[3] * e5;
                if (K->ifop[f] == 4) fcent = fcent + e6;
                const double fclog = fm_log10(fmax(fcent, K->small));
                // single-precision literals of cklib.f:6915-6917
                const double xn = 0.75 - (double)1.27f * fclog;
                const double cprlog = prlog - ((double)0.4f + (double)0.67f * fclog);
                const double q = cprlog / (xn - (double)0.14f * cprlog);
                fc = fm_exp10(fclog / (1.0 + q * q));
            }
            pcor = fc * pcor;
        }
        kf = kf * pcor;
        kr = kr * pcor;
    }
    kf = kf * ctb * 1.0;
    kr = kr * ctb * 1.0;
    Rop r{kf, kr, 0.0, 0.0};
    const int *nu = K->nu[i], *nk = K->nk[i];
    if (nu[0] != 0) {
        r.qf = kf * ipow(c[nk[0] * STRIDE], abs(nu[0]));
        r.qr = kr * ipow(c[nk[3] * STRIDE], nu[3]);
        if (nk[1] >= 0) {
            r.qf = r.qf * ipow(c[nk[1] * STRIDE], abs(nu[1]));
            if (nk[2] >= 0) r.qf = r.qf * ipow(c[nk[2] * STRIDE], abs(nu[2]));
        }
        if (nk[4] >= 0) {
            r.qr = r.qr * ipow(c[nk[4] * STRIDE], nu[4]);
            if (nk[5] >= 0) r.qr = r.qr * ipow(c[nk[5] * STRIDE], nu[5]);
        }
    }
    return r;
}

// x**m for 0 <= m <= 3 without branches, with ipow's product order (x**3 = x * (x * x)).
__device__ __forceinline__ double ipow3(double x, int m) {
    const double x2 = x * x;
    const double x3 = x * x2;
    return m == 1 ? x : (m == 2 ? x2 : (m == 3 ? x3 : 1.0));
}

// reaction() for a row without third body or fall-off whose coefficients and coefficient sum are at
// most 3 in magnitude (rows K->nplain0 .. ii - 1), written without branches: the conditional terms
// become selects of values computed anyway, and every select reproduces reaction()'s arithmetic (a
// factor of exactly 1.0 where reaction() skips a multiplication).  Straight-line code lets the
// compiler interleave the rate chains of several rows.
template <int STRIDE>
__device__ __forceinline__ Rop reaction_plain(const DevKin *__restrict__ K, int i, double T, double alogt, double pfac,
                                              const double *__restrict__ c, const double *__restrict__ smh) {
    // exp(0) = 1 exactly: rows without temperature dependence skip the exponential; irreversible
    // rows skip the equilibrium constant (with a wave-uniform row index both are scalar branches)
    double kf = K->A[i];
    if (!K->arr0[i]) kf = kf * fm_exp(K->B[i] * alogt - K->E[i] / T);
    const int *nu = K->nu[i], *nk = K->nk[i];
    int k[6], a[6];
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
        k[n] = nk[n];
        a[n] = nu[n];
        const double t = (double)a[n] * smh[(k[n] >= 0 ? k[n] : 0) * STRIDE];
        s = k[n] >= 0 ? s + t : s;
    }
    double ex = 1.0;
    if (K->rev[i]) ex = fm_exp(fmin(s, K->exparg));
    double eqk = s != 0.0 ? ex : 1.0;
    const int ns = K->nus[i];
    const double pp = ipow3(pfac, ns < 0 ? -ns : ns);
    eqk = eqk * (ns < 0 ? 1.0 / pp : pp);
    const double kr = K->rev[i] ? kf / fmax(eqk, K->small) : 0.0;
    double cc[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) cc[n] = ipow3(c[(k[n] >= 0 ? k[n] : 0) * STRIDE], a[n] < 0 ? -a[n] : a[n]);
    double qf = kf * cc[0];
    double qr = kr * cc[3];
    qf = qf * (k[1] >= 0 ? cc[1] : 1.0);
    qf = qf * (k[1] >= 0 && k[2] >= 0 ? cc[2] : 1.0);
    qr = qr * (k[4] >= 0 ? cc[4] : 1.0);
    qr = qr * (k[4] >= 0 && k[5] >= 0 ? cc[5] : 1.0);
    const bool live = a[0] != 0;
    return Rop{kf, kr, live ? qf : 0.0, live ? qr : 0.0};
}

}  // namespace fcfd
