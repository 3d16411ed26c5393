// fcfd_device.hpp -- device functions of the cell / interface update.
//
// Compiled with -ffp-contract=off: every expression below keeps the operation order of the
// reference's Fortran (left-to-right, x**2 as x*x), so on gfx950 the fluid path rounds exactly
// like the CPU reference (IEEE +,-,*,/ and sqrt are correctly rounded on both).
// Species loops run over a compile-time NSP and are fully unrolled so per-lane species arrays live
// in VGPRs; table reads use a wave-uniform species index and become scalar loads.
#pragma once

#include "fcfd_common.hpp"
#include "fcfd_math.h"

namespace fcfd {

// ---------------------------------------------------------------------------------------------
// Thermo (CKLIB semantics)
// ---------------------------------------------------------------------------------------------
// Temperature powers shared by all species at one T (cklib.f:1849-1852, 3537-3541).
struct TPow {
    double T, rut, tnh1, tnh2, tnh3, tnh4;   // T^n/(n+1) for CKHML
    double rtn0, rtn1, rtn2, rtn3, rtn4;     // RU*T^n for CKCPML
};

// Workgroup index that gives each of the GPU's eight XCDs one contiguous range of the grid: the
// dispatcher places workgroup b on XCD b mod 8, so consecutive workgroups -- neighbouring cells,
// whose stencils share cache lines -- would otherwise land in eight different L2 caches.  A
// bijection of [0, gridDim.x): XCD x takes ranks x*q + min(x, r) .. (q or q + 1 of them).
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned n = gridDim.x, b = blockIdx.x;
    const unsigned q = n / 8, r = n % 8, x = b % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

__device__ __forceinline__ TPow tpow(double T, double ru) {
    TPow p;
    const double t2 = T * T, t3 = T * t2, t4 = t2 * t2;
    p.T = T;
    p.rut = T * ru;
    p.tnh1 = T / 2.0; p.tnh2 = t2 / 3.0; p.tnh3 = t3 / 4.0; p.tnh4 = t4 / 5.0;
    p.rtn0 = ru * 1.0; p.rtn1 = ru * T; p.rtn2 = ru * t2; p.rtn3 = ru * t3; p.rtn4 = ru * t4;
    return p;
}

struct Nasa {
    double a0, a1, a2, a3, a4, a5, a6;
};

// Both coefficient sets are read (with a wave-uniform k these are scalar loads) and the range is
// selected per lane; selecting the pointer instead would turn every coefficient read into a
// per-lane memory access.
__device__ __forceinline__ Nasa nasa(const DevThermo *__restrict__ m, int k, double T) {
    const double *lo = m->alo[k], *hi = m->ahi[k];
    const bool up = T > m->tmid[k];
    return Nasa{up ? hi[0] : lo[0], up ? hi[1] : lo[1], up ? hi[2] : lo[2], up ? hi[3] : lo[3],
                up ? hi[4] : lo[4], up ? hi[5] : lo[5], up ? hi[6] : lo[6]};
}

// CKHML (cklib.f:3551-3555): molar enthalpy, erg/mol.
__device__ __forceinline__ double h_mol(const Nasa &a, const TPow &p) {
    double s = a.a0;                       // 0 + 1.0*a0
    s = s + p.tnh1 * a.a1;
    s = s + p.tnh2 * a.a2;
    s = s + p.tnh3 * a.a3;
    s = s + p.tnh4 * a.a4;
    return p.rut * (s + a.a5 / p.T);
}

// CKCPML (cklib.f:1862-1864): molar cp, erg/(mol K).
__device__ __forceinline__ double cp_mol(const Nasa &a, const TPow &p) {
    double c = p.rtn0 * a.a0;
    c = c + p.rtn1 * a.a1;
    c = c + p.rtn2 * a.a2;
    c = c + p.rtn3 * a.a3;
    c = c + p.rtn4 * a.a4;
    return c;
}

// Mixture sums at one T: CKUBML, CKCPBL, CKCVBL (cklib.f:8346-8351, 1746-1751, 2553-2558).
struct MixSums {
    double ubml, cpbml, cvbml;
};

template <int NSP>
__device__ __forceinline__ MixSums mix_sums(const DevThermo *__restrict__ m, const double (&x)[NSP],
                                            double T, bool want_u, bool want_cp, bool want_cv) {
    const TPow p = tpow(T, m->ru);
    MixSums s{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const Nasa a = nasa(m, k, T);
        if (want_u) s.ubml = s.ubml + x[k] * (h_mol(a, p) - p.rut);
        if (want_cp || want_cv) {
            const double c = cp_mol(a, p);
            if (want_cp) s.cpbml = s.cpbml + x[k] * c;
            if (want_cv) s.cvbml = s.cvbml + x[k] * (c - m->ru);
        }
    }
    return s;
}

// sum_k X_k W_k in species order (CKPX / thermo.f90 wtm loops).
template <int NSP>
__device__ __forceinline__ double mix_w(const DevThermo *__restrict__ m, const double (&x)[NSP]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s = s + x[k] * m->wt[k];
    return s;
}

// adjust_sum (SRC/thermo.f90:136-149): clip at zero, renormalise.
template <int NSP>
__device__ __forceinline__ void adjust_sum(double (&v)[NSP]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        v[k] = fmax(v[k], 0.0);
        s = s + v[k];
    }
#pragma unroll
    for (int k = 0; k < NSP; ++k) v[k] = v[k] / s;
}

// CKYTX (cklib.f:9626-9632), in place.
template <int NSP>
__device__ __forceinline__ void ytx(const DevThermo *__restrict__ m, double (&v)[NSP]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) s = s + v[k] / m->wt[k];
#pragma unroll
    for (int k = 0; k < NSP; ++k) v[k] = v[k] / (s * m->wt[k]);
}

// temperature_con (SRC/thermo.f90:211-257): Newton on molar internal energy from 1000 K.
template <int NSP>
__device__ double temperature_con(const DevThermo *__restrict__ m, double e, const double (&x)[NSP]) {
    const double wtm = mix_w<NSP>(m, x);
    const double spe = e * 1.e4 * wtm;
    double T = 1000.0, f = 1.e30, dT = 1.e30;
    int it = 0;
    while (fabs(f) >= fabs(spe) * 1.e-9 && fabs(dT) >= T * 1.e-6 && it < 10) {
        const MixSums s = mix_sums<NSP>(m, x, T, true, false, true);
        f = s.ubml - spe;
        dT = -f / s.cvbml;
        T = T + dT;
        T = fmax(20.0, fmin(6000.0, T));
        ++it;
    }
    return T;
}

// ---------------------------------------------------------------------------------------------
// Reconstruction (SRC/weno.f90, SRC/muscl.f90); s[] in the reference's stencil order.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double sq(double x) { return x * x; }

__device__ __forceinline__ void weno3z(const double *s, double &sl, double &sr) {
    const double eps = 1.0e-6;
    double b2 = (s[0] - s[1]) * (s[0] - s[1]);
    double b1 = (s[1] - s[2]) * (s[1] - s[2]);
    double tau = fabs(b2 - b1);
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    double a2 = 2.0 * b2, a1 = 1.0 * b1, ai = 1.0 / (a2 + a1);
    sl = 0.5 * ai * (a2 * (-s[0] + 3.0 * s[1]) + a1 * (s[1] + s[2]));
    b2 = (s[3] - s[2]) * (s[3] - s[2]);
    b1 = (s[2] - s[1]) * (s[2] - s[1]);
    tau = fabs(b2 - b1);
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    a2 = 2.0 * b2; a1 = 1.0 * b1; ai = 1.0 / (a2 + a1);
    sr = 0.5 * ai * (a2 * (3.0 * s[2] - s[3]) + a1 * (s[1] + s[2]));
}

template <int FLAG>
__device__ __forceinline__ void w5_weights(double &b3, double &b2, double &b1) {
    const double eps = 1.0e-6;
    if (FLAG == 2) {
        b3 = 1.0 / ((eps + b3) * (eps + b3));
        b2 = 1.0 / ((eps + b2) * (eps + b2));
        b1 = 1.0 / ((eps + b1) * (eps + b1));
    } else {
        const double tau = fabs(b3 - b1);
        b3 = 1.0 + (tau / (eps + b3)) * (tau / (eps + b3));
        b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
        b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    }
}

template <int FLAG>
__device__ __forceinline__ void weno5(const double *s, double &sl, double &sr) {
    const double c13 = 13.0 / 12.0;
    double b3 = c13 * sq(s[0] - 2.0 * s[1] + s[2]) + 0.25 * sq(s[0] - 4.0 * s[1] + 3.0 * s[2]);
    double b2 = c13 * sq(s[1] - 2.0 * s[2] + s[3]) + 0.25 * (s[1] - s[3]) * (s[1] - s[3]);
    double b1 = c13 * sq(s[2] - 2.0 * s[3] + s[4]) + 0.25 * sq(3.0 * s[2] - 4.0 * s[3] + s[4]);
    w5_weights<FLAG>(b3, b2, b1);
    double a3 = b3, a2 = 6.0 * b2, a1 = 3.0 * b1, ai = 1.0 / (a3 + a2 + a1);
    sl = (1.0 / 6.0) * ai *
         (a3 * (2.0 * s[0] - 7.0 * s[1] + 11.0 * s[2]) + a2 * (-s[1] + 5.0 * s[2] + 2.0 * s[3]) +
          a1 * (2.0 * s[2] + 5.0 * s[3] - s[4]));
    b3 = c13 * sq(s[5] - 2.0 * s[4] + s[3]) + 0.25 * sq(s[5] - 4.0 * s[4] + 3.0 * s[3]);
    b2 = c13 * sq(s[4] - 2.0 * s[3] + s[2]) + 0.25 * (s[4] - s[2]) * (s[4] - s[2]);
    b1 = c13 * sq(s[3] - 2.0 * s[2] + s[1]) + 0.25 * sq(3.0 * s[3] - 4.0 * s[2] + s[1]);
    w5_weights<FLAG>(b3, b2, b1);
    a3 = b3; a2 = 6.0 * b2; a1 = 3.0 * b1; ai = 1.0 / (a3 + a2 + a1);
    sr = (1.0 / 6.0) * ai *
         (a3 * (11.0 * s[3] - 7.0 * s[4] + 2.0 * s[5]) + a2 * (-s[4] + 5.0 * s[3] + 2.0 * s[2]) +
          a1 * (2.0 * s[3] + 5.0 * s[2] - s[1]));
}

// One side of WENO7-Z on q[0..6] in the orientation of the left-biased formula.
__device__ __forceinline__ double w7_side(double q0, double q1, double q2, double q3, double q4,
                                          double q5, double q6) {
    const double eps = 1.0e-6;
    double b4 = q0 * (547.0 * q0 - 3882.0 * q1 + 4642.0 * q2 - 1854.0 * q3) +
                q1 * (7043.0 * q1 - 17246.0 * q2 + 7042.0 * q3) + q2 * (11003.0 * q2 - 9402.0 * q3) +
                2107.0 * (q3 * q3);
    double b3 = q1 * (267.0 * q1 - 1642.0 * q2 + 1602.0 * q3 - 494.0 * q4) +
                q2 * (2843.0 * q2 - 5966.0 * q3 + 1922.0 * q4) + q3 * (3443.0 * q3 - 2522.0 * q4) +
                547.0 * (q4 * q4);
    double b2 = q2 * (547.0 * q2 - 2522.0 * q3 + 1922.0 * q4 - 494.0 * q5) +
                q3 * (3443.0 * q3 - 5966.0 * q4 + 1602.0 * q5) + q4 * (2843.0 * q4 - 1642.0 * q5) +
                267.0 * (q5 * q5);
    double b1 = q3 * (2107.0 * q3 - 9402.0 * q4 + 7042.0 * q5 - 1854.0 * q6) +
                q4 * (11003.0 * q4 - 17246.0 * q5 + 4642.0 * q6) + q5 * (7043.0 * q5 - 3882.0 * q6) +
                547.0 * (q6 * q6);
    const double tau = fabs(b4 - b1);
    b4 = 1.0 + (tau / (eps + b4)) * (tau / (eps + b4));
    b3 = 1.0 + (tau / (eps + b3)) * (tau / (eps + b3));
    b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
    b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
    const double a4 = (1.0 / 35.0) * b4, a3 = (12.0 / 35.0) * b3, a2 = (18.0 / 35.0) * b2,
                 a1 = (4.0 / 35.0) * b1;
    const double ai = 1.0 / (a4 + a3 + a2 + a1);
    return (1.0 / 12.0) * ai *
           (a4 * (-3.0 * q0 + 13.0 * q1 - 23.0 * q2 + 25.0 * q3) +
            a3 * (1.0 * q1 - 5.0 * q2 + 13.0 * q3 + 3.0 * q4) +
            a2 * (-1.0 * q2 + 7.0 * q3 + 7.0 * q4 - 1.0 * q5) +
            a1 * (3.0 * q3 + 13.0 * q4 - 5.0 * q5 + 1.0 * q6));
}

__device__ __forceinline__ void weno7z(const double *s, double &sl, double &sr) {
    sl = w7_side(s[0], s[1], s[2], s[3], s[4], s[5], s[6]);
    sr = w7_side(s[7], s[6], s[5], s[4], s[3], s[2], s[1]);
}

__device__ __forceinline__ double lim_minmod(double x, double y) {
    const double sx = copysign(1.0, x);
    return sx * fmax(0.0, fmin(fabs(x), sx * y));
}
__device__ __forceinline__ double lim_superbee(double x, double y) {
    return fmax(0.0, fmax(fmin(x, 2.0 * y), fmin(2.0 * x, y)));
}
__device__ __forceinline__ double lim_mc(double x, double y) {
    return fmax(0.0, fmin(fmin(2.0 * x, 2.0 * y), 0.5 * (x + y)));
}

__device__ __forceinline__ double limit(int flag, double x, double y) {
    return flag == 2 ? lim_minmod(x, y) : (flag == 3 ? lim_superbee(x, y) : lim_mc(x, y));
}

__device__ __forceinline__ void muscl3(const double *s, double &sl, double &sr, int flag) {
    const double k = 1.0 / 3.0;
    const double b = (3.0 - k) / (1.0 - k);
    const double dmi = s[1] - s[0], dpi = s[2] - s[1], dmip = s[2] - s[1], dpip = s[3] - s[2];
    if (flag == 1) {
        sl = s[1] + (1.0 / 4.0) * ((1.0 - k) * dmi + (1.0 + k) * dpi);
        sr = s[2] - (1.0 / 4.0) * ((1.0 - k) * dpip + (1.0 + k) * dmip);
        return;
    }
    const double lpi = limit(flag, dpi, b * dmi), lmi = limit(flag, dmi, b * dpi);
    const double lpip = limit(flag, dpip, b * dmip), lmip = limit(flag, dmip, b * dpip);
    sl = s[1] + (1.0 / 4.0) * ((1.0 - k) * lmi + (1.0 + k) * lpi);
    sr = s[2] - (1.0 / 4.0) * ((1.0 - k) * lpip + (1.0 + k) * lmip);
}

// Reconstruction kinds, selected at compile time because they fix the stencil width.
enum Recon { R_FIRST = 0, R_MUSCL = 1, R_WENO3Z = 2, R_WENO5JS = 3, R_WENO5Z = 4, R_WENO7Z = 5 };

template <int RK> struct Stencil;
template <> struct Stencil<R_FIRST> { static constexpr int lo = 0, n = 2; };
template <> struct Stencil<R_MUSCL> { static constexpr int lo = -1, n = 4; };
template <> struct Stencil<R_WENO3Z> { static constexpr int lo = -1, n = 4; };
template <> struct Stencil<R_WENO5JS> { static constexpr int lo = -2, n = 6; };
template <> struct Stencil<R_WENO5Z> { static constexpr int lo = -2, n = 6; };
template <> struct Stencil<R_WENO7Z> { static constexpr int lo = -3, n = 7; };

// Reconstructs one variable at the interface right of cell c0.  q points at the variable's row
// (SoA), c0 is the column of the interface's left cell.
template <int RK>
__device__ __forceinline__ void reconstruct(const double *__restrict__ q, long c0, int muscl_flag,
                                            double &l, double &r) {
    constexpr int lo = Stencil<RK>::lo, n = Stencil<RK>::n;
    double s[8];
#pragma unroll
    for (int j = 0; j < n; ++j) s[j] = q[c0 + lo + j];
    if (RK == R_FIRST) { l = s[0]; r = s[1]; }
    else if (RK == R_MUSCL) muscl3(s, l, r, muscl_flag);
    else if (RK == R_WENO3Z) weno3z(s, l, r);
    else if (RK == R_WENO5JS) weno5<2>(s, l, r);
    else if (RK == R_WENO5Z) weno5<3>(s, l, r);
    else { s[7] = s[6]; weno7z(s, l, r); }   // s(8) = qf(i+3) (SRC/flux.f90:232)
}

// One side of the reconstruction at the interface right of cell c0: side 0 the left state, side 1
// the right state, with the same operations as the two-sided routines above.  The right-state
// formulas of WENO3-Z, WENO5 and WENO7-Z are the left-state formulas on the mirrored stencil, so
// both lanes of an interface pair run the same instructions on mirrored loads; the one place
// where WENO5's two sides add in a different order (2 s0 - 7 s1 + 11 s2 against
// 11 s3 - 7 s4 + 2 s5) takes its operands through a select.
template <int RK>
__device__ __forceinline__ double reconstruct_side(const double *__restrict__ q, long c0, int muscl_flag, int side) {
    if (RK == R_FIRST) return q[c0 + side];
    if (RK == R_MUSCL) {
        double l, r;
        reconstruct<R_MUSCL>(q, c0, muscl_flag, l, r);
        return side ? r : l;
    }
    if (RK == R_WENO3Z) {
        // left: s0 s1 s2 = q[c0-1 .. c0+1]; right: s3 s2 s1 = q[c0+2 .. c0]
        const double u0 = q[c0 + (side ? 2 : -1)], u1 = q[c0 + (side ? 1 : 0)], u2 = q[c0 + (side ? 0 : 1)];
        const double eps = 1.0e-6;
        double b2 = (u0 - u1) * (u0 - u1);
        double b1 = (u1 - u2) * (u1 - u2);
        const double tau = fabs(b2 - b1);
        b2 = 1.0 + (tau / (eps + b2)) * (tau / (eps + b2));
        b1 = 1.0 + (tau / (eps + b1)) * (tau / (eps + b1));
        const double a2 = 2.0 * b2, a1 = 1.0 * b1, ai = 1.0 / (a2 + a1);
        return 0.5 * ai * (a2 * (-u0 + 3.0 * u1) + a1 * (u1 + u2));
    }
    if (RK == R_WENO5JS || RK == R_WENO5Z) {
        // left: s0..s4 = q[c0-2 .. c0+2]; right: s5..s1 = q[c0+3 .. c0-1]
        const long o = side ? c0 + 3 : c0 - 2;
        const long d = side ? -1 : 1;
        const double u0 = q[o], u1 = q[o + d], u2 = q[o + 2 * d], u3 = q[o + 3 * d], u4 = q[o + 4 * d];
        const double c13 = 13.0 / 12.0;
        double b3 = c13 * sq(u0 - 2.0 * u1 + u2) + 0.25 * sq(u0 - 4.0 * u1 + 3.0 * u2);
        double b2 = c13 * sq(u1 - 2.0 * u2 + u3) + 0.25 * (u1 - u3) * (u1 - u3);
        double b1 = c13 * sq(u2 - 2.0 * u3 + u4) + 0.25 * sq(3.0 * u2 - 4.0 * u3 + u4);
        w5_weights<RK == R_WENO5JS ? 2 : 3>(b3, b2, b1);
        const double a3 = b3, a2 = 6.0 * b2, a1 = 3.0 * b1, ai = 1.0 / (a3 + a2 + a1);
        const double x2 = 2.0 * u0, x11 = 11.0 * u2;
        const double t3 = ((side ? x11 : x2) - 7.0 * u1) + (side ? x2 : x11);
        return (1.0 / 6.0) * ai *
               (a3 * t3 + a2 * (-u1 + 5.0 * u2 + 2.0 * u3) + a1 * (2.0 * u2 + 5.0 * u3 - u4));
    }
    // WENO7-Z: left s0..s6 = q[c0-3 .. c0+3]; right s7 s6 .. s1 with s7 = s6 = q[c0+3]
    // (SRC/flux.f90:232)
    const long o = side ? c0 + 3 : c0 - 3;
    const long d = side ? -1 : 1;
    const double v0 = q[o];
    const double v1 = side ? v0 : q[o + d];
    const long o1 = side ? o : o + d;         // element 1 of the side's stencil
    return w7_side(v0, v1, q[o1 + d], q[o1 + 2 * d], q[o1 + 3 * d], q[o1 + 4 * d], q[o1 + 5 * d]);
}

}  // namespace fcfd
