// fluid.hip -- kernels of the hydrodynamic part of one time step.
//
// Device layout ("SoA, ghost-padded"): variable v of padded column c lives at q[v*P + c].  Block b
// owns columns [b*S, (b+1)*S), S = ncell + 2*lv; its interior cell i (0-based) is column
// b*S + lv + i, exactly the reference's b_obj(ib)%qf(:, ista-lv : iend+lv) column range.
// Interface j (0..ncell) of block b -- the reference's Fplus(:, ista-1+j) -- is F[v*PF + b*(ncell+1)+j].
#include "fcfd_device.hpp"
#include "fluid.hpp"

namespace fcfd {

// ---------------------------------------------------------------------------------------------
// Per-side interface state: what qfToQc + pressure + sonic_speed give for one reconstructed
// primitive vector (SRC/flux.f90:247-268, SRC/eos.f90:11-46).
// ---------------------------------------------------------------------------------------------
struct Side {
    double rho, u, T, sy, p, a, E;
};

// y[k * stride] holds the reconstructed Y_k (a register array with stride 1 or an LDS column).  Over
// an LDS column the species loops unroll by four only: fully unrolled, the advection-flux kernel
// needs more than 256 registers and spills to scratch.
template <int NSP, int STRIDE = 1>
__device__ __forceinline__ Side side_state(const DevThermo *__restrict__ m, double rho, double u,
                                           double T, const double *y) {
    Side s;
    s.rho = rho; s.u = u; s.T = T;
    // adjust_sum(Y) denominator (Y itself is renormalised lazily where QcL/R needs it).
    double sy = 0.0, syw = 0.0;
#pragma unroll (STRIDE == 1 ? NSP : 4)
    for (int k = 0; k < NSP; ++k) {
        sy = sy + fmax(y[k * STRIDE], 0.0);
        syw = syw + y[k * STRIDE] / m->wt[k];          // CKYTX on the raw reconstructed Y
    }
    s.sy = sy;
    // mole fractions, clipped and renormalised (SRC/flux.f90:250-253), formed on the fly in the
    // mixture sums instead of being held in an array: each sum keeps its species order, and the
    // per-species quotient is the same expression every time it is formed
    double sx = 0.0;
#pragma unroll (STRIDE == 1 ? NSP : 4)
    for (int k = 0; k < NSP; ++k) sx = sx + fmax(y[k * STRIDE] / (syw * m->wt[k]), 0.0);
    const TPow tp = tpow(T, m->ru);
    double wtm = 0.0;
    MixSums ms{0.0, 0.0, 0.0};
#pragma unroll (STRIDE == 1 ? NSP : 4)
    for (int k = 0; k < NSP; ++k) {
        const double xk = fmax(y[k * STRIDE] / (syw * m->wt[k]), 0.0) / sx;
        wtm = wtm + xk * m->wt[k];                                   // mix_w
        const Nasa a = nasa(m, k, T);                                // mix_sums
        ms.ubml = ms.ubml + xk * (h_mol(a, tp) - tp.rut);
        const double c = cp_mol(a, tp);
        ms.cpbml = ms.cpbml + xk * c;
        ms.cvbml = ms.cvbml + xk * (c - m->ru);
    }
    s.p = (rho * 1.e-3) * m->ru * T / wtm * 0.1;                 // pressure(), CKPX
    const double e = ms.ubml * 1.e-4 / wtm;                      // specific_energy()
    s.E = rho * (e + 0.5 * (u * u));
    const double cp = ms.cpbml * 1.e-4 / wtm, cv = ms.cvbml * 1.e-4 / wtm;
    s.a = sqrt(cp / cv * s.p / rho);                             // sonic_speed()
    return s;
}

// ---------------------------------------------------------------------------------------------
// calc_flux_advection: TWO LANES PER INTERFACE.  Lanes 2j and 2j+1 of a wave share interface j
// of the wave's range: lane parity s = 0 builds the left state, s = 1 the right state (each
// reconstructs the stencil, keeps its side's values and runs qfToQc / pressure / sonic speed on
// them); the two side records and, per species, the two mass fractions are exchanged between the
// pair (side records by a lane shuffle, mass fractions through LDS), both lanes evaluate the same
// Riemann branch, and the flux components
// are written alternately by the two lanes.  Every output is the same expression of the same
// operands as in a one-lane evaluation, so the rounding is unchanged.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double pair_other(double v) { return __shfl_xor(v, 1, 64); }

__device__ __forceinline__ Side pair_other(const Side &me) {
    Side o;
    o.rho = pair_other(me.rho); o.u = pair_other(me.u); o.T = pair_other(me.T);
    o.sy = pair_other(me.sy); o.p = pair_other(me.p); o.a = pair_other(me.a); o.E = pair_other(me.E);
    return o;
}

template <int NSP, int RK, bool SPH, int BS = 64, int WPE = 2>
__global__ void __launch_bounds__(BS, WPE) k_flux(FluxArgs a) {
    const long g2 = long(xcd_block()) * blockDim.x + threadIdx.x;
    const long g = g2 >> 1;                               // interface (pairs never straddle a wave)
    const int s = int(threadIdx.x & 1);                   // 0: left state, 1: right state
    const int nif = a.ncell + 1;
    if (g >= long(a.nblk) * nif) return;                  // both lanes of a pair leave together
    // a launch covers blocks b0 .. b0 + nblk - 1 (the coarse blocks of a refined mesh reconstruct
    // at first order in launches of their own, SRC/flux.f90:240)
    const int b = a.b0 + int(g / nif), j = int(g % nif);
    const long gi = long(b) * nif + j;                    // interface index over all blocks
    const long c0 = long(b) * a.S + a.lv + j - 1;      // left cell of the interface
    const DevThermo *__restrict__ m = a.thermo;
    const double *__restrict__ q = a.qf;

    // each lane reconstructs its own side only
    const double rho_s = reconstruct_side<RK>(q + 0 * a.P, c0, a.muscl, s);
    const double u_s = reconstruct_side<RK>(q + 1 * a.P, c0, a.muscl, s);
    const double T_s = reconstruct_side<RK>(q + 2 * a.P, c0, a.muscl, s);
    // reconstructed mass fractions of this lane's side, species-major in LDS (column = lane)
    __shared__ double ys[NSP * BS];
    double *y = ys + threadIdx.x;
    for (int k = 0; k < NSP; ++k) y[k * BS] = reconstruct_side<RK>(q + (3 + k) * a.P, c0, a.muscl, s);
    const double *yo_col = ys + (threadIdx.x ^ 1);     // the pair partner's side
    const Side me = side_state<NSP, BS>(m, rho_s, u_s, T_s, y);
    const Side ot = pair_other(me);
    const Side L = s ? ot : me;
    const Side R = s ? me : ot;

    // physical fluxes of mass, momentum, energy (SRC/flux.f90:270-280)
    const double fl0 = L.rho * L.u, fr0 = R.rho * R.u;
    // spherical: no pressure in the momentum flux; the update adds its gradient (SRC/flux.f90:272-278)
    const double fl1 = SPH ? L.rho * L.u * L.u : L.rho * L.u * L.u + L.p;
    const double fr1 = SPH ? R.rho * R.u * R.u : R.rho * R.u * R.u + R.p;
    const double fl2 = (L.E + L.p) * L.u, fr2 = (R.E + R.p) * R.u;
    double *__restrict__ F = a.F + gi;
    const long PF = a.PF;
    // Fplus accumulates (SRC/flux.f90:358): on top of the transport flux when there is one
    const bool acc = a.accumulate != 0;
    // component v is written by lane parity v & 1
    auto put = [acc, s](double &d, int v, double val) {
        if ((v & 1) == s) d = acc ? d + val : val;
    };

    if (a.riemann == 0) {                                         // LxF (SRC/flux.f90:291-294)
        const double lL = fmax(fmax(fabs(L.u - L.a), fabs(L.u)), fabs(L.u + L.a));
        const double lR = fmax(fmax(fabs(R.u - R.a), fabs(R.u)), fabs(R.u + R.a));
        const double al = fmax(lL, lR);
        put(F[0], 0, 0.5 * (fl0 + fr0 - al * (R.rho - L.rho)));
        put(F[PF], 1, 0.5 * (fl1 + fr1 - al * (R.rho * R.u - L.rho * L.u)));
        put(F[2 * PF], 2, 0.5 * (fl2 + fr2 - al * (R.E - L.E)));
#pragma unroll 4
        for (int k = 0; k < NSP; ++k) {
            const double ym = y[k * BS], yo = yo_col[k * BS];
            const double yl = s ? yo : ym, yr = s ? ym : yo;
            const double QL = L.rho * (fmax(yl, 0.0) / L.sy), QR = R.rho * (fmax(yr, 0.0) / R.sy);
            const double fL = L.rho * yl * L.u, fR = R.rho * yr * R.u;
            put(F[(3 + k) * PF], 3 + k, 0.5 * (fL + fR - al * (QR - QL)));
        }
    } else {
        const double rt = sqrt(R.rho / L.rho);
        const double uro = (L.u + rt * R.u) / (1.0 + rt);
        const double aro = (L.a + rt * R.a) / (1.0 + rt);
        const double sL = fmin(L.u - L.a, uro - aro), sR = fmax(R.u + R.a, uro + aro);
        if (a.riemann == 1) {                                     // HLL (SRC/flux.f90:300-315)
            int br = (0.0 <= sL) ? 0 : ((sL <= 0.0 && 0.0 <= sR) ? 1 : 2);
            auto hll = [&](double fL, double fR, double QL, double QR) {
                return br == 0 ? fL : (br == 1 ? (sR * fL - sL * fR + sL * sR * (QR - QL)) / (sR - sL) : fR);
            };
            put(F[0], 0, hll(fl0, fr0, L.rho, R.rho));
            put(F[PF], 1, hll(fl1, fr1, L.rho * L.u, R.rho * R.u));
            put(F[2 * PF], 2, hll(fl2, fr2, L.E, R.E));
#pragma unroll 4
            for (int k = 0; k < NSP; ++k) {
                const double ym = y[k * BS], yo = yo_col[k * BS];
                const double yl = s ? yo : ym, yr = s ? ym : yo;
                const double QL = L.rho * (fmax(yl, 0.0) / L.sy), QR = R.rho * (fmax(yr, 0.0) / R.sy);
                put(F[(3 + k) * PF], 3 + k, hll(L.rho * yl * L.u, R.rho * yr * R.u, QL, QR));
            }
        } else {                                                  // HLLC (SRC/flux.f90:321-354)
            const double sM = (L.p - R.p + R.rho * R.u * (sR - R.u) - L.rho * L.u * (sL - L.u)) /
                              (R.rho * (sR - R.u) - L.rho * (sL - L.u));
            // branch chain of SRC/flux.f90:332-354
            const int br = (0.0 <= sL) ? 0 : ((sL <= 0.0 && 0.0 <= sM) ? 1 : ((sM <= 0.0 && 0.0 <= sR) ? 2 : 3));
            const bool left = br <= 1;
            const Side S = left ? L : R;
            const double f0 = left ? fl0 : fr0, f1 = left ? fl1 : fr1, f2 = left ? fl2 : fr2;
            if (br == 0 || br == 3) {                             // supersonic: upwind flux
                put(F[0], 0, f0);
                put(F[PF], 1, f1);
                put(F[2 * PF], 2, f2);
#pragma unroll 4
                for (int k = 0; k < NSP; ++k) {
                    const double yk = (left == (s == 0)) ? y[k * BS] : yo_col[k * BS];
                    put(F[(3 + k) * PF], 3 + k, S.rho * yk * S.u);
                }
            } else {                                              // star state of the upwind side
                const double sK = left ? sL : sR;
                const double c = S.rho * (sK - S.u) / (sK - sM);
                const double v2 = S.E / S.rho + (sM - S.u) * (sM + S.p / (S.rho * (sK - S.u)));
                put(F[0], 0, f0 + sK * (c * 1.0 - S.rho));
                put(F[PF], 1, f1 + sK * (c * sM - S.rho * S.u));
                put(F[2 * PF], 2, f2 + sK * (c * v2 - S.E));
#pragma unroll 4
                for (int k = 0; k < NSP; ++k) {
                    const double yk = (left == (s == 0)) ? y[k * BS] : yo_col[k * BS];
                    const double Qk = S.rho * (fmax(yk, 0.0) / S.sy);
                    const double fk = S.rho * yk * S.u;
                    put(F[(3 + k) * PF], 3 + k, fk + sK * (c * (Qk / S.rho) - Qk));
                }
            }
        }
    }
    if (a.godunov && s == 0) {                                    // SRC/flux.f90:360-367
        const double rt = sqrt(R.rho / L.rho);
        double *G = a.godunov + gi;
        G[0] = sqrt(R.rho * L.rho);
        G[PF] = (L.u + rt * R.u) / (1.0 + rt);
        G[2 * PF] = (L.p + rt * R.p) / (1.0 + rt);
    }
}

// ---------------------------------------------------------------------------------------------
// QcToqf on one column, writing qf (SRC/eos.f90:49-83).
// ---------------------------------------------------------------------------------------------
template <int NSP>
__device__ __forceinline__ void qc_to_qf_store(const DevThermo *__restrict__ m, const double (&Qv)[NSP + 3],
                                               double *__restrict__ qf, long P, long c) {
    const double rho = Qv[0];
    const double u = Qv[1] / rho;
    double y[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = Qv[3 + k] / rho;
    adjust_sum<NSP>(y);
#pragma unroll
    for (int k = 0; k < NSP; ++k) qf[(3 + k) * P + c] = y[k];
    ytx<NSP>(m, y);
    adjust_sum<NSP>(y);
    const double e = Qv[2] / rho - 0.5 * (u * u);
    qf[c] = rho;
    qf[P + c] = u;
    qf[2 * P + c] = temperature_con<NSP>(m, e, y);
}

// Source column of padded column (b, i): interior -> itself; wall ghost -> mirrored interior cell
// with momentum negated (SRC/boundary.f90:152-199); block ghost -> the neighbour block's interior
// cell (SRC/block.f90:490-497).
__device__ __forceinline__ void ghost_source(const StageArgs &a, int b, int i, int &sb, int &si, double &sgn) {
    sb = b; si = i; sgn = 1.0;
    if (i < 0) {
        if (b == 0) { si = -i - 1; sgn = -1.0; }
        else { sb = b - 1; si = a.ncell + i; }
    } else if (i >= a.ncell) {
        if (b == a.nblk - 1) { si = 2 * a.ncell - 1 - i; sgn = -1.0; }
        else { sb = b + 1; si = i - a.ncell; }
    }
}

// Stage end: RK combine of the source cell, ghost fill, primitive recovery -- the interior update
// of SRC/timeintegration.f90:150-175 (or :48-56 for Euler), then update_boundary /
// interpolation_block / QcToqf on every column (:201-215).  mode 0: Q = Qin - dt/dx*dF;
// mode 1: Q = h1*Q0 + h2*(Qin - dt/dx*dF); mode 2: no update (ghost fill + QcToqf only).
template <int NSP, int BS = 64, int WPE = 2>
__global__ void __launch_bounds__(BS, WPE) k_stage_end(StageArgs a) {
    const long c = long(xcd_block()) * blockDim.x + threadIdx.x;
    if (c >= long(a.nblk) * a.S) return;
    const int b = int(c / a.S), i = int(c % a.S) - a.lv;
    const bool ghost = (i < 0 || i >= a.ncell);
    if (a.ghosts_only && !ghost) return;
    int sb, si;
    double sgn;
    ghost_source(a, b, i, sb, si, sgn);
    const long sc = long(sb) * a.S + a.lv + si;
    const long fj = long(sb) * (a.ncell + 1) + si;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double dx = a.dxb ? a.dxb[sb] : a.dx;         // the source cell's block spacing
    const double dtdx = dt / dx;
    // spherical (SRC/timeintegration.f90:58-72, 177-196): face areas (r -+ dx/2)^2 and volume
    // r^2 dx at the source cell's centre
    const bool sph = a.spherical && a.mode != 2;
    double Al = 0.0, Ar = 0.0, dtdV = 0.0;
    if (sph) {
        const double rm = a.xm[long(sb) * a.ncell + si];
        const double tl = rm - (dx / 2.0), tr = rm + (dx / 2.0);
        Al = tl * tl;
        Ar = tr * tr;
        dtdV = dt / (rm * rm * dx);
    }
    double Qv[NSP + 3];
#pragma unroll
    for (int v = 0; v < NSP + 3; ++v) {
        double Q;
        if (a.mode == 2) {
            Q = a.Qin[v * a.P + sc];
        } else {
            const double fr = a.F[v * a.PF + fj + 1], fl = a.F[v * a.PF + fj];
            const double qi = a.Qin[v * a.P + sc];
            const double base = sph ? qi - dtdV * (Ar * fr - Al * fl) : qi - dtdx * (fr - fl);
            if (a.mode == 0) Q = base;
            else Q = a.h1 * a.Q0[v * a.P + sc] + a.h2 * base;
            if (sph && v == 1) Q = Q - dtdx * (a.god[2 * a.PF + fj + 1] - a.god[2 * a.PF + fj]);
        }
        if (v == 1) Q = sgn < 0.0 ? -Q : Q;
        Qv[v] = Q;
        a.Qout[v * a.P + c] = Q;
    }
    qc_to_qf_store<NSP>(a.thermo, Qv, a.qf, a.P, c);
}

// RK4 stage (SRC/timeintegration.f90:222-356) with the reference's semantics: within the step the
// advection fluxes keep reading the primitives of the step start (only the left wall ghosts' are
// refreshed, by update_boundary), interior cells move by a quarter, a quarter and a full dt/dx
// flux difference in stages 1-3, k_s = -dt/dx dF_s accumulate as k1 + 2 k2 + 2 k3 (+ k4 at the end),
// the walls are mirrored after every stage and the block-interface ghosts are not refreshed; the
// last stage sets Q = Q0 + (1/6)(k1 + 2 k2 + 2 k3 + k4) and recovers the primitives everywhere.
template <int NSP>
__global__ void __launch_bounds__(64) k_rk4_stage(StageArgs a) {
    const long c = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= long(a.nblk) * a.S) return;
    const int b = int(c / a.S), i = int(c % a.S) - a.lv;
    const int st = a.mode;
    const bool interior = i >= 0 && i < a.ncell;
    const bool lwall = b == 0 && i < 0, rwall = b == a.nblk - 1 && i >= a.ncell;
    double Qv[NSP + 3];
    if (!interior && !lwall && !rwall) {                  // block-interface ghost: unchanged
#pragma unroll
        for (int v = 0; v < NSP + 3; ++v) {
            Qv[v] = a.Qin[v * a.P + c];
            a.Qout[v * a.P + c] = Qv[v];
        }
        if (st == 3) qc_to_qf_store<NSP>(a.thermo, Qv, a.qf, a.P, c);
        return;
    }
    int sb, si;
    double sgn;
    ghost_source(a, b, i, sb, si, sgn);
    const long sc = long(sb) * a.S + a.lv + si;
    const long fj = long(sb) * (a.ncell + 1) + si;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double dx = a.dxb ? a.dxb[sb] : a.dx;
#pragma unroll
    for (int v = 0; v < NSP + 3; ++v) {
        const double d = a.F[v * a.PF + fj + 1] - a.F[v * a.PF + fj];
        const double k = -(dt / dx) * d;
        const double qi = a.Qin[v * a.P + sc];
        double Q;
        if (st < 2) Q = qi - 0.5 * (0.5 * dt / dx) * d;
        else if (st == 2) Q = qi - (dt / dx) * d;
        else Q = a.Q0[v * a.P + sc] + (1.0 / 6.0) * (a.acc[v * a.P + sc] + k);
        if (interior && st < 3) {
            const double prev = st == 0 ? 0.0 : a.acc[v * a.P + c];
            a.acc[v * a.P + c] = st == 0 ? k : prev + 2.0 * k;
        }
        if (v == 1) Q = sgn < 0.0 ? -Q : Q;
        Qv[v] = Q;
        a.Qout[v * a.P + c] = Q;
    }
    if (st == 3 || lwall) qc_to_qf_store<NSP>(a.thermo, Qv, a.qf, a.P, c);
}

// ---------------------------------------------------------------------------------------------
// Halo between blocks of unequal cell size (interpolation_block, SRC/block.f90:499-563), after the
// kernel that filled every halo as a plain copy.  k_smr_fill: one thread per (block pair,
// variable): the coarse block's lv ghosts average pairs of fine cells, then the fine block's lv
// ghosts take the coarse cell they lie in, -+ a quarter of its minmod slope (the slope stencil
// reaches the coarse ghost just written).  k_smr_qf: QcToqf on those ghost columns.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double smr_slope(double m, double p) {
    return 0.5 * (copysign(1.0, m) + copysign(1.0, p)) * fmin(fabs(m), fabs(p));
}

__global__ void __launch_bounds__(64) k_smr_fill(SmrArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nif * a.nequ) return;
    const int ib = a.ifb[t / a.nequ], v = t % a.nequ;
    const long L = long(ib) * a.S + a.lv, R = long(ib + 1) * a.S + a.lv;   // interior cell 0 of each
    const int nc = a.ncell, lv = a.lv;
    double *__restrict__ Q = a.Q + long(v) * a.P;
    if (a.dxb[ib] > a.dxb[ib + 1]) {                   // coarse | fine
        for (int g = 1; g <= lv; ++g)
            Q[L + nc - 1 + g] = (Q[R + 2 * (g - 1)] + Q[R + 2 * (g - 1) + 1]) / 2.0;
        for (int g = 1; g <= lv; ++g) {
            const long c = L + nc - 1 - (g - 1) / 2;
            const double s = smr_slope(Q[c] - Q[c - 1], Q[c + 1] - Q[c]);
            Q[R - g] = (g % 2 == 1) ? Q[c] + 0.25 * s : Q[c] - 0.25 * s;
        }
    } else {                                            // fine | coarse
        for (int g = 1; g <= lv; ++g)
            Q[R - g] = (Q[L + nc - 1 - 2 * (g - 1)] + Q[L + nc - 1 - (2 * (g - 1) + 1)]) / 2.0;
        for (int g = 1; g <= lv; ++g) {
            const long c = R + (g - 1) / 2;
            const double s = smr_slope(Q[c] - Q[c - 1], Q[c + 1] - Q[c]);
            Q[L + nc - 1 + g] = (g % 2 == 1) ? Q[c] - 0.25 * s : Q[c] + 0.25 * s;
        }
    }
}

template <int NSP>
__global__ void __launch_bounds__(64) k_smr_qf(SmrArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nif * 2 * a.lv) return;
    const int ib = a.ifb[t / (2 * a.lv)], g = t % (2 * a.lv);
    // the left block's right ghosts, then the right block's left ghosts
    const long c = g < a.lv ? long(ib) * a.S + a.lv + a.ncell + g : long(ib + 1) * a.S + (g - a.lv);
    double Qv[NSP + 3];
#pragma unroll
    for (int v = 0; v < NSP + 3; ++v) Qv[v] = a.Q[v * a.P + c];
    qc_to_qf_store<NSP>(a.thermo, Qv, a.qf, a.P, c);
}

// ---------------------------------------------------------------------------------------------
// calc_dt (SRC/update.f90:13-59): max wave speed over interior cells -> dt on the device.
// ---------------------------------------------------------------------------------------------
template <int NSP>
__global__ void __launch_bounds__(256) k_wave_speed(DtArgs a) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    double aupc = 0.0;
    if (t < long(a.nblk) * a.ncell) {
        const int b = int(t / a.ncell), i = int(t % a.ncell);
        const long c = long(b) * a.S + a.lv + i;
        const DevThermo *__restrict__ m = a.thermo;
        double x[NSP];
#pragma unroll
        for (int k = 0; k < NSP; ++k) x[k] = a.qf[(3 + k) * a.P + c];
        ytx<NSP>(m, x);                                   // no adjust_sum here (update.f90:396)
        const double rho = a.qf[c], u = a.qf[a.P + c], T = a.qf[2 * a.P + c];
        const double wtm = mix_w<NSP>(m, x);
        const double p = (rho * 1.e-3) * m->ru * T / wtm * 0.1;
        const MixSums ms = mix_sums<NSP>(m, x, T, false, true, true);
        const double cp = ms.cpbml * 1.e-4 / wtm, cv = ms.cvbml * 1.e-4 / wtm;
        const double as = sqrt(cp / cv * p / rho);
        aupc = fmax(fmax(fabs(u - as), fabs(u)), fabs(u + as));
    }
    // wave64 max via cross-lane shuffles, then one LDS slot per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) aupc = fmax(aupc, __shfl_xor(aupc, off, 64));
    __shared__ double wmax[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wmax[w] = aupc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = wmax[0];
        for (int k = 1; k < int(blockDim.x >> 6); ++k) mx = fmax(mx, wmax[k]);
        // non-negative doubles order like their bit patterns
        atomicMax(reinterpret_cast<unsigned long long *>(a.cvel_bits), (unsigned long long)__double_as_longlong(mx));
    }
}

__global__ void k_dt_finalize(DtArgs a) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double cvel = __longlong_as_double((long long)*a.cvel_bits);
        const double dt = (a.cfl * a.dx) / cvel;
        *a.dt_out = dt;
        if (a.time) *a.time = *a.time + dt;
        *a.cvel_bits = 0ull;
    }
}

// ---------------------------------------------------------------------------------------------
// Initial state (SRC/params.f90:305-580 + SRC/block.f90:413-429): interior columns.
// ---------------------------------------------------------------------------------------------
template <int NSP>
__global__ void __launch_bounds__(64) k_init_cells(InitArgs a) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= long(a.nblk) * a.ncell) return;
    const int b = int(t / a.ncell), i = int(t % a.ncell);
    const long c = long(b) * a.S + a.lv + i;
    const DevThermo *__restrict__ m = a.thermo;
    const double *cell = a.cells + t * (3 + NSP);
    const double T = cell[0], p = cell[1], u = cell[2];
    double x[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) x[k] = cell[3 + k];
    adjust_sum<NSP>(x);
    const double wx = mix_w<NSP>(m, x);
    const double rho = wx * (p * 10.0) / (m->ru * T) * 1.e3;      // density(): CKRHOX
    const MixSums ms = mix_sums<NSP>(m, x, T, true, false, false);
    const double e = ms.ubml * 1.e-4 / wx;                        // specific_energy()
    a.qf[c] = rho; a.qf[a.P + c] = u; a.qf[2 * a.P + c] = T;
    a.Qc[c] = rho; a.Qc[a.P + c] = rho * u; a.Qc[2 * a.P + c] = rho * (e + 0.5 * (u * u));
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const double y = x[k] * m->wt[k] / wx;                    // CKXTY
        a.qf[(3 + k) * a.P + c] = y;
        a.Qc[(3 + k) * a.P + c] = rho * y;
    }
}

// ---------------------------------------------------------------------------------------------
// Conversions and reconstruction on host-array batches (stateless entry points).
// ---------------------------------------------------------------------------------------------
template <int NSP>
__global__ void k_qf_to_qc(const DevThermo *__restrict__ m, int n, const double *__restrict__ qf,
                           double *__restrict__ Qc, long P) {
    const long c = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double y[NSP];
    const double rho = qf[c], u = qf[P + c], T = qf[2 * P + c];
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = qf[(3 + k) * P + c];
    double ya[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) ya[k] = y[k];
    adjust_sum<NSP>(ya);
#pragma unroll
    for (int k = 0; k < NSP; ++k) Qc[(3 + k) * P + c] = rho * ya[k];
    ytx<NSP>(m, y);
    adjust_sum<NSP>(y);
    const double wtm = mix_w<NSP>(m, y);
    const MixSums ms = mix_sums<NSP>(m, y, T, true, false, false);
    const double e = ms.ubml * 1.e-4 / wtm;
    Qc[c] = rho;
    Qc[P + c] = rho * u;
    Qc[2 * P + c] = rho * (e + 0.5 * (u * u));
}

template <int NSP>
__global__ void k_qc_to_qf(const DevThermo *__restrict__ m, int n, const double *__restrict__ Qc,
                           double *__restrict__ qf, long P) {
    const long c = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double Qv[NSP + 3];
#pragma unroll
    for (int v = 0; v < NSP + 3; ++v) Qv[v] = Qc[v * P + c];
    qc_to_qf_store<NSP>(m, Qv, qf, P, c);
}

// qfToQc on the interior columns of a padded domain (reaction_macks' final loop,
// SRC/source.f90:78-82).
template <int NSP>
__global__ void __launch_bounds__(64) k_qf_to_qc_interior(const DevThermo *__restrict__ m, int nblk, int ncell, int lv, int S,
                                    const double *__restrict__ qf, double *__restrict__ Qc, long P) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= long(nblk) * ncell) return;
    const long c = long(t / ncell) * S + lv + (t % ncell);
    double y[NSP];
    const double rho = qf[c], u = qf[P + c], T = qf[2 * P + c];
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = qf[(3 + k) * P + c];
    double ya[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) ya[k] = y[k];
    adjust_sum<NSP>(ya);
#pragma unroll
    for (int k = 0; k < NSP; ++k) Qc[(3 + k) * P + c] = rho * ya[k];
    ytx<NSP>(m, y);
    adjust_sum<NSP>(y);
    const double wtm = mix_w<NSP>(m, y);
    const MixSums ms = mix_sums<NSP>(m, y, T, true, false, false);
    const double e = ms.ubml * 1.e-4 / wtm;
    Qc[c] = rho;
    Qc[P + c] = rho * u;
    Qc[2 * P + c] = rho * (e + 0.5 * (u * u));
}

__global__ void k_recon(int n, const double *__restrict__ s, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double st[8];
    for (int j = 0; j < 8; ++j) st[j] = s[r * 8 + j];
    double *o = out + r * 18;
    weno3z(st, o[0], o[1]);
    weno5<2>(st, o[2], o[3]);
    weno5<3>(st, o[4], o[5]);
    weno7z(st, o[6], o[7]);
    for (int f = 1; f <= 4; ++f) muscl3(st, o[6 + 2 * f], o[7 + 2 * f], f);
    o[16] = 0.0;
    o[17] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------------
static int recon_kind(int flag_acu, int flag_weno) {
    if (flag_acu == 0) return R_FIRST;
    if (flag_acu == 1) return R_MUSCL;
    switch (flag_weno) {
        case 1: return R_WENO3Z;
        case 2: return R_WENO5JS;
        case 3: return R_WENO5Z;
        case 4: return R_WENO7Z;
        default: return -1;
    }
}

template <int NSP, bool SPH>
static hipError_t launch_flux_nsp(const FluxArgs &a, int rk, hipStream_t st) {
    const long n = 2 * long(a.nblk) * (a.ncell + 1);            // two lanes per interface
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
    const dim3 grid2(unsigned((n + 127) / 128)), blk2(128);
    switch (rk) {
        case R_FIRST: hipLaunchKernelGGL((k_flux<NSP, R_FIRST, SPH>), grid, blk, 0, st, a); break;
        case R_MUSCL: hipLaunchKernelGGL((k_flux<NSP, R_MUSCL, SPH>), grid, blk, 0, st, a); break;
        case R_WENO3Z: hipLaunchKernelGGL((k_flux<NSP, R_WENO3Z, SPH>), grid, blk, 0, st, a); break;
        case R_WENO5JS:
            if (a.variant == 1) hipLaunchKernelGGL((k_flux<NSP, R_WENO5JS, SPH, 64, 1>), grid, blk, 0, st, a);
            else if (a.variant == 2) hipLaunchKernelGGL((k_flux<NSP, R_WENO5JS, SPH, 128, 1>), grid2, blk2, 0, st, a);
            else hipLaunchKernelGGL((k_flux<NSP, R_WENO5JS, SPH>), grid, blk, 0, st, a);
            break;
        case R_WENO5Z:
            if (a.variant == 1) hipLaunchKernelGGL((k_flux<NSP, R_WENO5Z, SPH, 64, 1>), grid, blk, 0, st, a);
            else if (a.variant == 2) hipLaunchKernelGGL((k_flux<NSP, R_WENO5Z, SPH, 128, 1>), grid2, blk2, 0, st, a);
            else hipLaunchKernelGGL((k_flux<NSP, R_WENO5Z, SPH>), grid, blk, 0, st, a);
            break;
        case R_WENO7Z: hipLaunchKernelGGL((k_flux<NSP, R_WENO7Z, SPH>), grid, blk, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_smr_halo(const SmrArgs &a, int nsp, hipStream_t st) {
    if (a.nif == 0) return hipSuccess;
    hipLaunchKernelGGL(k_smr_fill, dim3(unsigned((a.nif * a.nequ + 63) / 64)), dim3(64), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid(unsigned((a.nif * 2 * a.lv + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_smr_qf<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_flux(const FluxArgs &a, int nsp, int flag_acu, int flag_weno, hipStream_t st) {
    const int rk = recon_kind(flag_acu, flag_weno);
    if (rk < 0) return hipErrorInvalidValue;
#define FCFD_CASE(N) if (nsp == N) return a.spherical ? launch_flux_nsp<N, true>(a, rk, st) : launch_flux_nsp<N, false>(a, rk, st);
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

template <int NSP>
static hipError_t launch_stage_end_nsp(const StageArgs &a, hipStream_t st) {
    const long n = long(a.nblk) * a.S;
    const int bs = a.variant == 2 ? 128 : (a.variant == 3 ? 256 : 64);
    const dim3 grid(unsigned((n + bs - 1) / bs)), blk(bs);
    switch (a.variant) {
        case 1: hipLaunchKernelGGL((k_stage_end<NSP, 64, 1>), grid, blk, 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_stage_end<NSP, 128, 1>), grid, blk, 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_stage_end<NSP, 256, 1>), grid, blk, 0, st, a); break;
        default: hipLaunchKernelGGL((k_stage_end<NSP>), grid, blk, 0, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_stage_end(const StageArgs &a, int nsp, hipStream_t st) {
#define FCFD_CASE(N) if (nsp == N) return launch_stage_end_nsp<N>(a, st);
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_rk4_stage(const StageArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.S;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_rk4_stage<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_dt(const DtArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.ncell;
    const dim3 grid(unsigned((n + 255) / 256)), blk(256);
    bool launched = false;
#define FCFD_CASE(N) if (!launched && nsp == N) { hipLaunchKernelGGL((k_wave_speed<N>), grid, blk, 0, st, a); launched = true; }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    if (!launched) return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_init(const InitArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.ncell;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_init_cells<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_qf_to_qc(const DevThermo *m, int nsp, int n, const double *qf, double *Qc, long P, hipStream_t st) {
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_qf_to_qc<N>), grid, blk, 0, st, m, n, qf, Qc, P); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_qc_to_qf(const DevThermo *m, int nsp, int n, const double *Qc, double *qf, long P, hipStream_t st) {
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_qc_to_qf<N>), grid, blk, 0, st, m, n, Qc, qf, P); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_qf_to_qc_interior(const DevThermo *m, int nsp, int nblk, int ncell, int lv, int S,
                                    const double *qf, double *Qc, long P, hipStream_t st) {
    const long n = long(nblk) * ncell;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_qf_to_qc_interior<N>), grid, blk, 0, st, m, nblk, ncell, lv, S, qf, Qc, P); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_recon(int n, const double *s, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_recon, dim3(unsigned((n + 63) / 64)), dim3(64), 0, st, n, s, out);
    return hipGetLastError();
}

}  // namespace fcfd
