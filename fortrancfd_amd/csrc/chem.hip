// chem.hip -- chemistry source term: the reaction_macks call site (SRC/source.f90:17-84) and the
// stiff integrator behind macks(dt, nspe+1, yiT, rho) (SRC/source.f90:58).  The reference does not
// ship its integrator (module ODEs), so the method is this project's own:
//   state y = (Y_1..Y_KK, T) at fixed density, f from the CKWYR rates (cklib.f:9107-9204),
//   3-stage L-stable SDIRK of order 3 (Alexander 1977, gamma = 0.43586652150845900) with an
//   embedded order-2 solution, simplified Newton per stage with M = I - h*gamma*J (LU with partial
//   pivoting), filtered error M^{-1} h sum (b - b^)_i k_i, analytic mass-action Jacobian (third-body
//   and fall-off factors held fixed) with a finite-difference temperature column, and a final
//   temperature correction that keeps sum_k Y_k u_k(T) / W_k at its initial value.  Before SDIRK,
//   an explicit Heun step over dt (unless the screen already tried it) and a retry with two Heun
//   steps of dt/2; SDIRK's first step starts from the cell's step-size proposal of the last call.
//
// GPU mapping: a persistent grid pulls cells from a device work queue, so cells of very different
// stiffness (1 step vs. hundreds) balance themselves; one cell per wave for long lists, one cell per
// four-wave workgroup for short ones (k_chem<NSP, NW>).  Lanes run over reactions (rates, Jacobian
// scatter) and species (thermo, production rates); the 54x54 Newton matrix lives in LDS one row per
// lane and is factored by pivoted elimination across lanes; vectors of the state are distributed one
// component per lane.
#include <algorithm>

#include "chem.hpp"
#include "chem_device.hpp"
#include "fcfd_device.hpp"

namespace fcfd {

namespace {

constexpr double CH_G = 0.435866521508459;
constexpr double CH_A21 = 0.2820667392457705;
constexpr double CH_B1 = 1.20849664917601;
constexpr double CH_B2 = -0.6443631706844692;
constexpr double CH_E1 = 0.43586652150845895;
constexpr double CH_E2 = -0.8717330430169181;
constexpr double CH_E3 = 0.435866521508459;
constexpr double CH_KAPPA = 0.03;
constexpr int CH_MAXIT = 7;
constexpr int CH_MAXSTEPS = 5000;
constexpr int LDJ = 65;   // LDS row stride of the Jacobian (odd in 64-bit words, > NSP + 8)

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// One DPP move of a double (both halves with the same control); rows outside ROWMASK and lanes
// without a source keep their own value.
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// As dpp_f64, but lanes without a source and rows outside ROWMASK read 0.
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_f64_or0(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, true);
    return __hiloint2double(hi, lo);
}

// Sum over the wave, the same value in every lane: prefix sums within rows of 16 lanes (row_shr
// 1/2/4/8), then row totals carried across rows (row_bcast 15 / 31), lane 63 broadcast.  The order
// of the additions is fixed, so equal inputs give bitwise equal sums in every wave.
__device__ __forceinline__ double wsum(double v) {
    v = v + dpp_f64_or0<0x111>(v);
    v = v + dpp_f64_or0<0x112>(v);
    v = v + dpp_f64_or0<0x114>(v);
    v = v + dpp_f64_or0<0x118>(v);
    v = v + dpp_f64_or0<0x142, 0xA>(v);
    v = v + dpp_f64_or0<0x143, 0xC>(v);
    return bcast(v, 63);
}

// Largest value over the wave: row_shr 1/2/4/8 within each row of 16 lanes, then row_bcast 15 / 31
// across rows; lane 63 ends with the maximum.  fmax returns one of its operands, so the result is
// bitwise one of the inputs (NaN inputs are passed over).
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_f64<0x111>(v));
    v = fmax(v, dpp_f64<0x112>(v));
    v = fmax(v, dpp_f64<0x114>(v));
    v = fmax(v, dpp_f64<0x118>(v));
    v = fmax(v, dpp_f64<0x142, 0xA>(v));
    v = fmax(v, dpp_f64<0x143, 0xC>(v));
    return bcast(v, 63);
}

// Largest value over the wave of an unsigned key, the same scheme as wave_max with one DPP max per
// stage (lanes without a source and rows outside the mask contribute 0, the identity).  Every lane
// of the wave must be active.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, true));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, true));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// row[j] = fma(-l, prow[j], row[j]) for j in [j0, n): the pivot row is never the row updated, so
// the loads of a group of eight do not wait for the stores of the previous group.
__device__ __forceinline__ void eliminate(double *__restrict__ row, const double *__restrict__ prow, double l,
                                          int j0, int n) {
    int j = j0;
    for (; j + 8 <= n; j += 8) {
        double r[8], p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            r[u] = row[j + u];
            p[u] = prow[j + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) row[j + u] = fma(-l, p[u], r[u]);
    }
    for (; j < n; ++j) row[j] = fma(-l, prow[j], row[j]);
}

// LDS of one cell's workgroup of NW waves.
template <int NSP, int NW>
struct Shared {
    double J[NW][(NSP + 1) * LDJ];   // per-wave Jacobian partials; J[0] ends as the Jacobian and is
                                     // then overwritten by the LU factors of M
    double W[NW][64];                // per-wave production-rate partials
    double Cd[NW][64], Dd[NW][64];   // per-wave creation / destruction partials (heat release)
    double c[64], smh[64], u[64];
    double rdiag[64];                // 1 / pivot of each elimination step
    int perm[64];
    // the four-wave layout's alternative factorizations; a placeholder element with one wave, so
    // that five one-wave workgroups fit the LDS of a CU
    int rank[NW == 4 ? 64 : 1];                    // elimination step of each row (lu_blocked)
    double U12[NW == 4 ? 4 : 1][NW == 4 ? 1024 : 1];  // per-wave panel rows right of the panel (lu_blocked)
    double lmul[2][NW == 4 ? 64 : 1];              // multipliers of the last two elimination steps (lu_regs)
    int lu_ok;                                     // 0 once a step had no usable pivot (lu_regs)
    int cell;
};

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Rows are evaluated in blocks of 64, one row per lane.  With one wave per cell, wave 0 takes every
// block.  With NW > 1 waves, wave 0 takes block 0 alone -- rows 0..63, where the fall-off and
// third-body rows sit and every path of reaction() is taken -- and waves 1..NW-1 take the other
// blocks in turn, so the heavy block runs beside the plain ones instead of before one of them.
template <int NW>
__device__ __forceinline__ int first_block(int wave) { return wave; }
template <int NW>
__device__ __forceinline__ int block_step(int wave) { return NW == 1 ? 1 : (wave == 0 ? (1 << 20) : NW - 1); }

struct RhsOut {
    double f;      // this lane's component of dy/dt
    double cvm;    // mixture cv, erg/(g K)
    double uk;     // this lane's species internal energy, erg/mol (0 on other lanes)
    double cvk;    // this lane's species cv, erg/(mol K)
    double cd, dd; // with split: this lane's creation / destruction rate (CKCDYR)
};

// f = dy/dt at (y_l, T); y_l is this lane's component (lane < NSP: Y_lane), the same in every wave
// of the workgroup.  Leaves the concentrations in sh.c, S/R-H/RT in sh.smh and u_k in sh.u; with
// split, returns CKCDYR's creation / destruction rates (cklib.f:1497-1517) instead of using them.
// Each wave evaluates its blocks of rows (first_block / block_step) and scatters their rates of
// progress into its own partial sums with LDS atomics (one wave's atomics land in a fixed order);
// the partials are then added in wave order, so the result is reproducible.  In three pieces so
// that the Jacobian can share the reaction loop: rhs_begin (species terms, ends with a barrier),
// rhs_scatter per reaction, rhs_end (partial sums and f).
struct RhsPre {
    double uk, cpk, ctot;
};

template <int NSP, int NW>
__device__ __forceinline__ RhsPre rhs_begin(const DevThermo *__restrict__ th, Shared<NSP, NW> &sh, double rho,
                                            double y_l, double T, int wave) {
    const int lane = lane_id();
    double uk = 0.0, cpk = 0.0, ck = 0.0;
    if (lane < NSP) {
        ck = rho * y_l / th->wt[lane];
        const Nasa a = nasa(th, lane, T);
        const TPow p = tpow(T, th->ru);
        uk = h_mol(a, p) - p.rut;
        cpk = cp_mol(a, p);
        if (wave == 0) {
            sh.c[lane] = ck;
            sh.u[lane] = uk;
            sh.smh[lane] = smh_species(a, T);
        }
    }
    sh.W[wave][lane] = 0.0;
    sh.Cd[wave][lane] = 0.0;
    sh.Dd[wave][lane] = 0.0;
    const double ctot = wsum(ck);
    __syncthreads();
    return RhsPre{uk, cpk, ctot};
}

template <int NSP, int NW>
__device__ __forceinline__ void rhs_scatter(const DevKin *__restrict__ K, Shared<NSP, NW> &sh, int i, const Rop &r,
                                            bool split, int wave) {
    const int *nu = K->nu[i], *nk = K->nk[i];
    if (!split) {
        double *wdot_w = sh.W[wave];
        const double q = r.qf - r.qr;
        for (int n = 0; n < 6; ++n)
            if (nk[n] >= 0) atomicAdd(&wdot_w[nk[n]], q * (double)nu[n]);
    } else {
        double *cdot_w = sh.Cd[wave], *ddot_w = sh.Dd[wave];
        for (int n = 0; n < 3; ++n) {
            if (nk[n] >= 0) {
                const double a = (double)abs(nu[n]);
                atomicAdd(&cdot_w[nk[n]], r.qr * a);
                atomicAdd(&ddot_w[nk[n]], r.qf * a);
            }
            if (nk[n + 3] >= 0) {
                const double a = (double)nu[n + 3];
                atomicAdd(&cdot_w[nk[n + 3]], r.qf * a);
                atomicAdd(&ddot_w[nk[n + 3]], r.qr * a);
            }
        }
    }
}

template <int NSP, int NW>
__device__ __forceinline__ RhsOut rhs_end(const DevThermo *__restrict__ th, Shared<NSP, NW> &sh, double rho,
                                          double y_l, const RhsPre &pre, bool split) {
    const int lane = lane_id();
    __syncthreads();
    double wdot = 0.0, cd = 0.0, dd = 0.0;
    if (lane < NSP) {
        wdot = sh.W[0][lane];
        if (split) {
            cd = sh.Cd[0][lane];
            dd = sh.Dd[0][lane];
        }
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            wdot = wdot + sh.W[w][lane];
            if (split) {
                cd = cd + sh.Cd[w][lane];
                dd = dd + sh.Dd[w][lane];
            }
        }
    }
    const double cvk = pre.cpk - th->ru;
    const double cvm = wsum(lane < NSP ? y_l * cvk / th->wt[lane] : 0.0);
    const double su = wsum(lane < NSP ? pre.uk * wdot : 0.0);
    RhsOut o;
    o.cvm = cvm;
    o.uk = pre.uk;
    o.cvk = cvk;
    o.cd = cd;
    o.dd = dd;
    o.f = lane < NSP ? wdot * th->wt[lane] / rho : (lane == NSP ? -su / (rho * cvm) : 0.0);
    if (NW > 1) __syncthreads();   // every wave has read the partials before the next evaluation clears them
    return o;
}

template <int NSP, int NW>
__device__ __forceinline__ RhsOut rhs(const DevThermo *__restrict__ th, const DevKin *__restrict__ K, Shared<NSP, NW> &sh,
                      double rho, double y_l, double T, bool split, int wave) {
    const RhsPre pre = rhs_begin<NSP, NW>(th, sh, rho, y_l, T, wave);
    const double alogt = fm_log(T), pfac = th->pa / (th->ru * T);
    for (int b = first_block<NW>(wave); b * 64 < K->ii; b += block_step<NW>(wave)) {
        const int i = b * 64 + lane_id();
        if (i >= K->ii) continue;
        // blocks from K->nplain0 on take the branch-free evaluation (a uniform test)
        const Rop r = b * 64 >= K->nplain0 ? reaction_plain<1>(K, i, T, alogt, pfac, sh.c, sh.smh)
                                           : reaction_ilp<1>(K, i, T, alogt, pfac, pre.ctot, sh.c, sh.smh);
        rhs_scatter<NSP, NW>(K, sh, i, r, split, wave);
    }
    return rhs_end<NSP, NW>(th, sh, rho, y_l, pre, split);
}

// The integrator evaluates f at six places (explicit pre-step, Jacobian, Newton, heat release);
// one out-of-line copy serves them all (inlined at each, the rate code is some 80 kB of the
// kernel).  The shared block goes over as an LDS-address-space pointer so the callee still addresses
// it with LDS instructions.
#define FCFD_LDS __attribute__((address_space(3)))

#define FCFD_GLB __attribute__((address_space(1)))

template <int NSP, int NW>
__device__ __attribute__((noinline)) RhsOut rhs_call(FCFD_GLB const DevThermo *th, FCFD_GLB const DevKin *K,
                                                     FCFD_LDS Shared<NSP, NW> *shl, double rho, double y_l, double T,
                                                     bool split, int wave) {
    Shared<NSP, NW> &sh = *(Shared<NSP, NW> *)shl;
    return rhs<NSP, NW>((const DevThermo *)th, (const DevKin *)K, sh, rho, y_l, T, split, wave);
}

__device__ __forceinline__ FCFD_GLB const DevThermo *glb(const DevThermo *p) { return (FCFD_GLB const DevThermo *)p; }
__device__ __forceinline__ FCFD_GLB const DevKin *glb(const DevKin *p) { return (FCFD_GLB const DevKin *)p; }

template <int NSP, int NW>
__device__ __forceinline__ FCFD_LDS Shared<NSP, NW> *lds(Shared<NSP, NW> &sh) {
    return (FCFD_LDS Shared<NSP, NW> *)&sh;
}

// Jacobian df/dy into sh.J (rows/cols 0..NSP, NSP = temperature); out of line like rhs_call.
// Returns f at (y, T), which the first Newton iteration of the step's first stage reuses.
// Wave w scatters its reactions into its own partial sh.J[w]; the partials are added in wave order.
template <int NSP, int NW>
__device__ __attribute__((noinline)) RhsOut jacobian(FCFD_GLB const DevThermo *th1, FCFD_GLB const DevKin *K1,
                                                   FCFD_LDS Shared<NSP, NW> *shl, double rho, double y_l, double T,
                                                   int wave) {
    const DevThermo *__restrict__ th = (const DevThermo *)th1;
    const DevKin *__restrict__ K = (const DevKin *)K1;
    Shared<NSP, NW> &sh = *(Shared<NSP, NW> *)shl;
    const int lane = lane_id();
    constexpr int N = NSP + 1;
    // f at (y, T) and the concentration Jacobian from ONE pass over the reactions: each rate of
    // progress goes to the production rates and its derivatives to the wave's partial sh.J[wave]
    // (cleared after rhs_begin's barrier, when no wave still reads the previous LU factors)
    const RhsPre pre = rhs_begin<NSP, NW>(th, sh, rho, y_l, T, wave);
    for (int e = lane; e < N * LDJ; e += 64) sh.J[wave][e] = 0.0;
    double *Jw = sh.J[wave];
    double *J = sh.J[0];
    const double ctot = pre.ctot;
    const double alogt = fm_log(T), pfac = th->pa / (th->ru * T);
    // two of the wave's blocks per pass: their rate chains are independent and overlap; the
    // scatters follow in block order
    constexpr int RB = 2;
    const int bs = block_step<NW>(wave);
    for (int b0 = first_block<NW>(wave); b0 * 64 < K->ii; b0 += bs * RB) {
        double dq[RB][6];
        Rop rr[RB];
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const int bu = b0 + bs * u;
            const int i = min(bu * 64 + lane, K->ii - 1);
            const int *nu = K->nu[i], *nk = K->nk[i];
            const Rop r = bu * 64 >= K->nplain0
                              ? reaction_plain<1>(K, i, T, alogt, pfac, sh.c, sh.smh)
                              : reaction_ilp<1>(K, i, T, alogt, pfac, ctot, sh.c, sh.smh);
            rr[u] = r;
            // the slots' factors under selects instead of branches (a factor is applied only where
            // the slot is filled, as before)
            double pw[6];
#pragma unroll
            for (int s = 0; s < 6; ++s) pw[s] = ipow(sh.c[nk[s] >= 0 ? nk[s] : 0], abs(nu[s]));
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const int lo = s < 3 ? 0 : 3;
                const int a = abs(nu[s]);
                double d = (double)a * ipow(sh.c[nk[s] >= 0 ? nk[s] : 0], a - 1);
#pragma unroll
                for (int t = lo; t < lo + 3; ++t)
                    if (t != s) d = nk[t] >= 0 ? d * pw[t] : d;
                dq[u][s] = nk[s] < 0 ? 0.0 : (s < 3 ? r.kf * d : -(r.kr * d));
            }
        }
#pragma unroll
        for (int u = 0; u < RB; ++u) {
            const int bu = b0 + bs * u;
            if (bu * 64 >= K->ii) continue;
            const int i = bu * 64 + lane;
            if (i >= K->ii) continue;
            rhs_scatter<NSP, NW>(K, sh, i, rr[u], false, wave);
            const int *nu = K->nu[i], *nk = K->nk[i];
            if (nu[0] == 0) continue;
            for (int s = 0; s < 6; ++s) {
                if (nk[s] < 0) continue;
                for (int t = 0; t < 6; ++t)
                    if (nk[t] >= 0) atomicAdd(&Jw[nk[t] * LDJ + nk[s]], (double)nu[t] * dq[u][s]);
            }
        }
    }
    const RhsOut f0 = rhs_end<NSP, NW>(th, sh, rho, y_l, pre, false);   // its barrier also completes J
    if (NW > 1) {
        for (int e = threadIdx.x; e < N * LDJ; e += 64 * NW) {
            double v = sh.J[0][e];
#pragma unroll
            for (int w = 1; w < NW; ++w) v = v + sh.J[w][e];
            J[e] = v;
        }
        __syncthreads();
    }
    // temperature row from the concentration Jacobian, before rows are rescaled to Y
    double trow = 0.0;
    if (wave == 0 && lane < NSP) {
        double s = 0.0;
        for (int j = 0; j < NSP; ++j) s = s + sh.u[j] * J[j * LDJ + lane];
        s = s * rho / th->wt[lane];
        const double fT = bcast(f0.f, NSP);
        trow = -s / (rho * f0.cvm) - fT * (f0.cvk / th->wt[lane]) / f0.cvm;
    }
    __syncthreads();
    if (lane < NSP) {
        const double wr = th->wt[lane];
        for (int k = wave; k < NSP; k += NW) J[lane * LDJ + k] = J[lane * LDJ + k] * wr / th->wt[k];
        if (wave == 0) J[NSP * LDJ + lane] = trow;
    }
    // temperature column by a forward difference
    const double dT = T * 1.0e-6;
    const RhsOut f1 = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, y_l, T + dT, false, wave);
    if (wave == 0 && lane < N) J[lane * LDJ + NSP] = (f1.f - f0.f) / dT;
    __syncthreads();
    return f0;
}

// The pivot of an elimination step: the open row with the largest |entry| rk in the pivot column, the
// lowest such row on ties (NaN entries are passed over).  For |x| >= 0 the upper 32 bits of the
// double order like the value, so one 32-bit wave max finds the candidates; only when several rows
// share those bits are their full values compared.  If the column of every open row is NaN, the
// lowest open row, and ok = false.  Called by every lane of the wave (wave-wide reductions).
__device__ __forceinline__ int pivot_row(double rk, bool open, bool &ok) {
    const double av = fabs(rk);
    const bool num = open && av == av;
    const unsigned key = num ? (unsigned)__double2hiint(av) : 0u;
    const unsigned kmax = wave_max_u32(key);
    const bool top = num && key == kmax;
    unsigned long long hit = __ballot(top);
    if (__popcll(hit) > 1) {
        const double v = top ? av : -1.0;
        const double vmax = wave_max(v);
        hit = __ballot(top && v == vmax);
    }
    if (hit == 0ull) {
        hit = __ballot(open);
        ok = false;
    }
    return __builtin_amdgcn_readfirstlane(__ffsll((long long)hit) - 1);
}

// LU with partial pivoting of M = I - hg*J, in place in sh.J[0]: lane r owns row r (LDS row stride
// odd in 64-bit words, so the lanes of one column hit distinct banks).  Rows are not swapped:
// sh.perm[k] is the pivot row of step k; ties go to the lower row.  Every wave finds the same pivot
// and multipliers.  Column k + 1, the next pivot column, is updated by every wave in a register and
// not stored during step k; the update of columns k + 2 .. N - 1 is split into NW contiguous column
// ranges, one per wave, so between two barriers no wave writes a column that another wave reads.
// Column k gets its final values at the start of step k + 1, after the barrier: the multiplier in
// each row still open, the pivot itself in the pivot row.  Returns false if a pivot is zero.  The
// loops stay rolled (unrolled, the 55-row factorisation is some 66 kB of straight-line code).
template <int NSP, int NW>
__device__ __forceinline__ bool lu(Shared<NSP, NW> &sh, double hg, int wave) {
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    int rank = lane < N ? N : -1;   // elimination step of this lane's row (N: not pivoted yet)
    double *row = sh.J[0] + (lane < N ? lane : 0) * LDJ;
    if (lane < N) {
        for (int j = 1 + wave; j < N; j += NW) {
            double m = -hg * row[j];
            if (j == lane) m = m + 1.0;
            row[j] = m;
        }
    }
    // this row's entry in the current pivot column, carried in a register from step to step: every
    // wave applies the step's update to column k + 1 itself, so the pivot search never waits for LDS.
    // Column 0 is formed here, in registers only (read before the barrier, so no load of it is still
    // outstanding inside the loop), and gets its final values at step 1 like every pivot column.
    double rk = 0.0;
    if (lane < N) {
        rk = -hg * row[0];
        if (lane == 0) rk = rk + 1.0;
    }
    __syncthreads();
    bool ok = true;
    int kl = -1;          // column of this lane's final value still to be stored (-1: none)
    double l = 0.0;       // that value: the multiplier, or the pivot in the pivot row
    for (int k = 0; k < N; ++k) {
        if (wave == 0 && kl >= 0) row[kl] = l;
        kl = -1;
        // this row's entry in column k + 1 after the updates of steps < k (stored by the owner of
        // that column in step k - 1), read before the pivot search so that the read overlaps it
        const int k1 = k + 1 < N ? k + 1 : k;
        const double next_r = row[k1];
        const int p = pivot_row(rk, rank == N, ok);
        const double *prow = sh.J[0] + p * LDJ;
        const double piv = bcast(rk, p);
        const double next_p = bcast(next_r, p);   // the pivot row's entry in column k + 1
        if (wave == 0 && lane == 0) sh.perm[k] = p;
        if (!(piv != 0.0)) ok = false;
        if (lane == p) {
            rank = k;
            l = rk;
            kl = k;
        } else if (rank == N) {
            l = rk / piv;
            kl = k;
            const int len = max(N - (k + 2), 0);
            eliminate(row, prow, l, k + 2 + len * wave / NW, k + 2 + len * (wave + 1) / NW);
            rk = fma(-l, next_p, next_r);
        }
        __syncthreads();
    }
    if (wave == 0 && kl >= 0) row[kl] = l;
    // reciprocal pivots, all at once: the pivot of step k is entry k of row perm[k], which no later
    // step changes
    if (wave == 0 && lane < N) sh.rdiag[lane] = 1.0 / sh.J[0][sh.perm[lane] * LDJ + lane];
    __syncthreads();
    return ok;
}

typedef double fcfd_d4 __attribute__((ext_vector_type(4)));

// lu() for the four-wave workgroup, blocked: the same factors, bit for bit, with the elimination of
// each 16-column panel [c0, c1) in registers and the update of the columns right of it as matrix
// products on the FP64 matrix cores.  Per panel:
//  1. every wave factors the panel redundantly: lane r holds row r's panel entries and runs lu()'s
//     steps c0 .. c1-1 on them (the same pivots, multipliers and fused multiply-adds); wave 0 stores
//     the panel's final values, sh.perm and each row's step (sh.rank); barrier;
//  2. every wave forms the panel's pivot rows right of the panel (U12): row perm[k] after the updates
//     of steps c0 .. k-1, one column per lane, the fused multiply-adds in step order, into its own
//     block sh.U12[wave];
//  3. wave w updates the rows 16w .. 16w+15 that are still open, right of the panel, with
//     v_mfma_f64_16x16x4f64 (D = C + A B, A = minus the multipliers of four steps, B = those steps'
//     rows of U12), four instructions per 16-column tile in step order; barrier.
// v_mfma_f64_16x16x4f64 forms each element of D as fused multiply-adds in k order (checked on the
// GPU against sequential fma for every element of random 16x4x16 products), so each entry of M
// undergoes the same operations, in the same order, as under lu()'s rank-one updates.  Wave 0 writes
// a panel's U12 into its pivot rows after that panel's last barrier, when no wave reads them any
// more.  Columns N .. 63 of the open rows receive values that nothing reads.  N <= 64.
template <int NSP>
__device__ __forceinline__ bool lu_blocked(Shared<NSP, 4> &sh, double hg, int wave) {
    constexpr int N = NSP + 1;
    static_assert(N <= 64, "one row per lane");
    const int lane = lane_id();
    int rank = lane < N ? N : -1;
    double *const M = sh.J[0];
    double *row = M + (lane < N ? lane : 0) * LDJ;
    // M = I - hg*J as lu() forms it (column j by wave j mod 4)
    if (lane < N) {
        for (int j = wave; j < N; j += 4) {
            double m = -hg * row[j];
            if (j == lane) m = m + 1.0;
            row[j] = m;
        }
    }
    __syncthreads();
    bool ok = true;
    double *const ub = sh.U12[wave];
    int pend_c0 = -1;                // wave 0: first column of the panel whose U12 (in ub) is still to be
                                     // written into its pivot rows (-1: none)
    for (int c0 = 0; c0 < N; c0 += 16) {
        const int pw = min(16, N - c0), c1 = c0 + pw;
        if (wave == 0 && pend_c0 >= 0) {           // the previous panel's U12 into its pivot rows
            const int pc1 = pend_c0 + 16;
            if (lane < N - pc1) {
#pragma unroll
                for (int k = 0; k < 16; ++k) M[sh.perm[pend_c0 + k] * LDJ + pc1 + lane] = ub[k * 64 + lane];
            }
        }
        pend_c0 = -1;
        // ---- 1. the panel, in registers ----
        double P[16];
        const bool open0 = rank == N;
#pragma unroll
        for (int j = 0; j < 16; ++j) P[j] = (open0 && j < pw) ? row[c0 + j] : 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            if (j < pw) {
                const int k = c0 + j;
                const double rk = P[j];
                const int p = pivot_row(rk, rank == N, ok);
                const double piv = bcast(rk, p);
                if (wave == 0 && lane == 0) sh.perm[k] = p;
                if (!(piv != 0.0)) ok = false;
                const bool upd = rank == N && lane != p;
                if (lane == p) rank = k;
                const double l = rk / piv;
                if (upd) P[j] = l;
                // the pivot row's entries (v_readlane reads its lane whatever the branch), and the
                // updates of the open rows
#pragma unroll
                for (int jj = j + 1; jj < 16; ++jj)
                    if (jj < pw) {
                        const double pv = bcast(P[jj], p);
                        if (upd) P[jj] = fma(-l, pv, P[jj]);
                    }
            }
        }
        if (wave == 0) {
            if (lane < N && rank >= c0) {
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (j < pw) row[c0 + j] = P[j];
            }
            sh.rank[lane] = rank;
        }
        __syncthreads();
        if (c1 < N) {                              // then pw = 16
            // ---- 2. U12: the panel's pivot rows right of the panel, column c1 + lane ----
            const int ncol = N - c1;
            double uu[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double *prow = M + sh.perm[c0 + k] * LDJ;     // its multipliers: stored above
                double v = lane < ncol ? prow[c1 + lane] : 0.0;
#pragma unroll
                for (int j = 0; j < k; ++j) v = fma(-prow[c0 + j], uu[j], v);
                uu[k] = v;
                ub[k * 64 + lane] = lane < ncol ? v : 0.0;
            }
            if (wave == 0) pend_c0 = c0;
            // ---- 3. the open rows 16 wave .. 16 wave + 15 right of the panel ----
            const int ra = 16 * wave + (lane & 15);        // row of this lane's A entries
            const bool a_open = ra < N && sh.rank[ra] >= c1;
            double A[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) A[q] = a_open ? -M[ra * LDJ + c0 + 4 * q + (lane >> 4)] : 0.0;
            const int nct = (ncol + 15) / 16;
            for (int ct = 0; ct < nct; ++ct) {
                const int col = c1 + 16 * ct + (lane & 15);
                fcfd_d4 acc;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = 16 * wave + (lane >> 4) + 4 * v;
                    acc[v] = rr < N ? M[rr * LDJ + col] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[q], ub[(4 * q + (lane >> 4)) * 64 + 16 * ct + (lane & 15)],
                                                              acc, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = 16 * wave + (lane >> 4) + 4 * v;
                    if (rr < N && sh.rank[rr] >= c1) M[rr * LDJ + col] = acc[v];
                }
            }
        }
        __syncthreads();
    }
    if (wave == 0 && pend_c0 >= 0) {
        const int pc1 = pend_c0 + 16;
        if (lane < N - pc1) {
#pragma unroll
            for (int k = 0; k < 16; ++k) M[sh.perm[pend_c0 + k] * LDJ + pc1 + lane] = ub[k * 64 + lane];
        }
    }
    __syncthreads();
    if (wave == 0 && lane < N) sh.rdiag[lane] = 1.0 / M[sh.perm[lane] * LDJ + lane];
    __syncthreads();
    return ok;
}

// lu() for the four-wave workgroup with the matrix in registers: wave w holds columns
// w*NCW .. w*NCW + NCW - 1 of every row (lane r: row r), so an elimination step moves no matrix data
// through LDS.  At step k the wave owning column k finds the pivot (pivot_row, as lu()), forms the
// open rows' multipliers l = rk / piv and publishes them with the pivot row (sh.lmul[k % 2],
// sh.perm[k]); after one barrier every wave applies fma(-l, u_j, m_j) to its columns right of k, u_j
// being the pivot row's entry (read from that lane).  Each entry sees the operations of lu() in the
// same order, so the factors are bitwise lu()'s; they are written back to sh.J[0] for solve().  The
// multipliers alternate between two buffers: a wave writes step k + 2's only after step k + 1's
// barrier, which every wave passes after reading step k's.
template <int NSP>
__device__ __forceinline__ bool lu_regs(Shared<NSP, 4> &sh, double hg, int wave) {
    constexpr int N = NSP + 1;
    constexpr int NCW = (N + 3) / 4;      // columns per wave
    static_assert(N <= 64, "one row per lane");
    const int lane = lane_id();
    int rank = lane < N ? N : -1;
    double *const M = sh.J[0];
    const double *row = M + (lane < N ? lane : 0) * LDJ;
    const int j0 = wave * NCW;
    double col[NCW];
#pragma unroll
    for (int jj = 0; jj < NCW; ++jj) {
        const int j = j0 + jj;
        double m = 0.0;
        if (lane < N && j < N) {
            m = -hg * row[j];
            if (j == lane) m = m + 1.0;
        }
        col[jj] = m;
    }
    if (threadIdx.x == 0) sh.lu_ok = 1;
    __syncthreads();
    for (int blk = 0; blk < 4; ++blk) {
#pragma unroll
        for (int jj = 0; jj < NCW; ++jj) {
            const int k = blk * NCW + jj;
            if (k >= N) continue;
            double *lk = sh.lmul[k & 1];
            if (wave == blk) {
                bool ok = true;
                const double rk = col[jj];
                const int p = pivot_row(rk, rank == N, ok);
                const double piv = bcast(rk, p);
                if (!(piv != 0.0)) ok = false;
                if (!ok && lane == 0) sh.lu_ok = 0;
                if (lane == 0) sh.perm[k] = p;
                const bool upd = rank == N && lane != p;
                const double l = rk / piv;
                lk[lane] = upd ? l : 0.0;
                if (upd) col[jj] = l;
#pragma unroll
                for (int jn = jj + 1; jn < NCW; ++jn) {
                    const double u = bcast(col[jn], p);
                    if (upd) col[jn] = fma(-l, u, col[jn]);
                }
            }
            __syncthreads();
            const int p = __builtin_amdgcn_readfirstlane(sh.perm[k]);
            if (wave > blk) {
                const bool upd = rank == N && lane != p;
                const double l = lk[lane];
#pragma unroll
                for (int jn = 0; jn < NCW; ++jn) {
                    const double u = bcast(col[jn], p);
                    if (upd) col[jn] = fma(-l, u, col[jn]);
                }
            }
            if (lane == p) rank = k;
        }
    }
    if (lane < N) {
        double *w = M + lane * LDJ;
#pragma unroll
        for (int jj = 0; jj < NCW; ++jj)
            if (j0 + jj < N) w[j0 + jj] = col[jj];
    }
    __syncthreads();
    const bool ok = sh.lu_ok != 0;
    if (wave == 0 && lane < N) sh.rdiag[lane] = 1.0 / M[sh.perm[lane] * LDJ + lane];
    __syncthreads();
    return ok;
}

// dst[0 .. n) = src[0 .. n), by the whole workgroup.  The leading barrier makes every wave finish
// its reads of dst first (a rejected attempt restores J over the factors its last solve read).
__device__ __forceinline__ void copy_j(double *dst, const double *src, int n) {
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
    __syncthreads();
}

// Solves M x = b with the factors in sh.J[0]; b and x distributed one component per lane (every
// wave solves the same system).  Lane r carries the right-hand side of the row pivoted at step r
// (row sh.perm[r]), so each step's broadcast comes from the lane the loop counter names.  Each step
// of the two substitutions is one broadcast and one fused multiply-add (the back substitution
// multiplies by the stored reciprocal pivot), the whole of the chain from one step to the next.
template <int NSP, int NW>
__device__ __forceinline__ double solve(const Shared<NSP, NW> &sh, double b) {
    constexpr int N = NSP + 1;
    static_assert(N + 7 <= LDJ, "the eight-step loads read up to column N + 6 of a row");
    const int lane = lane_id();
    const int pr = lane < N ? sh.perm[lane] : 0;
    const double *row = sh.J[0] + pr * LDJ;
    double bq = __shfl(b, pr, 64);
    // the factors of eight steps are loaded together, ahead of the chain that consumes them (row
    // storage is LDJ > N + 7 wide, so the reads past column N - 1 stay inside the row)
    for (int k0 = 0; k0 < N; k0 += 8) {
        double rv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) rv[u] = row[k0 + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + u;
            if (k < N) {
                const double bp = bcast(bq, k);
                if (lane > k && lane < N) bq = fma(-rv[u], bp, bq);
            }
        }
    }
    double x = 0.0;
    for (int k1 = N - 1; k1 >= 0; k1 -= 8) {
        double rv[8], dv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k1 - u >= 0 ? k1 - u : 0;
            rv[u] = row[k];
            dv[u] = sh.rdiag[k];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k1 - u;
            if (k >= 0) {
                const double xk = bcast(bq, k) * dv[u];
                if (lane == k) x = xk;
                if (lane < k) bq = fma(-rv[u], xk, bq);
            }
        }
    }
    return x;
}

__device__ __forceinline__ double wnorm(double v, double w, int n) {
    const double t = v * w;
    return sqrt(wsum(t * t) / n);
}

// Mixture e = sum Y u / W and cv = sum Y cv / W at T (wave sums); y_l this lane's Y.
template <int NSP>
__device__ __forceinline__ void energy(const DevThermo *__restrict__ th, double y_l, double T, double &e, double &cv) {
    const int lane = lane_id();
    double eu = 0.0, ec = 0.0;
    if (lane < NSP) {
        const Nasa a = nasa(th, lane, T);
        const TPow p = tpow(T, th->ru);
        eu = y_l * (h_mol(a, p) - p.rut) / th->wt[lane];
        ec = y_l * (cp_mol(a, p) - th->ru) / th->wt[lane];
    }
    e = wsum(eu);
    cv = wsum(ec);
}

__device__ __forceinline__ double adjusted(double y_l, int lane, int nsp) {
    const double v = lane < nsp ? fmax(y_l, 0.0) : 0.0;
    return v / wsum(v);
}

// NW waves per cell.  NW = 1: one cell per wave, for long work lists (throughput).  NW = 4: the four
// waves of a workgroup split every reaction loop and the LU updates of one cell and run the rest of
// the integrator redundantly on identical values, so they take identical decisions and meet at the
// same barriers; for short lists, where the launch lasts as long as its slowest cell.  A launch
// runs only if its mode matches the list length (a.coop_max), read on the device.  The one-wave
// mode is held to two waves per SIMD (256 registers a lane, a few spilled): its LDS admits five
// workgroups per CU, which one wave per SIMD would cap at four.
template <int NSP, int NW>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 1 ? 2 : 1))) k_chem(ChemArgs a) {
    __shared__ Shared<NSP, NW> sh;
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    const int wave = NW > 1 ? wave_id() : 0;
    const DevThermo *__restrict__ th = a.thermo;
    const DevKin *__restrict__ K = a.kin;
    const int nitems = a.list_n ? int(*a.list_n) : a.ncells;
    if ((NW > 1) != (nitems <= a.coop_max)) return;
    for (;;) {
        if (threadIdx.x == 0) sh.cell = int(atomicAdd(a.counter, 1u));
        __syncthreads();
        const int item = sh.cell;
        __syncthreads();
        if (item >= nitems) break;
        const int cell = a.list ? a.list[item] : item;

        // ---- load (reaction_macks, SRC/source.f90:44-56) ----
        long col = 0;
        double rho_si, dt, T0, y_raw = 0.0;
        if (a.layout == 0) {
            const int b = cell / a.ncell, i = cell % a.ncell;
            col = long(b) * a.S + a.lv + i;
            rho_si = a.qf[col];
            T0 = a.qf[2 * a.P + col];
            dt = a.dt_dev ? *a.dt_dev : a.dt;
            if (lane < NSP) y_raw = a.qf[(3 + lane) * a.P + col];
        } else {
            rho_si = a.b_rho[cell];
            dt = a.b_dt[cell];
            T0 = a.b_T[cell];
            if (lane < NSP) y_raw = a.b_y[long(lane) * a.ncells + cell];
        }
        const double rho = rho_si * 1.e-3;
        double y = a.layout == 0 ? adjusted(y_raw, lane, NSP) : y_raw;   // adjust_sum(y_sp)
        double e0, cv0;
        energy<NSP>(th, adjusted(y, lane, NSP), T0, e0, cv0);
        if (lane == NSP) y = T0;

        double t = 0.0, h = dt;
        if (a.hmem) {   // the last call's proposal for this cell, if smaller than dt
            const double hp = a.hmem[cell];
            if (hp > 0.0 && hp < dt) h = hp;
        }
        bool expl = false;
        int steps = 0, rejects = 0, newton = 0;
        // ---- explicit pre-step: Heun over the whole dt, accepted when the embedded Euler error
        //      estimate is within tolerance (quiescent cells never build a Jacobian) ----
        if (!a.skip_explicit) {
            const double w = lane < NSP ? 1.0 / (a.atol + a.rtol * fabs(y))
                                        : (lane == NSP ? 1.0 / (a.rtol * fabs(y)) : 0.0);
            const RhsOut f0 = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, y, bcast(y, NSP), false, wave);
            const double y1 = lane < N ? y + dt * f0.f : 0.0;
            const RhsOut f1 = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, y1, bcast(y1, NSP), false, wave);
            const double err = wnorm(lane < N ? (dt * 0.5) * (f1.f - f0.f) : 0.0, w, N);
            if (err <= 1.0) {
                if (lane < N) y = y + (dt * 0.5) * (f0.f + f1.f);
                t = dt;
                steps = 1;
                expl = true;
            }
        }
        // ---- explicit retry: two Heun steps of dt/2, each accepted by its own embedded Euler
        //      estimate (weights at the step's start).  Trace radicals at the atol level in a cold
        //      mixture relax so fast that one explicit step of dt overshoots (dt times their
        //      consumption rate near 2) while a half step does not; such cells need no Jacobian ----
        if (t < dt && !(a.retry_done && *a.retry_done)) {
            const double hs = dt * 0.5;
            double yh = y;
            bool ok = true;
            for (int sub = 0; sub < 2 && ok; ++sub) {
                const double w = lane < NSP ? 1.0 / (a.atol + a.rtol * fabs(yh))
                                            : (lane == NSP ? 1.0 / (a.rtol * fabs(yh)) : 0.0);
                const RhsOut fa = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, yh, bcast(yh, NSP), false, wave);
                const double yb = lane < N ? yh + hs * fa.f : 0.0;
                const RhsOut fb = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, yb, bcast(yb, NSP), false, wave);
                const double err = wnorm(lane < N ? (hs * 0.5) * (fb.f - fa.f) : 0.0, w, N);
                if (err <= 1.0) {
                    if (lane < N) yh = yh + (hs * 0.5) * (fa.f + fb.f);
                } else {
                    ok = false;
                }
            }
            if (ok) {
                y = yh;
                t = dt;
                steps = 2;
                expl = true;
            }
        }
        // ---- SDIRK3 over [0, dt] ----
        while (t < dt && steps + rejects < CH_MAXSTEPS) {
            if (h > dt - t) h = dt - t;
            const double Tn = bcast(y, NSP);
            const double w = lane < NSP ? 1.0 / (a.atol + a.rtol * fabs(y))
                                        : (lane == NSP ? 1.0 / (a.rtol * fabs(y)) : 0.0);
            // J at (y, Tn) and f there, the first Newton evaluation of stage 0.  lu() overwrites J;
            // with several waves the spare partial sh.J[1] keeps a copy, so a rejected attempt (same
            // y) restores J instead of evaluating it again.
            RhsOut fy;
            bool have_j = false;
            for (;;) {
                if (!have_j) {
                    fy = jacobian<NSP, NW>(glb(th), glb(K), lds(sh), rho, y, Tn, wave);
                    if (NW > 1) copy_j(sh.J[1], sh.J[0], N * LDJ);
                } else {
                    copy_j(sh.J[0], sh.J[NW > 1 ? 1 : 0], N * LDJ);
                }
                have_j = NW > 1;
                const double hg = h * CH_G;
                bool ok;
                if constexpr (NW == 4)
                    ok = a.lu_blocked == 1 ? lu_blocked<NSP>(sh, hg, wave)
                                           : (a.lu_blocked == 2 ? lu_regs<NSP>(sh, hg, wave) : lu<NSP, NW>(sh, hg, wave));
                else ok = lu<NSP, NW>(sh, hg, wave);
                double k0 = 0.0, k1 = 0.0, k2 = 0.0, z = 0.0;
                for (int st = 0; st < 3 && ok; ++st) {
                    double base = y;
                    if (st == 1) base = base + h * CH_A21 * k0;
                    if (st == 2) base = base + h * CH_B1 * k0 + h * CH_B2 * k1;
                    z = st == 0 ? base : base + hg * (st == 1 ? k0 : k1);
                    double prev = 0.0;
                    bool conv = false;
                    for (int it = 0; it < CH_MAXIT; ++it) {
                        const double Tz = bcast(z, NSP);
                        RhsOut f;
                        if (st == 0 && it == 0) f = fy;      // z = y, Tz = Tn
                        else f = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, z, Tz, false, wave);
                        double r = lane < N ? -(z - base - hg * f.f) : 0.0;
                        r = solve<NSP, NW>(sh, r);
                        if (lane < N) z = z + r;
                        ++newton;
                        const double nrm = wnorm(lane < N ? r : 0.0, w, N);
                        if (!(nrm == nrm) || (it > 0 && nrm > 0.9 * prev)) break;
                        if (nrm <= CH_KAPPA) { conv = true; break; }
                        prev = nrm;
                    }
                    if (!conv) { ok = false; break; }
                    const double kv = (z - base) / hg;
                    if (st == 0) k0 = kv; else if (st == 1) k1 = kv; else k2 = kv;
                }
                if (!ok) {
                    ++rejects;
                    h = h * 0.25;
                    if (steps + rejects >= CH_MAXSTEPS) break;
                    continue;
                }
                double ev = lane < N ? h * (CH_E1 * k0 + CH_E2 * k1 + CH_E3 * k2) : 0.0;
                ev = solve<NSP, NW>(sh, ev);
                const double err = wnorm(lane < N ? ev : 0.0, w, N);
                const double fac = 0.9 * fm_pow(fmax(err, 1.0e-10), -1.0 / 3.0);
                if (err <= 1.0) {
                    if (lane < N) y = z;
                    t = t + h;
                    ++steps;
                    h = h * fmin(4.0, fmax(0.2, fac));
                    break;
                }
                ++rejects;
                h = h * fmax(0.2, fac);
                if (steps + rejects >= CH_MAXSTEPS) break;
            }
        }
        const int failed = t < dt ? 1 : 0;

        // ---- energy-consistent temperature ----
        const double ya = adjusted(lane < NSP ? y : 0.0, lane, NSP);
        double T = bcast(y, NSP);
        for (int it = 0; it < 20; ++it) {
            double ek, cvk;
            energy<NSP>(th, ya, T, ek, cvk);
            const double d = (e0 - ek) / cvk;
            T = T + d;
            if (fabs(d) <= 1.0e-12 * T) break;
        }

        // ---- store (SRC/source.f90:60-73); wave 0 writes ----
        if (a.layout == 0) {
            if (wave == 0 && lane < NSP) a.qf[(3 + lane) * a.P + col] = ya;
            if (wave == 0 && lane == 0) a.qf[2 * a.P + col] = T;
            if (a.hrr) {
                // heat_release_rate (SRC/thermo.f90:98-101)
                const double term = lane < NSP ? rho * th->h298[lane] * (ya - y_raw) / dt : 0.0;
                const double s1 = wsum(term);
                // heat_release_rate_omegadot via CKCDYR at the new state (SRC/thermo.f90:128-132)
                const RhsOut fc = rhs_call<NSP, NW>(glb(th), glb(K), lds(sh), rho, ya, T, true, wave);
                const double cd = lane < NSP ? fc.cd : 0.0;
                const double dd = lane < NSP ? fc.dd : 0.0;
                const double s2 = wsum(lane < NSP ? th->wt[lane] * th->h298[lane] * (cd - dd) : 0.0);
                if (wave == 0 && lane == 0) {
                    a.hrr[cell] = -s1 * 1.e-1;
                    a.hrr_dot[cell] = -s2 * 1.e-1;
                }
            }
        } else {
            if (wave == 0 && lane < NSP) a.b_y[long(lane) * a.ncells + cell] = y;
            if (wave == 0 && lane == 0) a.b_T[cell] = T;
        }
        if (a.stats && wave == 0 && lane == 0) a.stats[cell] = ChemCellStats{steps, rejects, newton, failed};
        // route of a cell the explicit attempts finished: 2 the step over dt, 3 the two half steps
        if (a.route && expl && wave == 0 && lane == 0) a.route[cell] = steps == 2 ? 3 : 2;
        if (a.hmem && wave == 0 && lane == 0) a.hmem[cell] = expl ? 0.0 : h;
    }
}

// FP64 verification of surrogate candidates, ONE CANDIDATE PER WAVE with the integrator's
// right-hand side: f at both ends of the step and the trapezoidal residual Y1 - Y0 - dt/2 (f0 + f1),
// weighted like the integrator's error norm (1 / (atol + rtol max(|Y0|, |Y1|)), 1 / (rtol T0)).  The
// candidate list is short (the stiff cells), so a grid of one-wave workgroups strides over it and
// every candidate is in flight at once.
struct VerifyArgs {
    const int *cand_list;
    const unsigned *ncand;
    const double *cand;
    int *fallback;
    unsigned *nfallback;
    double tol;
};

// (two waves per SIMD like k_chem's one-wave mode, whose RHS function it shares)
template <int NSP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_chem_verify(ChemArgs a, VerifyArgs v) {
    __shared__ Shared<NSP, 1> sh;
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    const DevThermo *__restrict__ th = a.thermo;
    const DevKin *__restrict__ K = a.kin;
    const int n = int(*v.ncand);
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int cell = v.cand_list[q];
        long col = 0;
        double rho_si, dt, T0, y0 = 0.0, y_raw = 0.0;
        if (a.layout == 0) {
            const int b = cell / a.ncell, i = cell % a.ncell;
            col = long(b) * a.S + a.lv + i;
            rho_si = a.qf[col];
            T0 = a.qf[2 * a.P + col];
            dt = a.dt_dev ? *a.dt_dev : a.dt;
            if (lane < NSP) y0 = a.qf[(3 + lane) * a.P + col];
            y_raw = y0;
            y0 = adjusted(y0, lane, NSP);                       // adjust_sum, as the integrator's load
        } else {
            rho_si = a.b_rho[cell];
            dt = a.b_dt[cell];
            T0 = a.b_T[cell];
            if (lane < NSP) y0 = a.b_y[long(lane) * a.ncells + cell];
        }
        const double rho = rho_si * 1.e-3;
        const double y1 = lane < N ? v.cand[long(q) * N + lane] : 0.0;   // lane NSP: T1
        const double T1 = bcast(y1, NSP);
        const double f0 = rhs_call<NSP, 1>(glb(th), glb(K), lds(sh), rho, y0, T0, false, 0).f;
        const double f1 = rhs_call<NSP, 1>(glb(th), glb(K), lds(sh), rho, y1, T1, false, 0).f;
        // heat-release diagnostics of an accepted candidate (SRC/thermo.f90:98-101, 128-132): HRR from
        // the raw composition before the call and the candidate; HRR_dot = -sum W_k h_k wdot_k at the
        // candidate, from the production rates just evaluated there (wdot_k W_k = f_k rho)
        double s1 = 0.0, s2 = 0.0;
        if (a.hrr) {
            s1 = wsum(lane < NSP ? rho * th->h298[lane] * (y1 - y_raw) / dt : 0.0);
            s2 = wsum(lane < NSP ? th->h298[lane] * f1 * rho : 0.0);
        }
        double r = 0.0;
        if (lane < NSP)
            r = (y1 - y0 - (dt * 0.5) * (f0 + f1)) * (1.0 / (a.atol + a.rtol * fmax(fabs(y0), fabs(y1))));
        else if (lane == NSP)
            r = (T1 - T0 - (dt * 0.5) * (f0 + f1)) * (1.0 / (a.rtol * fabs(T0)));
        const double err = sqrt(wsum(r * r) / double(N));
        if (err <= v.tol) {
            if (a.layout == 0) {
                if (lane < NSP) a.qf[(3 + lane) * a.P + col] = y1;
                if (lane == 0) a.qf[2 * a.P + col] = T1;
                if (a.hrr && lane == 0) {
                    a.hrr[cell] = -s1 * 1.e-1;
                    a.hrr_dot[cell] = -s2 * 1.e-1;
                }
            } else {
                if (lane < NSP) a.b_y[long(lane) * a.ncells + cell] = y1;
                if (lane == 0) a.b_T[cell] = T1;
            }
            if (lane == 0) {
                if (a.route) a.route[cell] = 1;
                if (a.stats) a.stats[cell] = ChemCellStats{1, 0, 0, 0};
                if (a.hmem) a.hmem[cell] = 0.0;
            }
        } else if (lane == 0) {
            v.fallback[atomicAdd(v.nfallback, 1u)] = cell;
        }
    }
}

// CKWYR and CKCDYR (cklib.f:9107-9204, 1422-1541) of n states through the integrator's own rate
// code: rhs_begin (concentrations, S/R - H/RT, ctot), reaction / reaction_plain per row and
// rhs_scatter into the production-rate sums, exactly as an integrator step evaluates them; the
// rates of progress are also stored per reaction in the mechanism's order.  One state per wave.
template <int NSP>
__global__ void __launch_bounds__(64) k_rates(const DevThermo *__restrict__ th, const DevKin *__restrict__ K, int n,
                                              const double *__restrict__ rho, const double *__restrict__ Tin,
                                              const double *__restrict__ Y, double *wdot, double *cdot,
                                              double *ddot, double *qf, double *qr) {
    __shared__ Shared<NSP, 1> sh;
    const int lane = lane_id();
    for (int c = blockIdx.x; c < n; c += gridDim.x) {
        const double r = rho[c], T = Tin[c];
        const double y_l = lane < NSP ? Y[long(c) * NSP + lane] : 0.0;
        const double alogt = fm_log(T), pfac = th->pa / (th->ru * T);
        for (int split = 0; split < 2; ++split) {
            const RhsPre pre = rhs_begin<NSP, 1>(th, sh, r, y_l, T, 0);
            for (int b = 0; b * 64 < K->ii; ++b) {
                const int i = b * 64 + lane;
                if (i >= K->ii) continue;
                const Rop q = b * 64 >= K->nplain0 ? reaction_plain<1>(K, i, T, alogt, pfac, sh.c, sh.smh)
                                                   : reaction_ilp<1>(K, i, T, alogt, pfac, pre.ctot, sh.c, sh.smh);
                rhs_scatter<NSP, 1>(K, sh, i, q, split == 1, 0);
                if (split == 0) {
                    qf[long(c) * K->ii + K->orig[i]] = q.qf;
                    qr[long(c) * K->ii + K->orig[i]] = q.qr;
                }
            }
            __syncthreads();
            if (lane < NSP) {
                if (split == 0) {
                    wdot[long(c) * NSP + lane] = sh.W[0][lane];
                } else {
                    cdot[long(c) * NSP + lane] = sh.Cd[0][lane];
                    ddot[long(c) * NSP + lane] = sh.Dd[0][lane];
                }
            }
            __syncthreads();
        }
    }
}

__global__ void k_zero_counter(unsigned int *c) {
    if (threadIdx.x == 0) *c = 0u;
}

__global__ void k_scale_scalar(const double *in, double *out, double f) {
    if (threadIdx.x == 0) *out = f * *in;
}

// Stable partition of a work list by the cells' attempt counts (steps + rejects) of the previous
// integration: first those with at least CH_HEAVY, then the rest.  A launch lasts as long as its
// slowest cell, and the cells that were slow in the last step are, near a flame, the slow ones in
// this step; started first, they do not wait behind a queue of cheap cells.  One workgroup.
constexpr int CH_HEAVY = 4;

__global__ void __launch_bounds__(1024) k_heavy_first(const int *list, const unsigned *list_n,
                                                      const ChemCellStats *stats, int *out) {
    __shared__ unsigned wcount[16], woff[16], base, total;
    const int n = int(*list_n);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) base = 0;
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
        for (int i0 = 0; i0 < n; i0 += 1024) {
            const int i = i0 + t;
            int cell = 0;
            bool take = false;
            if (i < n) {
                cell = list[i];
                const ChemCellStats st = stats[cell];
                take = (st.steps + st.rejects >= CH_HEAVY) == (pass == 0);
            }
            const unsigned long long b = __ballot(take);
            const unsigned rank = unsigned(__popcll(b & ((1ull << lane) - 1ull)));
            if (lane == 0) wcount[w] = unsigned(__popcll(b));
            __syncthreads();
            if (t == 0) {
                unsigned s = 0;
                for (int q = 0; q < 16; ++q) { woff[q] = s; s += wcount[q]; }
                total = s;
            }
            __syncthreads();
            if (take) out[base + woff[w] + rank] = cell;
            __syncthreads();
            if (t == 0) base += total;
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_scale_scalar(const double *in, double *out, double f, hipStream_t st) {
    hipLaunchKernelGGL(k_scale_scalar, dim3(1), dim3(64), 0, st, in, out, f);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned *p, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_rates(const DevThermo *th, const DevKin *K, int nsp, int n, const double *rho, const double *T,
                        const double *Y, double *wdot, double *cdot, double *ddot, double *qf, double *qr,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = unsigned(std::min(n, 4096));
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_rates<N>), dim3(nb), dim3(64), 0, st, th, K, n, rho, T, Y, \
                                                        wdot, cdot, ddot, qf, qr); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_chem_verify(const ChemArgs &a, int nsp, const int *cand_list, const unsigned *ncand,
                              const double *cand, int rows_max, int *fallback, unsigned *nfallback, double tol,
                              hipStream_t st) {
    // one-wave workgroups, at most 1024 (four per CU; two waves per SIMD and the LDS would admit five)
    const unsigned nb = unsigned(std::min(rows_max, 1024));
    if (nb == 0) return hipSuccess;
    VerifyArgs v{cand_list, ncand, cand, fallback, nfallback, tol};
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_chem_verify<N>), dim3(nb), dim3(64), 0, st, a, v); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_chem_full(ChemArgs a, int nsp, int n_workgroups, int *list, unsigned *list_n, hipStream_t st) {
    hipError_t e = launch_zero_u32(list_n, st);
    if (e != hipSuccess) return e;
    a.stiff = list;
    a.stiff_n = list_n;
    a.list = nullptr;
    a.list_n = nullptr;
    if ((e = launch_chem_screen(a, nsp, st)) != hipSuccess) return e;
    a.list = list;
    a.list_n = list_n;
    a.skip_explicit = 1;
    return launch_chem(a, nsp, n_workgroups, st);
}

// n_workgroups one-wave workgroups (five per CU by default) for long lists; a quarter as many
// four-wave workgroups for lists of at most a.coop_max cells (the tuned factor, default 8, times
// n_workgroups).  The four-wave workgroup's LDS admits one per CU: the first CU-count of them run,
// and the others start only once the queue has drained and return.  Up to coop_max the launch is set
// by its slowest cells, which run about twice as fast with four waves, more than the four-wave mode's
// lower throughput costs.  Both kernels are launched; the one whose mode does not match the list
// length on the device returns at once.
hipError_t launch_chem(const ChemArgs &a0, int nsp, int n_workgroups, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, a0.counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ChemArgs a = a0;
    a.coop_max = chem_coop_max(n_workgroups, nsp);
    a.lu_blocked = chem_lu_variant(nsp);
    if (a.order && a.list && a.list_n && a.stats && a.layout == 0) {
        hipLaunchKernelGGL(k_heavy_first, dim3(1), dim3(1024), 0, st, a.list, a.list_n, a.stats, a.order);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        a.list = a.order;
    }
    const int nwg1 = std::min(n_workgroups, a.ncells);
    const int nwg4 = std::min(std::max(n_workgroups / 4, 1), a.ncells);
    if (nwg1 <= 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) { \
        hipLaunchKernelGGL((k_chem<N, 4>), dim3(nwg4), dim3(256), 0, st, a); \
        if ((e = hipGetLastError()) != hipSuccess) return e; \
        hipLaunchKernelGGL((k_chem<N, 1>), dim3(nwg1), dim3(64), 0, st, a); \
        return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
