// chem_lane.hip -- explicit screen of the chemistry source: ONE CELL PER LANE.
//
// Every cell first tries the integrator's explicit pre-step (Heun over the whole dt with the
// embedded Euler error estimate, accepted when the weighted RMS error is <= 1, then the
// energy-consistent temperature).  Quiescent cells -- the cold bulk of a reacting domain -- end
// here; the others are appended to a work list for the wave-per-cell SDIRK kernel (chem.hip),
// which then starts from the unchanged state.  The arithmetic is the pre-step of that kernel; only
// the order of the species sums differs.  k_chem_retry runs the integrator's next explicit attempt
// (two half steps) the same way over a long work list, before the SDIRK kernel sees it.
//
// Mapping: a workgroup of NW = 8 waves owns NC = 64 cells (lane l of every wave is cell l; with
// NC < 64, lanes NC..63 would repeat cell NC - 1 and store nothing).  Concentrations and S/R - H/RT of
// the cells live in LDS as species-major [k][NC] blocks.  The reactions are taken in chunks of
// CH = 32 rows: each wave evaluates four rows of a chunk for all cells and puts their rates of
// progress in LDS; after a barrier every wave adds the chunk's contributions to the species it owns
// (k mod NW = wave) from a precomputed list, in row order, into one shared production-rate block
// whose rows no two waves write.  Each species' production rate is therefore one sequential sum over
// the rows, independent of the wave count.  Species-wise work (thermo, state update, norms) is split
// over the waves by k mod NW; cell-wise scalars (T, sums) are combined through LDS in wave order.
// The LDS footprint (~101 KB) holds one workgroup per compute unit: two waves per SIMD, and every
// lane a cell, so 65,536 cells make exactly four rounds of 1024 workgroups on 256 CUs.  Reaction
// tables are read with a wave-uniform row index (scalar loads).
#include <algorithm>

#include "chem.hpp"
#include "chem_device.hpp"
#include "fcfd_device.hpp"

namespace fcfd {

namespace {

constexpr int NW = LANE_NW;    // waves per workgroup
constexpr int NC = LANE_NC;    // cells per workgroup
constexpr int CH = LANE_CHUNK; // reaction rows per chunk, CH / NW per wave
constexpr int RB = CH / NW;

template <int NSP>
struct LaneShared {
    double C[NSP * NC];        // concentrations, mol/cm^3
    double G[NSP * NC];        // S/R - H/RT
    double W[NSP * NC];        // production rates; row k is written by wave k mod NW only
    double Q[CH * NC];         // rates of progress (forward minus reverse) of the current chunk; outside
                               // the reaction loop, the per-wave values of cross-wave sums
    double PF[7 * NC];         // pfac**m for m = -3 .. 3 (row m + 3), the equilibrium constants' factor
};

// reaction_plain (chem_device.hpp) for a row the whole wave evaluates (wave-uniform row index), as
// forward minus reverse rate of progress: the same operations, with the slot counts, coefficient
// powers and flags of K->lane_plain taken as scalar branches instead of selects, the species rows
// read at byte offsets from this lane's column of the C and G blocks, and pfac**nus from the lane's
// PF column.  Multiplications by exactly 1.0 that reaction_plain makes for empty slots are skipped
// (they do not change a value).
__device__ __forceinline__ double ipow_u(double x, int m) {      // ipow3 for a wave-uniform m >= 1
    if (m == 1) return x;
    const double x2 = x * x;
    return m == 2 ? x2 : x * x2;
}

__device__ __forceinline__ double lane_plain_q(const LanePlain *__restrict__ L, double T, double alogt,
                                               double small, double exparg, const char *Cb, const char *Gb,
                                               const double *Pf) {
    const int bits = L->bits;
    if (bits & LP_DEAD) return 0.0;
    const int nr = (bits >> LP_NR) & 3, np = (bits >> LP_NP) & 3;
    auto C = [&](int n) { return *reinterpret_cast<const double *>(Cb + L->off[n]); };
    auto G = [&](int n) { return *reinterpret_cast<const double *>(Gb + L->off[n]); };
    double kf = L->A;
    if (!(bits & LP_ARR0)) kf = kf * fm_exp(L->B * alogt - L->E / T);
    double kr = 0.0;
    if (bits & LP_REV) {
        // reaction_plain starts from 0.0 + t0; the two differ at most in the sign of a zero sum,
        // which the test s != 0 and exp(+-0) = 1 do not see
        double s = L->nud[0] * G(0);
        if (nr >= 2) s = s + L->nud[1] * G(1);
        if (nr >= 3) s = s + L->nud[2] * G(2);
        if (np >= 1) s = s + L->nud[3] * G(3);
        if (np >= 2) s = s + L->nud[4] * G(4);
        if (np >= 3) s = s + L->nud[5] * G(5);
        double eqk = s != 0.0 ? fm_exp(fmin(s, exparg)) : 1.0;
        eqk = eqk * Pf[((bits >> LP_NUS) & 7) * NC];
        kr = kf / fmax(eqk, small);
    }
    double qf, qr = kr;
    if (!(bits & LP_POW)) {
        qf = kf * C(0);
        if (nr >= 2) qf = qf * C(1);
        if (nr >= 3) qf = qf * C(2);
        if (np >= 1) qr = kr * C(3);
        if (np >= 2) qr = qr * C(4);
        if (np >= 3) qr = qr * C(5);
    } else {
        const int pw = bits >> LP_PW;
        qf = kf * ipow_u(C(0), pw & 3);
        if (nr >= 2) qf = qf * ipow_u(C(1), (pw >> 2) & 3);
        if (nr >= 3) qf = qf * ipow_u(C(2), (pw >> 4) & 3);
        if (np >= 1) qr = kr * ipow_u(C(3), (pw >> 6) & 3);
        if (np >= 2) qr = qr * ipow_u(C(4), (pw >> 8) & 3);
        if (np >= 3) qr = qr * ipow_u(C(5), (pw >> 10) & 3);
    }
    return qf - qr;
}

// Species slots of wave w: k = w + NW * j.
template <int NSP>
constexpr int slots() { return (NSP + NW - 1) / NW; }

// Sum over the NW waves of one value per cell, in wave order, through a block of NW * NC doubles
// (the start of the rate block Q, free outside the reaction loop).
__device__ __forceinline__ double sum_waves(double *slot, int w, int cl, double v) {
    slot[w * NC + cl] = v;
    __syncthreads();
    double s = slot[cl];
#pragma unroll
    for (int u = 1; u < NW; ++u) s = s + slot[u * NC + cl];
    __syncthreads();
    return s;
}

// The same for two values at once through a block of 2 * NW * NC doubles (the rate block Q, free
// outside the reaction loop): one pair of barriers instead of two.
__device__ __forceinline__ void sum_waves2(double *pair, int w, int cl, double va, double vb, double &ra, double &rb) {
    pair[w * NC + cl] = va;
    pair[(NW + w) * NC + cl] = vb;
    __syncthreads();
    const double *q = pair + NW * NC;
    ra = pair[cl];
    rb = q[cl];
#pragma unroll
    for (int u = 1; u < NW; ++u) {
        ra = ra + pair[u * NC + cl];
        rb = rb + q[u * NC + cl];
    }
    __syncthreads();
}

// f = dy/dt at (y + s d, T) for this wave's species slots (f[j] of species w + NW*j) and fT for
// all waves (SHIFT = false: at y).  rho in g/cm^3.  cl: this lane's LDS column.
template <int NSP, bool SHIFT>
__device__ __forceinline__ void rhs_lane(const DevThermo *__restrict__ th, const DevKin *__restrict__ K, LaneShared<NSP> &sh,
                         int w, int cl, double rho, const double (&y)[slots<NSP>()],
                         const double (&d)[slots<NSP>()], double s, double T, double (&f)[slots<NSP>()],
                         double &fT) {
    constexpr int NS = slots<NSP>();
    const TPow p = tpow(T, th->ru);
    double uk[NS];
    double ctot_w = 0.0, cvm_w = 0.0;
    const double lnT = fm_log(T);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int k = w + NW * j;
        uk[j] = 0.0;
        if (k < NSP) {
            const double yj = SHIFT ? y[j] + s * d[j] : y[j];
            const double ck = rho * yj / th->wt[k];
            sh.C[k * NC + cl] = ck;
            sh.W[k * NC + cl] = 0.0;
            ctot_w = ctot_w + ck;
            const Nasa a = nasa(th, k, T);
            uk[j] = h_mol(a, p) - p.rut;
            cvm_w = cvm_w + yj * (cp_mol(a, p) - th->ru) / th->wt[k];
            const double t2 = T * T, t3 = T * t2, t4 = t2 * t2;     // CKSMH, as smh_species
            double s = (lnT - 1.0) * a.a0;
            s = s + T / 2.0 * a.a1;
            s = s + t2 / 6.0 * a.a2;
            s = s + t3 / 12.0 * a.a3;
            s = s + t4 / 20.0 * a.a4;
            sh.G[k * NC + cl] = s + a.a6 - a.a5 / T;
        }
    }
    const double pfac = th->pa / (th->ru * T);
    {
        // pfac**m as reaction_plain forms it (x**3 = x * (x * x), x**-m = 1 / x**m); row m + 3 of PF,
        // row r written by wave r mod NW
        const double p2 = pfac * pfac, p3 = pfac * p2;
        const double pm[7] = {1.0 / p3, 1.0 / p2, 1.0 / pfac, 1.0, pfac, p2, p3};
#pragma unroll
        for (int r = 0; r < 7; ++r)
            if (r % NW == w) sh.PF[r * NC + cl] = pm[r];
    }
    static_assert(CH >= 2 * NW, "the rate block holds a pair of cross-wave sums");
    double ctot, cvm;
    sum_waves2(sh.Q, w, cl, ctot_w, cvm_w, ctot, cvm);     // its barrier also publishes C, G and PF
    const double *Cp = sh.C + cl, *Gp = sh.G + cl;
    const char *Cb = reinterpret_cast<const char *>(Cp), *Gb = reinterpret_cast<const char *>(Gp);
    const double *Pf = sh.PF + cl;
    double *Wp = sh.W + cl;
    char *Wb = reinterpret_cast<char *>(Wp);
    const double *Qp = sh.Q + cl;
    const char *Qb = reinterpret_cast<const char *>(Qp);
    for (int c0 = 0, chunk = 0; c0 < K->ii; c0 += CH, ++chunk) {
        // this wave's RB rows of the chunk: independent rate chains, so their table loads, LDS
        // reads and transcendental evaluations overlap
        const int i0 = c0 + w;
        double q[RB];
        if (i0 >= K->nplain0) {
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int i = i0 + NW * u;
                q[u] = i < K->ii ? lane_plain_q(&K->lane_plain[i], T, lnT, K->small, K->exparg, Cb, Gb, Pf) : 0.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int i = i0 + NW * u;
                q[u] = 0.0;
                if (i < K->ii) {
                    const Rop r = reaction<NC>(K, i, T, lnT, pfac, ctot, Cp, Gp);
                    q[u] = r.qf - r.qr;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < RB; ++u) sh.Q[(w + NW * u) * NC + cl] = q[u];
        __syncthreads();
        // the chunk's contributions to this wave's species, in row then slot order
        const int e1 = K->lane_ent0[chunk * NW + w + 1];
        for (int e = K->lane_ent0[chunk * NW + w]; e < e1; ++e) {
            const LaneEnt &en = K->lane_ent[e];
            double *wk = reinterpret_cast<double *>(Wb + en.offW);
            *wk = *wk + *reinterpret_cast<const double *>(Qb + en.offQ) * en.nu;
        }
        __syncthreads();                   // Q is rewritten by the next chunk
    }
    double su_w = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int k = w + NW * j;
        f[j] = 0.0;
        if (k < NSP) {
            const double wd = Wp[k * NC];  // a row this wave wrote
            f[j] = wd * th->wt[k] / rho;
            su_w = su_w + uk[j] * wd;
        }
    }
    const double su = sum_waves(sh.Q, w, cl, su_w);     // after the last chunk's closing barrier
    fT = -su / (rho * cvm);
}

// Mixture e = sum Y u / W and cv (erg/g) over all species at T for this cell.
template <int NSP>
__device__ __forceinline__ void energy_lane(const DevThermo *__restrict__ th, double *pair, int w, int lane,
                            const double (&y)[slots<NSP>()], double T, double &e, double &cv) {
    constexpr int NS = slots<NSP>();
    const TPow p = tpow(T, th->ru);
    double eu = 0.0, ec = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int k = w + NW * j;
        if (k < NSP) {
            const Nasa a = nasa(th, k, T);
            eu = eu + y[j] * (h_mol(a, p) - p.rut) / th->wt[k];
            ec = ec + y[j] * (cp_mol(a, p) - th->ru) / th->wt[k];
        }
    }
    sum_waves2(pair, w, lane, eu, ec, e, cv);
}

// adjust_sum (SRC/thermo.f90:136-149) of this wave's slots: clip at 0, divide by the cell's sum.
template <int NSP>
__device__ __forceinline__ void adjust_lane(double *slot, int w, int lane, double (&y)[slots<NSP>()]) {
    constexpr int NS = slots<NSP>();
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        y[j] = (w + NW * j < NSP) ? fmax(y[j], 0.0) : 0.0;
        s = s + y[j];
    }
    const double tot = sum_waves(slot, w, lane, s);
#pragma unroll
    for (int j = 0; j < NS; ++j) y[j] = y[j] / tot;
}

template <int NSP>
__global__ void __launch_bounds__(64 * NW) k_chem_screen(ChemArgs a) {
    constexpr int NS = slots<NSP>();
    __shared__ LaneShared<NSP> sh;
    // the wave index is uniform across the wave; saying so turns every table read indexed by it
    // (reactions i = w + NW n, species k = w + NW j) into a scalar load and every branch on the
    // reaction kind into a scalar branch
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = min(int(threadIdx.x & 63), NC - 1);   // the LDS column (lanes >= NC repeat NC - 1)
    const DevThermo *__restrict__ th = a.thermo;
    const DevKin *__restrict__ K = a.kin;
    const int cell0 = int(xcd_block()) * NC;   // neighbouring cell ranges share their XCD's L2
    const int cell = cell0 + lane;
    // every lane of a column takes the column's decisions (lanes >= NC share column NC - 1 and must
    // write the same LDS values as its owner); only the owner lane stores to global memory
    const bool live = cell < a.ncells;
    const bool owner = live && int(threadIdx.x & 63) < NC;
    const int c = live ? cell : cell0;          // columns past the end shadow a live cell
    double *slot = sh.Q;

    long col = 0;
    double rho_si, dt, T0;
    double y[NS];
    if (a.layout == 0) {
        const int b = c / a.ncell, i = c % a.ncell;
        col = long(b) * a.S + a.lv + i;
        rho_si = a.qf[col];
        T0 = a.qf[2 * a.P + col];
        dt = a.dt_dev ? *a.dt_dev : a.dt;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            y[j] = k < NSP ? a.qf[(3 + k) * a.P + col] : 0.0;
        }
        adjust_lane<NSP>(slot, w, lane, y);
    } else {
        rho_si = a.b_rho[c];
        dt = a.b_dt[c];
        T0 = a.b_T[c];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            y[j] = k < NSP ? a.b_y[long(k) * a.ncells + c] : 0.0;
        }
    }
    const double rho = rho_si * 1.e-3;
    double e0, cv0;
    {
        double ya[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) ya[j] = y[j];
        adjust_lane<NSP>(slot, w, lane, ya);
        energy_lane<NSP>(th, sh.Q, w, lane, ya, T0, e0, cv0);
    }

    // ---- Heun over dt with the embedded Euler error estimate ----
    double f0[NS], f0T;
    rhs_lane<NSP, false>(th, K, sh, w, lane, rho, y, y, 0.0, T0, f0, f0T);
    const double T1 = T0 + dt * f0T;
    double f1[NS], f1T;
    rhs_lane<NSP, true>(th, K, sh, w, lane, rho, y, f0, dt, T1, f1, f1T);   // at y + dt f0
    double err_w = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        if (w + NW * j < NSP) {
            const double wt = 1.0 / (a.atol + a.rtol * fabs(y[j]));
            const double t = ((dt * 0.5) * (f1[j] - f0[j])) * wt;
            err_w = err_w + t * t;
        }
    }
    if (w == 0) {
        const double t = ((dt * 0.5) * (f1T - f0T)) * (1.0 / (a.rtol * fabs(T0)));
        err_w = err_w + t * t;
    }
    const double err = sqrt(sum_waves(slot, w, lane, err_w) / double(NSP + 1));
    const bool ok = live && err <= 1.0;

    if (!ok && owner && a.f0) {   // f at the start state, for the retry of a long list
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            if (k < NSP) a.f0[long(k) * a.ncells + cell] = f0[j];
        }
        if (w == 0) a.f0[long(NSP) * a.ncells + cell] = f0T;
    }
    if (!ok) {
        if (owner && w == 0) {
            const unsigned long long bal = __ballot(1);
            const int leader = __ffsll((long long)bal) - 1;
            const int rank = __popcll(bal & ((1ull << lane) - 1ull));
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(a.stiff_n, unsigned(__popcll(bal)));
            base = __shfl(base, leader, 64);
            a.stiff[base + rank] = cell;
            if (a.route) a.route[cell] = 0;
        }
    }
    // ---- accepted cells: state update and energy-consistent temperature ----
#pragma unroll
    for (int j = 0; j < NS; ++j) y[j] = y[j] + (dt * 0.5) * (f0[j] + f1[j]);
    double T = T0 + (dt * 0.5) * (f0T + f1T);
    adjust_lane<NSP>(slot, w, lane, y);
    bool done = !ok;
    for (int it = 0; it < 20; ++it) {
        if (!__syncthreads_or(done ? 0 : 1)) break;
        double ek, cvk;
        energy_lane<NSP>(th, sh.Q, w, lane, y, T, ek, cvk);
        if (!done) {
            const double d = (e0 - ek) / cvk;
            T = T + d;
            if (fabs(d) <= 1.0e-12 * T) done = true;
        }
    }
    double hrr_s = 0.0, hrr_d = 0.0;
    if (a.hrr) {
        // heat_release_rate (SRC/thermo.f90:98-101) and heat_release_rate_omegadot
        // (SRC/thermo.f90:128-132): the creation minus destruction rates of CKCDYR are the net
        // production rates, evaluated at the new state.
        double fn[NS], fnT;
        rhs_lane<NSP, false>(th, K, sh, w, lane, rho, y, y, 0.0, T, fn, fnT);
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            if (k < NSP) {
                const double y_in = a.layout == 0 ? a.qf[(3 + k) * a.P + col] : a.b_y[long(k) * a.ncells + c];
                s1 = s1 + rho * th->h298[k] * (y[j] - y_in) / dt;
                s2 = s2 + th->h298[k] * fn[j] * rho;     // W_k h_k wdot_k = h_k (W_k wdot_k / rho) rho
            }
        }
        sum_waves2(sh.Q, w, lane, s1, s2, hrr_s, hrr_d);
    }
    if (!ok || !owner) return;
    if (a.layout == 0) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            if (k < NSP) a.qf[(3 + k) * a.P + col] = y[j];
        }
        if (w == 0) {
            a.qf[2 * a.P + col] = T;
            if (a.hrr) {
                a.hrr[cell] = -hrr_s * 1.e-1;
                a.hrr_dot[cell] = -hrr_d * 1.e-1;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            if (k < NSP) a.b_y[long(k) * a.ncells + cell] = y[j];
        }
        if (w == 0) a.b_T[cell] = T;
    }
    if (w == 0) {
        if (a.stats) a.stats[cell] = ChemCellStats{1, 0, 0, 0};
        if (a.route) a.route[cell] = 2;
        if (a.hmem) a.hmem[cell] = 0.0;
    }
}

// Explicit retry of a LONG stiff list, ONE CELL PER LANE like k_chem_screen: two Heun steps of dt/2,
// each accepted by its own embedded Euler estimate with weights at its start -- the retry k_chem
// makes before its implicit steps, with the same formulas.  In: cells a.list[0 .. *a.list_n), and
// (when a.f0 is given) f at their start states, which the screen computed at the same states.
// Accepted cells are final (state, heat release, stats {2, 0, 0, 0}, step-size memory 0); the
// others go to a.stiff (*a.stiff_n zero before the launch) for the implicit integrator, which
// *a.retry_done = 1 tells to skip its own retry.  A list of at most a.coop_max cells is copied to
// a.stiff unchanged with *a.retry_done = 0: the integrator then runs it four waves per cell and
// retries each cell itself.  Domain layout only.
template <int NSP>
__global__ void __launch_bounds__(64 * NW) k_chem_retry(ChemArgs a) {
    constexpr int NS = slots<NSP>();
    __shared__ LaneShared<NSP> sh;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = min(int(threadIdx.x & 63), NC - 1);   // the LDS column (lanes >= NC repeat NC - 1)
    const DevThermo *__restrict__ th = a.thermo;
    const DevKin *__restrict__ K = a.kin;
    const int n = int(*a.list_n);
    if (n <= a.coop_max) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *a.retry_done = 0;
            *a.stiff_n = unsigned(n);
        }
        for (int i = int(blockIdx.x) * (64 * NW) + int(threadIdx.x); i < n; i += int(gridDim.x) * (64 * NW)) a.stiff[i] = a.list[i];
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.retry_done = 1;
    const int item0 = int(blockIdx.x) * NC;
    if (item0 >= n) return;                     // the whole workgroup: no barrier is skipped
    const int item = item0 + lane;
    // every lane of a column takes the column's decisions (lanes >= NC share column NC - 1); only the
    // owner lane stores to global memory
    const bool live = item < n;
    const bool owner = live && int(threadIdx.x & 63) < NC;
    const int cell = a.list[live ? item : item0];   // columns past the end shadow a live cell
    double *slot = sh.Q;

    const int b = cell / a.ncell, i = cell % a.ncell;
    const long col = long(b) * a.S + a.lv + i;
    const double rho_si = a.qf[col];
    const double T0 = a.qf[2 * a.P + col];
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    double y[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int k = w + NW * j;
        y[j] = k < NSP ? a.qf[(3 + k) * a.P + col] : 0.0;
    }
    adjust_lane<NSP>(slot, w, lane, y);
    const double rho = rho_si * 1.e-3;
    double e0, cv0;
    {
        double ya[NS];
#pragma unroll
        for (int j = 0; j < NS; ++j) ya[j] = y[j];
        adjust_lane<NSP>(slot, w, lane, ya);
        energy_lane<NSP>(th, sh.Q, w, lane, ya, T0, e0, cv0);
    }

    // ---- two Heun steps of dt/2; a lane keeps going only while its half steps are accepted ----
    const double hs = dt * 0.5;
    double Th = T0;
    bool ok = live;
    for (int sub = 0; sub < 2; ++sub) {
        double fa[NS], faT, fb[NS], fbT;
        if (sub == 0 && a.f0) {   // f at the start state, stored by the screen
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int k = w + NW * j;
                fa[j] = k < NSP ? a.f0[long(k) * a.ncells + cell] : 0.0;
            }
            faT = a.f0[long(NSP) * a.ncells + cell];
        } else {
            rhs_lane<NSP, false>(th, K, sh, w, lane, rho, y, y, 0.0, Th, fa, faT);
        }
        rhs_lane<NSP, true>(th, K, sh, w, lane, rho, y, fa, hs, Th + hs * faT, fb, fbT);   // at y + hs fa
        double err_w = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (w + NW * j < NSP) {
                const double wt = 1.0 / (a.atol + a.rtol * fabs(y[j]));
                const double t = ((hs * 0.5) * (fb[j] - fa[j])) * wt;
                err_w = err_w + t * t;
            }
        }
        if (w == 0) {
            const double t = ((hs * 0.5) * (fbT - faT)) * (1.0 / (a.rtol * fabs(Th)));
            err_w = err_w + t * t;
        }
        const double err = sqrt(sum_waves(slot, w, lane, err_w) / double(NSP + 1));
        ok = ok && err <= 1.0;
        if (ok) {
#pragma unroll
            for (int j = 0; j < NS; ++j) y[j] = y[j] + (hs * 0.5) * (fa[j] + fb[j]);
            Th = Th + (hs * 0.5) * (faT + fbT);
        }
        // no lane of the workgroup passed the first half step: the second would be wasted work.
        // ok is the same in all waves of a lane (err comes from sum_waves), and every wave reaches
        // this barrier, so the exit is uniform
        if (sub == 0 && !__syncthreads_or(ok ? 1 : 0)) break;
    }
    if (!ok) {
        if (owner && w == 0) {
            const unsigned long long bal = __ballot(1);
            const int leader = __ffsll((long long)bal) - 1;
            const int rank = __popcll(bal & ((1ull << lane) - 1ull));
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(a.stiff_n, unsigned(__popcll(bal)));
            base = __shfl(base, leader, 64);
            a.stiff[base + rank] = cell;
        }
    }
    // ---- accepted cells: energy-consistent temperature ----
    adjust_lane<NSP>(slot, w, lane, y);
    double T = Th;
    bool done = !ok;
    for (int it = 0; it < 20; ++it) {
        if (!__syncthreads_or(done ? 0 : 1)) break;
        double ek, cvk;
        energy_lane<NSP>(th, sh.Q, w, lane, y, T, ek, cvk);
        if (!done) {
            const double d = (e0 - ek) / cvk;
            T = T + d;
            if (fabs(d) <= 1.0e-12 * T) done = true;
        }
    }
    double hrr_s = 0.0, hrr_d = 0.0;
    if (a.hrr) {
        // heat_release_rate and heat_release_rate_omegadot (SRC/thermo.f90:98-101, 128-132), as
        // in k_chem_screen
        double fn[NS], fnT;
        rhs_lane<NSP, false>(th, K, sh, w, lane, rho, y, y, 0.0, T, fn, fnT);
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int k = w + NW * j;
            if (k < NSP) {
                const double y_in = a.qf[(3 + k) * a.P + col];
                s1 = s1 + rho * th->h298[k] * (y[j] - y_in) / dt;
                s2 = s2 + th->h298[k] * fn[j] * rho;
            }
        }
        sum_waves2(sh.Q, w, lane, s1, s2, hrr_s, hrr_d);
    }
    if (!ok || !owner) return;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const int k = w + NW * j;
        if (k < NSP) a.qf[(3 + k) * a.P + col] = y[j];
    }
    if (w == 0) {
        a.qf[2 * a.P + col] = T;
        if (a.hrr) {
            a.hrr[cell] = -hrr_s * 1.e-1;
            a.hrr_dot[cell] = -hrr_d * 1.e-1;
        }
        if (a.stats) a.stats[cell] = ChemCellStats{2, 0, 0, 0};
        if (a.route) a.route[cell] = 3;
        if (a.hmem) a.hmem[cell] = 0.0;
    }
}

}  // namespace

hipError_t launch_chem_retry(const ChemArgs &a, int nsp, hipStream_t st) {
    const unsigned nb = unsigned((a.ncells + NC - 1) / NC);
    if (nb == 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_chem_retry<N>), dim3(nb), dim3(64 * NW), 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_chem_screen(const ChemArgs &a, int nsp, hipStream_t st) {
    const unsigned nb = unsigned((a.ncells + NC - 1) / NC);
    if (nb == 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_chem_screen<N>), dim3(nb), dim3(64 * NW), 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
