// aux.hip -- open boundaries and the auxiliary per-cell sources of solve_source.
//
//   k_open_bc: the inflow / outflow ghost cells of update_boundary (SRC/boundary.f90:70-149), one
//     thread per ghost cell of the domain's two ends.  It runs after the kernel that refilled the
//     ghost columns as walls (the stage update or the refresh), reads the interior cell the
//     reference reads (already updated in that launch's output) and overwrites the ghost's
//     conserved and primitive columns.
//   k_aux: radiation (SRC/source.f90:85-132), micro-flow-reactor wall heat (:135-182) and the
//     ignition deposit (:185-236), one thread per interior cell, in the reference's order.  Each
//     enabled source adds dt * source to rho E and recovers the primitives with QcToqf before the
//     next one reads them, as the reference's per-source QcToqf loops do.
//
// Arithmetic follows the reference's expressions term by term (no contraction), with the shared
// exp / log / tanh of fcfd_math.h, so the CPU restatement reproduces it bit for bit.
#include "aux.hpp"
#include "fcfd_device.hpp"

namespace fcfd {

namespace {

constexpr double PI_IG = 3.14159265359;              // SRC/ignitionsource.f90:27
constexpr double PI_IG_15 = 0x1.645f7c63f2ed6p+2;    // PI_IG**1.5, correctly rounded

__device__ __forceinline__ double mceval(const double *c, int no, double x) {   // MCEVAL tranlib.f:1335
    if (no == 4) return ((c[3] * x + c[2]) * x + c[1]) * x + c[0];              // MCEVAL4 :1388
    double v = c[no - 1];
    for (int i = 1; i < no; ++i) v = v * x + c[no - 1 - i];
    return v;
}

// MCACON (tranlib.f:876-891) in species order, cgs.
template <int NSP>
__device__ __forceinline__ double mcacon(const DevTran *__restrict__ tr, double T, const double (&x)[NSP]) {
    const double lt = fm_log(T);
    double s = 0.0, sr = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const double c = fm_exp(mceval(tr->coflam[k], tr->no, lt));
        s = s + x[k] * c;
        sr = sr + x[k] / c;
    }
    return 0.5 * (s + 1.0 / sr);
}

// QcToqf's mass fractions (SRC/eos.f90:49-83): adjust_sum(rho Y / rho), rho Y_k at rY[k * P].
template <int NSP>
__device__ __forceinline__ void qc_y(const double *__restrict__ rY, long P, double rho, double (&y)[NSP]) {
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = rY[k * P] / rho;
    adjust_sum<NSP>(y);
}

template <int NSP>
__device__ __forceinline__ void qc_y(const double (&Qv)[NSP + 3], double (&y)[NSP]) {
#pragma unroll
    for (int k = 0; k < NSP; ++k) y[k] = Qv[3 + k] / Qv[0];
    adjust_sum<NSP>(y);
}

template <int NSP>
__global__ void __launch_bounds__(64) k_open_bc(OpenBCArgs a) {
    const int t = threadIdx.x;
    if (t >= 2 * a.lv) return;
    const bool left = t < a.lv;
    if (left ? a.lbound != 1 : a.rbound != 1) return;
    const int g = t % a.lv + 1;
    const long gc = left ? long(a.lv - g) : long(a.nblk - 1) * a.S + a.lv + a.ncell - 1 + g;
    const long ic = left ? long(a.lv + g - 1) : long(a.nblk - 1) * a.S + a.lv + a.ncell - g;
    const DevThermo *__restrict__ m = a.thermo;
    const FlowBC *__restrict__ bc = a.bc;
    const long P = a.P;
    double Qv[NSP + 3];
    if (left) {                                        // lbound_inlet, SRC/boundary.f90:70-99
        const double rho_in = a.Q[ic];
        double x[NSP];
#pragma unroll
        for (int k = 0; k < NSP; ++k) x[k] = bc->x_in[k];
        const double u_in = bc->rho_u_in / rho_in;
        const MixSums ms = mix_sums<NSP>(m, x, bc->T_in, true, false, false);
        const double e_in = ms.ubml * 1.e-4 / mix_w<NSP>(m, x);   // specific_energy (thermo.f90:369-374)
        Qv[0] = rho_in;
        Qv[1] = bc->rho_u_in;
        Qv[2] = rho_in * (e_in + 0.5 * (u_in * u_in));
#pragma unroll
        for (int k = 0; k < NSP; ++k) Qv[3 + k] = rho_in * bc->y_in[k];
    } else {                                           // rbound_outlet, SRC/boundary.f90:102-149
        const double rho = a.Q[ic];
        const double u = a.Q[P + ic] / rho;
        const double e = a.Q[2 * P + ic] / rho - 0.5 * (u * u);
        double x[NSP];
#pragma unroll
        for (int k = 0; k < NSP; ++k) x[k] = a.Q[(3 + k) * P + ic] / rho;
#pragma unroll
        for (int k = 0; k < NSP; ++k) Qv[3 + k] = x[k];            // y_out, kept until rho_out is known
        ytx<NSP>(m, x);                                             // Yi2Xi, no adjust_sum
        const double T = temperature_con<NSP>(m, e, x);
        const double wx = mix_w<NSP>(m, x);
        const double rho_out = wx * (bc->p_out * 10.0) / (m->ru * T) * 1.e3;   // density (CKRHOX)
        const MixSums ms = mix_sums<NSP>(m, x, T, true, false, false);
        const double e2 = ms.ubml * 1.e-4 / wx;
        Qv[0] = rho_out;
        Qv[1] = rho_out * u;
        Qv[2] = rho_out * (e2 + 0.5 * (u * u));
#pragma unroll
        for (int k = 0; k < NSP; ++k) Qv[3 + k] = rho_out * Qv[3 + k];
    }
#pragma unroll
    for (int v = 0; v < NSP + 3; ++v) a.Q[v * P + gc] = Qv[v];
    // QcToqf on the ghost (update_boundary's own loop, SRC/boundary.f90:38-45, 55-62)
    double y[NSP];
    qc_y<NSP>(Qv, y);
#pragma unroll
    for (int k = 0; k < NSP; ++k) a.qf[(3 + k) * P + gc] = y[k];
    ytx<NSP>(m, y);
    adjust_sum<NSP>(y);
    const double u = Qv[1] / Qv[0];
    const double e = Qv[2] / Qv[0] - 0.5 * (u * u);
    a.qf[gc] = Qv[0];
    a.qf[P + gc] = u;
    a.qf[2 * P + gc] = temperature_con<NSP>(m, e, y);
}

template <int NSP>
__global__ void __launch_bounds__(64) k_aux(AuxArgs a) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= long(a.nblk) * a.ncell) return;
    const long c = long(t / a.ncell) * a.S + a.lv + (t % a.ncell);
    const DevThermo *__restrict__ m = a.thermo;
    const AuxParams *__restrict__ p = a.par;
    const long P = a.P;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double time = a.time_dev ? *a.time_dev : a.time;
    const double xm = a.xm[t];
    // the sources change rho E only: rho, rho u and rho Y_k stay in memory
    const double Qr = a.Q[c], Qm = a.Q[P + c];
    double QE = a.Q[2 * P + c];
    const double *rY = a.Q + 3 * P + c;
    const double rho = a.qf[c];
    double T = a.qf[2 * P + c];
    double x[NSP];
    bool fresh = false;            // true once the state has been through QcToqf (Y is then recomputed)
    auto mole_fractions = [&]() {   // Yi2Xi of the current primitive Y, no adjust_sum
        if (fresh) {
            qc_y<NSP>(rY, P, Qr, x);
        } else {
#pragma unroll
            for (int k = 0; k < NSP; ++k) x[k] = a.qf[(3 + k) * P + c];
        }
        ytx<NSP>(m, x);
    };
    // one pass per enabled source, in the reference's order; the QcToqf that ends each pass is
    // the same code for all three
    for (int src = 0; src < 3; ++src) {
        if (src == 0 ? !a.rad : src == 1 ? !a.wall : !a.ig) continue;
        double add;
        if (src == 0) {                                 // heat_transfer_radiation
            mole_fractions();
            const double pres = (rho * 1.e-3) * m->ru * T / mix_w<NSP>(m, x) * 0.1;   // pressure (CKPX)
            const double T1 = T, T2 = T1 * T1, T3 = T2 * T1, T4 = T3 * T1, T5 = T4 * T1;
            const double T0 = 300.0, T04 = T0 * T0 * T0 * T0;
            double Kp = 0.0;
            for (int r = 0; r < p->nrad; ++r) {         // SRC/radiation.f90:110-116
                const double *cf = p->pmac[r];
                const double kpi = cf[0] + cf[1] * T1 + cf[2] * T2 + cf[3] * T3 + cf[4] * T4 + cf[5] * T5;
                double xr = 0.0;
#pragma unroll
                for (int k = 0; k < NSP; ++k)
                    if (k == p->rad_idx[r]) xr = x[k];
                Kp = Kp + kpi * xr * (pres / 101325.0);
            }
            add = -(4.0 * Kp * 5.6703744191844294e-08 * (T4 - T04));   // :119
            a.qrad[t] = add;
        } else if (src == 1) {                          // heat_transfer_mfr
            mole_fractions();
            const double lambda = mcacon<NSP>(a.tran, T, x) * 1.e-5;     // SRC/trans.f90:74-76
            const double A = 0.5 * (p->Tw_max - p->Tw_min);               // Tw, SRC/mfr.f90:63-78
            const double B = 0.5 * (p->Tw_max + p->Tw_min);
            const double Tw = A * fm_tanh(p->alpha * (xm - p->beta)) + B;
            add = -(4.0 * lambda * p->Nu * (T - Tw) / (p->d_m * p->d_m));   // :57
            a.qw[t] = add;
        } else {                                        // energy_source
            add = 0.0;
            for (int s = 0; s < p->n_ig; ++s) {         // a later active source overwrites
                if (!(p->t0_s[s] < time && time < p->t0_s[s] + p->t_s[s])) continue;
                const double q = xm / p->r_m[s];
                if (p->ig_type == 1) {                  // SRC/ignitionsource.f90:165-167
                    const double coef = p->E_j[s] / (p->vol[s] * p->t_s[s]);
                    const double q3 = q * q * q;
                    add = coef * fm_exp(-(PI_IG / 4.0) * (q3 * q3));
                } else {                                // :218-219
                    const double r = p->r_m[s];
                    const double coef = p->E_j[s] / (PI_IG_15 * (r * r * r) * p->t_s[s]);
                    add = coef * fm_exp(-(q * q));
                }
            }
        }
        QE = QE + dt * add;
        // QcToqf's temperature: Newton on e = E/rho - u^2/2 with X = adjust_sum(CKYTX(Y))
        qc_y<NSP>(rY, P, Qr, x);
        ytx<NSP>(m, x);
        adjust_sum<NSP>(x);
        const double u = Qm / Qr;
        T = temperature_con<NSP>(m, QE / Qr - 0.5 * (u * u), x);
        fresh = true;
    }
    a.Q[2 * P + c] = QE;
    qc_y<NSP>(rY, P, Qr, x);
#pragma unroll
    for (int k = 0; k < NSP; ++k) a.qf[(3 + k) * P + c] = x[k];
    a.qf[c] = Qr;
    a.qf[P + c] = Qm / Qr;
    a.qf[2 * P + c] = T;
}

// set_eos_params (SRC/block.f90:433-470) and the dataset rows of write_hdf5_b
// (SRC/write_hdf5.f90:138-160), one thread per interior cell.
template <int NSP>
__global__ void __launch_bounds__(64) k_eos_fields(EosArgs a) {
    const long n = long(a.nblk) * a.ncell;
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int b = int(t / a.ncell);
    const long c = long(b) * a.S + a.lv + (t % a.ncell);
    const DevThermo *__restrict__ m = a.thermo;
    const long P = a.P;
    double x[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) x[k] = a.qf[(3 + k) * P + c];
    ytx<NSP>(m, x);                                     // Yi2Xi, no adjust_sum
    const double rho = a.qf[c], u = a.qf[P + c], T = a.qf[2 * P + c];
    const double wx = mix_w<NSP>(m, x);
    const MixSums ms = mix_sums<NSP>(m, x, T, true, false, false);
    double *o = a.out;
    o[0 * n + t] = rho;
    o[1 * n + t] = u;
    o[2 * n + t] = T;
    o[3 * n + t] = (rho * 1.e-3) * m->ru * T / wx * 0.1;                // pressure (CKPX)
    o[4 * n + t] = a.Q[2 * P + c] / rho;
    o[5 * n + t] = 0.5 * (u * u);
    o[6 * n + t] = ms.ubml * 1.e-4 / wx;                                 // specific_energy
    o[7 * n + t] = a.hrr ? a.hrr[t] : 0.0;
    o[8 * n + t] = a.hrd ? a.hrd[t] : 0.0;
    o[9 * n + t] = a.qrad ? a.qrad[t] : 0.0;
    double Tw = 0.0;
    if (a.wall_heat) {                                  // SRC/params.f90:495-498 -> SRC/mfr.f90:63-78
        const AuxParams *__restrict__ p = a.par;
        const double A = 0.5 * (p->Tw_max - p->Tw_min), B = 0.5 * (p->Tw_max + p->Tw_min);
        Tw = A * fm_tanh(p->alpha * (a.xm[t] - p->beta)) + B;
    }
    o[10 * n + t] = Tw;
    o[11 * n + t] = a.qw ? a.qw[t] : 0.0;
    o[12 * n + t] = a.xm[t];
    o[13 * n + t] = a.dxb ? a.dxb[b] : a.dx;
    o[14 * n + t] = double(a.idx0 + t);
#pragma unroll
    for (int k = 0; k < NSP; ++k) o[(15 + k) * n + t] = x[k];
}

// Restart state (read_hdf5_restart's X -> Y, SRC/read_hdf5.f90:415-417, then set_block_params,
// SRC/block.f90:413-429) from rows rho, u, T, e_total, X_1..X_KK.
template <int NSP>
__global__ void __launch_bounds__(64) k_restart(EosArgs a) {
    const long n = long(a.nblk) * a.ncell;
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long c = long(t / a.ncell) * a.S + a.lv + (t % a.ncell);
    const DevThermo *__restrict__ m = a.thermo;
    const long P = a.P;
    const double *f = a.out;
    double s = 0.0;                                     // CKXTY (cklib.f:9461-9468)
#pragma unroll
    for (int k = 0; k < NSP; ++k) s = s + f[(4 + k) * n + t] * m->wt[k];
    const double rho = f[t], u = f[n + t], T = f[2 * n + t], et = f[3 * n + t];
    a.qf[c] = rho; a.qf[P + c] = u; a.qf[2 * P + c] = T;
    a.Q[c] = rho; a.Q[P + c] = rho * u; a.Q[2 * P + c] = rho * et;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const double y = f[(4 + k) * n + t] * m->wt[k] / s;
        a.qf[(3 + k) * P + c] = y;
        a.Q[(3 + k) * P + c] = rho * y;
    }
}

template <int NSP>
__global__ void k_flow_bc_setup(const DevThermo *__restrict__ m, FlowBC *bc, double u_in) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;                                     // CKXTY (cklib.f:9461-9468)
    for (int k = 0; k < NSP; ++k) s = s + bc->x_in[k] * m->wt[k];
    for (int k = 0; k < NSP; ++k) bc->y_in[k] = bc->x_in[k] * m->wt[k] / s;
    // density(p_OUT, T_IN, x_sp_IN) * u_in (SRC/flow_bound.f90:54)
    const double rho = s * (bc->p_out * 10.0) / (m->ru * bc->T_in) * 1.e3;
    bc->rho_u_in = rho * u_in;
}

}  // namespace

hipError_t launch_open_bc(const OpenBCArgs &a, int nsp, hipStream_t st) {
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_open_bc<N>), dim3(1), dim3(64), 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_aux(const AuxArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.ncell;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_aux<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_eos_fields(const EosArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.ncell;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_eos_fields<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_restart(const EosArgs &a, int nsp, hipStream_t st) {
    const long n = long(a.nblk) * a.ncell;
    const dim3 grid(unsigned((n + 63) / 64)), blk(64);
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_restart<N>), grid, blk, 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_flow_bc_setup(const DevThermo *m, FlowBC *bc, double u_in, int nsp, hipStream_t st) {
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_flow_bc_setup<N>), dim3(1), dim3(64), 0, st, m, bc, u_in); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
