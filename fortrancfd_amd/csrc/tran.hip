// tran.hip -- calc_flux_tran (SRC/flux.f90:18-127) with the TRANLIB mixture rules it calls.
//
// Two kernels per flux evaluation:
//   k_tran_props: ONE CELL PER WAVE (lanes over species) for the cells on both sides of every
//     interface (interior cells plus one ghost layer per block end): T from the conserved state
//     (temperature_con on the unnormalised Y, SRC/flux.f90:33-55), p, the Wilke mixture viscosity
//     (MCAVIS, tranlib.f:734-762), the mixture conductivity (MCACON, tranlib.f:876-891) and the
//     mixture-averaged diffusion coefficients (MCADIF -> MCEDIF, tranlib.f:1022-1024, 1095-1148),
//     then dif_k = (1 - X_k) / (1 - Y_k) D_km (SRC/flux.f90:62-72).  The O(K^2) sums run one row
//     per lane with the other species' values broadcast from LDS; every O(K) sum runs in species
//     order, as TRANLIB's loops do.
//   k_flux_tran: one thread per interface: central differences of u, T and Y between the two
//     cells and arithmetic means of the coefficients (SRC/flux.f90:74-123); F = 0 + transport
//     flux, and the advection kernel then adds its flux on top, as Fplus accumulates in the
//     reference (SRC/timeintegration.f90:138-146).
//
// SRC/flux.f90:56-57 takes mu from viscosity_ea -> MCAVIS_EA, a routine the reference does not
// ship; the Wilke mixture rule of MCAVIS (the routine SRC/trans.f90:46-59 wraps) stands in for it.
#include "fcfd_device.hpp"
#include "tran.hpp"

namespace fcfd {

namespace {

__device__ __forceinline__ double mceval(const double *c, int no, double x) {   // MCEVAL tranlib.f:1335
    double v = c[no - 1];
    for (int i = 1; i < no; ++i) v = v * x + c[no - 1 - i];
    return v;
}

template <int NSP>
__global__ void __launch_bounds__(64) k_tran_props(TranArgs a) {
    __shared__ double sy[64], sx[64], svis[64], sxx[64], s_so[64], s_lam[64];
    const int lane = threadIdx.x;
    const int nc = a.ncell + 2;
    const int b = blockIdx.x / nc, c = blockIdx.x % nc;
    const long col = long(b) * a.S + (a.lv - 1) + c;
    const DevThermo *__restrict__ th = a.thermo;
    const DevTran *__restrict__ tr = a.tran;
    const double rho = a.Qc[col];
    const double u = a.Qc[a.P + col] / rho;
    const double et = a.Qc[2 * a.P + col] / rho;
    double yk = 0.0;
    if (lane < NSP) yk = a.Qc[(3 + lane) * a.P + col] / rho;
    sy[lane] = yk;
    __syncthreads();
    // Yi2Xi (CKYTX) and temperature_con, redundantly in every lane (identical results)
    double x[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) x[k] = sy[k];
    ytx<NSP>(th, x);
    const double ek = 0.5 * (u * u);
    const double T = temperature_con<NSP>(th, et - ek, x);
    const double p = (rho * 1.e-3) * th->ru * T / mix_w<NSP>(th, x) * 0.1;   // pressure (CKPX, SI)
    double xk = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k)
        if (lane == k) xk = x[k];
    const double lt = fm_log(T);
    double visk = 1.0, lamk = 1.0;
    if (lane < NSP) {
        visk = fm_exp(mceval(tr->cofeta[lane], tr->no, lt));
        lamk = fm_exp(mceval(tr->coflam[lane], tr->no, lt));
    }
    sx[lane] = xk;
    svis[lane] = visk;
    sxx[lane] = fmax(xk, 1.0e-50);
    s_lam[lane] = lamk;
    __syncthreads();
    double so_k = 0.0, sxw = 0.0, sxod = 0.0;
    if (lane < NSP) {
        // MCAVIS: row `lane` of the Wilke sum
        double si = 0.0;
        for (int j = 0; j < NSP; ++j) {
            const double aa = 1.0 + sqrt(visk / svis[j] * sqrt(tr->twt[j] / tr->twt[lane]));
            si = si + sx[j] * (aa * aa) / sqrt(1.0 + tr->twt[lane] / tr->twt[j]);
        }
        so_k = xk * visk / si;
        // MCEDIF: row `lane` of the binary-diffusion sums (D_jk symmetric, D_kk term zero)
        sxw = -sxx[lane] * tr->twt[lane];
        for (int j = 0; j < NSP; ++j) {
            double djk = 0.0;
            if (j != lane) {
                const int hi = j > lane ? j : lane, lo = j > lane ? lane : j;
                djk = fm_exp(-mceval(tr->cofd[hi][lo], tr->no, lt));
            }
            sxw = sxw + sxx[j] * tr->twt[j];
            sxod = sxod + sxx[j] * djk;
        }
    }
    s_so[lane] = so_k;
    __syncthreads();
    double so = 0.0, sl = 0.0, slr = 0.0, wtm = 0.0;
    for (int k = 0; k < NSP; ++k) {
        so = so + s_so[k];
        sl = sl + sx[k] * s_lam[k];
        slr = slr + sx[k] / s_lam[k];
        wtm = wtm + tr->twt[k] * sx[k];
    }
    const double mu = so * sqrt(8.0) * 0.1;                   // SRC/trans.f90:57-59 (cgs -> SI)
    const double kap = 0.5 * (sl + 1.0 / slr) * 1.e-5;        // SRC/trans.f90:74-76
    const long PP = a.PP;
    const long pc = long(b) * nc + c;
    if (lane == 0) {
        a.props[0 * PP + pc] = rho;
        a.props[1 * PP + pc] = u;
        a.props[2 * PP + pc] = T;
        a.props[3 * PP + pc] = mu;
        a.props[4 * PP + pc] = kap;
    }
    if (lane < NSP) {
        // MCADIF: D_km = sxw / (wtm sxod) * patmos / p (p in dyn/cm^2), cm^2/s -> m^2/s
        double D = sxw / (wtm * sxod);
        D = D * tr->patmos / (p * 10.0);
        D = D * 1.e-4;
        double dif = (1.0 - xk) / (1.0 - yk) * D;
        if (yk == 1.0) dif = 0.0;
        a.props[(5 + lane) * PP + pc] = yk;
        a.props[(5 + NSP + lane) * PP + pc] = dif;
    }
}

template <int NSP>
__global__ void __launch_bounds__(64) k_flux_tran(TranArgs a) {
    const int n1 = a.ncell + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nblk * n1) return;
    const int b = t / n1, j = t % n1;
    const int nc = a.ncell + 2;
    const long PP = a.PP;
    const long c0 = long(b) * nc + j, c1 = c0 + 1;    // cells ista-1+j and ista+j
    const double *P = a.props;
    const double rho0 = P[c0], rho1 = P[c1];
    const double u0 = P[PP + c0], u1 = P[PP + c1];
    const double T0 = P[2 * PP + c0], T1 = P[2 * PP + c1];
    const double rho_a = 0.5 * (rho0 + rho1), u_a = 0.5 * (u0 + u1);
    const double T_a = 0.5 * (T0 + T1), mu_a = 0.5 * (P[3 * PP + c0] + P[3 * PP + c1]);
    const double kap_a = 0.5 * (P[4 * PP + c0] + P[4 * PP + c1]);
    const double dx = a.dxb ? a.dxb[b] : a.dx;                // b_obj(ib)%smr_dx
    const double tau = -(2.0 / 3.0 * mu_a * (2.0 * ((u1 - u0) / dx)));
    const DevThermo *__restrict__ th = a.thermo;
    const TPow tp = tpow(T_a, th->ru);
    double hsum = 0.0;
    double *F = a.F + long(b) * n1 + j;
    const long PF = a.PF;
#pragma unroll 4
    for (int k = 0; k < NSP; ++k) {
        const double hs = h_mol(nasa(th, k, T_a), tp) / th->wt[k] * 1.e-4;      // CKHMS, SI
        const double da = 0.5 * (P[(5 + NSP + k) * PP + c0] + P[(5 + NSP + k) * PP + c1]);
        const double dy = (P[(5 + k) * PP + c1] - P[(5 + k) * PP + c0]) / dx;
        hsum = hsum - hs * da * dy;
        F[(3 + k) * PF] = 0.0 - rho_a * da * dy;
    }
    F[0] = 0.0 + 0.0;
    F[PF] = 0.0 + tau;
    F[2 * PF] = 0.0 + (tau * u_a + rho_a * hsum - kap_a * ((T1 - T0) / dx));
}

}  // namespace

hipError_t launch_flux_tran(const TranArgs &a, int nsp, hipStream_t st) {
    const unsigned ncells = unsigned(a.nblk * (a.ncell + 2));
    const unsigned nif = unsigned(a.nblk * (a.ncell + 1));
#define FCFD_CASE(N)                                                                               \
    if (nsp == N) {                                                                                \
        hipLaunchKernelGGL((k_tran_props<N>), dim3(ncells), dim3(64), 0, st, a);                  \
        hipError_t e = hipGetLastError();                                                          \
        if (e != hipSuccess) return e;                                                             \
        hipLaunchKernelGGL((k_flux_tran<N>), dim3((nif + 63) / 64), dim3(64), 0, st, a);          \
        return hipGetLastError();                                                                  \
    }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

void build_dev_tran(const Mech &m, DevTran &t) {
    t.no = m.no;
    t.patmos = m.patmos;
    for (int k = 0; k < m.kk; ++k) {
        t.twt[k] = m.twt[k];
        for (int i = 0; i < m.no && i < 4; ++i) {
            t.cofeta[k][i] = m.cofeta[size_t(k) * m.no + i];
            t.coflam[k][i] = m.coflam[size_t(k) * m.no + i];
        }
        for (int j = 0; j < m.kk; ++j)
            for (int i = 0; i < m.no && i < 4; ++i)
                t.cofd[k][j][i] = m.cofd[(size_t(k) * m.kk + j) * m.no + i];
    }
}

}  // namespace fcfd
