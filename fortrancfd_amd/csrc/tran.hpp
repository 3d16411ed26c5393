// tran.hpp -- transport tables and the launcher of calc_flux_tran (tran.hip).
#pragma once

#include "fcfd_common.hpp"

namespace fcfd {

// TRANLIB fit coefficients (MCINIT's RMCWRK arrays, tranlib.f:93): polynomials in log T of degree
// no-1 (no <= 4) for log viscosity, log conductivity and log binary diffusion (cofd[k][j], k > j).
struct DevTran {
    int no;
    double patmos;
    double twt[MAXK];
    double cofeta[MAXK][4];
    double coflam[MAXK][4];
    double cofd[MAXK][MAXK][4];
};

void build_dev_tran(const Mech &m, DevTran &t);

struct TranArgs {
    const DevThermo *thermo;
    const DevTran *tran;
    const double *Qc;      // [nequ][P] conserved, ghost-padded
    double *props;         // [5 + 2 KK][PP]: rho, u, T, mu, kappa, Y_k, dif_k of cells ista-1 .. iend+1
    double *F;             // [nequ][PF] interface fluxes (overwritten)
    int nblk, ncell, lv, S;
    long P, PF, PP;
    double dx;
    const double *dxb;     // [nblk] per-block cell size (nullptr: dx everywhere)
};

hipError_t launch_flux_tran(const TranArgs &a, int nsp, hipStream_t st);

}  // namespace fcfd
