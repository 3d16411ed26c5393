// fluid.hpp -- argument blocks and launchers of the hydrodynamic kernels (fluid.hip).
#pragma once

#include "fcfd_common.hpp"

namespace fcfd {

struct FluxArgs {
    const DevThermo *thermo;
    const double *qf;     // [nequ][P] primitive, ghost-padded
    double *F;            // [nequ][PF] interface fluxes
    double *godunov;      // [3][PF] or nullptr
    int nblk, ncell, lv, S;
    long P, PF;
    int muscl, riemann;
    int accumulate;       // 1: F += advection flux (transport flux already in F), 0: F = flux
    int spherical;        // 1: momentum flux rho u u without p (SRC/flux.f90:272-278)
    int b0;               // first block of this launch (nblk blocks from there)
    int variant;          // launch shape (WENO5 only): 0 64 threads, 2 waves/EU; 1 64, 1; 2 128, 1
};

struct StageArgs {
    const DevThermo *thermo;
    const double *Q0, *Qin, *F;
    double *Qout, *qf;
    double *acc;           // RK4: running k1 + 2 k2 + 2 k3
    const double *dt_dev;  // device dt (nullptr: use dt)
    double dt, dx, h1, h2;
    int mode;              // 0 Qin - dt/dx dF, 1 h1*Q0 + h2*(Qin - dt/dx dF), 2 copy; RK4: stage 0..3
    int ghosts_only;
    int spherical;         // 1: r^2-weighted update with the pressure-gradient term (needs xm, god)
    const double *xm;      // [nblk * ncell] cell centres (spherical)
    const double *god;     // [3][PF] Roe states of the interfaces; [2] the pressure (spherical)
    const double *dxb;     // [nblk] per-block cell size (nullptr: dx everywhere)
    int variant;           // launch shape of k_stage_end: 0 64 threads, 2 waves/EU; 1 64, 1; 2 128, 1; 3 256, 1
    int nblk, ncell, lv, S;
    long P, PF;
};

struct DtArgs {
    const DevThermo *thermo;
    const double *qf;
    unsigned long long *cvel_bits;
    double *dt_out;
    double *time;          // optional running time, advanced by dt
    double cfl, dx;
    int nblk, ncell, lv, S;
    long P;
};

struct InitArgs {
    const DevThermo *thermo;
    const double *cells;   // [ntot][3+KK]: T, p, u, X
    double *Qc, *qf;
    int nblk, ncell, lv, S;
    long P;
};

// Halo of blocks with unequal cell sizes (interpolation_block's refinement branches,
// SRC/block.f90:499-563): ifb lists the left block of each such block pair.
struct SmrArgs {
    const DevThermo *thermo;
    double *Q, *qf;
    const int *ifb;
    const double *dxb;
    int nif, nblk, ncell, lv, S, nequ;
    long P;
};

hipError_t launch_flux(const FluxArgs &a, int nsp, int flag_acu, int flag_weno, hipStream_t st);
hipError_t launch_smr_halo(const SmrArgs &a, int nsp, hipStream_t st);
hipError_t launch_stage_end(const StageArgs &a, int nsp, hipStream_t st);
hipError_t launch_rk4_stage(const StageArgs &a, int nsp, hipStream_t st);
hipError_t launch_dt(const DtArgs &a, int nsp, hipStream_t st);
hipError_t launch_init(const InitArgs &a, int nsp, hipStream_t st);
hipError_t launch_qf_to_qc(const DevThermo *m, int nsp, int n, const double *qf, double *Qc, long P, hipStream_t st);
hipError_t launch_qc_to_qf(const DevThermo *m, int nsp, int n, const double *Qc, double *qf, long P, hipStream_t st);
hipError_t launch_qf_to_qc_interior(const DevThermo *m, int nsp, int nblk, int ncell, int lv, int S,
                                    const double *qf, double *Qc, long P, hipStream_t st);
hipError_t launch_recon(int n, const double *s, double *out, hipStream_t st);

}  // namespace fcfd
