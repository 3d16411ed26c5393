// aux.hpp -- open boundaries and auxiliary per-cell sources (aux.hip).
#pragma once

#include "fcfd_common.hpp"
#include "tran.hpp"

namespace fcfd {

constexpr int MAX_RAD = 16;   // radiating species
constexpr int MAX_IG = 16;    // ignition sources

// Inflow / outflow state of update_boundary (SRC/boundary.f90:70-149), as set_flow_boundary
// derives it from its input file (SRC/flow_bound.f90:15-57), SI units.  Lives in device memory.
struct FlowBC {
    int lbound, rbound;          // 0 wall, 1 inflow (left) / outflow (right)
    double T_in, rho_u_in, p_out;
    double x_in[MAXK], y_in[MAXK];
};

// Parameters of the auxiliary sources of solve_source (SRC/update.f90:127-147).  Device memory.
struct AuxParams {
    int nrad;                              // radiation_heat_loss, SRC/radiation.f90:74-121
    int rad_idx[MAX_RAD];
    double pmac[MAX_RAD][6];
    int wall_heat;                         // heat_transfer, SRC/mfr.f90:41-78
    double Nu, d_m, Tw_min, Tw_max, alpha, beta;
    int ig_type, n_ig;                     // add_igsource (1) / add_igsource_WZhang (2)
    double E_j[MAX_IG], r_m[MAX_IG], t_s[MAX_IG], t0_s[MAX_IG], vol[MAX_IG];
};

struct OpenBCArgs {
    const DevThermo *thermo;
    const FlowBC *bc;
    double *Q, *qf;                        // the state the preceding stage wrote
    int lbound, rbound;
    int nblk, ncell, lv, S;
    long P;
};

struct AuxArgs {
    const DevThermo *thermo;
    const DevTran *tran;                   // wall heat only
    const AuxParams *par;
    double *Q, *qf;
    const double *xm;                      // [nblk * ncell] cell centres
    double *qrad, *qw;                     // [nblk * ncell] outputs (nullptr when that source is off)
    const double *dt_dev, *time_dev;       // device dt / time (nullptr: the host values)
    double dt, time;
    int rad, wall, ig;                     // which sources run
    int nblk, ncell, lv, S;
    long P;
};

// set_eos_params + the write_hdf5_b field rows, and the restart state of set_block_params.
struct EosArgs {
    const DevThermo *thermo;
    const AuxParams *par;                  // for Tw (wall_heat)
    int wall_heat;
    double *Q, *qf;
    const double *xm, *dxb;                // dxb nullptr: dx
    double dx;
    const double *hrr, *hrd, *qrad, *qw;   // nullptr: zeros
    double *out;                           // [15 + KK][n]; restart: the input [4 + KK][n]
    int idx0;
    int nblk, ncell, lv, S;
    long P;
};

hipError_t launch_open_bc(const OpenBCArgs &a, int nsp, hipStream_t st);
hipError_t launch_eos_fields(const EosArgs &a, int nsp, hipStream_t st);
hipError_t launch_restart(const EosArgs &a, int nsp, hipStream_t st);
hipError_t launch_aux(const AuxArgs &a, int nsp, hipStream_t st);
// y_in = CKXTY(x_in) and rho_u_in = density(p_out, T_in, x_in) * u_in on the device (one thread).
hipError_t launch_flow_bc_setup(const DevThermo *m, FlowBC *bc, double u_in, int nsp, hipStream_t st);

}  // namespace fcfd
