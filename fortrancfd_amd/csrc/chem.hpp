// chem.hpp -- kinetics tables and launchers of the chemistry source kernels (chem.hip).
#pragma once

#include "fcfd_common.hpp"

namespace fcfd {

constexpr int LANE_NW = 8;       // waves per workgroup of the lane-per-cell kernels
constexpr int LANE_CHUNK = 32;   // reaction rows per chunk of their scatter (four per wave)
constexpr int LANE_NC = 64;      // cells per workgroup of the lane-per-cell kernels

// A row of the branch-free range (nplain0 .. ii-1) as the lane-per-cell kernels read it, with one
// row index for the whole wave: coefficients as doubles, and the species of each present slot as the
// byte offset k * LANE_NC * 8 of its row in a species-major [k][LANE_NC] LDS block.  Reactant slots
// are 0 .. nr-1, product slots 3 .. 3+np-1 (build_dev_kin admits only rows stored that way).
struct LanePlain {
    double A, B, E;
    double nud[6];
    unsigned short off[6];
    int bits;                    // LP_* fields
};
constexpr int LP_NR = 0;         // bits 0-1: reactant slots (1-3)
constexpr int LP_NP = 2;         // bits 2-3: product slots (0-3)
constexpr int LP_REV = 1 << 4;   // reverse rate from the equilibrium constant
constexpr int LP_ARR0 = 1 << 5;  // B = E = 0
constexpr int LP_NUS = 6;        // bits 6-8: coefficient sum + 3
constexpr int LP_POW = 1 << 9;   // some slot has |coefficient| > 1
constexpr int LP_PW = 10;        // bits 10-21: |coefficient| of slot n in bits 10 + 2n .. 11 + 2n
constexpr int LP_DEAD = 1 << 22; // coefficient of slot 0 is 0: no rate
// A scatter entry of the lane-per-cell kernels: W[k] += Q[r] * nu with the byte offsets of row k of
// the production-rate block and row r of the chunk's rate block.
struct LaneEnt {
    int offW, offQ;
    double nu;
};

// Reaction data in the form CKRATT / CKRATX / CKWYR consume it (cklib.f:6630-6993, 9107-9204),
// re-indexed 0-based with one row per reaction.
struct DevKin {
    int ii, kk, nthb, nfal;
    int nplain0;                        // rows nplain0 .. ii-1 suit reaction_plain (chem_device.hpp)
    double small, exparg;
    double A[MAXR], B[MAXR], E[MAXR];   // PAR(1..3): A, temperature exponent, E/R
    int rev[MAXR];                      // NSPEC > 0: reverse rate from equilibrium
    int arr0[MAXR];                     // B = E = 0: the rate constant is A (exp(0) = 1)
    int nus[MAXR];                      // sum of the six stoichiometric coefficients
    int nu[MAXR][6], nk[MAXR][6];       // coefficients and 0-based species (-1: empty slot)
    int tb[MAXR], fo[MAXR];             // third-body / fall-off table row, -1 if none
    int ntbs[MAXR], tbk[MAXR][MAXTB];   // enhanced third bodies (species, 0-based)
    double tba[MAXR][MAXTB];            // their efficiencies
    int ifop[MAXR], kfal[MAXR];         // fall-off form (1 Lindemann, 2 SRI, 3/4 Troe), [M] species
    double fpar[MAXR][8];
    int orig[MAXR];                     // the mechanism's (0-based) index of each device row
    // scatter lists of the lane-per-cell kernels (chem_lane.hip): rows are taken in chunks of
    // LANE_CHUNK; within chunk c, wave w adds the entries lane_ent[lane_ent0[c * LANE_NW + w] ..
    // lane_ent0[c * LANE_NW + w + 1]) to the production rates of the species it owns (k mod LANE_NW
    // = w), in row then slot order.
    int lane_ent0[(MAXR + LANE_CHUNK - 1) / LANE_CHUNK * LANE_NW + 1];
    LaneEnt lane_ent[MAXR * 6];
    LanePlain lane_plain[MAXR];         // rows nplain0 .. ii-1 (other entries unused)
};


void build_dev_kin(const Mech &m, DevKin &k);

// Per-cell integrator counters (optional output).
struct ChemCellStats {
    int steps, rejects, newton, failed;
};

struct ChemArgs {
    const DevThermo *thermo;
    const DevKin *kin;
    unsigned int *counter;       // work-queue head, zeroed before launch
    int ncells;
    double rtol, atol;
    // layout A (solve_source): primitive state of a padded SoA domain, updated in place
    double *qf;
    long P;
    int ncell, lv, S;
    double dt;
    const double *dt_dev;        // device dt (overrides dt when non-null)
    double *hrr, *hrr_dot;       // per interior cell (may be null)
    // layout B (batched cells): arrays of length ncells; y is [KK][ncells]
    const double *b_rho, *b_dt;
    double *b_T, *b_y;
    ChemCellStats *stats;        // may be null
    // optional work list: cells list[0 .. *list_n) instead of 0 .. ncells (both device memory)
    const int *list;
    const unsigned *list_n;
    int skip_explicit;           // 1: the explicit pre-step already failed for every listed cell
    // explicit screen output: cells that need the implicit integrator (list + length)
    int *stiff;
    unsigned *stiff_n;
    int *route;                  // optional per-cell route: 2 explicit step over dt, 3 two explicit half
                                 // steps, 1 surrogate, 0 implicit (SDIRK) integrator
    int layout;                  // 0: domain, 1: batch
    int coop_max;                // set by launch_chem: lists up to this length run four waves per cell
    // optional (domain layout with stats): buffer of the list's capacity that receives the list
    // reordered with the cells whose previous integration took the most attempts first
    int *order;
    // optional (domain layout): per interior cell, the integrator's last step-size proposal, carried
    // from one call to the next (0: none; the screen and the surrogate reset it for cells they finish)
    double *hmem;
    // optional (domain layout): set by k_chem_retry to 1 when every listed cell has already failed
    // the explicit half-step retry (k_chem then goes straight to its implicit steps), else 0
    int *retry_done;
    // optional (domain layout): f at the start state of each cell the explicit screen rejects,
    // [NSP + 1][ncells] (species, then T); k_chem_retry starts from it instead of evaluating it again
    double *f0;
    int lu_blocked;              // set by launch_chem: the four-wave mode's LU (0 lu, 1 lu_blocked, 2 lu_regs)
};

// SDIRK integrator over all cells or over a work list (persistent grid): one cell per wave for long
// lists, one cell per four-wave workgroup for short ones.
hipError_t launch_chem(const ChemArgs &a, int nsp, int n_workgroups, hipStream_t st);
// Lane-per-cell explicit screen over cells 0 .. ncells: accepted cells are final, the others are
// appended to a.stiff (a.stiff_n must be zero before the launch).
hipError_t launch_chem_screen(const ChemArgs &a, int nsp, hipStream_t st);
// Lane-per-cell explicit half-step retry of the list a.list[0 .. *a.list_n) when it is longer than
// a.coop_max: accepted cells are final, the others go to a.stiff (a.stiff_n must be zero before
// the launch) and *a.retry_done = 1; a shorter list is copied to a.stiff unchanged, *a.retry_done = 0.
hipError_t launch_chem_retry(const ChemArgs &a, int nsp, hipStream_t st);
// Length up to which a stiff list runs four waves per cell in launch_chem.
// Longest list that the four-wave integrator mode takes: the tuned factor (default 8) times the
// workgroup count (api.hip).
int chem_coop_max(int n_workgroups, int nsp);
// The four-wave integrator's LU (the tuned "chem_lu" entry, api.hip): 0 step by step through LDS,
// 1 blocked with matrix-core trailing updates, 2 with the matrix in registers; the same factors.
int chem_lu_variant(int nsp);
hipError_t launch_zero_u32(unsigned *p, hipStream_t st);
// *out = f * *in on the device (one thread), e.g. the half step of Strang splitting.
hipError_t launch_scale_scalar(const double *in, double *out, double f, hipStream_t st);
// CKWYR / CKCDYR (cklib.f:9107-9204, 1422-1541) on n states with the integrator's own rate code:
// rho [g/cm^3], T [K], Y [n][KK] in; wdot, cdot, ddot [n][KK] (mol/(cm^3 s)) and the rates of
// progress qf, qr [n][II] (mechanism order) out.  One state per wave.
hipError_t launch_rates(const DevThermo *th, const DevKin *K, int nsp, int n, const double *rho, const double *T,
                        const double *Y, double *wdot, double *cdot, double *ddot, double *qf, double *qr,
                        hipStream_t st);
// FP64 verification of surrogate answers for the cells cand_list[0 .. *ncand): accepted when the
// trapezoidal-rule residual Y1 - Y0 - dt/2 (f(Y0) + f(Y1)) of the candidate (Y1, T1) = cand[i],
// weighted by the integrator's tolerances, has an RMS <= tol (1: the integrator's own bar); accepted
// answers are written to the state, the others appended to the fallback list.  One candidate per
// wave (chem.hip).
hipError_t launch_chem_verify(const ChemArgs &a, int nsp, const int *cand_list, const unsigned *ncand,
                              const double *cand, int rows_max, int *fallback, unsigned *nfallback, double tol,
                              hipStream_t st);
// Screen, then SDIRK on the cells the screen rejected; list / list_n are scratch for ncells cells.
hipError_t launch_chem_full(ChemArgs a, int nsp, int n_workgroups, int *list, unsigned *list_n,
                            hipStream_t st);

}  // namespace fcfd
