/* fcfd.h -- C ABI of libfcfd_hip.so, the MI355X (gfx950) implementation of fortranCFD's
 * per-timestep cell and interface update.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a negative FCFD_E*
 * code; fcfd_last_error() then holds a message (per calling thread).  Nothing throws across the
 * ABI.  Host arrays use the reference's Fortran layout: per cell the (nequ = 3 + KK) column
 * rho, rho*u, rho*E, rho*Y_k (conserved, "Qc") or rho, u, T, Y_k (primitive, "qf"), equation index
 * fastest, cells consecutive (SRC/variables.f90:6-16, SRC/block.f90:409-413).
 *
 * Threading: one host thread per context; each context owns one HIP stream.  Entry points that
 * take a context are synchronous at return unless stated otherwise.
 *
 * Reference call sites each entry point replaces are cited as file:line (paths relative to the
 * reference root).
 */
#ifndef FCFD_H
#define FCFD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FCFD_OK = 0,
    FCFD_E_ARG = -1,      /* invalid argument / unsupported option          */
    FCFD_E_IO = -2,       /* link file unreadable or malformed              */
    FCFD_E_HIP = -3,      /* HIP runtime error (no device, launch failure)  */
    FCFD_E_UNSUPPORTED = -4
};

typedef struct fcfd_mech fcfd_mech;
typedef struct fcfd_ctx fcfd_ctx;

/* Solver switches, same names and codes as the reference's &params namelist
 * (SRC/params.f90:106-111) and &problem/&boundary groups. */
typedef struct fcfd_opts {
    int do_hydro;        /* 1: advection fluxes                                   */
    int do_tran;         /* 1: viscous/heat/diffusive fluxes (needs tran.bin)     */
    int do_reaction;     /* 1: chemistry source                                   */
    int select_time;     /* 0 explicit Euler, 1 TVD-RK3, 2 RK4                    */
    int select_riemann;  /* 0 LxF, 1 HLL, 2 HLLC                                  */
    int flag_acu;        /* 0 first order, 1 MUSCL, 2 WENO                        */
    int flag_muscl;      /* 1 unlimited, 2 minmod, 3 superbee, 4 MC               */
    int flag_weno;       /* 1 WENO3-Z, 2 WENO5-JS, 3 WENO5-Z, 4 WENO7-Z           */
    int is_spherical;    /* must be 0 (Cartesian)                                 */
    int flag_lbound;     /* 0 reflective wall, 1 inflow (fcfd_set_flow_boundary)  */
    int flag_rbound;     /* 0 reflective wall, 1 outflow (fcfd_set_flow_boundary) */
    double atol, rtol;   /* chemistry tolerances (solver.inp atol/rtol)          */
} fcfd_opts;

/* ---- mechanism ---------------------------------------------------------------------------
 * Reads the CHEMKIN-II link files the reference opens in set_thermo / set_trans
 * (SRC/thermo.f90:43-52 -> CKINIT cklib.f:3972; SRC/trans.f90:35-42 -> MCINIT tranlib.f:93).
 * tran_bin may be NULL when transport is not used. */
int fcfd_mech_load(const char *chem_bin, const char *tran_bin, fcfd_mech **out);
/* The same mechanism from its CHEMKIN-II text instead of the link file: chem.inp (ELEMENTS,
 * SPECIES, optional THERMO, REACTIONS with LOW / TROE / SRI / DUPLICATE / third-body
 * efficiencies), therm.dat (may be NULL when chem.inp holds all thermo), tran.bin (may be NULL).
 * Replaces the interpreter step ckinterp.f -> chem.bin of WORK/KineticModels/Chem_bin/GRI3.0_CH4/
 * comp.sh:3-4 and gives the arrays CKINIT (cklib.f:3972) would read from its output. */
int fcfd_mech_load_text(const char *chem_inp, const char *therm_dat, const char *tran_bin,
                        fcfd_mech **out);
/* Mechanism tables, per species k / reaction i (0-based), as CKINIT holds them: NASA-7 temperatures
 * (Tlow, Tmid, Thigh) and coefficients (7 low-range, then 7 high-range); Arrhenius A, b, E/R,
 * coefficients nu[6] and 1-based species nk[6] (reactants in slots 0-2), NSPEC (negative when
 * irreversible) and NREAC; fall-off form (0: none, 1 Lindemann, 2 SRI, 3/4 Troe), its [M] species
 * (1-based, 0 for +M) and 8 parameters; third-body efficiency count (-1: no third body), species
 * and values (at most 10). */
int fcfd_mech_thermo(const fcfd_mech *m, int k, double *temps, double *coef);
int fcfd_mech_reaction(const fcfd_mech *m, int i, double *par, int *nu, int *nk, int *ns, int *nr);
int fcfd_mech_falloff(const fcfd_mech *m, int i, int *ifo, int *kfl, double *fpar);
int fcfd_mech_thirdbody(const fcfd_mech *m, int i, int *ntbs, int *nktb, double *aik);
void fcfd_mech_free(fcfd_mech *m);
int fcfd_mech_dims(const fcfd_mech *m, int *kk, int *ii);
/* Species k (0-based): symbol (NUL-terminated, at most len-1 chars), molar mass g/mol and element
 * counts (ncf[MM]), as CKSYMS / CKWT / CKNCF return them (cklib.f CKINIT species record). Each
 * output may be NULL. */
int fcfd_mech_species(const fcfd_mech *m, int k, char *name, int len, double *wt, int *ncf);
/* Number of elements MM. */
int fcfd_mech_elements(const fcfd_mech *m, int *mm);

/* ---- domain context -----------------------------------------------------------------------
 * nblk equal blocks of ncell interior cells and spacing dx, all on the current HIP device
 * (replaces set_one_block_size / set_block_size + set_block_params, SRC/block.f90:38,91,394). */
int fcfd_create(const fcfd_mech *m, const fcfd_opts *o, int nblk, int ncell, double dx,
                fcfd_ctx **out);
/* The reference's stretched-mesh layout, its &blocks namelist (set_block_size,
 * SRC/block.f90:91-392 on one rank): nblk_comp blocks of ncell cells at dx_comp, then
 * (lev_smr_0 - 1) * num_smr_0 coarser blocks on the left and (lev_smr_n - 1) * num_smr_n on the
 * right, num_smr blocks per level, the cell size doubling per level away from the fine region
 * (dx_comp * 2^(level - 1)).  Blocks are numbered left to right (fcfd_nblocks of them; host arrays
 * cover all of them).  Coarse blocks reconstruct at first order (SRC/flux.f90:240); the halo
 * between unequal blocks averages pairs of fine cells and interpolates coarse cells with a
 * minmod slope (interpolation_block, SRC/block.f90:499-563); the cell centres follow
 * set_block_size (x = 0 where the left refinement region ends).  ncell >= 2 lvir; select_time 0
 * or 1.  calc_dt keeps dx_comp (SRC/main.f90:134). */
int fcfd_create_smr(const fcfd_mech *m, const fcfd_opts *o, int nblk_comp, int ncell, double dx_comp,
                    int lev_smr_0, int num_smr_0, int lev_smr_n, int num_smr_n, fcfd_ctx **out);
void fcfd_destroy(fcfd_ctx *c);
/* Ghost layers per block end for the configured scheme (SRC/params.f90:161-173). */
int fcfd_lvir(const fcfd_ctx *c);
/* Number of blocks; each block's cell size b_obj(ib)%smr_dx (nblocks values); the interior cell
 * centres b_obj(ib)%xm_b (nblocks * ncell values, m). */
int fcfd_nblocks(const fcfd_ctx *c);
int fcfd_block_dx(const fcfd_ctx *c, double *dx_block);
int fcfd_get_cell_centres(fcfd_ctx *c, double *xm);

/* Host <-> device state of block ib (0-based): (ncell + 2*lvir) cells x nequ, ghosts included,
 * i.e. b_obj(ib)%Qc / %qf (SRC/block.f90:409-413).  qf may be NULL on upload, in which case it
 * is rebuilt from Qc on every cell. */
int fcfd_upload(fcfd_ctx *c, int ib, const double *Qc, const double *qf);
int fcfd_download(fcfd_ctx *c, int ib, double *Qc, double *qf);

/* Initial state from (T [K], p [Pa], u [m/s], X_1..X_KK) per interior cell, all blocks in order,
 * as initialize_variables + set_block_params + the ghost fill of main_sim do
 * (SRC/params.f90:305-580, SRC/block.f90:394-431, SRC/main.f90:94-114). */
int fcfd_init_cells(fcfd_ctx *c, const double *cells);

/* calc_dt (SRC/main.f90:134 -> SRC/update.f90:13-59): CFL time step over all interior cells. */
int fcfd_calc_dt(fcfd_ctx *c, double cfl, double *dt);

/* solve_fluid (SRC/main.f90:142 -> SRC/update.f90:62-97 -> SRC/timeintegration.f90:12,99). */
int fcfd_solve_fluid(fcfd_ctx *c, double dt);

/* solve_source (SRC/main.f90:145 -> SRC/update.f90:100-150 -> reaction_macks,
 * SRC/source.f90:17-84): integrates every interior cell's (Y, T) over dt at fixed density,
 * renormalises Y, writes T and Y into the primitive state and rebuilds the interior conserved
 * state (qfToQc).  hrr / hrr_dot (may be NULL) receive eos_obj(:)%HRR and %HRR_dot, one value per
 * interior cell, blocks in order (do_reaction = 1).  Then the auxiliary sources that are on
 * (fcfd_set_radiation, fcfd_set_wall_heat, fcfd_set_ignition) in the reference's order
 * (SRC/update.f90:127-147); time is the driver's time after the step, which the ignition source's
 * window reads.  With do_reaction = 0 and no auxiliary source the call does nothing.  The context keeps,
 * per cell, the implicit integrator's last step-size proposal and starts the next call's first step
 * there (when below dt); fcfd_upload and fcfd_init_cells clear it. */
int fcfd_solve_source(fcfd_ctx *c, double dt, double time, double *hrr, double *hrr_dot);
/* Per-cell integrator counters of the last chemistry call: 4 ints per interior cell
 * (accepted steps, rejected attempts, Newton iterations, failed flag). */
int fcfd_chem_stats(fcfd_ctx *c, int *per_cell);

/* Post-step boundary / halo / primitive refresh (SRC/main.f90:149-163). */
int fcfd_refresh(fcfd_ctx *c);

/* Device-resident stepping for production loops: nsteps x {calc_dt, solve_fluid, solve_source
 * (when do_reaction = 1 or an auxiliary source is on), refresh} with dt kept on the device and the
 * step replayed from a
 * captured HIP graph.  *time is advanced by the sum of the steps' dt.  Equivalent to the driver
 * loop SRC/main.f90:130-163 without output.  fcfd_set_step chooses the splitting (Lie by default)
 * and whether the source calls compute HRR / HRR_dot (the reference does on every call,
 * SRC/source.f90:70-73; off by default here). */
int fcfd_advance(fcfd_ctx *c, double cfl, int nsteps, double *time);
/* Step composition of fcfd_advance (no counterpart in the reference, whose driver splits Lie-style,
 * SRC/main.f90:142-145).  strang = 0: calc_dt, solve_fluid(dt), solve_source(dt), refresh (the
 * default, the reference's order and the parity mode); strang = 1: calc_dt, solve_source(dt/2),
 * refresh, solve_fluid(dt), solve_source(dt/2), refresh (the Strang splitting the reference's
 * documentation describes, DOCS/README.md:142,181).  hrr_each_step = 1: the last source call of
 * every step computes HRR / HRR_dot (SRC/source.f90:70-73), read with fcfd_get_hrr; 0: it skips
 * them (one rate evaluation per cell less). */
int fcfd_set_step(fcfd_ctx *c, int strang, int hrr_each_step);
/* ---- open boundaries and auxiliary per-cell sources ---------------------------------------- */
/* Inflow / outflow state for flag_lbound = 1 / flag_rbound = 1 (update_boundary,
 * SRC/boundary.f90:12-149), from the values set_flow_boundary reads from its input file
 * (SRC/flow_bound.f90:15-57): inflow temperature T_in (K), velocity u_in (cm/s) and mole fractions
 * x_in[KK] (may be NULL without an inflow), outflow pressure p_out (atm).  The inflow ghost cells
 * take the density of the interior cell they mirror, the mass flux density(p_out, T_in, x_in) * u_in,
 * T_in and Y_in; the outflow ghost cells take the mirrored cell's velocity, composition and
 * temperature at p_out.  Refills the ghost cells of the current state; must precede
 * fcfd_init_cells when either flag is 1.  Supported with select_time 0 and 1. */
int fcfd_set_flow_boundary(fcfd_ctx *c, double T_in, double u_in_cm_s, const double *x_in, double p_out_atm);
/* Optically thin radiative heat loss (heat_transfer_radiation, SRC/source.f90:85-132 ->
 * radiation_heat_loss, SRC/radiation.f90:74-121): n radiating species (0-based indices, at most 16)
 * with six Planck-mean absorption coefficients a_0..a_5 each (row-major n x 6, the columns of the
 * reference's radiation parameter file, set_radiation_params :21-71).  n = 0 turns it off. */
int fcfd_set_radiation(fcfd_ctx *c, int n, const int *species, const double *pmac);
/* Wall heat loss of the micro flow reactor (heat_transfer_mfr, SRC/source.f90:135-182 ->
 * heat_transfer, SRC/mfr.f90:41-60): Qw = -4 lambda Nu (T - Tw(x)) / d^2 with the mixture
 * conductivity lambda (MCACON, needs the transport link file) and the wall temperature
 * Tw(x) = (Tw_max - Tw_min)/2 tanh(alpha (x - beta)) + (Tw_max + Tw_min)/2 (:63-78); d in mm, the
 * &mfrwalltemp namelist's values.  on = 0 turns it off. */
int fcfd_set_wall_heat(fcfd_ctx *c, int on, double Nu, double d_mm, double Tw_min, double Tw_max,
                       double alpha, double beta);
/* Ignition energy deposit (energy_source, SRC/source.f90:185-236): type 1 A. Frendi
 * (add_igsource), 2 W. Zhang (add_igsource_WZhang), SRC/ignitionsource.f90:150-264; n sources
 * (at most 16; 1 for the single-source input, more for set_multi_igsource) with energy (mJ),
 * radius (um), duration (us) and start time (us), the units of the reference's input files.  A
 * source deposits while start < time < start + duration; where several are active, the last one's
 * power is the cell's (the reference overwrites, :180-195).  type = 0 turns it off. */
int fcfd_set_ignition(fcfd_ctx *c, int type, int n, const double *E_mJ, const double *r_um,
                      const double *t_us, const double *t0_us);
/* Cell centres x_m (m) of the interior cells, blocks in order (b_obj(ib)%xm_b).  fcfd_create sets
 * the uniform block layout (SRC/block.f90:264-305): a running sum of dx from 0 plus dx/2. */
int fcfd_set_cell_centres(fcfd_ctx *c, const double *xm);
/* Qrad and Qw (W/m^3) of the last auxiliary-source call, one value per interior cell, blocks in order
 * (eos_obj(:)%Qrad, %qTw); zeros for a source that is off.  Either pointer may be NULL. */
int fcfd_get_aux(fcfd_ctx *c, double *qrad, double *qw);

/* ---- snapshot fields and restart ------------------------------------------------------------
 * The per-cell output of write_hdf5_b (set_eos_params, SRC/block.f90:433-470, then
 * SRC/write_hdf5.f90:138-160): out holds 15 + KK rows of nblocks * ncell values, in the reference's
 * dataset order rho, u_x, T, p, e_total, e_kinetic, e_internal, HRR, HRR_dot, Qrad, Tw, qTw, x_m, dx,
 * idx (the reference's cell index, 1 at the first fine cell), X_1..X_KK.  HRR / HRR_dot are those of
 * the last source call that computed them, Qrad / qTw of the last auxiliary-source call, Tw the wall
 * temperature when the wall heat loss is on (zeros otherwise). */
int fcfd_eos_fields(fcfd_ctx *c, double *out);
/* Start from a snapshot (read_hdf5_restart, SRC/read_hdf5.f90:13-423, then set_block_params,
 * SRC/block.f90:394-431): fields holds 4 + KK rows of nblocks * ncell values, rho, u, T, e_total,
 * X_1..X_KK; qf = (rho, u, T, Y = CKXTY(X)), Qc = (rho, rho u, rho e_total, rho Y), then the start-up
 * ghost fill.  Clears the integrator's step-size memory. */
int fcfd_init_restart(fcfd_ctx *c, const double *fields);

/* HRR and HRR_dot of the last source call that computed them (one value per interior cell, blocks
 * in order; either pointer may be NULL). */
int fcfd_get_hrr(fcfd_ctx *c, double *hrr, double *hrr_dot);

/* Kernel timing with HIP events on the context's stream (0 = off; while on, steps are launched
 * one kernel at a time instead of from the captured graph).  Kernel ids: 0 advection flux,
 * 1 stage update + primitive recovery, 2 CFL reduction, 3 chemistry implicit integrator,
 * 4 interior qfToQc after the source, 5 surrogate (features, network, acceptance), 6 chemistry
 * explicit screen, 7 transport flux, 8 chemistry explicit half-step retry of a long stiff list,
 * 9 auxiliary sources. */
int fcfd_set_timing(fcfd_ctx *c, int enable);
/* Launch shapes (no counterpart in the reference, whose CPU loops have none).  Read once per process
 * from a table of "<arch> <species count> <key> <value>" lines written by the tuner
 * (python -m fortrancfd_amd.tune): data/tuned_launch.txt beside the library directory, or the file
 * FCFD_TUNE_FILE names; entries of other GPU architectures are ignored ("none" stands for a host
 * without a GPU), absent keys keep their
 * defaults.  Keys: "flux_variant" (WENO5 advection flux: 0 = 64 threads and two waves per SIMD,
 * 1 = 64 / one, 2 = 128 / one), "stage_variant" (stage update: 0 = 64 / two, 1 = 64 / one,
 * 2 = 128 / one, 3 = 256 / one), "chem_wg_per_cu" (implicit-integrator workgroups per CU, default 5),
 * "chem_four_wave_factor" (four-wave integrator mode for lists of at most factor x workgroups cells,
 * default 8), "chem_lu" (how the four-wave integrator factors its Newton matrix: 0 = one elimination
 * step at a time through LDS, 1 = in 16-column panels with the FP64 matrix cores updating the rest,
 * 2 = with each wave's columns held in registers; the factors are bitwise the same, default 2).  fcfd_set_launch overrides one entry for the mechanisms of nsp species (every context
 * re-captures its step graph); fcfd_load_tuning reloads the table (NULL: the default file). */
int fcfd_set_launch(const char *key, int nsp, int value);
int fcfd_get_launch(const char *key, int nsp, int *value);
int fcfd_load_tuning(const char *path);
int fcfd_get_timing(fcfd_ctx *c, int kernel_id, double *ms_total, long long *launches);

/* ---- stateless entry points (host arrays in, host arrays out; run on the current device) ----
 * calc_flux_advection on one block (SRC/flux.f90:130-373; called at
 * SRC/timeintegration.f90:146): qf (n + 2*lv) cells x nequ -> Fplus (n+1) x nequ and
 * godunov_state (n+1) x 3 (overwritten, not accumulated). */
int fcfd_flux_advection_host(const fcfd_mech *m, const fcfd_opts *o, int n, int lv, double dx,
                             const double *qf, double *Fplus, double *godunov);
/* calc_flux_tran (SRC/flux.f90:18-127; called at SRC/timeintegration.f90:140) on one block:
 * Qc (n + 2*lv) cells x nequ (conserved, ghosts included) -> Fplus (n+1) x nequ, the viscous,
 * heat-conduction and diffusive fluxes at the n+1 interfaces (overwritten).  Needs a mechanism
 * loaded with its transport link file.  The viscosity is TRANLIB's Wilke mixture rule (MCAVIS):
 * the reference calls MCAVIS_EA there (SRC/trans.f90:148), which it does not ship. */
int fcfd_flux_tran_host(const fcfd_mech *m, int n, int lv, double dx, const double *Qc,
                        double *Fplus);
/* qfToQc / QcToqf (SRC/eos.f90:11,49) on n cells. */
int fcfd_qf_to_qc_host(const fcfd_mech *m, int n, const double *qf, double *Qc);
int fcfd_qc_to_qf_host(const fcfd_mech *m, int n, const double *Qc, double *qf);
/* Reconstruction kernels (SRC/weno.f90, SRC/muscl.f90) on n 8-value stencils: out (n x 18) =
 * weno3z(s1..4), weno5 JS(s1..6), weno5 Z, weno7z(s1..8), muscl3(s1..4) flags 1..4, (l, r) pairs,
 * last two columns zero. */
int fcfd_recon_host(int n, const double *s, double *out);
/* CKWYR (WORK/KineticModels/Chem_bin/GRI3.0_CH4/cklib.f:9107-9204) and CKCDYR (:1422-1541), the
 * rates SRC/thermo.f90:105-134 and the integrator's right-hand side call, on n states: rho
 * [g/cm^3], T [K], Y (n x KK mass fractions) in; wdot, cdot, ddot (n x KK, mol/(cm^3 s)) and the
 * forward / reverse rates of progress qf, qr (n x II, mechanism order) out.  Evaluated by the
 * device rate code the integrator uses (one state per wavefront). */
int fcfd_rates_host(const fcfd_mech *m, int n, const double *rho, const double *T, const double *Y,
                    double *wdot, double *cdot, double *ddot, double *qf, double *qr);

/* The macks call (SRC/source.f90:58, macks(dt, nspe+1, yiT, rho)) on n independent cells:
 * yT is n rows of (Y_1..Y_KK, T), updated in place; dt and rho (kg/m^3) per cell; stats
 * (may be NULL) receives 4 ints per cell as fcfd_chem_stats.  Uses o->atol / o->rtol. */
int fcfd_chem_step_host(const fcfd_mech *m, const fcfd_opts *o, int n, const double *dt,
                        const double *rho, double *yT, int *stats);
/* Same on DEVICE arrays of the current device (no copies): rho[n], dt[n], T[n] in/out,
 * Y[KK][n] in/out (species-major), stats[4n] or NULL; stream is a hipStream_t or NULL.
 * Calls that share a mechanism must not overlap (one work-queue counter per mechanism). */
int fcfd_chem_batch(const fcfd_mech *m, const fcfd_opts *o, int n, const double *rho,
                    const double *dt, double *T, double *Y, int *stats, void *stream);
/* Device copies of the mechanism's tables on the current device, and the address of this device's
 * scratch of fcfd_chem_batch_noalloc: the set-up call (it allocates) that the allocation-free entry
 * point needs once per device.  Other entry points do the same on first use. */
int fcfd_mech_prepare(const fcfd_mech *m);
/* fcfd_chem_batch without any device allocation: no counterpart in the reference.  For callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The work-queue counter and the stiff-cell list live in a static device array of the
 * library (4 MiB + 256 B per GPU, held until the process ends), so n is limited to 2^20 cells
 * (larger n fails with FCFD_E_ARG).  Needs fcfd_mech_prepare on this device first.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams or threads) race and must
 * not be made.  Results are bitwise those of fcfd_chem_batch. */
int fcfd_chem_batch_noalloc(const fcfd_mech *m, const fcfd_opts *o, int n, const double *rho,
                            const double *dt, double *T, double *Y, int *stats, void *stream);

/* ---- chemistry surrogate (no reference counterpart: stands in for the macks call of
 * SRC/source.f90:58 where its answer is accepted) ----------------------------------------------
 * A 56 -> 4 x 256 GELU -> 54 multilayer perceptron with BF16 weights and activations and FP32
 * accumulation.  Inputs per cell: b(Y_k) (KK = 53), T, log(p / 1 atm), log dt, z-scored, with the
 * Box-Cox transform b(Y) = 10 (Y^0.1 - 1); outputs (b(Y1) - b(Y0), T1 - T0) z-scored.  Only cells
 * the explicit screen cannot finish reach the network.  A network answer becomes a candidate when
 * every raw input lies in the training range, the element-conservation residual of the final state
 * (after projection onto element conservation and clipping of negative mass fractions) is at most
 * resid_tol relative, and the network temperature agrees with the energy-consistent temperature to
 * t_tol relative; a candidate is accepted when its FP64 trapezoidal-rule residual
 * Y1 - Y0 - dt/2 (f(Y0) + f(Y1)), weighted like the integrator's error norm, has an RMS of at most
 * verify_tol.  Every other cell is integrated by the FP64 integrator.  The final T of an accepted
 * cell comes from conservation of the specific internal energy, as in the FP64 path.
 *
 * Parameters (fcfd_sur_nparams() floats) are ordered W0 (64 x 256, row-major, k-major), b0 (256),
 * W1 (256 x 256), b1, W2, b2, W3, b3, W4 (256 x 64), b4 (64); input rows 56..63 of W0 and output
 * columns 54..63 of W4 / b4 are padding.  The norm record is six arrays of 64 doubles: input mean,
 * input scale, training-range low, training-range high (raw units), output mean, output scale.
 * Device pointers are arrays on the mechanism's device; X is BF16 [rows][64], Tg and Yh FP32
 * [rows][64]; stream is a hipStream_t or NULL.  Calls on one surrogate must not overlap. */
typedef struct fcfd_surrogate fcfd_surrogate;
int fcfd_sur_create(const fcfd_mech *m, fcfd_surrogate **out);
void fcfd_sur_free(fcfd_surrogate *s);
int fcfd_sur_nparams(void);
int fcfd_sur_set_params(fcfd_surrogate *s, const float *params);   /* host; resets Adam state */
int fcfd_sur_get_params(fcfd_surrogate *s, float *params);
int fcfd_sur_get_grads(fcfd_surrogate *s, float *grads);           /* of the last train step */
int fcfd_sur_set_norm(fcfd_surrogate *s, const double *norm);
/* Acceptance bounds: element residual (default 1e-5), network-vs-energy T (2e-3, relative) and the
 * FP64 verification residual (1: the integrator's own tolerance). */
int fcfd_sur_set_thresholds(fcfd_surrogate *s, double resid_tol, double t_tol, double verify_tol);
/* Features of n cells (device rho[n] kg/m^3, dt[n], T[n], Y[KK][n]) into X (rows padded to a
 * multiple of 64 with zeros); in_range[n] (may be NULL) = 1 where every input is in range. */
int fcfd_sur_features(fcfd_surrogate *s, int n, const double *rho, const double *dt,
                      const double *T, const double *Y, void *X, int *in_range, void *stream);
/* Normalised targets of n (state, result) pairs into Tg (rows padded like X). */
int fcfd_sur_targets(fcfd_surrogate *s, int n, const double *T0, const double *Y0,
                     const double *T1, const double *Y1, float *Tg, void *stream);
/* Network output for rows (multiple of 64) rows of X. */
int fcfd_sur_forward(fcfd_surrogate *s, int rows, const void *X, float *Yh, void *stream);
/* One Adam step (b1 0.9, b2 0.999, eps 1e-8) on the mean squared error over the first 54 output
 * columns of rows idx[0..batch) of (X_all, Tg_all); lr = 0 computes the gradients only.  loss
 * (may be NULL; then the call does not wait for the device) receives the batch's MSE. */
int fcfd_sur_train_step(fcfd_surrogate *s, const void *X_all, const float *Tg_all, const int *idx,
                        int batch, float lr, float *loss, void *stream);
/* Chemistry of n independent cells (arrays as fcfd_chem_batch) through the explicit screen, the
 * surrogate for the cells the screen rejects, and the FP64 implicit integrator for the cells the
 * surrogate does not accept.  route[n] (may be NULL) receives 2 (explicit step over dt), 3 (two
 * explicit half steps), 1 (surrogate) or 0 (implicit integrator) per cell; stats[4n] may be NULL;
 * *n_fallback (may be NULL) receives the number of implicit-integrator cells. */
int fcfd_sur_infer_batch(fcfd_surrogate *s, const fcfd_opts *o, int n, const double *rho,
                         const double *dt, double *T, double *Y, int *route, int *n_fallback,
                         int *stats, void *stream);
/* solve_source of a context through the surrogate (NULL: the FP64 integrator on every cell).  HRR /
 * HRR_dot of cells the surrogate answers are computed from its answer like those of the integrator
 * (HRR_dot from the production rates at the answer). */
int fcfd_set_surrogate(fcfd_ctx *c, fcfd_surrogate *s);
/* Per interior cell of the last solve_source: 2 explicit step over dt (the screen), 3 two explicit
 * half steps (the retry), 1 surrogate, 0 FP64 implicit integrator; *n_fallback = number of cells
 * with 0. */
int fcfd_source_accepted(fcfd_ctx *c, int *per_cell, int *n_fallback);
/* Grid of the FP64 implicit integrator in this context's solve_source: n one-wave workgroups for
 * lists of stiff cells longer than 8 n, else n / 4 four-wave workgroups (0: the default, four per
 * compute unit; at most 2^24).  No counterpart in the reference: a tuning knob (a smaller grid leaves compute
 * units to concurrent work).  A cell's answer can differ in the last bits with the workgroup kind
 * that integrates it. */
int fcfd_set_chem_workgroups(fcfd_ctx *c, int n);

/* Plain BF16 GEMM C (FP32) = A B on device arrays with element (m, k) of A at A[m*sam + k*sak]
 * and (k, n) of B at B[k*sbk + n*sbn]; M, N multiples of 64, K / ksplit a multiple of 16;
 * split s of K writes C + s*M*ldc.  The kernel the surrogate's layers run on. */
int fcfd_bf16_gemm(int M, int N, int K, const void *A, long sam, long sak, const void *B,
                   long sbk, long sbn, float *C, long ldc, int ksplit, void *stream);

const char *fcfd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
