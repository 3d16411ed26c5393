/* fcfd_oracle.h -- CPU restatement of the reference's per-timestep cell and interface update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (fortrancfd_amd/) never does.
 *
 * Every function states the reference routine it restates (paths relative to the reference root;
 * CKLIB/TRANLIB are WORK/KineticModels/Chem_bin/GRI3.0_CH4/{cklib,tranlib}.f).  Units follow the
 * reference: the thermo.f90-level functions are SI, the CK* / MC* level functions are cgs.
 * Layout of per-cell vectors: q[0]=rho, q[1]=u (or rho*u), q[2]=T (or rho*E), q[3+k]=Y_k (rho*Y_k),
 * i.e. the reference's (nequ) column, equation index fastest (SRC/variables.f90:6-16).
 */
#ifndef FCFD_ORACLE_H
#define FCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_mech fo_mech;

/* Mechanism from CHEMKIN-II link files (records as read by CKINIT cklib.f:3972-4170 and MCINIT
 * tranlib.f:93-326).  tran_bin may be NULL.  Returns NULL on failure (message on stderr). */
fo_mech *fo_mech_load(const char *chem_bin, const char *tran_bin);
void fo_mech_free(fo_mech *m);
int fo_mech_kk(const fo_mech *m);
int fo_mech_ii(const fo_mech *m);
int fo_mech_mm(const fo_mech *m);
void fo_mech_ncf(const fo_mech *m, int *ncf);               /* [KK*MM] element counts */
/* Export of the parsed tables, in CKINIT's own units and orders, for pinning against CKINIT. */
void fo_mech_weights(const fo_mech *m, double *wt);               /* [KK] g/mol          */
void fo_mech_nasa(const fo_mech *m, double *tt, double *aa);      /* [KK*3], [KK*2*7]    */
void fo_mech_arrhenius(const fo_mech *m, double *par, int *nu, int *nk, int *ns, int *nr);

/* ---- CKLIB species / mixture thermo (cgs, molar) ---------------------------------------- */
void fo_ckhml(const fo_mech *m, double T, double *hml);   /* CKHML  cklib.f:3500 */
void fo_ckuml(const fo_mech *m, double T, double *uml);   /* CKUML  cklib.f:8409 */
void fo_ckcpml(const fo_mech *m, double T, double *cpml); /* CKCPML cklib.f:1811 */
void fo_cksmh(const fo_mech *m, double T, double *smh);   /* CKSMH  cklib.f:7501 */
void fo_ckhms(const fo_mech *m, double T, double *hms);   /* CKHMS  cklib.f:3562 */
void fo_ckytx(const fo_mech *m, const double *Y, double *X); /* CKYTX cklib.f:9587 */
void fo_ckxty(const fo_mech *m, const double *X, double *Y); /* CKXTY cklib.f:9421 */
void fo_adjust_sum(int n, double *v);                     /* SRC/thermo.f90:136-149 */

/* ---- thermo.f90 wrappers (SI) -------------------------------------------------------------- */
double fo_pressure(const fo_mech *m, double rho, double T, const double *X);     /* :166 */
double fo_density(const fo_mech *m, double p, double T, const double *X);        /* :151 */
double fo_specific_energy(const fo_mech *m, double T, const double *X);          /* :356 */
double fo_cp(const fo_mech *m, double T, const double *X);                       /* :324 */
double fo_cv(const fo_mech *m, double T, const double *X);                       /* :304 */
double fo_sonic_speed(const fo_mech *m, double rho, double p, double T, const double *X); /* :421 */
double fo_temperature_con(const fo_mech *m, double e, const double *X);          /* :211 */
double fo_heat_release_rate(const fo_mech *m, const double *Y0, const double *Y, double rho,
                            double dt);                                          /* :83  */
double fo_heat_release_rate_omegadot(const fo_mech *m, const double *Y, double rho, double T);

/* ---- eos.f90 --------------------------------------------------------------------------------- */
void fo_qf_to_qc(const fo_mech *m, const double *qf, double *Qc);   /* SRC/eos.f90:11-46 */
void fo_qc_to_qf(const fo_mech *m, const double *Qc, double *qf);   /* SRC/eos.f90:49-83 */

/* ---- reconstruction (SRC/weno.f90, SRC/muscl.f90) ------------------------------------------- */
void fo_weno3z(const double *s, double *sl, double *sr);
void fo_weno5(const double *s, double *sl, double *sr, int flag_weno);
void fo_weno7z(const double *s, double *sl, double *sr);
void fo_muscl3(const double *s, double *sl, double *sr, int flag_muscl);

typedef struct {
    int flag_acu, flag_weno, flag_muscl, select_riemann, select_time;
    int is_spherical, do_hydro, do_tran;
    double dx_comp;
} fo_opts;

/* Number of ghost layers for a scheme (SRC/params.f90:161-173). */
int fo_lvir(const fo_opts *o);

/* calc_flux_advection (SRC/flux.f90:130-373) over one block of n interior cells.
 * qf: (n + 2*lv) cells x nequ, first cell = ista-lv.  Fplus/godunov: (n+1) interfaces
 * (ista-1 .. iend) x nequ / x 3.  Fplus is ADDED to (caller zeroes it, as TVDRK3 does). */
void fo_flux_advection(const fo_mech *m, const fo_opts *o, int n, int lv, double dx_block,
                       const double *qf, double *Fplus, double *godunov);

/* calc_flux_tran (SRC/flux.f90:18-127) over one block, with the Wilke viscosity MCAVIS in place of
 * the reference's undefined MCAVIS_EA (SRC/trans.f90:148).  Qc: (n+2*lv) x nequ; Fplus added to. */
void fo_flux_tran(const fo_mech *m, int n, int lv, double dx_block, const double *Qc,
                  double *Fplus);

/* TRANLIB mixture-averaged properties (cgs). */
double fo_mcavis(const fo_mech *m, double T, const double *X);                  /* :681 */
double fo_mcacon(const fo_mech *m, double T, const double *X);                  /* :827 */
void fo_mcadif(const fo_mech *m, double p, double T, const double *X, double *D); /* :973 */

/* ---- auxiliary per-cell sources and open boundaries ------------------------------------------ */
#define FO_MAX_RAD 16
#define FO_MAX_IG 16
typedef struct {
    /* update_boundary (SRC/boundary.f90:12-67): 0 reflective wall, 1 inflow (left, lbound_inlet
     * :70-99) / outflow (right, rbound_outlet :102-149); inflow state as set_flow_boundary
     * derives it (SRC/flow_bound.f90:15-57), SI */
    int lbound, rbound;
    double T_in, u_in, rho_u_in, p_out;
    double x_in[64], y_in[64];
    /* radiation_heat_loss (SRC/radiation.f90:74-121): radiating species (0-based) and their six
     * Planck-mean absorption coefficients */
    int nrad;
    int rad_idx[FO_MAX_RAD];
    double pmac[FO_MAX_RAD][6];
    /* heat_transfer (SRC/mfr.f90:41-78): Nusselt number, diameter (m), wall-temperature profile */
    int wall_heat;
    double Nu, d_m, Tw_min, Tw_max, alpha, beta;
    /* add_igsource (1) / add_igsource_WZhang (2) (SRC/ignitionsource.f90:150-264), SI */
    int ig_type, n_ig;
    double E_j[FO_MAX_IG], r_m[FO_MAX_IG], t_s[FO_MAX_IG], t0_s[FO_MAX_IG], vol[FO_MAX_IG];
} fo_aux;

/* radiation_heat_loss: Qrad (W/m^3) at p (Pa), T, mole fractions X. */
double fo_radiation_heat_loss(const fo_aux *a, double p, double T, const double *X);
/* Tw (SRC/mfr.f90:63-78) and heat_transfer (:41-60): Qw (W/m^3) of a cell at x_m (m). */
double fo_wall_temperature(const fo_aux *a, double x);
double fo_wall_heat(const fo_mech *m, const fo_aux *a, double T, const double *X, double x_m);
/* add_igsource / add_igsource_WZhang: Eadd (W/m^3) at time t of a cell at x_m. */
double fo_ignition_power(const fo_aux *a, double t, double x_m);

/* ---- multi-block wall-bounded domain (reference storage: per block (ncell+2lv) x nequ) ------ */
typedef struct {
    const fo_mech *m;
    fo_opts o;
    int nblk, ncell, lv, nequ;
    double dx;
    double *Qc, *qf;     /* nblk * (ncell+2lv) * nequ each */
    double *Fplus;       /* nblk * (ncell+1) * nequ        */
    double *godunov;     /* nblk * (ncell+1) * 3           */
    double *Qc0, *k1, *k2, *k3, *k4;  /* interior work   */
    double *hmem;        /* nblk * ncell: the integrator's last step-size proposal (0: none) */
    int *cstats;         /* nblk * ncell * 4: steps, rejects, newton, failed of the last solve_source */
    fo_aux aux;          /* open boundaries and auxiliary sources (all off after create)          */
    double *xm;          /* nblk * ncell cell centres x_m (m)                                      */
    double *qrad, *qw;   /* nblk * ncell: Qrad and Qw of the last auxiliary-source call            */
    double time;         /* driver time, advanced by fo_step                                        */
    double *dxb;         /* nblk: each block's cell size (smr_dx)                                  */
} fo_domain;

/* Allocates and initialises from (T, p, u, X) per interior cell the way initialize_variables /
 * set_block_params / main_sim do (SRC/params.f90:305-580, SRC/block.f90:394-431,
 * SRC/main.f90:94-114). cells: nblk*ncell rows of (3+KK). */
fo_domain *fo_domain_create(const fo_mech *m, const fo_opts *o, int nblk, int ncell, double dx,
                            const double *cells);
/* The reference's stretched-mesh layout (&blocks: set_block_size, SRC/block.f90:91-392 on one
 * rank): nblk_comp blocks at dx_comp, (lev_smr_0 - 1) * num_smr_0 coarser blocks on the left and
 * (lev_smr_n - 1) * num_smr_n on the right, the spacing doubling per level away from the fine
 * region; cells covers all nblk_comp + ... blocks in order.  Coarse blocks reconstruct at first
 * order (SRC/flux.f90:240), block halos between unequal blocks are averaged / minmod-interpolated
 * (SRC/block.f90:499-563).  NULL for invalid levels. */
fo_domain *fo_domain_create_smr(const fo_mech *m, const fo_opts *o, int nblk_comp, int ncell,
                                double dx_comp, int lev_smr_0, int num_smr_0, int lev_smr_n,
                                int num_smr_n, const double *cells);
void fo_domain_free(fo_domain *d);
double fo_calc_dt(fo_domain *d, double cfl);   /* SRC/update.f90:13-59, single rank */
void fo_solve_fluid(fo_domain *d, double dt);  /* SRC/update.f90:62-97 -> timeintegration.f90 */
void fo_refresh(fo_domain *d);                 /* post-step BC/halo/QcToqf, SRC/main.f90:149-163 */

/* Inflow / outflow boundaries: flag_lbound 0 wall, 1 inflow at T_in (K), u_in (cm/s) and mole
 * fractions x_in[KK]; flag_rbound 0 wall, 1 outflow at p_out (atm) -- the inputs of
 * set_flow_boundary (SRC/flow_bound.f90:34-54).  Re-fills the ghost cells the way the driver does
 * at start-up.  Returns 0, or -1 for an invalid flag. */
int fo_domain_set_flow_boundary(fo_domain *d, int flag_lbound, int flag_rbound, double T_in,
                                double u_in_cm_s, const double *x_in, double p_out_atm);
/* Radiation (set_radiation_params, SRC/radiation.f90:21-71): n species (0-based) with six
 * coefficients each (row-major n x 6); n = 0 turns it off.  Returns 0 or -1 (n too large). */
int fo_domain_set_radiation(fo_domain *d, int n, const int *species, const double *pmac);
/* Wall heat of the micro flow reactor (set_wallT_boundary, SRC/mfr.f90:21-38): d in mm. */
void fo_domain_set_wall_heat(fo_domain *d, int on, double Nu, double d_mm, double Tw_min,
                             double Tw_max, double alpha, double beta);
/* Ignition sources (set_igsource / set_multi_igsource, SRC/ignitionsource.f90:40-147): type 1
 * Frendi, 2 Zhang, 0 off; energy mJ, radius um, duration and start us.  Returns 0 or -1. */
int fo_domain_set_ignition(fo_domain *d, int type, int n, const double *E_mJ, const double *r_um,
                           const double *t_us, const double *t0_us);
/* Cell centres (nblk * ncell, m); create sets the uniform block layout (SRC/block.f90:264-305). */
void fo_domain_set_xm(fo_domain *d, const double *xm);
/* The auxiliary sources of solve_source (SRC/update.f90:127-147) on the interior cells, in the
 * reference's order: radiation, wall heat, ignition, each adding dt * source to rho E and
 * recovering the primitives (SRC/source.f90:85-239).  time: the driver's time after the step. */
void fo_aux_sources(fo_domain *d, double dt, double time);

/* The output fields of set_eos_params + write_hdf5_b (SRC/block.f90:433-470,
 * SRC/write_hdf5.f90:138-160) per interior cell, blocks in order: rho, u, T, p, e_total, e_kinetic,
 * e_internal, HRR, HRR_dot, Qrad, Tw, qTw, x_m, dx, idx, X_1..X_KK; out is (15 + KK) rows of
 * nblk * ncell values.  hrr / hrr_dot (may be NULL: zeros) are the last source call's; idx0 is the
 * reference's index of the first cell. */
void fo_eos_fields(const fo_domain *d, const double *hrr, const double *hrr_dot, int idx0, double *out);
/* Restart from such fields (read_hdf5_restart + set_block_params, SRC/read_hdf5.f90:404-419,
 * SRC/block.f90:413-429): qf = (rho, u, T, CKXTY(X)), Qc = (rho, rho u, rho e_total, rho Y), then the
 * start-up ghost fill. */
void fo_domain_restart(fo_domain *d, const double *fields);

/* ---- CKLIB kinetics -------------------------------------------------------------------------- */
/* CKWYR (cklib.f:9107-9204): rho g/cm^3, T K, Y -> wdot mol/(cm^3 s).  qf_/qr_ (may be NULL):
 * per-reaction forward / reverse rates of progress as left in RCKWRK(NcI1/NcI2). */
void fo_ckwyr(const fo_mech *m, double rho, double T, const double *Y, double *wdot,
              double *qf_, double *qr_);
/* CKCDYR (cklib.f:1422-1541). */
void fo_ckcdyr(const fo_mech *m, double rho, double T, const double *Y, double *cdot,
               double *ddot);

/* ---- chemistry source (own stiff integrator; see fcfd_oracle.c) -------------------------------- */
typedef struct { int steps, rejects, newton, jacobians, failed; } fo_chem_stats;
/* f = d(Y_1..Y_KK, T)/dt at fixed density rho (g/cm^3). */
void fo_chem_rhs(const fo_mech *m, double rho, const double *y, double *f);
/* J = df/dy (n x n row-major, n = KK+1). */
void fo_chem_jac(const fo_mech *m, double rho, const double *y, double *J);
/* macks(dt, nspe+1, yiT, rho) (SRC/source.f90:58): yT = (Y, T) in/out, rho kg/m^3.
 * Returns 0, or -1 when the step budget ran out (state = last accepted). */
int fo_chem_step(const fo_mech *m, double dt, double rho, double *yT, double rtol, double atol,
                 fo_chem_stats *st);
/* reaction_macks (SRC/source.f90:17-84) over a domain's interior cells. */
void fo_solve_source(fo_domain *d, double dt, double rtol, double atol, double *hrr,
                     double *hrr_dot, int *nfail);

/* exp (0), log (1), log10 (2), pow (3), 10^x (4), tanh (5), expm1 (6) of the build's math on n
 * arguments */
void fo_math_eval(int which, int n, const double *x, const double *y, double *out);

/* One driver time step (SRC/main.f90:130-163): Lie (strang = 0, the reference's order) or Strang
 * splitting of the chemistry, then the auxiliary sources that are on (with the full dt, after the
 * chemistry); advances d->time by dt and returns dt. */
double fo_step(fo_domain *d, double cfl, int react, int strang, double rtol, double atol, double *hrr,
               double *hrr_dot, int *nfail);
/* OpenMP threads of the parallel loops (1 without OpenMP). */
int fo_threads(void);

#ifdef __cplusplus
}
#endif
#endif
