! fcfd.f90 -- Fortran 2003 binding (ISO_C_BINDING) of libfcfd_hip.so, for the fortranCFD driver.
!
! The reference driver's call sites keep their shape:
!   calc_dt      (SRC/main.f90:134)  -> fcfd_calc_dt(ctx, cfl, dt)
!   solve_fluid  (SRC/main.f90:142)  -> fcfd_solve_fluid(ctx, dt)
!   solve_source (SRC/main.f90:145)  -> fcfd_solve_source(ctx, dt, time, c_loc(hrr), c_loc(hrr_dot))
!                                       (c_null_ptr for either diagnostic skips it)
!   refresh      (SRC/main.f90:149-163) -> fcfd_refresh(ctx)
! with the block state moved to the device once (fcfd_upload from b_obj(ib)%Qc / %qf, whose storage
! order -- equation index fastest, ghosts included -- the library takes as is) and back for output
! (fcfd_download).  Every function returns 0 or a negative error code; fcfd_check turns a failure
! into `error stop` with the library's message, the reference's convention for fatal errors.
module fcfd
    use iso_c_binding
    implicit none
    private
    public :: fcfd_opts, fcfd_check, fcfd_error_message
    public :: fcfd_mech_load, fcfd_mech_free, fcfd_mech_dims
    public :: fcfd_create, fcfd_destroy, fcfd_lvir, fcfd_upload, fcfd_download, fcfd_init_cells
    public :: fcfd_calc_dt, fcfd_solve_fluid, fcfd_solve_source, fcfd_refresh, fcfd_advance
    public :: fcfd_flux_advection_host, fcfd_flux_tran_host, fcfd_chem_step_host, fcfd_rates_host
    public :: fcfd_chem_stats, fcfd_source_accepted, fcfd_set_chem_workgroups
    public :: fcfd_sur_create, fcfd_sur_free, fcfd_sur_nparams, fcfd_sur_set_params, fcfd_sur_get_params
    public :: fcfd_sur_set_norm, fcfd_sur_set_thresholds, fcfd_set_surrogate
    public :: fcfd_create_smr, fcfd_nblocks, fcfd_block_dx, fcfd_get_cell_centres, fcfd_set_cell_centres
    public :: fcfd_set_flow_boundary, fcfd_set_radiation, fcfd_set_wall_heat, fcfd_set_ignition
    public :: fcfd_get_aux, fcfd_set_step, fcfd_get_hrr, fcfd_eos_fields, fcfd_init_restart

    ! Same fields and codes as the reference's &params switches (SRC/params.f90:106-111).
    type, bind(C) :: fcfd_opts
        integer(c_int) :: do_hydro = 1, do_tran = 0, do_reaction = 0
        integer(c_int) :: select_time = 1, select_riemann = 2, flag_acu = 2
        integer(c_int) :: flag_muscl = 0, flag_weno = 2, is_spherical = 0
        integer(c_int) :: flag_lbound = 0, flag_rbound = 0
        real(c_double) :: atol = 1.0d-13, rtol = 1.0d-5
    end type fcfd_opts

    interface
        integer(c_int) function fcfd_mech_load(chem_bin, tran_bin, mech) bind(C, name="fcfd_mech_load")
            import :: c_int, c_char, c_ptr
            character(kind=c_char), dimension(*), intent(in) :: chem_bin, tran_bin
            type(c_ptr), intent(out) :: mech
        end function
        subroutine fcfd_mech_free(mech) bind(C, name="fcfd_mech_free")
            import :: c_ptr
            type(c_ptr), value :: mech
        end subroutine
        integer(c_int) function fcfd_mech_dims(mech, kk, ii) bind(C, name="fcfd_mech_dims")
            import :: c_int, c_ptr
            type(c_ptr), value :: mech
            integer(c_int), intent(out) :: kk, ii
        end function
        integer(c_int) function fcfd_create(mech, opts, nblk, ncell, dx, ctx) bind(C, name="fcfd_create")
            import :: c_int, c_double, c_ptr, fcfd_opts
            type(c_ptr), value :: mech
            type(fcfd_opts), intent(in) :: opts
            integer(c_int), value :: nblk, ncell
            real(c_double), value :: dx
            type(c_ptr), intent(out) :: ctx
        end function
        ! &blocks layout with refinement levels (set_block_size, SRC/block.f90:91-392)
        integer(c_int) function fcfd_create_smr(mech, opts, nblk_comp, ncell, dx_comp, lev_smr_0, num_smr_0, &
                                                lev_smr_n, num_smr_n, ctx) bind(C, name="fcfd_create_smr")
            import :: c_int, c_double, c_ptr, fcfd_opts
            type(c_ptr), value :: mech
            type(fcfd_opts), intent(in) :: opts
            integer(c_int), value :: nblk_comp, ncell, lev_smr_0, num_smr_0, lev_smr_n, num_smr_n
            real(c_double), value :: dx_comp
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function fcfd_nblocks(ctx) bind(C, name="fcfd_nblocks")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fcfd_block_dx(ctx, dx_block) bind(C, name="fcfd_block_dx")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: dx_block(*)
        end function
        integer(c_int) function fcfd_get_cell_centres(ctx, xm) bind(C, name="fcfd_get_cell_centres")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: xm(*)
        end function
        integer(c_int) function fcfd_set_cell_centres(ctx, xm) bind(C, name="fcfd_set_cell_centres")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: xm(*)
        end function
        ! set_flow_boundary's inputs (SRC/flow_bound.f90:15-57); x_in: c_null_ptr without an inflow
        integer(c_int) function fcfd_set_flow_boundary(ctx, T_in, u_in_cm_s, x_in, p_out_atm) &
                                                       bind(C, name="fcfd_set_flow_boundary")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), value :: T_in, u_in_cm_s, p_out_atm
            type(c_ptr), value :: x_in
        end function
        integer(c_int) function fcfd_set_radiation(ctx, n, species, pmac) bind(C, name="fcfd_set_radiation")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            integer(c_int), intent(in) :: species(*)     ! 0-based species indices
            real(c_double), intent(in) :: pmac(*)        ! 6 coefficients per species
        end function
        integer(c_int) function fcfd_set_wall_heat(ctx, on, Nu, d_mm, Tw_min, Tw_max, alpha, beta) &
                                                   bind(C, name="fcfd_set_wall_heat")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
            real(c_double), value :: Nu, d_mm, Tw_min, Tw_max, alpha, beta
        end function
        integer(c_int) function fcfd_set_ignition(ctx, kind, n, E_mJ, r_um, t_us, t0_us) &
                                                  bind(C, name="fcfd_set_ignition")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: kind, n
            real(c_double), intent(in) :: E_mJ(*), r_um(*), t_us(*), t0_us(*)
        end function
        ! eos_obj(:)%Qrad and %qTw of the last auxiliary-source call (c_null_ptr skips either)
        integer(c_int) function fcfd_get_aux(ctx, qrad, qw) bind(C, name="fcfd_get_aux")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            type(c_ptr), value :: qrad, qw
        end function
        integer(c_int) function fcfd_set_step(ctx, strang, hrr_each_step) bind(C, name="fcfd_set_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: strang, hrr_each_step
        end function
        integer(c_int) function fcfd_get_hrr(ctx, hrr, hrr_dot) bind(C, name="fcfd_get_hrr")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            type(c_ptr), value :: hrr, hrr_dot
        end function
        ! write_hdf5_b's rows (15 + KK of nblocks*ncell) and the restart of set_block_params
        integer(c_int) function fcfd_eos_fields(ctx, out) bind(C, name="fcfd_eos_fields")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function fcfd_init_restart(ctx, fields) bind(C, name="fcfd_init_restart")
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: fields(*)
        end function
        subroutine fcfd_destroy(ctx) bind(C, name="fcfd_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        integer(c_int) function fcfd_lvir(ctx) bind(C, name="fcfd_lvir")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        ! ib is 0-based; Qc / qf are b_obj(ib+1)%Qc(:, ista-lvir:iend+lvir) etc.
        integer(c_int) function fcfd_upload(ctx, ib, Qc, qf) bind(C, name="fcfd_upload")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ib
            real(c_double), intent(in) :: Qc(*), qf(*)
        end function
        integer(c_int) function fcfd_download(ctx, ib, Qc, qf) bind(C, name="fcfd_download")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ib
            real(c_double), intent(out) :: Qc(*), qf(*)
        end function
        integer(c_int) function fcfd_init_cells(ctx, cells) bind(C, name="fcfd_init_cells")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: cells(*)
        end function
        integer(c_int) function fcfd_calc_dt(ctx, cfl, dt) bind(C, name="fcfd_calc_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: cfl
            real(c_double), intent(out) :: dt
        end function
        integer(c_int) function fcfd_solve_fluid(ctx, dt) bind(C, name="fcfd_solve_fluid")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        ! hrr / hrr_dot: c_loc of one value per interior cell (eos_obj%HRR / %HRR_dot, blocks in
        ! order), or c_null_ptr
        integer(c_int) function fcfd_solve_source(ctx, dt, time, hrr, hrr_dot) bind(C, name="fcfd_solve_source")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt, time
            type(c_ptr), value :: hrr, hrr_dot
        end function
        ! steps, rejects, Newton iterations, failed flag of the last chemistry call, per interior cell
        integer(c_int) function fcfd_chem_stats(ctx, per_cell) bind(C, name="fcfd_chem_stats")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: per_cell(*)
        end function
        integer(c_int) function fcfd_source_accepted(ctx, per_cell, n_fallback) bind(C, name="fcfd_source_accepted")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: per_cell(*), n_fallback
        end function
        integer(c_int) function fcfd_set_chem_workgroups(ctx, n) bind(C, name="fcfd_set_chem_workgroups")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
        end function
        ! the chemistry surrogate: create, load parameters / normalisation, attach to a context
        integer(c_int) function fcfd_sur_create(mech, sur) bind(C, name="fcfd_sur_create")
            import :: c_int, c_ptr
            type(c_ptr), value :: mech
            type(c_ptr), intent(out) :: sur
        end function
        subroutine fcfd_sur_free(sur) bind(C, name="fcfd_sur_free")
            import :: c_ptr
            type(c_ptr), value :: sur
        end subroutine
        integer(c_int) function fcfd_sur_nparams() bind(C, name="fcfd_sur_nparams")
            import :: c_int
        end function
        integer(c_int) function fcfd_sur_set_params(sur, params) bind(C, name="fcfd_sur_set_params")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: sur
            real(c_float), intent(in) :: params(*)
        end function
        integer(c_int) function fcfd_sur_get_params(sur, params) bind(C, name="fcfd_sur_get_params")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: sur
            real(c_float), intent(out) :: params(*)
        end function
        ! norm(64, 6): input mean, input std, training-range low, high, output mean, output std
        integer(c_int) function fcfd_sur_set_norm(sur, norm) bind(C, name="fcfd_sur_set_norm")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sur
            real(c_double), intent(in) :: norm(*)
        end function
        integer(c_int) function fcfd_sur_set_thresholds(sur, resid_tol, t_tol, verify_tol) &
                bind(C, name="fcfd_sur_set_thresholds")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sur
            real(c_double), value :: resid_tol, t_tol, verify_tol
        end function
        ! sur = c_null_ptr: back to the FP64 integrator on every cell
        integer(c_int) function fcfd_set_surrogate(ctx, sur) bind(C, name="fcfd_set_surrogate")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, sur
        end function
        integer(c_int) function fcfd_refresh(ctx) bind(C, name="fcfd_refresh")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function fcfd_advance(ctx, cfl, nsteps, time) bind(C, name="fcfd_advance")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: cfl
            integer(c_int), value :: nsteps
            real(c_double), intent(inout) :: time
        end function
        integer(c_int) function fcfd_flux_advection_host(mech, opts, n, lv, dx, qf, Fplus, godunov) &
                bind(C, name="fcfd_flux_advection_host")
            import :: c_int, c_ptr, c_double, fcfd_opts
            type(c_ptr), value :: mech
            type(fcfd_opts), intent(in) :: opts
            integer(c_int), value :: n, lv
            real(c_double), value :: dx
            real(c_double), intent(in) :: qf(*)
            real(c_double), intent(out) :: Fplus(*), godunov(*)
        end function
        integer(c_int) function fcfd_flux_tran_host(mech, n, lv, dx, Qc, Fplus) bind(C, name="fcfd_flux_tran_host")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: mech
            integer(c_int), value :: n, lv
            real(c_double), value :: dx
            real(c_double), intent(in) :: Qc(*)
            real(c_double), intent(out) :: Fplus(*)
        end function
        ! macks(dt, nspe+1, yiT, rho) (SRC/source.f90:58) on n cells
        integer(c_int) function fcfd_chem_step_host(mech, opts, n, dt, rho, yT, stats) bind(C, name="fcfd_chem_step_host")
            import :: c_int, c_ptr, c_double, fcfd_opts
            type(c_ptr), value :: mech
            type(fcfd_opts), intent(in) :: opts
            integer(c_int), value :: n
            real(c_double), intent(in) :: dt(*), rho(*)
            real(c_double), intent(inout) :: yT(*)
            integer(c_int), intent(out) :: stats(*)
        end function
        ! CKWYR / CKCDYR on n states: rho [g/cm^3], T [K], Y(KK, n) -> wdot, cdot, ddot (KK, n),
        ! qf, qr (II, n)
        integer(c_int) function fcfd_rates_host(mech, n, rho, T, Y, wdot, cdot, ddot, qf, qr) &
                bind(C, name="fcfd_rates_host")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: mech
            integer(c_int), value :: n
            real(c_double), intent(in) :: rho(*), T(*), Y(*)
            real(c_double), intent(out) :: wdot(*), cdot(*), ddot(*), qf(*), qr(*)
        end function
        type(c_ptr) function fcfd_last_error_c() bind(C, name="fcfd_last_error")
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
        end function
    end interface

contains

    function fcfd_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = fcfd_last_error_c()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function fcfd_error_message

    subroutine fcfd_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= 0) then
            write(*, '(a, a, i0, a, a)') where, ': fcfd error ', rc, ': ', fcfd_error_message()
            error stop 1
        end if
    end subroutine fcfd_check

end module fcfd
