! sod_driver.f90 -- test program: the reference driver loop (SRC/main.f90:130-163, without output)
! through the Fortran binding of libfcfd_hip.so: calc_dt, solve_fluid, solve_source (when
! do_reaction = 1, with the heat-release diagnostics), refresh.
!   sod_driver <mech_dir> <in.bin> <out.bin>
! in.bin (stream): int32 nblk, ncell, nsteps, select_time, select_riemann, flag_acu, flag_muscl,
!   flag_weno, cfl_flag, do_reaction; float64 dx, cfl, dt_fixed; float64 cells(3+KK, nblk*ncell)
!   (T, p, u, X).
! out.bin (stream): float64 dts(nsteps); float64 Qc and qf of every block (nequ x (ncell+2 lv));
!   with do_reaction: float64 HRR, HRR_dot of the last step (nblk*ncell each) and int32 counters
!   (4 x nblk*ncell) of the last chemistry call.
program sod_driver
    use iso_c_binding
    use fcfd
    implicit none
    character(len=512) :: mdir, fin, fout
    type(c_ptr) :: mech, ctx
    type(fcfd_opts) :: o
    integer(c_int) :: ints(10), kk, ii, nblk, ncell, nsteps, lv, nequ, ib, s, rc
    real(c_double) :: dbl(3), dt, time
    real(c_double), allocatable, target :: cells(:, :), Qc(:, :), qf(:, :), dts(:), hrr(:), hrr_dot(:)
    integer(c_int), allocatable :: stats(:)
    call get_command_argument(1, mdir)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    call fcfd_check(fcfd_mech_load(trim(mdir)//'/chem.bin'//c_null_char, trim(mdir)//'/tran.bin'//c_null_char, mech), &
                    'fcfd_mech_load')
    call fcfd_check(fcfd_mech_dims(mech, kk, ii), 'fcfd_mech_dims')
    nequ = 3 + kk
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) ints, dbl
    nblk = ints(1); ncell = ints(2); nsteps = ints(3)
    o%select_time = ints(4); o%select_riemann = ints(5); o%flag_acu = ints(6)
    o%flag_muscl = ints(7); o%flag_weno = ints(8); o%do_reaction = ints(10)
    allocate(cells(nequ, nblk * ncell))
    read(10) cells
    close(10)
    call fcfd_check(fcfd_create(mech, o, nblk, ncell, dbl(1), ctx), 'fcfd_create')
    lv = fcfd_lvir(ctx)
    call fcfd_check(fcfd_init_cells(ctx, cells), 'fcfd_init_cells')
    allocate(dts(nsteps), hrr(nblk * ncell), hrr_dot(nblk * ncell), stats(4 * nblk * ncell))
    time = 0.0d0
    do s = 1, nsteps
        if (ints(9) /= 0) then
            call fcfd_check(fcfd_calc_dt(ctx, dbl(2), dt), 'calc_dt')
        else
            dt = dbl(3)
        end if
        dts(s) = dt
        call fcfd_check(fcfd_solve_fluid(ctx, dt), 'solve_fluid')
        if (o%do_reaction == 1) then
            call fcfd_check(fcfd_solve_source(ctx, dt, time, c_loc(hrr), c_loc(hrr_dot)), 'solve_source')
        end if
        call fcfd_check(fcfd_refresh(ctx), 'refresh')
        time = time + dt
    end do
    allocate(Qc(nequ, ncell + 2 * lv), qf(nequ, ncell + 2 * lv))
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) dts
    do ib = 0, nblk - 1
        rc = fcfd_download(ctx, ib, Qc, qf)
        call fcfd_check(rc, 'fcfd_download')
        write(11) Qc, qf
    end do
    if (o%do_reaction == 1) then
        call fcfd_check(fcfd_chem_stats(ctx, stats), 'fcfd_chem_stats')
        write(11) hrr, hrr_dot, stats
    end if
    close(11)
    call fcfd_destroy(ctx)
    call fcfd_mech_free(mech)
end program sod_driver
