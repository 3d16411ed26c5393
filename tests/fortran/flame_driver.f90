! flame_driver.f90 -- test program: the spherical-flame configuration through the Fortran binding:
! a stretched mesh (fcfd_create_smr), spherical geometry, outflow on the right, radiation and the
! ignition source, and the reference driver loop with the time advanced before the fluid step
! (SRC/main.f90:136-146) so that the ignition window reads the time after the step.
!   flame_driver <mech_dir> <in.bin> <out.bin>
! in.bin (stream): int32 nblk_comp, ncell, lev_smr_0, num_smr_0, lev_smr_n, num_smr_n, nsteps,
!   do_reaction, nrad; float64 dx_comp, cfl, p_out_atm, E_mJ, r_um, t_us, t0_us;
!   int32 rad species(nrad); float64 pmac(6, nrad); float64 cells(3+KK, nblocks*ncell) (T, p, u, X).
! out.bin (stream): int32 nblocks; float64 dts(nsteps); float64 fcfd_eos_fields rows
!   (nblocks*ncell, 15+KK).
program flame_driver
    use iso_c_binding
    use fcfd
    implicit none
    character(len=512) :: mdir, fin, fout
    type(c_ptr) :: mech, ctx
    type(fcfd_opts) :: o
    integer(c_int) :: ints(9), kk, ii, nequ, nblk, nsteps, s, ncells
    real(c_double) :: dbl(7), dt, time
    real(c_double) :: E(1), r(1), tig(1), t0(1)
    integer(c_int), allocatable :: rad(:)
    real(c_double), allocatable, target :: cells(:, :), pmac(:, :), dts(:), fields(:, :)
    call get_command_argument(1, mdir)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    call fcfd_check(fcfd_mech_load(trim(mdir)//'/chem.bin'//c_null_char, trim(mdir)//'/tran.bin'//c_null_char, mech), &
                    'fcfd_mech_load')
    call fcfd_check(fcfd_mech_dims(mech, kk, ii), 'fcfd_mech_dims')
    nequ = 3 + kk
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) ints, dbl
    allocate(rad(ints(9)), pmac(6, ints(9)))
    read(10) rad, pmac
    o%is_spherical = 1
    o%flag_rbound = 1
    o%do_reaction = ints(8)
    call fcfd_check(fcfd_create_smr(mech, o, ints(1), ints(2), dbl(1), ints(3), ints(4), ints(5), ints(6), ctx), &
                    'fcfd_create_smr')
    nblk = fcfd_nblocks(ctx)
    ncells = nblk * ints(2)
    nsteps = ints(7)
    allocate(cells(nequ, ncells))
    read(10) cells
    close(10)
    call fcfd_check(fcfd_set_flow_boundary(ctx, 300.0d0, 0.0d0, c_null_ptr, dbl(3)), 'fcfd_set_flow_boundary')
    call fcfd_check(fcfd_set_radiation(ctx, ints(9), rad, pmac), 'fcfd_set_radiation')
    E(1) = dbl(4); r(1) = dbl(5); tig(1) = dbl(6); t0(1) = dbl(7)
    call fcfd_check(fcfd_set_ignition(ctx, 1, 1, E, r, tig, t0), 'fcfd_set_ignition')
    call fcfd_check(fcfd_init_cells(ctx, cells), 'fcfd_init_cells')
    allocate(dts(nsteps), fields(ncells, 15 + kk))
    time = 0.0d0
    do s = 1, nsteps
        call fcfd_check(fcfd_calc_dt(ctx, dbl(2), dt), 'calc_dt')
        dts(s) = dt
        time = time + dt
        call fcfd_check(fcfd_solve_fluid(ctx, dt), 'solve_fluid')
        call fcfd_check(fcfd_solve_source(ctx, dt, time, c_null_ptr, c_null_ptr), 'solve_source')
        call fcfd_check(fcfd_refresh(ctx), 'refresh')
    end do
    call fcfd_check(fcfd_eos_fields(ctx, fields), 'fcfd_eos_fields')
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) nblk, dts, fields
    close(11)
    call fcfd_destroy(ctx)
    call fcfd_mech_free(mech)
end program flame_driver
