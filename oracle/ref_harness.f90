! Test infrastructure only (checker side): never part of the product.
!
! Driver around the reference's own Fortran modules (compiled in place from the reference tree by
! oracle/Makefile).  Each sub-command reads one binary input file and writes one binary output file
! so that tests/golden/make_golden.py can turn reference results into committed fixtures, and so
! bench.py can time the reference CPU path.
!
! File format (stream, little endian): int32 ni, int32 nd, int32 iv(ni), float64 dv(nd).
!
! usage: ref_harness <cmd> <chem_dir> <in.bin> <out.bin>
!   cmd = mech | thermo | eos | rates | hrr | trans | recon | flux | run
!
! set_thermo (SRC/thermo.f90:25-67) calls CKLEN/CKINIT with one argument too few
! (SRC/thermo.f90:47,52 vs cklib.f:4920,3972); h_init below performs the same initialisation with the
! library's real argument lists and fills the same module variables.
program ref_harness
    use thermo
    use trans, only: set_trans, viscosity, thermal_conductivity, set_diffusivity
    use variables
    use weno
    use muscl
    use eos
    use module_params, only: flag_acu, flag_weno, flag_muscl, dx_comp, is_spherical, do_hydro, &
                             do_tran, select_riemann, select_time, flag_lbound, flag_rbound
    use block_t, only: block_st, b_obj, bsta, bend, block_all_0, block_all_n, interpolation_block
    use module_boundary, only: update_boundary
    use flux, only: calc_flux_advection
    use timeintegration, only: explicit_euler, TVDRK3, RK4
    use parallel, only: nprocs, myrank
    implicit none

    character(len=256) :: cmd, chem_dir, fin, fout
    integer, allocatable :: iv(:)
    real(8), allocatable :: dv(:)
    integer :: KK, II

    call get_command_argument(1, cmd)
    call get_command_argument(2, chem_dir)
    call get_command_argument(3, fin)
    call get_command_argument(4, fout)

    nprocs = 1
    myrank = 0
    call h_init(chem_dir, KK, II)

    select case (trim(cmd))
    case ('mech')
        call do_mech()
    case default
        call read_input(fin, iv, dv)
        select case (trim(cmd))
        case ('thermo');  call do_thermo()
        case ('eos');     call do_eos()
        case ('rates');   call do_rates()
        case ('hrr');     call do_hrr()
        case ('trans');   call do_trans()
        case ('recon');   call do_recon()
        case ('flux');    call do_flux()
        case ('run');     call do_run()
        case default
            write(0,*) 'ref_harness: unknown command ', trim(cmd)
            stop 2
        end select
    end select

contains

    subroutine h_init(cdir, kk_out, ii_out)
        character(len=*), intent(in) :: cdir
        integer, intent(out) :: kk_out, ii_out
        integer :: leniwk, lenrwk, lencwk, mm, kk_, ii_, nfit, iflag
        logical :: kerr
        character(len=128) :: cdir128
        integer, parameter :: linc = 25, lout = 99

        open(lout, file='/dev/null')
        open(linc, file=trim(cdir)//'/chem.bin', form='UNFORMATTED', status='OLD')
        call CKLEN(linc, lout, leniwk, lenrwk, lencwk, iflag)
        if (iflag /= 0) stop 'CKLEN failed'
        allocate(ICKWRK(leniwk), RCKWRK(lenrwk), CCKWRK(lencwk))
        ICKWRK = 0
        RCKWRK = 0.d0
        call CKINIT(leniwk, lenrwk, lencwk, linc, lout, ICKWRK, RCKWRK, CCKWRK, iflag)
        if (iflag /= 0) stop 'CKINIT failed'
        call CKINDX(ICKWRK, RCKWRK, mm, kk_, ii_, nfit)
        n_species = kk_
        allocate(species_name(kk_), molar_weight(kk_), std_heat_formation(kk_))
        call CKSYMS(CCKWRK, lout, species_name, kerr)
        call CKWT(ICKWRK, RCKWRK, molar_weight)
        call CKHMS(298.15d0, ICKWRK, RCKWRK, std_heat_formation)
        call CKRP(ICKWRK, RCKWRK, RU, RUC, PA)
        universal_gas_constant = RU * 1.d-4
        close(linc)
        cdir128 = cdir
        call set_trans(cdir128)
        kk_out = kk_
        ii_out = ii_
    end subroutine h_init

    subroutine read_input(f, ivv, dvv)
        character(len=*), intent(in) :: f
        integer, allocatable, intent(out) :: ivv(:)
        real(8), allocatable, intent(out) :: dvv(:)
        integer :: u, hdr(2)
        open(newunit=u, file=trim(f), access='stream', form='unformatted', status='old')
        read(u) hdr
        allocate(ivv(hdr(1)), dvv(hdr(2)))
        if (hdr(1) > 0) read(u) ivv
        if (hdr(2) > 0) read(u) dvv
        close(u)
    end subroutine read_input

    subroutine write_output(f, ivv, dvv)
        character(len=*), intent(in) :: f
        integer, intent(in) :: ivv(:)
        real(8), intent(in) :: dvv(:)
        integer :: u
        open(newunit=u, file=trim(f), access='stream', form='unformatted', status='replace')
        write(u) size(ivv), size(dvv)
        if (size(ivv) > 0) write(u) ivv
        if (size(dvv) > 0) write(u) dvv
        close(u)
    end subroutine write_output

    ! CKINIT / MCINIT work arrays and pointer blocks, for checking our own link-file readers.
    subroutine do_mech()
        integer :: ck(85), mci(32), u, k
        real(8) :: mach(3), mcd(3)
        integer, allocatable :: io(:)
        real(8), allocatable :: do_(:)
        use_trans: block
            use trans, only: IMCWRK, RMCWRK
            call RHCKPT(ck)
            call RHMCPT(mcd, mci)
            call RHMACH(mach)
            io = [ck, size(ICKWRK), size(RCKWRK), ICKWRK, mci, size(IMCWRK), size(RMCWRK), IMCWRK]
            do_ = [RCKWRK, mach, mcd, RMCWRK, molar_weight, std_heat_formation]
        end block use_trans
        call write_output(fout, io, do_)
        open(newunit=u, file=trim(fout)//'.names', status='replace')
        do k = 1, KK
            write(u, '(a)') trim(species_name(k))
        end do
        close(u)
    end subroutine do_mech

    ! Thermo wrappers (SRC/thermo.f90) and CKLIB species arrays on (T, rho, Y) rows.
    subroutine do_thermo()
        integer :: n, r, off, w
        real(8) :: T, rho, p, e, cpv, cvv, a, Tn
        real(8) :: Y(KK), X(KK), Xa(KK), Ya(KK), hms(KK), smh(KK), cpml(KK), uml(KK)
        real(8), allocatable :: o(:)
        n = iv(1)
        w = 6 + 7*KK
        allocate(o(n*w))
        do r = 1, n
            off = (r-1)*(2+KK)
            T = dv(off+1)
            rho = dv(off+2)
            Y = dv(off+3:off+2+KK)
            call Yi2Xi(KK, Y, X)
            Xa = X
            call adjust_sum(KK, Xa)
            Ya = Y
            call adjust_sum(KK, Ya)
            p = pressure(rho, T, KK, Xa)
            e = specific_energy(rho, p, T, KK, Xa)
            cpv = Cp(rho, T, KK, Xa)
            cvv = Cv(rho, T, KK, Xa)
            a = sonic_speed(rho, p, T, KK, Xa)
            Tn = temperature_con(e, KK, Xa)
            call CKHMS(T, ICKWRK, RCKWRK, hms)
            call CKSMH(T, ICKWRK, RCKWRK, smh)
            call CKCPML(T, ICKWRK, RCKWRK, cpml)
            call CKUML(T, ICKWRK, RCKWRK, uml)
            o((r-1)*w+1:r*w) = [X, Xa, Ya, p, e, cpv, cvv, a, Tn, hms, smh, cpml, uml]
        end do
        call write_output(fout, [n], o)
    end subroutine do_thermo

    ! qfToQc / QcToqf (SRC/eos.f90) on n primitive rows and n conserved rows.
    subroutine do_eos()
        integer :: n, r, nequ
        real(8), allocatable :: o(:)
        n = iv(1)
        nequ = 3 + KK
        allocate(o(2*n*nequ))
        do r = 1, n
            o((r-1)*nequ+1:r*nequ) = qfToQc(dv((r-1)*nequ+1:r*nequ), KK, nequ)
        end do
        do r = 1, n
            o(n*nequ+(r-1)*nequ+1:n*nequ+r*nequ) = &
                QcToqf(dv(n*nequ+(r-1)*nequ+1:n*nequ+r*nequ), KK, nequ)
        end do
        call write_output(fout, [n], o)
    end subroutine do_eos

    ! CKWYR / CKCDYR rates plus per-reaction rates of progress and the omega-dot HRR.
    subroutine do_rates()
        integer :: n, r, off, w, ck(85), nci1, nci2, nckf, nckr
        real(8) :: rho, T, hrrd
        real(8) :: Y(KK), wdot(KK), cdot(KK), ddot(KK)
        real(8) :: q_f(II), q_r(II), k_f(II), k_r(II)
        real(8), allocatable :: o(:)
        call RHCKPT(ck)
        nckf = ck(74); nckr = ck(75); nci1 = ck(82); nci2 = ck(83)
        n = iv(1)
        w = 3*KK + 4*II + 1
        allocate(o(n*w))
        do r = 1, n
            off = (r-1)*(2+KK)
            rho = dv(off+1)
            T = dv(off+2)
            Y = dv(off+3:off+2+KK)
            call CKWYR(rho*1.d-3, T, Y, ICKWRK, RCKWRK, wdot)
            q_f = RCKWRK(nci1:nci1+II-1)
            q_r = RCKWRK(nci2:nci2+II-1)
            k_f = RCKWRK(nckf:nckf+II-1)
            k_r = RCKWRK(nckr:nckr+II-1)
            call CKCDYR(rho*1.d-3, T, Y, ICKWRK, RCKWRK, cdot, ddot)
            hrrd = heat_release_rate_omegadot(KK, Y, rho, T)
            o((r-1)*w+1:r*w) = [wdot, cdot, ddot, q_f, q_r, k_f, k_r, hrrd]
        end do
        call write_output(fout, [n], o)
    end subroutine do_rates

    ! heat_release_rate (SRC/thermo.f90:83-103) on rows (rho, dt, Y0, Y).
    subroutine do_hrr()
        integer :: n, r, off
        real(8), allocatable :: o(:)
        n = iv(1)
        allocate(o(n))
        do r = 1, n
            off = (r-1)*(2+2*KK)
            o(r) = heat_release_rate(KK, dv(off+3:off+2+KK), dv(off+3+KK:off+2+2*KK), &
                                     dv(off+1), dv(off+2))
        end do
        call write_output(fout, [n], o)
    end subroutine do_hrr

    ! TRANLIB mixture-averaged properties through SRC/trans.f90 on rows (T, p, X).
    subroutine do_trans()
        integer :: n, r, off, w
        real(8) :: T, p, mu, kap
        real(8) :: X(KK), dkm(KK)
        real(8), allocatable :: o(:)
        n = iv(1)
        w = 2 + KK
        allocate(o(n*w))
        do r = 1, n
            off = (r-1)*(2+KK)
            T = dv(off+1)
            p = dv(off+2)
            X = dv(off+3:off+2+KK)
            mu = viscosity(0.d0, p, T, KK, X)
            kap = thermal_conductivity(0.d0, p, T, KK, X)
            call set_diffusivity(0.d0, p, T, KK, X, dkm)
            o((r-1)*w+1:r*w) = [mu, kap, dkm]
        end do
        call write_output(fout, [n], o)
    end subroutine do_trans

    ! Reconstruction kernels (SRC/weno.f90, SRC/muscl.f90) on 8-value stencils.
    subroutine do_recon()
        integer :: n, r, f
        real(8) :: s(8), l, rr
        real(8), allocatable :: o(:)
        n = iv(1)
        allocate(o(n*18))
        do r = 1, n
            s = dv((r-1)*8+1:r*8)
            call weno3z(s(1:4), l, rr);      o((r-1)*18+1) = l;  o((r-1)*18+2) = rr
            call weno5(s(1:6), l, rr, 2);    o((r-1)*18+3) = l;  o((r-1)*18+4) = rr
            call weno5(s(1:6), l, rr, 3);    o((r-1)*18+5) = l;  o((r-1)*18+6) = rr
            call weno7z(s, l, rr);           o((r-1)*18+7) = l;  o((r-1)*18+8) = rr
            do f = 1, 4
                call muscl3(s(1:4), l, rr, f)
                o((r-1)*18+7+2*f) = l
                o((r-1)*18+8+2*f) = rr
            end do
        end do
        call write_output(fout, [n], o)
    end subroutine do_recon

    subroutine set_flags(acu, wen, mus, riem)
        integer, intent(in) :: acu, wen, mus, riem
        flag_acu = acu
        flag_weno = wen
        flag_muscl = mus
        select_riemann = riem
        is_spherical = 0
        do_hydro = 1
        do_tran = 0
        flag_lbound = 0
        flag_rbound = 0
    end subroutine set_flags

    ! calc_flux_advection (SRC/flux.f90:130-373) on one block whose primitive state (ghosts
    ! included) is given.  iv = [n, lvir, flag_acu, flag_weno, flag_muscl, select_riemann]
    subroutine do_flux()
        integer :: n, lv, nequ
        type(block_st), allocatable :: bo(:)
        n = iv(1)
        lv = iv(2)
        nequ = 3 + KK
        call set_flags(iv(3), iv(4), iv(5), iv(6))
        dx_comp = dv(1)
        allocate(bo(1:1))
        bo(1)%smr_level = 1
        bo(1)%smr_dx = dv(1)
        bo(1)%ista_b = 1
        bo(1)%iend_b = n
        allocate(bo(1)%qf(nequ, 1-lv:n+lv), bo(1)%Qc(nequ, 1-lv:n+lv))
        allocate(bo(1)%Fplus(nequ, 0:n), bo(1)%godunov_state(nprim, 0:n))
        bo(1)%qf = reshape(dv(2:1+nequ*(n+2*lv)), [nequ, n+2*lv])
        bo(1)%Qc = 0.d0
        bo(1)%Fplus = 0.d0
        bo(1)%godunov_state = 0.d0
        call calc_flux_advection(1, 1, bo, lv, KK, nequ, select_riemann)
        call write_output(fout, [n], [reshape(bo(1)%Fplus, [nequ*(n+1)]), &
                                      reshape(bo(1)%godunov_state, [nprim*(n+1)])])
    end subroutine do_flux

    ! Multi-block wall-bounded run with the reference's own block, boundary, halo and
    ! time-integration routines, initialised the way initialize_variables / set_block_params /
    ! main_sim do it (SRC/params.f90:305-580, SRC/block.f90:394-431, SRC/main.f90:94-114) and
    ! stepped like the main loop (SRC/main.f90:130-163) with the source term off.
    ! iv = [nblk, ncell, nsteps, lvir, flag_acu, flag_weno, flag_muscl, select_riemann,
    !       select_time, flag_cfl]
    ! dv = [dx, cfl, dt_fixed, (T, p, u, X(KK)) per interior cell]
    ! out = [seconds in the step loop, dt per step, Qc and qf after initialisation (ghosts included,
    !        block by block), Qc and qf after the last step]
    subroutine do_run()
        integer :: nblk, ncell, nsteps, lv, nequ, ib, i, g, step, off, ncfl
        integer :: ista_rank_, iend_rank_
        integer(8) :: c0, c1, crate
        real(8) :: dx, cflv, dtv, rho, p, T, u, e
        real(8) :: X(KK), Y(KK)
        real(8), allocatable :: dts(:), o(:), init_state(:)
        nblk = iv(1); ncell = iv(2); nsteps = iv(3); lv = iv(4)
        call set_flags(iv(5), iv(6), iv(7), iv(8))
        select_time = iv(9)
        ncfl = iv(10)
        dx = dv(1); cflv = dv(2); dtv = dv(3)
        dx_comp = dx
        nequ = 3 + KK
        bsta = 1; bend = nblk; block_all_0 = 1; block_all_n = nblk
        ista_rank_ = 1; iend_rank_ = nblk*ncell
        allocate(b_obj(1:nblk))
        do ib = 1, nblk
            b_obj(ib)%smr_level = 1
            b_obj(ib)%smr_dx = dx
            b_obj(ib)%ista_b = (ib-1)*ncell + 1
            b_obj(ib)%iend_b = ib*ncell
            allocate(b_obj(ib)%Qc(nequ, b_obj(ib)%ista_b-lv:b_obj(ib)%iend_b+lv))
            allocate(b_obj(ib)%qf(nequ, b_obj(ib)%ista_b-lv:b_obj(ib)%iend_b+lv))
            allocate(b_obj(ib)%Fplus(nequ, b_obj(ib)%ista_b-1:b_obj(ib)%iend_b))
            allocate(b_obj(ib)%godunov_state(nprim, b_obj(ib)%ista_b-1:b_obj(ib)%iend_b))
            allocate(b_obj(ib)%xm_b(b_obj(ib)%ista_b:b_obj(ib)%iend_b))
            b_obj(ib)%Qc = 0.d0
            b_obj(ib)%qf = 0.d0
            b_obj(ib)%Fplus = 0.d0
            b_obj(ib)%godunov_state = 0.d0
            do i = b_obj(ib)%ista_b, b_obj(ib)%iend_b
                b_obj(ib)%xm_b(i) = dx*i - dx/2.d0
                off = 3 + (i-1)*(3+KK)
                T = dv(off+1); p = dv(off+2); u = dv(off+3)
                X = dv(off+4:off+3+KK)
                call adjust_sum(KK, X)
                call Xi2Yi(KK, X, Y)
                rho = density(p, T, KK, X)
                e = specific_energy(rho, p, T, KK, X)
                b_obj(ib)%qf(iqdens, i) = rho
                b_obj(ib)%qf(iqxvel, i) = u
                b_obj(ib)%qf(iqtemp, i) = T
                b_obj(ib)%qf(nprim+1:nequ, i) = Y
                b_obj(ib)%Qc(iudens, i) = rho
                b_obj(ib)%Qc(iumomx, i) = rho*u
                b_obj(ib)%Qc(iuener, i) = rho*(e + 0.5d0*u**2)
                b_obj(ib)%Qc(ncons+1:nequ, i) = rho*Y
            end do
        end do
        call update_boundary(bsta, bend, b_obj, lv, KK, nequ, block_all_0, block_all_n)
        call interpolation_block(nequ, lv)
        call update_boundary(bsta, bend, b_obj, lv, KK, nequ, block_all_0, block_all_n)
        do ib = bsta, bend
            do g = 1, lv
                i = b_obj(ib)%ista_b - g
                b_obj(ib)%qf(:, i) = QcToqf(b_obj(ib)%Qc(:, i), KK, nequ)
                i = b_obj(ib)%iend_b + g
                b_obj(ib)%qf(:, i) = QcToqf(b_obj(ib)%Qc(:, i), KK, nequ)
            end do
        end do

        allocate(init_state(2*nblk*nequ*(ncell+2*lv)))
        off = 0
        do ib = 1, nblk
            init_state(off+1:off+nequ*(ncell+2*lv)) = reshape(b_obj(ib)%Qc, [nequ*(ncell+2*lv)])
            off = off + nequ*(ncell+2*lv)
        end do
        do ib = 1, nblk
            init_state(off+1:off+nequ*(ncell+2*lv)) = reshape(b_obj(ib)%qf, [nequ*(ncell+2*lv)])
            off = off + nequ*(ncell+2*lv)
        end do

        allocate(dts(nsteps))
        call system_clock(c0, crate)
        do step = 1, nsteps
            if (ncfl == 1) dtv = harness_calc_dt(cflv, dx, nequ)
            dts(step) = dtv
            if (select_time == 0) then
                call explicit_euler(bsta, bend, b_obj, lv, dtv, KK, nequ, block_all_0, block_all_n)
            else if (select_time == 1) then
                call TVDRK3(ista_rank_, iend_rank_, bsta, bend, b_obj, lv, dtv, KK, nequ, &
                            block_all_0, block_all_n)
            else
                call RK4(ista_rank_, iend_rank_, bsta, bend, b_obj, lv, dtv, KK, nequ, &
                         block_all_0, block_all_n)
            end if
            call update_boundary(bsta, bend, b_obj, lv, KK, nequ, block_all_0, block_all_n)
            call interpolation_block(nequ, lv)
            call update_boundary(bsta, bend, b_obj, lv, KK, nequ, block_all_0, block_all_n)
            do ib = bsta, bend
                do i = b_obj(ib)%ista_b-lv, b_obj(ib)%iend_b+lv
                    b_obj(ib)%qf(:, i) = QcToqf(b_obj(ib)%Qc(:, i), KK, nequ)
                end do
            end do
        end do
        call system_clock(c1)

        allocate(o(1 + nsteps + 4*nblk*nequ*(ncell+2*lv)))
        o(1) = dble(c1 - c0) / dble(crate)
        o(2:1+nsteps) = dts
        o(2+nsteps:1+nsteps+size(init_state)) = init_state
        off = 1 + nsteps + size(init_state)
        do ib = 1, nblk
            o(off+1:off+nequ*(ncell+2*lv)) = reshape(b_obj(ib)%Qc, [nequ*(ncell+2*lv)])
            off = off + nequ*(ncell+2*lv)
        end do
        do ib = 1, nblk
            o(off+1:off+nequ*(ncell+2*lv)) = reshape(b_obj(ib)%qf, [nequ*(ncell+2*lv)])
            off = off + nequ*(ncell+2*lv)
        end do
        call write_output(fout, [nsteps], o)
    end subroutine do_run

    ! calc_dt (SRC/update.f90:13-59) lives in a module that also needs the missing ODEs module, so
    ! the harness evaluates the same per-cell expression with the reference's thermo functions
    ! (single rank: the MPI_ALLREDUCE(MIN) is the identity).
    function harness_calc_dt(cflv, dx, nequ) result(dtv)
        real(8), intent(in) :: cflv, dx
        integer, intent(in) :: nequ
        real(8) :: dtv, cvel, aupc, rho, temp, pres, a_sound, x_sp(KK)
        integer :: ib, i
        cvel = 0.d0
        do ib = bsta, bend
            do i = b_obj(ib)%ista_b, b_obj(ib)%iend_b
                call Yi2Xi(KK, b_obj(ib)%qf(nprim+1:nequ, i), x_sp)
                rho = b_obj(ib)%qf(iqdens, i)
                temp = b_obj(ib)%qf(iqtemp, i)
                pres = pressure(rho, temp, KK, x_sp)
                a_sound = sonic_speed(rho, pres, temp, KK, x_sp)
                aupc = dmax1(dabs(b_obj(ib)%qf(iqxvel, i) - a_sound), dabs(b_obj(ib)%qf(iqxvel, i)), &
                             dabs(b_obj(ib)%qf(iqxvel, i) + a_sound))
                if (aupc > cvel) cvel = aupc
            end do
        end do
        dtv = (cflv*dx)/cvel
    end function harness_calc_dt

end program ref_harness
