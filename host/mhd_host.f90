! mhd_host.f90 -- Fortran driver of the MI355X LAPS step.
!
! Mirrors the reference driver `program mhd` (src_compressible/mhd.f90:1-287):
! the same namelist groups read in the same order from mhd.input
! (mhd.f90:30-53), the background / perturbation hooks (mhdinit.f90:183-1036,
! the cases listed below), the restart path (restart.f90:17-63), the output
! files outNNN.dat, grid.dat, parallel_info.dat, rms.dat, EBM_info.dat and
! log (mhdoutput.f90, mhdrms.f90:39-51, AEBmod.f90:75-85, mhd.f90:431-457),
! and the main loop (mhd.f90:170-287). The time step itself runs on the GPU
! through the C ABI (module laps_c): laps_step performs the dt < 1e-8 test,
! evolve, time/radius update and the next vardt on the device.
!
! One GPU replaces the MPI decomposition: &prl is read and ignored, and
! parallel_info.dat records one task. An optional group &laps read after
! &Hall selects nstep (stop after that many steps; < 0: no limit), the
! precision ('fp64' or 'mixed') and the HIP device.
program mhd_host
    use, intrinsic :: iso_c_binding
    use laps_c
    implicit none

    integer, parameter :: dp = kind(1.0d0)
    real(dp), parameter :: pi = 3.141592653589793_dp

    ! &genr
    real(dp) :: tmax = 0, dtout = 1, dtrms = 1
    logical :: output_primitive = .true., if_restart = .false.
    integer :: n_start = 0
    ! &numerical
    real(dp) :: cfl = 0.5_dp, afx = 0.495_dp, afy = 0.495_dp, afz = 0.495_dp
    integer :: dealias_option = 2
    logical :: if_resis_exp = .false., if_visc_exp = .false., if_conserve_background = .false.
    ! &prl
    integer :: ndim_parallel = 2
    ! &grid
    integer :: nx = 128, ny = 128, nz = 64
    real(dp) :: Lx = 2 * pi, Ly = 2 * pi, Lz = 2 * pi
    ! &field
    integer :: ifield = 0
    real(dp) :: B0 = 1.0_dp, a0 = 0.05_dp, shear_width = 0, press0 = 0, U_fast = 0, U_slow = 0
    real(dp) :: n_fast = 0, n_slow = 0, Bx0 = 0, By0 = 0, Bz0 = 0, jet_width = 1.0_dp, Vjet = 0.3_dp
    ! &pert
    integer :: ipert = 0, nmode = 16, nmodex = 16, nmodey = 16, nmodez = 16, wave_number_jet = 1
    real(dp) :: db0 = 0, dbx = 0, dbz = 0, in_out_ratio = 1.0_dp, dv0 = 0, drho0 = 0
    real(dp) :: B0_SB = 0, dB_SB = 0, R0_SB = 0, R1_SB = 0, H_SB = 0, Rm_SB = 0, Rd_SB = 0
    ! &phys
    real(dp) :: adiabatic_index = 5.0_dp / 3.0_dp, resistivity = 0, viscosity = 0
    logical :: if_resis = .false., if_visc = .false.
    ! &AEB
    logical :: if_AEB = .false., if_corotating = .false.
    real(dp) :: radius0 = 30.0_dp, Ur0 = 0, corotating_angle = 0
    ! &Hall
    logical :: if_hall = .false.
    real(dp) :: ion_inertial_length = 0
    ! &laps (optional; no counterpart in the reference)
    integer :: nstep = -1, device = 0
    character(len=8) :: precision = 'fp64'

    namelist /genr/ tmax, dtout, dtrms, output_primitive, if_restart, n_start
    namelist /numerical/ cfl, dealias_option, afx, afy, afz, if_resis_exp, if_visc_exp, if_conserve_background
    namelist /prl/ ndim_parallel
    namelist /grid/ nx, ny, nz, Lx, Ly, Lz
    namelist /phys/ adiabatic_index, if_resis, resistivity, if_visc, viscosity
    namelist /field/ ifield, B0, a0, shear_width, press0, U_fast, U_slow, &
        n_fast, n_slow, Bx0, By0, Bz0, jet_width, Vjet
    namelist /pert/ ipert, db0, nmode, nmodex, nmodey, nmodez, dbx, dbz, in_out_ratio, &
        dv0, drho0, wave_number_jet, B0_SB, dB_SB, R0_SB, R1_SB, H_SB, Rm_SB, Rd_SB
    namelist /AEB/ if_AEB, radius0, Ur0, if_corotating, corotating_angle
    namelist /Hall/ if_hall, ion_inertial_length
    namelist /laps/ nstep, precision, device

    real(dp), allocatable :: uu(:, :, :, :), uu_prim(:, :, :, :), xgrid(:), ygrid(:), zgrid(:)
    type(laps_params) :: p
    type(c_ptr) :: ctx
    real(dp) :: time, dt, radius, dtmin_raw, tout, toutrms, tlog, dtlog, max_divb, t_restart
    real(dp) :: clocktime, clocktime_output
    real(dp), parameter :: delta_clocktime_output = 60 * 50
    integer(c_int64_t) :: istep
    integer(c_int32_t) :: status
    integer :: iout, ios, i
    integer :: pending_out = -1          ! output staged on the device, written after the next step is queued
    real(dp) :: pending_t
    integer(8) :: clock_start, clock_now, clock_rate, clock_max
    integer(8), parameter :: dstep_checksave = 40

    open(unit=4, status='old', file='mhd.input')
    read(4, genr)
    read(4, numerical)
    read(4, prl)
    read(4, grid)
    read(4, field)
    read(4, pert)
    read(4, phys)
    read(4, AEB)
    read(4, Hall)
    read(4, laps, iostat=ios)
    close(unit=4)

    call system_clock(clock_start, clock_rate, clock_max)
    if (.not. if_AEB) Ur0 = 0.0_dp   ! mhd.f90:88-90
    if (.not. if_corotating) corotating_angle = 0.0_dp   ! AEBmod.f90:28-30

    allocate(uu(nx, ny, nz, 8), uu_prim(nx, ny, nz, 4), xgrid(nx), ygrid(ny), zgrid(nz))
    do i = 1, nx
        xgrid(i) = (i - 1) * (Lx / nx)
    enddo
    do i = 1, ny
        ygrid(i) = (i - 1) * (Ly / ny)
    enddo
    do i = 1, nz
        zgrid(i) = (i - 1) * (Lz / nz)
    enddo

    call laps_default_params(p)
    p%nx = nx; p%ny = ny; p%nz = nz
    p%Lx = Lx; p%Ly = Ly; p%Lz = Lz
    p%adiabatic_index = adiabatic_index
    p%cfl = cfl
    p%dealias_option = dealias_option
    p%if_resis = merge(1, 0, if_resis); p%if_resis_exp = merge(1, 0, if_resis_exp)
    p%if_visc = merge(1, 0, if_visc); p%if_visc_exp = merge(1, 0, if_visc_exp)
    p%if_conserve_background = merge(1, 0, if_conserve_background)
    p%afx = afx; p%afy = afy; p%afz = afz
    p%resistivity = resistivity; p%viscosity = viscosity
    p%if_hall = merge(1, 0, if_hall); p%ion_inertial_length = ion_inertial_length
    p%if_aeb = merge(1, 0, if_AEB); p%if_corotating = merge(1, 0, if_corotating)
    p%radius0 = radius0; p%Ur0 = Ur0; p%corotating_angle = corotating_angle
    p%precision = merge(LAPS_MIXED, LAPS_FP64, trim(precision) == 'mixed')
    p%device = device
    call laps_check(laps_create(p, ctx), 'laps_create')
    write(*, '(3x,a)') 'laps_create finished......'

    t_restart = 0
    if (if_restart) then
        call read_restart(n_start, t_restart)
        write(*, '(3x,a)') 'read_restart finished......'
    else
        call background_fields_initialize
        call perturbation_initialize
        write(*, '(3x,a)') 'initial fields finished......'
    endif
    call initial_calc_conserve_variable
    call laps_check(laps_set_state(ctx, uu, uu_prim), 'laps_set_state')
    if (if_restart) call laps_check(laps_set_time(ctx, t_restart, 0.0_dp), 'laps_set_time')

    call system_clock(clock_now)
    write(*, '(1x,a,f15.2,a)') 'Initialization takes ', real(clock_now - clock_start, dp) / clock_rate, ' seconds.'
    call system_clock(clock_start)

    call laps_check(laps_cfl(ctx, dtmin_raw, dt), 'laps_cfl')
    call write_initial
    call output_grid
    call output_parallel_info

    dtlog = min(dtout, dtrms) / 10.0_dp
    time = t_restart
    iout = n_start
    tout = time + dtout
    toutrms = time + dtrms
    tlog = time + dtlog
    istep = 0

    write(*, '(2x,a,2x,f10.4,2x,a,2x,1pe12.4)') 'OUTPUT at time:', time, ',dt = ', dt
    call output_uu(iout, time)
    iout = iout + 1
    call output_rms_all(time)
    clocktime_output = delta_clocktime_output

    principal: do
        if (time >= tout) then
            write(*, '(2x,a,2x,f10.4,2x,a,2x,1pe12.4)') 'OUTPUT at time:', time, ',dt = ', dt
            ! staged: the state is copied to pinned memory while the next step runs, and the
            ! file is written from there once that step is queued
            call laps_check(laps_stage_output(ctx, merge(1_c_int, 0_c_int, output_primitive)), 'laps_stage_output')
            pending_out = iout
            pending_t = time
            iout = iout + 1
            tout = tout + dtout
        endif
        if (time >= toutrms) then
            call output_rms_all(time)
            toutrms = toutrms + dtrms
        endif
        if (istep > 0 .and. modulo(istep, dstep_checksave) == 0) then
            call system_clock(clock_now)
            clocktime = real(clock_now - clock_start, dp) / clock_rate
            if (clocktime >= clocktime_output) then
                write(*, '(3x,a,f15.2,2x,a,2x,f10.4)') 'OUTPUT for backup at real time (sec):', clocktime, &
                    ', time = ', time
                call output_uu(999, time)
                clocktime_output = clocktime_output + delta_clocktime_output
            endif
        endif

        ! dt < 1e-8 test, evolve, time += dt, evolve_radius, next vardt (on the device)
        call laps_check(laps_step(ctx, 1), 'laps_step')
        if (pending_out >= 0) then
            call write_staged(pending_out, pending_t)
            pending_out = -1
        endif
        call laps_check(laps_get_scalars(ctx, time, dt, radius, istep, status), 'laps_get_scalars')
        if (status == LAPS_DT_STOP) then
            write(*, *) '***********************************'
            write(*, *) 'STOP: dt  < 1.e-8    '
            write(*, '(1x,a14,1pe12.4,a14,i6)') '  time = ', time, 'istep =', istep
            write(*, '(1x,a14,1pe12.4)') 'dt =', dt
            write(*, *) '***********************************'
            write(*, '(2x,a,2x,f10.4,2x,a,2x,1pe12.4)') 'OUTPUT at time:', time, ',dt = ', dt
            call output_uu(iout, time)
            call output_rms_all(time)
            exit principal
        endif

        if (time >= tmax .or. (nstep >= 0 .and. istep >= nstep)) then
            write(*, '(2x,a,2x,f10.4,2x,a,2x,1pe12.4)') 'OUTPUT at time:', time, ',dt = ', dt
            call output_uu(iout, time)
            call output_rms_all(time)
            write(*, *) '-------------------------------------------------'
            if (time >= tmax) then
                write(*, *) 'Simulation ends because tmax is reached.'
            else
                write(*, *) 'Simulation ends because nstep is reached.'
            endif
            write(*, *) '-------------------------------------------------'
            exit principal
        endif

        if (time >= tlog) then
            call write_log
            tlog = tlog + dtlog
        endif
    enddo principal

    call laps_destroy(ctx)

contains

    ! background_fields_initialize (mhdinit.f90:183-299), ifield = 1..5
    subroutine background_fields_initialize
        real(dp), parameter :: T0 = 1.0_dp, rho0 = 1.0_dp, ninf = 1.0_dp
        real(dp) :: zlow, zcent, zup, ylow, ycent, yup, r_yz, aa
        integer :: iy, iz
        uu = 0
        uu(:, :, :, 1) = rho0
        uu(:, :, :, 8) = rho0 * T0
        select case (ifield)
        case (1)
            zlow = 0.25_dp * Lz; zup = 0.75_dp * Lz; zcent = 0.5_dp * Lz
            do iz = 1, nz
                if (zgrid(iz) < zcent) then
                    uu(:, :, iz, 5) = B0 * tanh((zgrid(iz) - zlow) / a0)
                    uu(:, :, iz, 1) = ninf + B0**2 / 2.0_dp / T0 / (cosh((zgrid(iz) - zlow) / a0))**2
                else
                    uu(:, :, iz, 5) = -B0 * tanh((zgrid(iz) - zup) / a0)
                    uu(:, :, iz, 1) = ninf + B0**2 / 2.0_dp / T0 / (cosh((zgrid(iz) - zup) / a0))**2
                endif
            enddo
            uu(:, :, :, 8) = uu(:, :, :, 1) * T0
        case (2)
            ylow = 0.25_dp * Ly; yup = 0.75_dp * Ly; ycent = 0.5_dp * Ly
            aa = Ly * 0.5_dp * shear_width
            do iy = 1, ny
                if (ygrid(iy) < ycent) then
                    uu(:, iy, :, 2) = 0.5_dp * (U_fast - U_slow) * tanh((ygrid(iy) - ylow) / aa) + 0.5_dp * (U_fast + U_slow)
                    uu(:, iy, :, 1) = 0.5_dp * (n_fast - n_slow) * tanh((ygrid(iy) - ylow) / aa) + 0.5_dp * (n_fast + n_slow)
                else
                    uu(:, iy, :, 2) = -0.5_dp * (U_fast - U_slow) * tanh((ygrid(iy) - yup) / aa) + 0.5_dp * (U_fast + U_slow)
                    uu(:, iy, :, 1) = -0.5_dp * (n_fast - n_slow) * tanh((ygrid(iy) - yup) / aa) + 0.5_dp * (n_fast + n_slow)
                endif
            enddo
            uu(:, :, :, 8) = press0
            uu(:, :, :, 5) = bx0
            uu(:, :, :, 6) = by0   ! Bz is left as it is (mhdinit.f90:245-249)
        case (3)
            uu(:, :, :, 5) = bx0
            uu(:, :, :, 6) = by0
            uu(:, :, :, 7) = bz0
            uu(:, :, :, 8) = press0
        case (4)
            ylow = 0.25_dp * Ly; yup = 0.75_dp * Ly; ycent = 0.5_dp * Ly
            do iy = 1, ny
                if (ygrid(iy) < ycent) then
                    uu(:, iy, :, 5) = B0 * tanh((ygrid(iy) - ylow) / a0)
                else
                    uu(:, iy, :, 5) = -B0 * tanh((ygrid(iy) - yup) / a0)
                endif
                uu(:, iy, :, 8) = press0 - 0.5_dp * (uu(:, iy, :, 5))**2
                uu(:, iy, :, 1) = ninf
            enddo
        case (5)
            uu(:, :, :, 5) = bx0
            uu(:, :, :, 6) = by0
            uu(:, :, :, 7) = bz0
            uu(:, :, :, 8) = press0
            do iz = 1, nz
                do iy = 1, ny
                    r_yz = sqrt((zgrid(iz) - Lz / 2)**2 + (ygrid(iy) - Ly / 2)**2)
                    uu(:, iy, iz, 2) = Vjet / 2 * (1 - tanh((r_yz - 1) / 0.1_dp))
                enddo
            enddo
        case default
            continue
        end select
    end subroutine background_fields_initialize

    ! perturbation_initialize (mhdinit.f90:302-1036): ipert = 1, 8 and 999. The
    ! other hooks draw from the compiler's RNG; start those runs from an
    ! out000.dat written by laps_amd.ic and if_restart = .true.
    subroutine perturbation_initialize
        real(dp) :: kx, s, c
        integer :: ix
        s = sin(corotating_angle)
        c = cos(corotating_angle)
        select case (ipert)
        case (0)
            continue
        case (1)
            kx = 2 * pi / Lx * wave_number_jet
            do ix = 1, nx
                uu(ix, :, :, 7) = uu(ix, :, :, 7) - db0 * sin(kx * xgrid(ix))
                uu(ix, :, :, 4) = uu(ix, :, :, 4) + db0 / sqrt(uu(ix, :, :, 1)) * sin(kx * xgrid(ix))
                uu(ix, :, :, 2) = uu(ix, :, :, 2) + db0 / sqrt(uu(ix, :, :, 1)) * cos(kx * xgrid(ix)) * s
                uu(ix, :, :, 5) = uu(ix, :, :, 5) - db0 * cos(kx * xgrid(ix)) * s
                uu(ix, :, :, 3) = uu(ix, :, :, 3) + db0 / sqrt(uu(ix, :, :, 1)) * cos(kx * xgrid(ix)) * c
                uu(ix, :, :, 6) = uu(ix, :, :, 6) - db0 * cos(kx * xgrid(ix)) * c
            enddo
        case (2:4)
            call random_modes_1d(ipert)
        case (5)
            call tearing_modes
        case (6)
            call alfvenic_modes
        case (7)
            call isotropic_modes
        case (8)
            call switchback
        case (999)
            block
                integer :: iy, iz, v
                do iz = 1, nz
                    do iy = 1, ny
                        do v = 1, 8
                            uu(:, iy, iz, v) = sqrt((ygrid(iy) - Ly / 2.0_dp)**2 + (zgrid(iz) - Lz / 2.0_dp)**2)
                        enddo
                    enddo
                enddo
            end block
        case default
            write(*, '(1x,a,i0,a)') 'ipert = ', ipert, ' is not a perturbation of the reference: write out000.dat ', &
                'with laps_amd.ic and restart from it (if_restart = .true., n_start = 0)'
            error stop 3
        end select
    end subroutine perturbation_initialize

    ! The random-phase hooks below draw their phases from the compiler's random_number, in the
    ! order and with the seeding calls of the reference (mhdinit.f90:344-829), so a run matches
    ! the reference built with the same compiler. Each adds plane waves a cos(k.x + ph) (or sin),
    ! evaluated as the real part of e^{i kx x} e^{i ky y} e^{i kz z} e^{i ph}.

    ! w(:,:,:) += amp * Re(e^{i(k.x + ph)} * rot) for integer wave numbers (mx, my, mz) of the
    ! box; rot = 1 gives cos, rot = -i gives sin
    subroutine add_wave(w, mx, my, mz, amp, ph, rot)
        real(dp), intent(inout) :: w(:, :, :)
        integer, intent(in) :: mx, my, mz
        real(dp), intent(in) :: amp, ph
        complex(dp), intent(in) :: rot
        complex(dp) :: ex(nx), ey(ny), ez(nz), cyz
        integer :: ix, iy, iz
        do ix = 1, nx
            ex(ix) = exp(cmplx(0.0_dp, 2 * pi * mx / Lx * xgrid(ix), dp))
        enddo
        do iy = 1, ny
            ey(iy) = exp(cmplx(0.0_dp, 2 * pi * my / Ly * ygrid(iy), dp))
        enddo
        do iz = 1, nz
            ez(iz) = exp(cmplx(0.0_dp, 2 * pi * mz / Lz * zgrid(iz) + ph, dp)) * rot * amp
        enddo
        do iz = 1, nz
            do iy = 1, ny
                cyz = ey(iy) * ez(iz)
                w(:, iy, iz) = w(:, iy, iz) + real(ex(:) * cyz, dp)
            enddo
        enddo
    end subroutine add_wave

    ! ipert = 2, 3, 4 (mhdinit.f90:344-392): nmode waves of amplitude db0 and random phase along
    ! x in Bz, along y in Bz, along z in Bx, as sin(k s + ph)
    subroutine random_modes_1d(kind)
        integer, intent(in) :: kind
        real(dp), allocatable :: phs(:)
        integer :: ik
        allocate(phs(nmode))
        call random_number(phs)
        phs = phs * 2 * pi
        do ik = 1, nmode
            select case (kind)
            case (2)
                call add_wave(uu(:, :, :, 7), ik, 0, 0, db0, phs(ik), (0.0_dp, -1.0_dp))
            case (3)
                call add_wave(uu(:, :, :, 7), 0, ik, 0, db0, phs(ik), (0.0_dp, -1.0_dp))
            case (4)
                call add_wave(uu(:, :, :, 5), 0, 0, ik, db0, phs(ik), (0.0_dp, -1.0_dp))
            end select
        enddo
    end subroutine random_modes_1d

    ! ipert = 5 (mhdinit.f90:394-483): 3D tearing perturbations of two current sheets at y = Ly/4
    ! and 3Ly/4, Gaussian profiles of width a0: for each (kx, kz), ikx = 1..nmodex, ikz =
    ! 0..nmodez, bx, bz += (dbx, dbz) cos(ph) m'(y) and by += (kx dbx + kz dbz) sin(ph) m(y), m =
    ! a0 exp(-((y - yc)/a0)^2) (divergence free), each sheet on its own half of the box. The
    ! phase of (ikx, ikz) is entry (ikx - 1) nmodez + ikz of each sheet's draw of nmodex (nmodez + 1)
    ! numbers. Entry 0 (mode (1, 0)) lies before the reference's array; here it is drawn after
    ! both sheets' draws.
    subroutine tearing_modes
        real(dp), allocatable :: phs(:, :), my(:), dmy(:)
        real(dp) :: kx, kz, yc
        integer :: sheet, ikx, ikz, iy, nsz, nph, idx
        nph = nmodex * (nmodez + 1)
        allocate(phs(0:nph, 2), my(ny), dmy(ny))
        ! the two sheets' draws in the reference's order, then the entries it reads before them
        do sheet = 1, 2
            call random_seed(size=nsz)   ! the reference's call random_seed(ir) queries the size
            call random_number(phs(1:nph, sheet))
        enddo
        call random_number(phs(0, :))
        phs = phs * 2 * pi
        do sheet = 1, 2
            yc = merge(0.25_dp, 0.75_dp, sheet == 1) * Ly
            do iy = 1, ny
                if ((sheet == 1) .eqv. (ygrid(iy) <= 0.5_dp * Ly)) then
                    my(iy) = a0 * exp(-((ygrid(iy) - yc) / a0)**2)
                    dmy(iy) = -2 * (ygrid(iy) - yc) / a0 * exp(-((ygrid(iy) - yc) / a0)**2)
                else
                    my(iy) = 0
                    dmy(iy) = 0
                endif
            enddo
            do ikx = 1, nmodex
                kx = ikx * 2 * pi / Lx
                do ikz = 0, nmodez
                    kz = ikz * 2 * pi / Lz
                    idx = (ikx - 1) * nmodez + ikz
                    call add_profiled(5, ikx, ikz, dbx, phs(idx, sheet), dmy, (1.0_dp, 0.0_dp))
                    call add_profiled(7, ikx, ikz, dbz, phs(idx, sheet), dmy, (1.0_dp, 0.0_dp))
                    call add_profiled(6, ikx, ikz, kx * dbx + kz * dbz, phs(idx, sheet), my, (0.0_dp, -1.0_dp))
                enddo
            enddo
        enddo
    end subroutine tearing_modes

    ! uu(:,:,:,v) += amp * prof(y) * Re(e^{i(kx x + kz z + ph)} rot)
    subroutine add_profiled(v, ikx, ikz, amp, ph, prof, rot)
        integer, intent(in) :: v, ikx, ikz
        real(dp), intent(in) :: amp, ph, prof(:)
        complex(dp), intent(in) :: rot
        complex(dp) :: ex(nx), ez
        integer :: ix, iy, iz
        do ix = 1, nx
            ex(ix) = exp(cmplx(0.0_dp, 2 * pi * ikx / Lx * xgrid(ix), dp))
        enddo
        do iz = 1, nz
            ez = exp(cmplx(0.0_dp, 2 * pi * ikz / Lz * zgrid(iz) + ph, dp)) * rot * amp
            do iy = 1, ny
                uu(:, iy, iz, v) = uu(:, iy, iz, v) + prof(iy) * real(ex(:) * ez, dp)
            enddo
        enddo
    end subroutine add_profiled

    ! ipert = 6 (mhdinit.f90:486-692): Alfven waves propagating along +x (outward) and -x
    ! (inward, amplitude times in_out_ratio), modes ikx = 1..nmodex, iky, ikz in -nmode..nmode,
    ! amplitude db0 / sqrt(|ikx| |iky| |ikz|) (a zero index counts as 1). (ky, kz) = 0: a
    ! circularly polarised wave, by = cos, bz = sin; otherwise b along (kz, -ky)/|(ky, kz)|, cos.
    ! u = -/+ sign(bx0) b for the outward / inward waves. One draw of phases per direction.
    subroutine alfvenic_modes
        real(dp), allocatable :: phs(:)
        real(dp) :: ky, kz, amp, dby, dbz, sgn, s_uv
        integer :: dir, ikx, iky, ikz, nsz, idx, nph
        nph = nmodex * (2 * nmodez + 1) * (2 * nmodey + 1)
        allocate(phs(nph))
        sgn = sign(1.0_dp, bx0)
        do dir = 1, 2
            call random_seed(size=nsz)
            call random_number(phs)
            phs = phs * 2 * pi
            s_uv = merge(-sgn, sgn, dir == 1)
            do ikx = 1, nmodex
                do iky = -nmodey, nmodey
                    ky = iky * 2 * pi / Ly
                    do ikz = -nmodez, nmodez
                        kz = ikz * 2 * pi / Lz
                        amp = db0 * sqrt(1 / real(ikx, dp)) * merge(1.0_dp, sqrt(1 / real(abs(iky), dp)), iky == 0) &
                            * merge(1.0_dp, sqrt(1 / real(abs(ikz), dp)), ikz == 0)
                        if (dir == 2) amp = amp * in_out_ratio
                        idx = 1 + ((ikx - 1) * (2 * nmodey + 1) + iky + nmodey) * (2 * nmodez + 1) + ikz + nmodez
                        if (iky == 0 .and. ikz == 0) then
                            call add_wave4(merge(ikx, -ikx, dir == 1), 0, 0, phs(idx), amp, amp, &
                                           (0.0_dp, -1.0_dp), s_uv)
                        else
                            dby = kz / sqrt(ky**2 + kz**2)
                            dbz = -ky / sqrt(ky**2 + kz**2)
                            call add_wave4(merge(ikx, -ikx, dir == 1), iky, ikz, phs(idx), amp * dby, amp * dbz, &
                                           (1.0_dp, 0.0_dp), s_uv)
                        endif
                    enddo
                enddo
            enddo
        enddo
    end subroutine alfvenic_modes

    ! one mode of ipert = 6: by += ay cos(ph'), bz += az Re(rot e^{i ph'}), and s times the same
    ! in uy, uz (ph' = k.x + ph)
    subroutine add_wave4(mx, my, mz, ph, ay, az, rot, s)
        integer, intent(in) :: mx, my, mz
        real(dp), intent(in) :: ph, ay, az, s
        complex(dp), intent(in) :: rot
        call add_wave(uu(:, :, :, 6), mx, my, mz, ay, ph, (1.0_dp, 0.0_dp))
        call add_wave(uu(:, :, :, 7), mx, my, mz, az, ph, rot)
        call add_wave(uu(:, :, :, 3), mx, my, mz, s * ay, ph, (1.0_dp, 0.0_dp))
        call add_wave(uu(:, :, :, 4), mx, my, mz, s * az, ph, rot)
    end subroutine add_wave4

    ! ipert = 7 (mhdinit.f90:695-829): isotropic random modes in the half space ikx >= 0 (ikx = 0:
    ! iky > 0, or iky = 0 and ikz > 0) with |k| <= max(nmodex, nmodey, nmodez), amplitude
    ! |k|^(-3/2), along k x B0 / (|k| |B0|): u (dv0, phase ph0, and the part correlated with b with
    ! coefficient -0.05 and phase ph1), b (db0, phase ph1), rho (drho0, phase ph2). The three
    ! phase draws follow random_seed(put = ...) with every seed word 101, 116 and 132.
    subroutine isotropic_modes
        real(dp), allocatable :: ph0(:), ph1(:), ph2(:)
        integer, allocatable :: seed(:)
        real(dp), parameter :: corr = -0.05_dp
        real(dp) :: kx, ky, kz, kmod, b0mod, kr, dperp(3), a
        integer :: ikx, iky, ikz, idx, nph, nsz, c
        nph = (2 * nmodex + 1) * (2 * nmodez + 1) * (2 * nmodey + 1)
        allocate(ph0(nph), ph1(nph), ph2(nph))
        call random_seed(size=nsz)
        allocate(seed(nsz))
        seed = 101
        call random_seed(put=seed)
        call random_number(ph0)
        seed = 116
        call random_seed(put=seed)
        call random_number(ph1)
        seed = 132
        call random_seed(put=seed)
        call random_number(ph2)
        ph0 = ph0 * 2 * pi
        ph1 = ph1 * 2 * pi
        ph2 = ph2 * 2 * pi
        b0mod = sqrt(bx0**2 + by0**2 + bz0**2)
        do ikx = 0, nmodex
            kx = ikx * 2 * pi / Lx
            do iky = -nmodey, nmodey
                ky = iky * 2 * pi / Ly
                do ikz = -nmodez, nmodez
                    kz = ikz * 2 * pi / Lz
                    if ((ikx == 0 .and. iky < 0) .or. (ikx == 0 .and. iky == 0 .and. ikz <= 0)) cycle
                    kr = sqrt(real(ikx**2 + iky**2 + ikz**2, dp))
                    if (kr > max(nmodex, nmodey, nmodez)) cycle
                    kmod = sqrt(kx**2 + ky**2 + kz**2)
                    if (kmod == 0 .or. b0mod == 0) then
                        write(*, *) 'When intializing perturbation, |k|=0 or |B0|=0!!!'
                        error stop 0
                    endif
                    dperp = [ky * bz0 - kz * by0, kz * bx0 - kx * bz0, kx * by0 - ky * bx0] / kmod / b0mod
                    idx = 1 + ((ikx + nmodex) * (2 * nmodey + 1) + iky + nmodey) * (2 * nmodez + 1) + ikz + nmodez
                    a = 1 / sqrt(kr**3)
                    do c = 1, 3
                        call add_wave(uu(:, :, :, 1 + c), ikx, iky, ikz, sqrt(1 - corr**2) * dv0 * a * dperp(c), &
                                      ph0(idx), (1.0_dp, 0.0_dp))
                        call add_wave(uu(:, :, :, 4 + c), ikx, iky, ikz, db0 * a * dperp(c), ph1(idx), (1.0_dp, 0.0_dp))
                        call add_wave(uu(:, :, :, 1 + c), ikx, iky, ikz, corr * dv0 * a * dperp(c), ph1(idx), &
                                      (1.0_dp, 0.0_dp))
                    enddo
                    call add_wave(uu(:, :, :, 1), ikx, iky, ikz, drho0 * a, ph2(idx), (1.0_dp, 0.0_dp))
                enddo
            enddo
        enddo
    end subroutine isotropic_modes

    ! ipert = 8 (mhdinit.f90:931-1020): a switchback, a localised reversal of the field along x
    ! centred in the box, added to the background (which should carry no field of its own).
    ! The x component (the reference's "Bz" of the structure) is B0 + dB g(r) h(x) with the
    ! Gaussian ring g(r) = exp(-((r - R0)/R1)^2) in the y-z radius r and h(x) = exp(-(x/H)^2);
    ! the radial component Br follows from div B = 0, and the azimuthal one fills |B| up to
    ! B0 where it can (elsewhere the pressure is lowered to keep p + B^2/2 uniform). Beyond
    ! r = Rm both are tapered with exp(-((r - Rm)/Rd)^2), the x component corrected with q so
    ! that the field stays solenoidal. The velocity is then anti-correlated, u -= B/sqrt(rho).
    subroutine switchback
        integer :: ix, iy, iz
        real(dp) :: xs, ys, zs, r, gring, gx, bx_core, bx, br, bphi, cconst, phi, taper, dtaper, q, sflux
        real(dp) :: b_y, b_z
        cconst = -(exp(-(R0_SB / R1_SB)**2) + sqrt(pi) * R0_SB / R1_SB * erf(R0_SB / R1_SB))
        do iz = 1, nz
            zs = zgrid(iz) - Lz / 2
            do iy = 1, ny
                ys = ygrid(iy) - Ly / 2
                r = sqrt(ys**2 + zs**2)
                gring = exp(-((r - R0_SB) / R1_SB)**2)
                do ix = 1, nx
                    xs = xgrid(ix) - Lx / 2
                    gx = exp(-(xs / H_SB)**2)
                    bx_core = B0_SB + dB_SB * gring
                    bx = B0_SB + dB_SB * exp(-((r - R0_SB) / R1_SB)**2 - (xs / H_SB)**2)
                    if (r < 1e-8_dp) then
                        phi = 0
                        br = 0
                        bphi = 0
                    else
                        phi = atan2(zs, ys)
                        ! the radial flux function of the ring, r Br = -dB R1^2 x/H^2 gx sflux
                        sflux = gring + sqrt(pi) * R0_SB / R1_SB * erf((R0_SB - r) / R1_SB) + cconst
                        br = -dB_SB * (xs * R1_SB**2) / (r * H_SB**2) * gx * sflux
                        if (r < Rm_SB) then
                            taper = 1
                            dtaper = 0
                            q = 1
                        else
                            taper = exp(-((r - Rm_SB) / Rd_SB)**2)
                            dtaper = -2 * (r - Rm_SB) / Rd_SB**2 * taper
                            q = bx_core / bx * (1 + 1 / bx_core * (taper * (bx - bx_core) - &
                                (dtaper * dB_SB * R1_SB**2 / r * sflux * 0.5_dp * (gx - 1))))
                        endif
                        br = br * taper
                        bx = bx * q
                        if (bx**2 + br**2 < B0_SB**2) then
                            bphi = sqrt(B0_SB**2 - br**2 - bx**2)
                        else
                            bphi = 0
                            uu(ix, iy, iz, 8) = uu(ix, iy, iz, 8) - 0.5_dp * (bx**2 + br**2 - B0_SB**2)
                        endif
                    endif
                    b_y = br * cos(phi) - bphi * sin(phi)
                    b_z = br * sin(phi) + bphi * cos(phi)
                    uu(ix, iy, iz, 5) = uu(ix, iy, iz, 5) + bx
                    uu(ix, iy, iz, 6) = uu(ix, iy, iz, 6) + b_y
                    uu(ix, iy, iz, 7) = uu(ix, iy, iz, 7) + b_z
                enddo
            enddo
        enddo
        uu(:, :, :, 2) = uu(:, :, :, 2) - uu(:, :, :, 5) / sqrt(uu(:, :, :, 1))
        uu(:, :, :, 3) = uu(:, :, :, 3) - uu(:, :, :, 6) / sqrt(uu(:, :, :, 1))
        uu(:, :, :, 4) = uu(:, :, :, 4) - uu(:, :, :, 7) / sqrt(uu(:, :, :, 1))
    end subroutine switchback

    ! initial_calc_conserve_variable (mhdinit.f90:1038-1056)
    subroutine initial_calc_conserve_variable
        uu_prim(:, :, :, 1:3) = uu(:, :, :, 2:4)
        uu_prim(:, :, :, 4) = uu(:, :, :, 8)
        uu(:, :, :, 2) = uu(:, :, :, 1) * uu_prim(:, :, :, 1)
        uu(:, :, :, 3) = uu(:, :, :, 1) * uu_prim(:, :, :, 2)
        uu(:, :, :, 4) = uu(:, :, :, 1) * uu_prim(:, :, :, 3)
        uu(:, :, :, 8) = uu_prim(:, :, :, 4) / (adiabatic_index - 1) + &
            0.5_dp * (uu(:, :, :, 1) * ((uu_prim(:, :, :, 1))**2 + &
            (uu_prim(:, :, :, 2))**2 + (uu_prim(:, :, :, 3))**2) + &
            (uu(:, :, :, 5))**2 + (uu(:, :, :, 6))**2 + (uu(:, :, :, 7))**2)
    end subroutine initial_calc_conserve_variable

    character(len=10) function out_name(n)
        integer, intent(in) :: n
        write(out_name, '(a3,i3.3,a4)') 'out', n, '.dat'
    end function out_name

    ! output_uu (mhdoutput.f90:72-131): [int32 4][float32 t][int32 4] then the
    ! raw (nx,ny,nz,8) FP64 array at byte 13, primitives unless output_primitive is false
    ! The payload (rho, rho u or u, B, e or p) is formed on the device and copied into pinned
    ! memory by laps_stage_output; output_uu stages and writes at once.
    subroutine output_uu(n, t)
        integer, intent(in) :: n
        real(dp), intent(in) :: t
        call laps_check(laps_stage_output(ctx, merge(1_c_int, 0_c_int, output_primitive)), 'laps_stage_output')
        call write_staged(n, t)
    end subroutine output_uu

    subroutine write_staged(n, t)
        integer, intent(in) :: n
        real(dp), intent(in) :: t
        type(c_ptr) :: pw
        real(dp), pointer :: w(:, :, :, :)
        real(dp) :: ts, dts
        integer(c_int64_t) :: ist
        integer :: u
        call laps_check(laps_staged_output(ctx, pw, ts, dts, ist), 'laps_staged_output')
        call c_f_pointer(pw, w, [nx, ny, nz, 8])
        open(newunit=u, file=out_name(n), access='stream', form='unformatted', status='replace', action='write')
        write(u) 4_c_int32_t, real(t, 4), 4_c_int32_t
        write(u) w
        close(u)
    end subroutine write_staged

    ! read_restart (restart.f90:17-63): primitives of out{n_start}.dat; time is float32
    subroutine read_restart(n, t)
        integer, intent(in) :: n
        real(dp), intent(out) :: t
        integer(c_int32_t) :: m1, m2
        real(4) :: t4
        integer :: u
        open(newunit=u, file=out_name(n), access='stream', form='unformatted', status='old', action='read')
        read(u) m1, t4, m2
        read(u) uu
        close(u)
        t = real(t4, dp)
        write(*, '(1x,a,f10.4)') 'Restart from time = ', t
    end subroutine read_restart

    subroutine fortran_record(u, vals)
        integer, intent(in) :: u
        real(4), intent(in) :: vals(:)
        write(u) int(4 * size(vals), c_int32_t), vals, int(4 * size(vals), c_int32_t)
    end subroutine fortran_record

    ! output_grid (mhdoutput.f90:51-60)
    subroutine output_grid
        integer :: u
        open(newunit=u, file='grid.dat', access='stream', form='unformatted', status='replace', action='write')
        call fortran_record(u, [real(nx, 4), real(ny, 4), real(nz, 4)])
        call fortran_record(u, [real(xgrid, 4), real(ygrid, 4), real(zgrid, 4)])
        close(u)
    end subroutine output_grid

    ! output_parallel_info (mhdoutput.f90:62-69): one task on one GPU
    subroutine output_parallel_info
        integer :: u
        open(newunit=u, file='parallel_info.dat', access='stream', form='unformatted', status='replace', &
            action='write')
        call fortran_record(u, [1.0, 1.0, 1.0, 8.0])
        close(u)
    end subroutine output_parallel_info

    ! calc_max_divB + output_rms + output_AEB (mhd.f90:183-192; mhdrms.f90:39-51; AEBmod.f90:75-85)
    subroutine output_rms_all(t)
        real(dp), intent(in) :: t
        real(dp) :: ave(8), var(8), ru2(3), tt, ddt, rr, bud(7)
        integer(c_int64_t) :: ist
        integer(c_int32_t) :: st
        integer :: u
        logical :: ex
        call laps_check(laps_max_divb(ctx, max_divb), 'laps_max_divb')
        write(*, '(6x,a,2x,f10.4,a,1pe10.2,a,1pe12.4)') 'OUTPUT RMS at time:', t, ', max(div B) = ', max_divb, &
            ', dt = ', dt
        call laps_check(laps_moments(ctx, ave, var, ru2), 'laps_moments')
        inquire(file='rms.dat', exist=ex)
        if (ex) then
            open(newunit=u, file='rms.dat', status='old', action='write', position='append')
        else
            open(newunit=u, file='rms.dat', status='replace', action='write')
        endif
        write(u, '(f12.6,2x,19(1pe16.8))') t, ave, var, ru2
        close(u)
        call laps_check(laps_get_scalars(ctx, tt, ddt, rr, ist, st), 'laps_get_scalars')
        inquire(file='EBM_info.dat', exist=ex)
        if (ex) then
            open(newunit=u, file='EBM_info.dat', status='old', action='write', position='append')
        else
            open(newunit=u, file='EBM_info.dat', status='replace', action='write')
        endif
        write(u, '(3(1pe16.8))') t, rr, Ur0
        close(u)
        ! budgets (no reference counterpart): <e>, kinetic, magnetic, internal energy,
        ! cross, magnetic and current helicity, one line per rms output
        call laps_check(laps_budgets(ctx, bud), 'laps_budgets')
        inquire(file='budgets.dat', exist=ex)
        if (ex) then
            open(newunit=u, file='budgets.dat', status='old', action='write', position='append')
        else
            open(newunit=u, file='budgets.dat', status='replace', action='write')
        endif
        write(u, '(f12.6,2x,7(1pe20.12))') t, bud
        close(u)
    end subroutine output_rms_all

    ! write_log (mhd.f90:431-457)
    subroutine write_log
        real(dp) :: ct
        integer :: hour, minute, second, u
        call system_clock(clock_now)
        ct = real(clock_now - clock_start, dp) / clock_rate
        hour = floor(ct / 3600.0_dp)
        minute = floor((ct / 3600.0_dp - hour) * 60)
        second = floor(((ct / 3600.0_dp - hour) * 60 - minute) * 60)
        open(newunit=u, file='log', status='replace', action='write')
        write(u, '(3x,a,f8.4)') 'Simulation time:', time
        write(u, *) 'dt:', dt
        write(u, '(3x,a,f15.2)') 'Real time (sec):', ct
        write(u, '(3x,a,i3,a,i3,a,i3,a)') 'Real time (hh,mm,ss):', hour, 'h', minute, 'm', second, 's'
        write(u, '(3x,a,i8)') 'Iterations     :', istep
        write(u, *) 'tasks: ', 1
        close(u)
    end subroutine write_log

    ! write_initial (mhd.f90:459-519)
    subroutine write_initial
        write(*, *) '--------------------------------------------------'
        write(*, '(5x,a8,f10.2)') 'cfl = ', cfl
        write(*, '(5x,a8,1pe10.2)') 'dt = ', dt
        write(*, '(3(3x,a6,i6))') 'nx = ', nx, 'ny = ', ny, 'nz = ', nz
        write(*, '(3x,a,a)') 'device: one MI355X, precision = ', trim(precision)
        if (output_primitive) then
            write(*, *) 'Output primitive variables...'
        else
            write(*, *) 'Output conserved variables...'
        endif
        if (if_resis) then
            if (if_resis_exp) then
                write(*, '(1x,a,1pe10.2)') 'Resistivity: ON. Explicit method. eta = ', resistivity
            else
                write(*, '(1x,a,1pe10.2)') 'Resistivity: ON. Implicit method. eta = ', resistivity
            endif
        else
            write(*, '(1x,a)') 'Resistivity: OFF.'
        endif
        if (if_visc) then
            if (if_visc_exp) then
                write(*, '(1x,a,1pe10.2)') 'Viscosity: ON. Explicit method. nu = ', viscosity
            else
                write(*, '(1x,a,1pe10.2)') 'Viscosity: ON. Implicit method. nu = ', viscosity
            endif
        else
            write(*, '(1x,a)') 'Viscosity: OFF.'
        endif
        if (if_AEB) then
            write(*, '(1x,a,1pe10.2,a,1pe10.2)') 'Expanding Box: ON. Ur0 = ', Ur0, ', R0 = ', radius0
        else
            write(*, '(1x,a)') 'Expanding Box: OFF.'
        endif
        if (if_hall) then
            write(*, '(1x,a,1pe10.2)') 'Hall term: ON. di = ', ion_inertial_length
        else
            write(*, '(1x,a)') 'Hall term: OFF.'
        endif
        if (dealias_option == 1) then
            write(*, '(1x,a)') 'Use circular-padding for dealiasing.'
        else if (dealias_option == 2) then
            write(*, '(1x,3(a,f10.4))') 'Use smoothing function for dealiasing. afx = ', afx, ', afy = ', afy, &
                ', afz = ', afz
        endif
        write(*, *) '--------------------------------------------------'
    end subroutine write_initial
end program mhd_host
