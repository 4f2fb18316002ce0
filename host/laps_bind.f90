! laps_bind.f90 -- iso_c_binding interface to the C ABI of include/laps.h.
!
! Each interface binds one entry point of liblaps_amd.so; the derived type
! laps_params mirrors the C struct field by field (bind(C) gives the same
! layout: 4-byte integers and 8-byte reals in declaration order).
module laps_c
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: LAPS_OK = 0
    integer(c_int), parameter :: LAPS_FP64 = 0, LAPS_MIXED = 1
    integer(c_int32_t), parameter :: LAPS_RUNNING = 0, LAPS_DT_STOP = 1

    type, bind(C) :: laps_params
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: precision
        real(c_double) :: Lx, Ly, Lz
        real(c_double) :: adiabatic_index
        real(c_double) :: cfl
        integer(c_int32_t) :: dealias_option
        integer(c_int32_t) :: if_resis, if_resis_exp
        integer(c_int32_t) :: if_visc, if_visc_exp
        integer(c_int32_t) :: if_conserve_background
        real(c_double) :: afx, afy, afz
        real(c_double) :: resistivity, viscosity
        integer(c_int32_t) :: if_hall, if_aeb
        integer(c_int32_t) :: if_corotating, device
        real(c_double) :: ion_inertial_length
        real(c_double) :: radius0, Ur0
        real(c_double) :: corotating_angle
    end type laps_params

    interface
        subroutine laps_default_params(p) bind(C, name='laps_default_params')
            import :: laps_params
            type(laps_params), intent(out) :: p
        end subroutine

        integer(c_int) function laps_create(p, ctx) bind(C, name='laps_create')
            import :: laps_params, c_int, c_ptr
            type(laps_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function

        subroutine laps_destroy(ctx) bind(C, name='laps_destroy')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine

        type(c_ptr) function laps_last_error() bind(C, name='laps_last_error')
            import :: c_ptr
        end function

        integer(c_int) function laps_set_state(ctx, uu, uu_prim) bind(C, name='laps_set_state')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: uu(*), uu_prim(*)
        end function

        integer(c_int) function laps_set_time(ctx, time, dt_prev) bind(C, name='laps_set_time')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: time, dt_prev
        end function

        integer(c_int) function laps_cfl(ctx, dtmin_raw, dt_next) bind(C, name='laps_cfl')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: dtmin_raw, dt_next
        end function

        integer(c_int) function laps_step(ctx, nsteps) bind(C, name='laps_step')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function

        integer(c_int) function laps_get_scalars(ctx, time, dt, radius, istep, status) &
                bind(C, name='laps_get_scalars')
            import :: c_ptr, c_int, c_double, c_int64_t, c_int32_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: time, dt, radius
            integer(c_int64_t), intent(out) :: istep
            integer(c_int32_t), intent(out) :: status
        end function

        integer(c_int) function laps_get_state(ctx, uu, uu_prim) bind(C, name='laps_get_state')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: uu(*), uu_prim(*)
        end function

        integer(c_int) function laps_stage_output(ctx, primitive) bind(C, name='laps_stage_output')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: primitive
        end function

        integer(c_int) function laps_staged_output(ctx, payload, time, dt, istep) bind(C, name='laps_staged_output')
            import :: c_ptr, c_int, c_double, c_int64_t
            type(c_ptr), value :: ctx
            type(c_ptr), intent(out) :: payload
            real(c_double), intent(out) :: time, dt
            integer(c_int64_t), intent(out) :: istep
        end function

        integer(c_int) function laps_moments(ctx, ave, var, rho_u2) bind(C, name='laps_moments')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: ave(8), var(8), rho_u2(3)
        end function

        integer(c_int) function laps_budgets(ctx, out) bind(C, name='laps_budgets')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(7)
        end function

        integer(c_int) function laps_max_divb(ctx, out) bind(C, name='laps_max_divb')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out
        end function

        ! ensembles of independent boxes (no counterpart in the reference): params is an
        ! array of nbox laps_params; laps_ens_box returns box i's context
        integer(c_int) function laps_ens_create(nbox, params, ens) bind(C, name='laps_ens_create')
            import :: c_int, c_ptr, laps_params
            integer(c_int), value :: nbox
            type(laps_params), intent(in) :: params(*)
            type(c_ptr), intent(out) :: ens
        end function

        subroutine laps_ens_destroy(ens) bind(C, name='laps_ens_destroy')
            import :: c_ptr
            type(c_ptr), value :: ens
        end subroutine

        type(c_ptr) function laps_ens_box(ens, i) bind(C, name='laps_ens_box')
            import :: c_ptr, c_int
            type(c_ptr), value :: ens
            integer(c_int), value :: i
        end function

        integer(c_int) function laps_ens_step(ens, nsteps) bind(C, name='laps_ens_step')
            import :: c_ptr, c_int
            type(c_ptr), value :: ens
            integer(c_int), value :: nsteps
        end function

        integer(c_int) function laps_ens_sync(ens) bind(C, name='laps_ens_sync')
            import :: c_ptr, c_int
            type(c_ptr), value :: ens
        end function

        ! learned closure (no counterpart in the reference): params holds
        ! LAPS_CLOSURE_NPARAM = 7062 reals; a null pointer switches it off
        integer(c_int) function laps_closure_set(ctx, params, nparams) bind(C, name='laps_closure_set')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            type(c_ptr), value :: params
            integer(c_int), value :: nparams
        end function
    end interface

contains

    ! stop with the library's message when a call fails
    subroutine laps_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == LAPS_OK) return
        call c_f_pointer(laps_last_error(), msg, [512])
        n = 0
        do while (n < 512)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        enddo
        write(*, '(1x,a,a,a,i0,a,512a)') 'laps error in ', what, ' (', rc, '): ', msg(1:n)
        error stop 2
    end subroutine laps_check
end module laps_c
