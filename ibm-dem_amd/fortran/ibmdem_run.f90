!------------------------------------------------------------------------------
! ibmdem_run: the reference's time loop (main.F90:51-56) over a snapshot.
!
!   ibmdem_run IN OUT NSTEPS [ops|step]
!       read the state from snapshot IN, advance NSTEPS coupled steps and
!       write the state to snapshot OUT. "ops" (default) calls
!       FirstOrderForwardEuler(), which issues the reference's subroutines
!       one by one; "step" calls the library's whole-step entry point.
!   ibmdem_run --sizes
!       print c_sizeof of the bind(C) types, one per line.
!------------------------------------------------------------------------------
program ibmdem_run
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: output_unit, error_unit, int64
    use ibmdem_api
    use ibmdem_host
    implicit none
    character(len=4096) :: inPath, outPath, arg
    character(len=8) :: mode
    integer :: nsteps, i, nargs
    integer(int64) :: t0, t1, rate
    type(ibmdem_case) :: cs
    type(ibmdem_markers) :: mk
    type(ibmdem_bodies) :: bd
    type(ibmdem_solve_info) :: si

    nargs = command_argument_count()
    if (nargs >= 1) then
        call get_command_argument(1, arg)
        if (trim(arg) == "--sizes") then
            write(output_unit, '(i0)') c_sizeof(cs), c_sizeof(mk), c_sizeof(bd), c_sizeof(si)
            stop
        end if
    end if
    if (nargs < 3) then
        write(error_unit, '(a)') "usage: ibmdem_run IN OUT NSTEPS [ops|step] | --sizes"
        error stop 2
    end if
    call get_command_argument(1, inPath)
    call get_command_argument(2, outPath)
    call get_command_argument(3, arg)
    read(arg, *) nsteps
    mode = "ops"
    if (nargs >= 4) call get_command_argument(4, mode)
    if (mode /= "ops" .and. mode /= "step") then
        write(error_unit, '(a,a)') "unknown mode ", trim(mode)
        error stop 2
    end if

    call ReadSnapshot(trim(inPath))
    call check(ibmdem_sync(ibmCtx), "ibmdem_sync")
    call system_clock(t0, rate)
    do i = 1, nsteps
        if (mode == "ops") then
            call FirstOrderForwardEuler()
        else
            call check(ibmdem_step(ibmCtx, solveInfo), "ibmdem_step")
        end if
    end do
    call check(ibmdem_sync(ibmCtx), "ibmdem_sync")
    call system_clock(t1)
    call WriteSnapshot(trim(outPath))
    write(output_unit, '(a,i0,a,a,a,f0.3,a,i0,a,es10.3)') "ibmdem_run: ", nsteps, " steps (", trim(mode), &
        ") in ", real(t1 - t0) / real(rate), " s; last solve ", solveInfo%iterations, &
        " CG iterations, residual ", solveInfo%rel_residual
    call FinalizeIBMDEM()
end program ibmdem_run
