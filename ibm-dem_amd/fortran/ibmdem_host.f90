!------------------------------------------------------------------------------
! ibmdem_host: the Fortran host side of the MI355X coupled IBM-DEM step.
!
! The subroutines keep the reference's names and argument lists (fluid.F90,
! poisson.F90, boundary.F90, ZJS_VOLUMEFRACTION_DEM.F90, ZJS_MPI_DEM.F90,
! ZJS_COLLISION_DEM.F90, ZJS_MS_DEM.F90, ZJS_EVOLUTION_DEM.F90,
! couplesolverCylinder.F90) and call the HIP kernels through ibmdem_api, so
! a reference driver such as main.F90:51's time loop keeps calling
! FirstOrderForwardEuler() unchanged. Fields, markers and bodies live in
! device memory owned by the context; the module's host arrays are mirrors
! that UploadState/DownloadState copy in and out.
!
! Host mirror layouts (they are the C ABI's layouts read column-major):
!   u, v, w, P         (nx, ny, nz)   interior cells, i fastest
!   partX, partV, ...  (M, 3)         component-major marker vectors
!   msX, msV, ...      (B, 3)         component-major body vectors
!   msRot, ...         (9, B)         row-major 3x3 per body (mat(1:9))
!   outflowX/Y/Z       (ny,nz) (nx,nz) (nx,ny)  outflow planes of sides 2,
!                                     4, 6 (size 0 where the side is not
!                                     an outflow)
!
! Errors: the reference checks no return code. Here every call's status is
! kept in ibmStatus; with stopOnError (the default) a failing call prints
! ibmdem_last_error and stops with the status as its code.
!------------------------------------------------------------------------------
module ibmdem_host
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: error_unit
    use ibmdem_api
    implicit none

    type(c_ptr), save :: ibmCtx = c_null_ptr
    type(ibmdem_case), save, target :: ibmCase
    type(ibmdem_solve_info), save :: solveInfo
    integer(c_int), save :: ibmStatus = IBMDEM_OK
    logical, save :: stopOnError = .true.

    ! host mirrors
    real(c_double), allocatable, target, save :: u(:,:,:), v(:,:,:), w(:,:,:), P(:,:,:)
    real(c_double), allocatable, target, save :: partX(:,:), partV(:,:), partFdrag(:,:), partFcol(:,:)
    real(c_double), allocatable, target, save :: partVol(:), partM(:), partRadEff(:)
    integer(c_int32_t), allocatable, target, save :: partMS(:)
    real(c_double), allocatable, target, save :: msX(:,:), msV(:,:), msVold(:,:)
    real(c_double), allocatable, target, save :: msOmega(:,:), msAngMom(:,:), msAngMomOld(:,:)
    real(c_double), allocatable, target, save :: msRot(:,:), msRotGlo(:,:), msInertia(:,:)
    real(c_double), allocatable, target, save :: msM(:), msRho(:)
    integer(c_int32_t), allocatable, target, save :: msFixed(:)
    real(c_double), allocatable, target, save :: msFdrag(:,:), msFmom(:,:), msFcol(:,:)
    real(c_double), allocatable, target, save :: msFmomCol(:,:), msFhydro(:,:), msFmomHydro(:,:)
    real(c_double), allocatable, target, save :: outflowX(:,:), outflowY(:,:), outflowZ(:,:)

    character(len=8), parameter :: SNAPSHOT_MAGIC = "IBMDEMS1"

contains

    !--------------------------------------------------------------------------
    ! status handling
    !--------------------------------------------------------------------------
    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        ibmStatus = rc
        if (rc == IBMDEM_OK) return
        if (stopOnError) then
            write(error_unit, '(a,a,i0,a,a)') where, ": status ", rc, ": ", ibmdem_error_text(ibmCtx)
            error stop int(rc)
        end if
    end subroutine

    !--------------------------------------------------------------------------
    ! lifetime: replaces allocate.F90, AllocateDEM (ZJS_INIT_DEM.F90:80-114)
    ! and the state set-up of Fluid/InitializationDEM
    !--------------------------------------------------------------------------
    subroutine InitializeIBMDEM(cs)
        type(ibmdem_case), intent(in) :: cs
        if (c_associated(ibmCtx)) call FinalizeIBMDEM()
        ibmCase = cs
        call AllocateHostMirrors()
        call check(ibmdem_create(ibmCase, ibmCtx), "ibmdem_create")
    end subroutine

    subroutine FinalizeIBMDEM()
        integer(c_int) :: rc
        if (c_associated(ibmCtx)) then
            rc = ibmdem_destroy(ibmCtx)
            ibmCtx = c_null_ptr
            call check(rc, "ibmdem_destroy")
        end if
    end subroutine

    subroutine AllocateHostMirrors()
        integer :: nx, ny, nz, m, b
        nx = ibmCase%nx; ny = ibmCase%ny; nz = ibmCase%nz
        m = ibmCase%n_markers; b = ibmCase%n_bodies
        if (allocated(u)) call FreeHostMirrors()
        allocate(u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), P(nx,ny,nz))
        u = 0d0; v = 0d0; w = 0d0; P = 0d0
        allocate(partX(m,3), partV(m,3), partFdrag(m,3), partFcol(m,3))
        allocate(partVol(m), partM(m), partRadEff(m), partMS(m))
        partX = 0d0; partV = 0d0; partFdrag = 0d0; partFcol = 0d0
        partVol = 0d0; partM = 0d0; partRadEff = 0d0; partMS = 0
        allocate(msX(b,3), msV(b,3), msVold(b,3), msOmega(b,3), msAngMom(b,3), msAngMomOld(b,3))
        allocate(msRot(9,b), msRotGlo(9,b), msInertia(9,b), msM(b), msRho(b), msFixed(b))
        allocate(msFdrag(b,3), msFmom(b,3), msFcol(b,3), msFmomCol(b,3), msFhydro(b,3), msFmomHydro(b,3))
        msX = 0d0; msV = 0d0; msVold = 0d0; msOmega = 0d0; msAngMom = 0d0; msAngMomOld = 0d0
        msRot = 0d0; msRotGlo = 0d0; msInertia = 0d0; msM = 0d0; msRho = 0d0; msFixed = 0
        msFdrag = 0d0; msFmom = 0d0; msFcol = 0d0; msFmomCol = 0d0; msFhydro = 0d0; msFmomHydro = 0d0
        allocate(outflowX(merge(ny, 0, ibmCase%btype(2) == IBMDEM_BC_OUTFLOW), nz))
        allocate(outflowY(merge(nx, 0, ibmCase%btype(4) == IBMDEM_BC_OUTFLOW), nz))
        allocate(outflowZ(merge(nx, 0, ibmCase%btype(6) == IBMDEM_BC_OUTFLOW), ny))
        outflowX = 0d0; outflowY = 0d0; outflowZ = 0d0
    end subroutine

    subroutine FreeHostMirrors()
        deallocate(u, v, w, P)
        deallocate(partX, partV, partFdrag, partFcol, partVol, partM, partRadEff, partMS)
        deallocate(msX, msV, msVold, msOmega, msAngMom, msAngMomOld)
        deallocate(msRot, msRotGlo, msInertia, msM, msRho, msFixed)
        deallocate(msFdrag, msFmom, msFcol, msFmomCol, msFhydro, msFmomHydro)
        deallocate(outflowX, outflowY, outflowZ)
    end subroutine

    ! C-ABI descriptors over the host mirrors. An empty array (no markers or
    ! no bodies) still has a valid address in Fortran; the library reads
    ! nothing from it.
    function MarkerView() result(mk)
        type(ibmdem_markers) :: mk
        if (size(partVol) == 0) return
        mk%x = c_loc(partX); mk%v = c_loc(partV)
        mk%fdrag = c_loc(partFdrag); mk%fcol = c_loc(partFcol)
        mk%vol = c_loc(partVol); mk%mass = c_loc(partM)
        mk%radeff = c_loc(partRadEff); mk%body = c_loc(partMS)
    end function

    function BodyView() result(bd)
        type(ibmdem_bodies) :: bd
        if (size(msM) == 0) return
        bd%x = c_loc(msX); bd%v = c_loc(msV); bd%vold = c_loc(msVold)
        bd%omega = c_loc(msOmega); bd%angmom = c_loc(msAngMom)
        bd%angmom_old = c_loc(msAngMomOld)
        bd%rot = c_loc(msRot); bd%rot_glo = c_loc(msRotGlo)
        bd%inertia = c_loc(msInertia)
        bd%mass = c_loc(msM); bd%rho = c_loc(msRho); bd%fixed = c_loc(msFixed)
        bd%fdrag = c_loc(msFdrag); bd%fmom = c_loc(msFmom)
        bd%fcol = c_loc(msFcol); bd%fmomcol = c_loc(msFmomCol)
        bd%fhydro = c_loc(msFhydro); bd%fmomhydro = c_loc(msFmomHydro)
    end function

    ! host mirrors -> device (bodies first: markers are validated against them)
    subroutine UploadState()
        if (ibmCase%n_bodies > 0) call check(ibmdem_upload_bodies(ibmCtx, BodyView()), "ibmdem_upload_bodies")
        if (ibmCase%n_markers > 0) call check(ibmdem_upload_markers(ibmCtx, MarkerView()), "ibmdem_upload_markers")
        call check(ibmdem_upload_fields(ibmCtx, c_loc(u), c_loc(v), c_loc(w), c_loc(P)), &
                   "ibmdem_upload_fields")
        ! after the fields, whose upload resets the outflow planes
        if (size(outflowX) > 0) call check(ibmdem_upload_outflow(ibmCtx, 1, c_loc(outflowX)), "ibmdem_upload_outflow")
        if (size(outflowY) > 0) call check(ibmdem_upload_outflow(ibmCtx, 3, c_loc(outflowY)), "ibmdem_upload_outflow")
        if (size(outflowZ) > 0) call check(ibmdem_upload_outflow(ibmCtx, 5, c_loc(outflowZ)), "ibmdem_upload_outflow")
    end subroutine

    ! device -> host mirrors (the reference's output copies, io.F90)
    subroutine DownloadState()
        if (ibmCase%n_bodies > 0) call check(ibmdem_download_bodies(ibmCtx, BodyView()), "ibmdem_download_bodies")
        if (ibmCase%n_markers > 0) call check(ibmdem_download_markers(ibmCtx, MarkerView()), "ibmdem_download_markers")
        call check(ibmdem_download_fields(ibmCtx, c_loc(u), c_loc(v), c_loc(w), c_loc(P)), &
                   "ibmdem_download_fields")
        if (size(outflowX) > 0) call check(ibmdem_download_outflow(ibmCtx, 1, c_loc(outflowX)), "ibmdem_download_outflow")
        if (size(outflowY) > 0) call check(ibmdem_download_outflow(ibmCtx, 3, c_loc(outflowY)), "ibmdem_download_outflow")
        if (size(outflowZ) > 0) call check(ibmdem_download_outflow(ibmCtx, 5, c_loc(outflowZ)), "ibmdem_download_outflow")
    end subroutine

    !--------------------------------------------------------------------------
    ! the coupled step in the reference's order (fluid.F90:252-327)
    !--------------------------------------------------------------------------
    subroutine FirstOrderForwardEuler()
        integer :: i, Ns
        Ns = ibmCase%n_mdf
        ! uold = u; fdrag = 0; fx = fy = fz = 0 (fluid.F90:258-263)
        call check(ibmdem_begin_step(ibmCtx), "ibmdem_begin_step")
        do i = 1, Ns
            call UpdateFluidVelocityOnLagrangianPoints()
            call UpdateParcelInfoInGhostRegion()
            call UpdateMomentumSourceTerm()
            call VelocityPostGuess()
            call FillGhostRegionForFluid()
            call Boundary()
        end do
        call VelocityGuess(1d0, 0d0, 1d0)
        call FillGhostRegionForFluid()
        call Boundary()

        call ModifyBulkVelocity()
        call FillGhostRegionForFluid()
        call Boundary()

        call VelocityDivergence(1d0)
        call FillGhostRegionForFluid()
        call Boundary()

        call SetPressureVector()
        call PoissonSolver()
        call PressureCorrection()
        call FillGhostRegionForFluid()
        call Boundary()

        call VelocityCorrection(1d0)
        call FillGhostRegionForFluid()
        call Boundary()

        call CloudEvolutionMS()
    end subroutine

    ! UpdateFluidVelocityOnLagrangianPoints (ZJS_VOLUMEFRACTION_DEM.F90:2-36),
    ! weight lists (ZJS_UPDATE_LIST_DEM.F90:323-416) computed inline
    subroutine UpdateFluidVelocityOnLagrangianPoints()
        call check(ibmdem_ibm_interpolate(ibmCtx), "ibmdem_ibm_interpolate")
    end subroutine

    ! UpdateParcelInfoInGhostRegion (ZJS_UPDATE_LIST_DEM.F90:179-246,
    ! ZJS_MPI_DEM.F90:173-269): periodic images and wall mirrors are
    ! generated inside the spreading kernel, so there is nothing to refresh.
    subroutine UpdateParcelInfoInGhostRegion()
    end subroutine

    ! UpdateMomentumSourceTerm (ZJS_VOLUMEFRACTION_DEM.F90:38-67): f_old <- f,
    ! f <- -sum fdrag vol delta, as a deterministic gather; the footprint
    ! velocity update of VelocityPostGuess is fused into the same kernel.
    subroutine UpdateMomentumSourceTerm()
        call check(ibmdem_ibm_spread(ibmCtx), "ibmdem_ibm_spread")
    end subroutine

    ! VelocityPostGuess (fluid.F90:135-153): u += dt (f - f_old) is applied
    ! by UpdateMomentumSourceTerm on the cells the markers touch (f and
    ! f_old are zero elsewhere).
    subroutine VelocityPostGuess()
    end subroutine

    ! VelocityGuess(aRK,bRK,cRK) (fluid.F90:107-133) with ConvectionTensor,
    ! ViscousTensor and PressureGradient (fluid.F90:2-87) fused
    subroutine VelocityGuess(aRK, bRK, cRK)
        real(c_double), intent(in) :: aRK, bRK, cRK
        call check(ibmdem_fluid_predictor(ibmCtx, aRK, bRK, cRK), "ibmdem_fluid_predictor")
    end subroutine

    ! ModifyBulkVelocity (fluid.F90:155-197)
    subroutine ModifyBulkVelocity()
        call check(ibmdem_bulk_forcing(ibmCtx), "ibmdem_bulk_forcing")
    end subroutine

    ! VelocityDivergence(cRK) (fluid.F90:89-105)
    subroutine VelocityDivergence(cRK)
        real(c_double), intent(in) :: cRK
        call check(ibmdem_divergence(ibmCtx, cRK), "ibmdem_divergence")
    end subroutine

    ! SetPressureVector (poisson.F90:107-131): the solver reads velDiv
    ! directly; the pinned row b(1,1,1) is handled inside the solve.
    subroutine SetPressureVector()
    end subroutine

    ! PoissonSolver + GetLocalArray (poisson.F90:144-186): multigrid-
    ! preconditioned CG in place of PETSc KSPSolve; the last solve's
    ! iteration count and residual are kept in solveInfo.
    subroutine PoissonSolver()
        call check(ibmdem_poisson_solve(ibmCtx, solveInfo), "ibmdem_poisson_solve")
    end subroutine

    ! PressureCorrection (fluid.F90:199-211): P += P' is applied together
    ! with the velocity correction by VelocityCorrection.
    subroutine PressureCorrection()
    end subroutine

    ! VelocityCorrection(cRK) (fluid.F90:213-238), with PressureCorrection
    subroutine VelocityCorrection(cRK)
        real(c_double), intent(in) :: cRK
        call check(ibmdem_project(ibmCtx, cRK), "ibmdem_project")
    end subroutine

    ! FillGhostRegionForFluid (boundary.F90:122-147) and Boundary
    ! (fluid.F90:240-250): every kernel applies the periodic and wall rules
    ! by index arithmetic; these calls only validate the context.
    subroutine FillGhostRegionForFluid()
        call check(ibmdem_apply_bc(ibmCtx), "ibmdem_apply_bc")
    end subroutine

    subroutine Boundary()
        call check(ibmdem_apply_bc(ibmCtx), "ibmdem_apply_bc")
    end subroutine

    !--------------------------------------------------------------------------
    ! particles (couplesolverCylinder.F90)
    !--------------------------------------------------------------------------
    ! CloudEvolutionMS (couplesolverCylinder.F90:189-217)
    subroutine CloudEvolutionMS()
        call check(ibmdem_cloud_evolution(ibmCtx), "ibmdem_cloud_evolution")
    end subroutine

    ! CloudEvolutionStepMS (couplesolverCylinder.F90:99-187): one sub-step
    subroutine CloudEvolutionStepMS()
        call check(ibmdem_dem_substep(ibmCtx), "ibmdem_dem_substep")
    end subroutine

    ! CalculateCollisionForce + CalculateSolidBoundaryCollisionForce
    ! (ZJS_COLLISION_DEM.F90:21-148)
    subroutine CalculateCollisionForce()
        call check(ibmdem_dem_contact(ibmCtx), "ibmdem_dem_contact")
    end subroutine

    ! AssembleForceMS / AssembleLocalForceMS (ZJS_MS_DEM.F90:494-519)
    subroutine AssembleForceMS()
        call check(ibmdem_body_assemble(ibmCtx), "ibmdem_body_assemble")
    end subroutine

    ! EvolutionMS (couplesolverCylinder.F90:2-42)
    subroutine EvolutionMS()
        call check(ibmdem_body_integrate(ibmCtx), "ibmdem_body_integrate")
    end subroutine

    ! ResetVelocityAndPositionLagrangianPoints + RotateLagrangianPoints
    ! (couplesolverCylinder.F90:44-74), AdjustParcelPosition
    ! (ZJS_EVOLUTION_DEM.F90:121-150)
    subroutine ResetVelocityAndPositionLagrangianPoints()
        call check(ibmdem_markers_follow(ibmCtx), "ibmdem_markers_follow")
    end subroutine

    ! RebuildAllListOrNot (ZJS_EVOLUTION_DEM.F90:34-41)
    subroutine RebuildAllListOrNot()
        call check(ibmdem_dem_rebuild_lists(ibmCtx), "ibmdem_dem_rebuild_lists")
    end subroutine

    !--------------------------------------------------------------------------
    ! snapshots: the whole state in one little-endian stream file
    !   "IBMDEMS1", int32 sizeof(case), case bytes,
    !   markers x v fdrag fcol (3M each) vol mass radeff (M) body (M int32),
    !   bodies x v vold omega angmom angmom_old (3B each),
    !          rot rot_glo inertia (9B each), mass rho (B), fixed (B int32),
    !          fdrag fmom fcol fmomcol fhydro fmomhydro (3B each),
    !   fields u v w p (nx*ny*nz each),
    !   solve info: int32 iterations, int32 converged, rel_residual, rhs_norm
    !--------------------------------------------------------------------------
    subroutine ReadSnapshot(path)
        character(len=*), intent(in) :: path
        integer :: un, ios
        character(len=8) :: magic
        integer(c_int32_t) :: nbytes
        type(ibmdem_case), target :: cs
        integer(c_int8_t), pointer :: raw(:)
        open(newunit=un, file=path, access="stream", form="unformatted", status="old", &
             action="read", iostat=ios)
        if (ios /= 0) then
            write(error_unit, '(a,a)') "cannot open snapshot ", path
            error stop 2
        end if
        read(un) magic, nbytes
        if (magic /= SNAPSHOT_MAGIC .or. nbytes /= int(c_sizeof(cs), c_int32_t)) then
            write(error_unit, '(a,a)') "not an ibmdem snapshot of this layout: ", path
            error stop 2
        end if
        call c_f_pointer(c_loc(cs), raw, [int(nbytes)])
        read(un) raw
        call InitializeIBMDEM(cs)
        read(un) partX, partV, partFdrag, partFcol, partVol, partM, partRadEff, partMS
        read(un) msX, msV, msVold, msOmega, msAngMom, msAngMomOld
        read(un) msRot, msRotGlo, msInertia, msM, msRho, msFixed
        read(un) msFdrag, msFmom, msFcol, msFmomCol, msFhydro, msFmomHydro
        read(un) u, v, w, P
        read(un) outflowX, outflowY, outflowZ
        read(un) solveInfo%iterations, solveInfo%converged, solveInfo%rel_residual, solveInfo%rhs_norm
        close(un)
        call UploadState()
    end subroutine

    subroutine WriteSnapshot(path)
        character(len=*), intent(in) :: path
        integer :: un
        integer(c_int8_t), pointer :: raw(:)
        call DownloadState()
        call c_f_pointer(c_loc(ibmCase), raw, [int(c_sizeof(ibmCase))])
        open(newunit=un, file=path, access="stream", form="unformatted", status="replace", &
             action="write")
        write(un) SNAPSHOT_MAGIC, int(c_sizeof(ibmCase), c_int32_t)
        write(un) raw
        write(un) partX, partV, partFdrag, partFcol, partVol, partM, partRadEff, partMS
        write(un) msX, msV, msVold, msOmega, msAngMom, msAngMomOld
        write(un) msRot, msRotGlo, msInertia, msM, msRho, msFixed
        write(un) msFdrag, msFmom, msFcol, msFmomCol, msFhydro, msFmomHydro
        write(un) u, v, w, P
        write(un) outflowX, outflowY, outflowZ
        write(un) solveInfo%iterations, solveInfo%converged, solveInfo%rel_residual, solveInfo%rhs_norm
        close(un)
    end subroutine

end module ibmdem_host
