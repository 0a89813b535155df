!------------------------------------------------------------------------------
! ibmdem_api: ISO_C_BINDING declarations of the C ABI in include/ibmdem.h.
!
! Every type here is bind(C) and laid out like its C struct; every interface
! binds one C entry point by its exact name. Nothing in this module holds
! state: the driver module ibmdem_host keeps the context, the way the
! reference keeps its state in module globals (vars.F90, ZJS_VARS.F90).
!------------------------------------------------------------------------------
module ibmdem_api
    use, intrinsic :: iso_c_binding
    implicit none

    ! status codes
    integer(c_int), parameter :: IBMDEM_OK = 0
    integer(c_int), parameter :: IBMDEM_ERR_HIP = 1
    integer(c_int), parameter :: IBMDEM_ERR_ARG = 2
    integer(c_int), parameter :: IBMDEM_ERR_CAPACITY = 3
    integer(c_int), parameter :: IBMDEM_ERR_OUTSIDE = 4
    integer(c_int), parameter :: IBMDEM_ERR_NOT_CONVERGED = 5
    integer(c_int), parameter :: IBMDEM_ERR_NAN = 6
    integer(c_int), parameter :: IBMDEM_ERR_ALLOC = 7
    integer(c_int), parameter :: IBMDEM_ERR_UNSUPPORTED = 8
    ! boundary types: the reference's bType (0 periodic, 1 solid wall)
    integer(c_int32_t), parameter :: IBMDEM_BC_PERIODIC = 0
    integer(c_int32_t), parameter :: IBMDEM_BC_WALL = 1
    integer(c_int32_t), parameter :: IBMDEM_BC_INFLOW = 2
    integer(c_int32_t), parameter :: IBMDEM_BC_OUTFLOW = 3
    ! precision forms
    integer(c_int32_t), parameter :: IBMDEM_FP64 = 0
    integer(c_int32_t), parameter :: IBMDEM_MIXED = 1
    ! time schemes
    integer(c_int32_t), parameter :: IBMDEM_SCHEME_EULER = 0
    integer(c_int32_t), parameter :: IBMDEM_SCHEME_RK3 = 1

    ! every component defaults to 0: the library reads 0 as "use the default"
    ! for the multigrid and tile parameters
    type, bind(C) :: ibmdem_case
        integer(c_int32_t) :: nx = 0, ny = 0, nz = 0
        integer(c_int32_t) :: btype(6) = 0
        real(c_double)     :: dx = 0d0, dy = 0d0, dz = 0d0
        real(c_double)     :: visc = 0d0, rhof = 0d0, dt = 0d0
        real(c_double)     :: cp(3) = 0d0
        integer(c_int32_t) :: bulk_switch(3) = 0
        real(c_double)     :: bulk_target(3) = 0d0
        real(c_double)     :: gravity(3) = 0d0
        integer(c_int32_t) :: n_mdf = 0
        integer(c_int32_t) :: couple_interval = 0
        real(c_double)     :: kn_coeff = 0d0
        real(c_double)     :: restitution = 0d0
        real(c_double)     :: collision_dist = 0d0
        real(c_double)     :: collision_buffer = 0d0
        integer(c_int32_t) :: eta_textbook = 0
        integer(c_int32_t) :: n_markers = 0
        integer(c_int32_t) :: n_bodies = 0
        integer(c_int32_t) :: max_pairs = 0
        integer(c_int32_t) :: precision = 0
        real(c_double)     :: poisson_rtol = 0d0
        integer(c_int32_t) :: poisson_max_iter = 0
        integer(c_int32_t) :: mg_pre = 0, mg_post = 0
        integer(c_int32_t) :: mg_coarse_sweeps = 0
        integer(c_int32_t) :: mg_max_levels = 0
        integer(c_int32_t) :: stencil_tile = 0
        integer(c_int32_t) :: time_scheme = 0
        real(c_double)     :: inflow(3) = 0d0
        integer(c_int32_t) :: reserved(6) = 0
    end type

    ! markers (reference type Parcel), component-major SoA
    type, bind(C) :: ibmdem_markers
        type(c_ptr) :: x = c_null_ptr, v = c_null_ptr
        type(c_ptr) :: fdrag = c_null_ptr, fcol = c_null_ptr
        type(c_ptr) :: vol = c_null_ptr, mass = c_null_ptr
        type(c_ptr) :: radeff = c_null_ptr, body = c_null_ptr
    end type

    ! bodies (reference type MultiSphere), SoA
    type, bind(C) :: ibmdem_bodies
        type(c_ptr) :: x = c_null_ptr, v = c_null_ptr, vold = c_null_ptr
        type(c_ptr) :: omega = c_null_ptr, angmom = c_null_ptr
        type(c_ptr) :: angmom_old = c_null_ptr
        type(c_ptr) :: rot = c_null_ptr, rot_glo = c_null_ptr
        type(c_ptr) :: inertia = c_null_ptr
        type(c_ptr) :: mass = c_null_ptr, rho = c_null_ptr
        type(c_ptr) :: fixed = c_null_ptr
        type(c_ptr) :: fdrag = c_null_ptr, fmom = c_null_ptr
        type(c_ptr) :: fcol = c_null_ptr, fmomcol = c_null_ptr
        type(c_ptr) :: fhydro = c_null_ptr, fmomhydro = c_null_ptr
    end type

    type, bind(C) :: ibmdem_solve_info
        integer(c_int32_t) :: iterations = 0
        integer(c_int32_t) :: converged = 0
        real(c_double)     :: rel_residual = 0d0
        real(c_double)     :: rhs_norm = 0d0
    end type

    interface
        ! ---- lifetime and state --------------------------------------------
        integer(c_int) function ibmdem_create(cs, ctx) bind(C, name="ibmdem_create")
            import :: c_int, c_ptr, ibmdem_case
            type(ibmdem_case), intent(in) :: cs
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function ibmdem_destroy(ctx) bind(C, name="ibmdem_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function ibmdem_last_error(ctx) bind(C, name="ibmdem_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_sync(ctx) bind(C, name="ibmdem_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_upload_fields(ctx, u, v, w, p) &
                bind(C, name="ibmdem_upload_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p
        end function
        integer(c_int) function ibmdem_download_fields(ctx, u, v, w, p) &
                bind(C, name="ibmdem_download_fields")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p
        end function
        integer(c_int) function ibmdem_download_outflow(ctx, side, plane) &
                bind(C, name="ibmdem_download_outflow")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, plane
            integer(c_int32_t), value :: side
        end function
        integer(c_int) function ibmdem_upload_outflow(ctx, side, plane) &
                bind(C, name="ibmdem_upload_outflow")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx, plane
            integer(c_int32_t), value :: side
        end function
        integer(c_int) function ibmdem_download_force(ctx, fx, fy, fz) &
                bind(C, name="ibmdem_download_force")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, fx, fy, fz
        end function
        integer(c_int) function ibmdem_download_poisson(ctx, pcorr, rhs) &
                bind(C, name="ibmdem_download_poisson")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, pcorr, rhs
        end function
        integer(c_int) function ibmdem_upload_poisson_rhs(ctx, rhs) &
                bind(C, name="ibmdem_upload_poisson_rhs")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, rhs
        end function
        integer(c_int) function ibmdem_upload_markers(ctx, m) bind(C, name="ibmdem_upload_markers")
            import :: c_int, c_ptr, ibmdem_markers
            type(c_ptr), value :: ctx
            type(ibmdem_markers), intent(in) :: m
        end function
        integer(c_int) function ibmdem_download_markers(ctx, m) bind(C, name="ibmdem_download_markers")
            import :: c_int, c_ptr, ibmdem_markers
            type(c_ptr), value :: ctx
            type(ibmdem_markers), intent(in) :: m
        end function
        integer(c_int) function ibmdem_upload_bodies(ctx, b) bind(C, name="ibmdem_upload_bodies")
            import :: c_int, c_ptr, ibmdem_bodies
            type(c_ptr), value :: ctx
            type(ibmdem_bodies), intent(in) :: b
        end function
        integer(c_int) function ibmdem_download_bodies(ctx, b) bind(C, name="ibmdem_download_bodies")
            import :: c_int, c_ptr, ibmdem_bodies
            type(c_ptr), value :: ctx
            type(ibmdem_bodies), intent(in) :: b
        end function

        ! ---- the coupled step, operator by operator ---------------------------
        integer(c_int) function ibmdem_begin_step(ctx) bind(C, name="ibmdem_begin_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_ibm_interpolate(ctx) bind(C, name="ibmdem_ibm_interpolate")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_ibm_spread(ctx) bind(C, name="ibmdem_ibm_spread")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_fluid_predictor(ctx, a, b, c) &
                bind(C, name="ibmdem_fluid_predictor")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: a, b, c
        end function
        integer(c_int) function ibmdem_bulk_forcing(ctx) bind(C, name="ibmdem_bulk_forcing")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_divergence(ctx, c) bind(C, name="ibmdem_divergence")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: c
        end function
        integer(c_int) function ibmdem_poisson_solve(ctx, info) bind(C, name="ibmdem_poisson_solve")
            import :: c_int, c_ptr, ibmdem_solve_info
            type(c_ptr), value :: ctx
            type(ibmdem_solve_info), intent(inout) :: info
        end function
        integer(c_int) function ibmdem_project(ctx, c) bind(C, name="ibmdem_project")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: c
        end function
        integer(c_int) function ibmdem_apply_bc(ctx) bind(C, name="ibmdem_apply_bc")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! ---- DEM and rigid bodies --------------------------------------------
        integer(c_int) function ibmdem_body_hydro(ctx) bind(C, name="ibmdem_body_hydro")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_dem_contact(ctx) bind(C, name="ibmdem_dem_contact")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_body_assemble(ctx) bind(C, name="ibmdem_body_assemble")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_body_integrate(ctx) bind(C, name="ibmdem_body_integrate")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_markers_follow(ctx) bind(C, name="ibmdem_markers_follow")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_dem_rebuild_lists(ctx) bind(C, name="ibmdem_dem_rebuild_lists")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_dem_substep(ctx) bind(C, name="ibmdem_dem_substep")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function ibmdem_cloud_evolution(ctx) bind(C, name="ibmdem_cloud_evolution")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! ---- whole step ---------------------------------------------------------
        integer(c_int) function ibmdem_step(ctx, info) bind(C, name="ibmdem_step")
            import :: c_int, c_ptr, ibmdem_solve_info
            type(c_ptr), value :: ctx
            type(ibmdem_solve_info), intent(inout) :: info
        end function
        integer(c_int) function ibmdem_run(ctx, n_steps, info) bind(C, name="ibmdem_run")
            import :: c_int, c_ptr, c_int32_t, ibmdem_solve_info
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n_steps
            type(ibmdem_solve_info), intent(inout) :: info
        end function

        ! ---- integer lists -------------------------------------------------------
        integer(c_int) function ibmdem_get_cells(ctx, cell_ijk) bind(C, name="ibmdem_get_cells")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, cell_ijk
        end function
        integer(c_int) function ibmdem_get_pairs(ctx, cap, count, partner, shift) &
                bind(C, name="ibmdem_get_pairs")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: cap
            type(c_ptr), value :: count, partner, shift
        end function

        ! ---- multigrid cycle and the batched form ------------------------------------
        integer(c_int) function ibmdem_get_mg_params(ctx, mg) bind(C, name="ibmdem_get_mg_params")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: mg(3)
        end function
        integer(c_int) function ibmdem_set_mg_params(ctx, mg) bind(C, name="ibmdem_set_mg_params")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(in) :: mg(3)
        end function
        integer(c_int) function ibmdem_batch_create(n, cases, batch) bind(C, name="ibmdem_batch_create")
            import :: c_int, c_ptr, c_int32_t, ibmdem_case
            integer(c_int32_t), value :: n
            type(ibmdem_case), intent(in) :: cases(*)
            type(c_ptr), intent(out) :: batch
        end function
        integer(c_int) function ibmdem_batch_destroy(batch) bind(C, name="ibmdem_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function
        integer(c_int) function ibmdem_batch_context(batch, q, ctx) bind(C, name="ibmdem_batch_context")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: q
            type(c_ptr), intent(out) :: ctx
        end function
        type(c_ptr) function ibmdem_batch_last_error(batch) bind(C, name="ibmdem_batch_last_error")
            import :: c_ptr
            type(c_ptr), value :: batch
        end function
        integer(c_int) function ibmdem_batch_run(batch, n_steps, info) bind(C, name="ibmdem_batch_run")
            import :: c_int, c_ptr, c_int32_t, ibmdem_solve_info
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: n_steps
            type(ibmdem_solve_info), intent(inout) :: info(*)
        end function
        integer(c_int) function ibmdem_batch_autotune(batch, trial_steps) bind(C, name="ibmdem_batch_autotune")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: trial_steps
        end function
        integer(c_int) function ibmdem_batch_set_fused(batch, mode) bind(C, name="ibmdem_batch_set_fused")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: mode
        end function
        integer(c_int) function ibmdem_batch_set_solve_groups(batch, groups) bind(C, name="ibmdem_batch_set_solve_groups")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: groups
        end function
        integer(c_int) function ibmdem_batch_get_solve_groups(batch, groups) bind(C, name="ibmdem_batch_get_solve_groups")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), intent(out) :: groups
        end function
        integer(c_int) function ibmdem_batch_stats(batch, fused_solves, ms) bind(C, name="ibmdem_batch_stats")
            import :: c_int, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: batch
            integer(c_int64_t), intent(out) :: fused_solves
            real(c_double), intent(out) :: ms(5)
        end function
        integer(c_int) function ibmdem_batch_profile(batch, on) bind(C, name="ibmdem_batch_profile")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: on
        end function
        integer(c_int) function ibmdem_batch_profile_read(batch, kind, total_ms, launches, bytes) &
                bind(C, name="ibmdem_batch_profile_read")
            import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: kind
            real(c_double), intent(out) :: total_ms, bytes
            integer(c_int64_t), intent(out) :: launches
        end function

        ! ---- predictor schedule, probes, force download without allocation ---------
        integer(c_int) function ibmdem_get_stencil_tile(ctx, tile) bind(C, name="ibmdem_get_stencil_tile")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: tile
        end function
        integer(c_int) function ibmdem_set_stencil_tile(ctx, tile) bind(C, name="ibmdem_set_stencil_tile")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: tile
        end function
        integer(c_int) function ibmdem_probe_bodies(ctx, n, bodies, capacity) bind(C, name="ibmdem_probe_bodies")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n, capacity
            integer(c_int32_t), intent(in) :: bodies(*)
        end function
        integer(c_int) function ibmdem_probe_read(ctx, n_steps, first_step, records) bind(C, name="ibmdem_probe_read")
            import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: n_steps
            integer(c_int64_t), intent(out) :: first_step
            real(c_double), intent(inout) :: records(*)
        end function
        integer(c_int) function ibmdem_download_force_noalloc(ctx, fx, fy, fz) &
                bind(C, name="ibmdem_download_force_noalloc")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: fx(*), fy(*), fz(*)
        end function

        ! ---- measurement ------------------------------------------------------------
        integer(c_int) function ibmdem_profile_kernel(ctx, kernel_id) bind(C, name="ibmdem_profile_kernel")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: kernel_id
        end function
        integer(c_int) function ibmdem_profile_read(ctx, total_ms, launches, bytes_per_launch) &
                bind(C, name="ibmdem_profile_read")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: total_ms, bytes_per_launch
            integer(c_int64_t), intent(out) :: launches
        end function
        integer(c_int) function ibmdem_memory_bytes(ctx, bytes) bind(C, name="ibmdem_memory_bytes")
            import :: c_int, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: bytes
        end function
    end interface

contains

    ! Text of ibmdem_last_error as a Fortran string.
    function ibmdem_error_text(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n
        interface
            integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
                import :: c_ptr, c_size_t
                type(c_ptr), value :: s
            end function
        end interface
        p = ibmdem_last_error(ctx)
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        msg = transfer(chars(1:n), msg)
    end function

end module ibmdem_api
