! Fortran bindings of the gfx950 Bio-CFD runtime (include/biocfd_hip.h).
!
! A module a Fortran driver `use`s to call the C ABI through iso_c_binding:
! one interface per entry point, the enumerations as named constants, and
! biocfd_check, which turns a non-zero return code into `error stop` with the
! library's message (the reference's procedures have no status; a failure
! there is a crash, so stopping keeps that contract with a readable message).
!
! Arrays cross as assumed-size dummies, so a Block_t component such as
! blk%u (real(real64), (nx+2,ny+2,nz+2)) or blk%fluidIndexPtr
! (integer(int64), (count,3)) is passed directly.
module biocfd_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  ! enum biocfd_field
  integer(c_int), parameter :: BF_U = 0, BF_V = 1, BF_W = 2, BF_UT = 3, BF_VT = 4, BF_WT = 5, BF_P = 6, &
                               BF_B = 7, BF_PC = 8, BF_PCO = 9
  ! enum biocfd_tag
  integer(c_int), parameter :: BT_CELL = 0, BT_CELL2 = 1, BT_NODEIDTAG = 2, BT_CELL_N = 3, BT_CELL_PR = 4
  ! enum biocfd_grid (Blocks.f90:26-27 order)
  integer(c_int), parameter :: BG_X1 = 0, BG_Y1 = 1, BG_Z1 = 2, BG_DX = 3, BG_DY = 4, BG_DZ = 5, &
                               BG_XU = 6, BG_YU = 7, BG_ZU = 8, BG_XV = 9, BG_YV = 10, BG_ZV = 11, &
                               BG_XW = 12, BG_YW = 13, BG_ZW = 14, BG_XP = 15, BG_YP = 16, BG_ZP = 17
  ! enum biocfd_list
  integer(c_int), parameter :: BL_FLUID = 0, BL_RED = 1, BL_BLACK = 2, BL_SOLID = 3, BL_INTERCEPTED = 4, &
                               BL_TS = 5
  ! enum biocfd_nel
  integer(c_int), parameter :: BN_P = 0, BN_U1 = 1, BN_U2 = 2, BN_V1 = 3, BN_V2 = 4, BN_W1 = 5, BN_W2 = 6
  ! enum biocfd_elem
  integer(c_int), parameter :: BE_XCENT = 0, BE_YCENT = 1, BE_ZCENT = 2, BE_COSA = 3, BE_COSB = 4, &
                               BE_COSG = 5, BE_ELEN = 6
  ! enum biocfd_ghost (Blocks.f90:55-58 order)
  integer(c_int), parameter :: BGH_P = 0, BGH_PT = 1, BGH_U1 = 2, BGH_U1T = 3, BGH_U2 = 4, BGH_U2T = 5, &
                               BGH_V1 = 6, BGH_V1T = 7, BGH_V2 = 8, BGH_V2T = 9, BGH_W1 = 10, &
                               BGH_W1T = 11, BGH_W2 = 12, BGH_W2T = 13
  ! enum biocfd_count
  integer(c_int), parameter :: BC_FLUID = 0, BC_SOLID = 1, BC_IB = 2, BC_RED = 3, BC_BLACK = 4, BC_TS = 5, &
                               BC_N = 6
  ! enum biocfd_kin: kin(0:BK_N-1)
  integer(c_int), parameter :: BK_XDOT = 0, BK_YDOT = 1, BK_YDDOT = 2, BK_THETADOT1 = 3, &
                               BK_THETADDOT1 = 4, BK_THETADOT2 = 5, BK_THETADDOT2 = 6, BK_ALPHADOT = 7, &
                               BK_PIVX = 8, BK_PIVY = 9, BK_PIVZ = 10, BK_N = 11
  ! enum biocfd_table (Blocks.f90:29-31 order)
  integer(c_int), parameter :: BTB_UU = 0, BTB_VV = 1, BTB_WW = 2, BTB_UV = 3, BTB_UW = 4, BTB_VU = 5, &
                               BTB_VW = 6, BTB_WU = 7, BTB_WV = 8
  integer(c_int), parameter :: BKS_N = 35

  interface
    type(c_ptr) function biocfd_last_error() bind(C, name="biocfd_last_error")
      import :: c_int, c_int64_t, c_double, c_ptr
    end function

    integer(c_int) function biocfd_init(device) bind(C, name="biocfd_init")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: device
    end function

    integer(c_int) function biocfd_finalize() bind(C, name="biocfd_finalize")
      import :: c_int, c_int64_t, c_double, c_ptr
    end function

    integer(c_int) function biocfd_synchronize() bind(C, name="biocfd_synchronize")
      import :: c_int, c_int64_t, c_double, c_ptr
    end function

    integer(c_int) function biocfd_timing_enable(on) bind(C, name="biocfd_timing_enable")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: on
    end function

    integer(c_int) function biocfd_timing_read(sor_ms, sor_iters, sor_cell_iters) bind(C, &
        name="biocfd_timing_read")
      import :: c_int, c_int64_t, c_double, c_ptr
      real(c_double) :: sor_ms(*)
      integer(c_int64_t) :: sor_iters(*)
      integer(c_int64_t) :: sor_cell_iters(*)
    end function

    integer(c_int) function biocfd_copy_probe(bytes, reps) bind(C, name="biocfd_copy_probe")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: bytes
      integer(c_int), value :: reps
    end function

    integer(c_int) function biocfd_block_create(blk, nx, ny, nz, omega, cintp, dx, dy, dz) bind(C, &
        name="biocfd_block_create")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: nx
      integer(c_int64_t), value :: ny
      integer(c_int64_t), value :: nz
      real(c_double), value :: omega
      integer(c_int64_t), value :: cintp
      real(c_double), value :: dx
      real(c_double), value :: dy
      real(c_double), value :: dz
    end function

    integer(c_int) function biocfd_block_destroy(blk) bind(C, name="biocfd_block_destroy")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_grid_upload(blk, which, host, n) bind(C, name="biocfd_grid_upload")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
    end function

    integer(c_int) function biocfd_field_upload(blk, field, host) bind(C, name="biocfd_field_upload")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: field
      real(c_double) :: host(*)
    end function

    integer(c_int) function biocfd_field_download(blk, field, host) bind(C, name="biocfd_field_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: field
      real(c_double) :: host(*)
    end function

    integer(c_int) function biocfd_tag_upload(blk, tag, host) bind(C, name="biocfd_tag_upload")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: tag
      integer(c_int64_t) :: host(*)
    end function

    integer(c_int) function biocfd_tag_download(blk, tag, host) bind(C, name="biocfd_tag_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: tag
      integer(c_int64_t) :: host(*)
    end function

    integer(c_int) function biocfd_counts(blk, out) bind(C, name="biocfd_counts")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t) :: out(*)
    end function

    integer(c_int) function biocfd_list_download(blk, which, host, count) bind(C, name="biocfd_list_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      integer(c_int64_t) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_nel_download(blk, which, host, count) bind(C, name="biocfd_nel_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      integer(c_int64_t) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_normdis_download(blk, which, host, count) bind(C, &
        name="biocfd_normdis_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_ghost_download(blk, which, host, count) bind(C, &
        name="biocfd_ghost_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_elem_download(blk, which, host, count) bind(C, name="biocfd_elem_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_index_ts_download(blk, host, count) bind(C, &
        name="biocfd_index_ts_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_coef_download(blk, axis, host, n) bind(C, name="biocfd_coef_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: axis
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
    end function

    integer(c_int) function biocfd_table_download(blk, which, is_ck, host, ncols) bind(C, &
        name="biocfd_table_download")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: which
      integer(c_int), value :: is_ck
      real(c_double) :: host(*)
      integer(c_int64_t), value :: ncols
    end function

    integer(c_int) function biocfd_sor_state(blk, n_iter, derr2, derr_std_st) bind(C, name="biocfd_sor_state")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t) :: n_iter(*)
      real(c_double) :: derr2(*)
      real(c_double) :: derr_std_st(*)
    end function

    integer(c_int) function biocfd_sor_exhausted(blk, count) bind(C, name="biocfd_sor_exhausted")
      import :: c_int
      integer(c_int), value :: blk
      integer(c_int) :: count
    end function

    integer(c_int) function biocfd_health_check(blk, n_nonfinite, vel_max, where) bind(C, name="biocfd_health_check")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int64_t) :: n_nonfinite
      real(c_double) :: vel_max
      integer(c_int64_t) :: where(4)
    end function

    integer(c_int) function biocfd_health_config(check_after_solve, vel_limit, exhausted_is_error) &
        bind(C, name="biocfd_health_config")
      import :: c_int, c_double
      integer(c_int), value :: check_after_solve
      real(c_double), value :: vel_limit
      integer(c_int), value :: exhausted_is_error
    end function

    integer(c_int) function biocfd_inject_alloc_failure(nth) bind(C, name="biocfd_inject_alloc_failure")
      import :: c_int
      integer(c_int), value :: nth
    end function

    integer(c_int) function biocfd_body_create(blk, n_nodes, n_elems, el_p1, el_p2, el_p3, surf_id, &
        search_box) bind(C, name="biocfd_body_create")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: n_nodes
      integer(c_int64_t), value :: n_elems
      integer(c_int64_t) :: el_p1(*)
      integer(c_int64_t) :: el_p2(*)
      integer(c_int64_t) :: el_p3(*)
      integer(c_int64_t) :: surf_id(*)
      integer(c_int64_t) :: search_box(*)
    end function

    integer(c_int) function biocfd_coefficient_matrix(blk, is_coarse) bind(C, &
        name="biocfd_coefficient_matrix")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: is_coarse
    end function

    integer(c_int) function biocfd_non_uni_coeff(blk) bind(C, name="biocfd_non_uni_coeff")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_ns_momentum(blk, dt, re) bind(C, name="biocfd_ns_momentum")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: dt
      real(c_double), value :: re
    end function
    integer(c_int) function biocfd_ns_momentum_set(first, last, dt, re) bind(C, name="biocfd_ns_momentum_set")
      import :: c_int, c_double
      integer(c_int), value :: first, last
      real(c_double), value :: dt, re
    end function

    integer(c_int) function biocfd_velocity_bc(blk, dt, uc) bind(C, name="biocfd_velocity_bc")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: dt
      real(c_double), value :: uc
    end function

    integer(c_int) function biocfd_solid_cell_bc(blk) bind(C, name="biocfd_solid_cell_bc")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_solid_cell_bc_move(blk, move_check) bind(C, &
        name="biocfd_solid_cell_bc_move")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_update_velocity_newv(blk, move_check) bind(C, &
        name="biocfd_update_velocity_newv")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_poisson_solver(first, last, pc_ita_max, epsi, dt, uc) bind(C, &
        name="biocfd_poisson_solver")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: first
      integer(c_int), value :: last
      integer(c_int64_t), value :: pc_ita_max
      real(c_double), value :: epsi
      real(c_double), value :: dt
      real(c_double), value :: uc
    end function

    integer(c_int) function biocfd_compute_div(blk) bind(C, name="biocfd_compute_div")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_redblack_sor(blk, pc_ita_max, dt, epsi) bind(C, name="biocfd_redblack_sor")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: pc_ita_max
      real(c_double), value :: dt
      real(c_double), value :: epsi
    end function

    integer(c_int) function biocfd_correct(blk, dt) bind(C, name="biocfd_correct")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: dt
    end function

    integer(c_int) function biocfd_sor_config(variant, planes_per_workgroup) bind(C, name="biocfd_sor_config")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: variant
      integer(c_int), value :: planes_per_workgroup
    end function
    integer(c_int) function biocfd_sor_block_config(blk, variant, planes_per_workgroup) &
        bind(C, name="biocfd_sor_block_config")
      import :: c_int
      integer(c_int), value :: blk, variant, planes_per_workgroup
    end function

    integer(c_int) function biocfd_sor_stop_rule(rule, cadence) bind(C, name="biocfd_sor_stop_rule")
      import :: c_int
      integer(c_int), value :: rule
      integer(c_int), value :: cadence
    end function

    integer(c_int) function biocfd_sor_variant_info(variant, info) bind(C, name="biocfd_sor_variant_info")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: variant
      integer(c_int) :: info(*)
    end function

    integer(c_int) function biocfd_sor_time(blk, variant, planes_per_workgroup, iters, dt, &
        ms) bind(C, name="biocfd_sor_time")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: variant
      integer(c_int), value :: planes_per_workgroup
      integer(c_int), value :: iters
      real(c_double), value :: dt
      real(c_double) :: ms(*)
    end function

    integer(c_int) function biocfd_set_precision(mode) bind(C, name="biocfd_set_precision")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: mode
    end function

    integer(c_int) function biocfd_mixed_config(variant, planes_per_workgroup, &
        sweeps_per_correction, residual_tol) bind(C, name="biocfd_mixed_config")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: variant
      integer(c_int), value :: planes_per_workgroup
      integer(c_int), value :: sweeps_per_correction
      real(c_double), value :: residual_tol
    end function

    integer(c_int) function biocfd_mixed_variant_info(variant, info) bind(C, name="biocfd_mixed_variant_info")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: variant
      integer(c_int) :: info(*)
    end function

    integer(c_int) function biocfd_mixed_time(blk, variant, planes_per_workgroup, iters, dt, &
        ms) bind(C, name="biocfd_mixed_time")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int), value :: variant
      integer(c_int), value :: planes_per_workgroup
      integer(c_int), value :: iters
      real(c_double), value :: dt
      real(c_double) :: ms(*)
    end function

    integer(c_int) function biocfd_residual_norm(blk, dt, out) bind(C, name="biocfd_residual_norm")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: dt
      real(c_double) :: out(*)
    end function

    integer(c_int) function biocfd_projection(blk, n_iter, dt) bind(C, name="biocfd_projection")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt
    end function

    integer(c_int) function biocfd_projection_adjoint(blk, n_iter, dt, ubar, vbar, wbar, pbar, utbar, vtbar, &
        wtbar) bind(C, name="biocfd_projection_adjoint")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt
      real(c_double) :: ubar(*), vbar(*), wbar(*), pbar(*), utbar(*), vtbar(*), wtbar(*)
    end function

    ! ---- tangent-linear and adjoint time step (geometry frozen) ----
    integer(c_int) function biocfd_ad_reset(blk) bind(C, name="biocfd_ad_reset")
      import :: c_int
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_ad_upload(blk, field, host) bind(C, name="biocfd_ad_upload")
      import :: c_int, c_double
      integer(c_int), value :: blk, field
      real(c_double) :: host(*)
    end function

    integer(c_int) function biocfd_ad_download(blk, field, host) bind(C, name="biocfd_ad_download")
      import :: c_int, c_double
      integer(c_int), value :: blk, field
      real(c_double) :: host(*)
    end function

    integer(c_int) function biocfd_ad_kin(blk, kin_bar, clear) bind(C, name="biocfd_ad_kin")
      import :: c_int, c_double
      integer(c_int), value :: blk, clear
      real(c_double) :: kin_bar(*)
    end function

    integer(c_int) function biocfd_ad_ghost_upload(blk, which, host, count) bind(C, name="biocfd_ad_ghost_upload")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk, which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_ad_ghost_download(blk, which, host, count) bind(C, &
        name="biocfd_ad_ghost_download")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk, which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function biocfd_tlm_momentum(blk, dt, re) bind(C, name="biocfd_tlm_momentum")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: dt, re
    end function

    integer(c_int) function biocfd_adj_momentum(blk, dt, re) bind(C, name="biocfd_adj_momentum")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: dt, re
    end function

    integer(c_int) function biocfd_tlm_velocity_bc(blk, dt, uc) bind(C, name="biocfd_tlm_velocity_bc")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: dt, uc
    end function

    integer(c_int) function biocfd_adj_velocity_bc(blk, dt, uc) bind(C, name="biocfd_adj_velocity_bc")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: dt, uc
    end function

    integer(c_int) function biocfd_tlm_solid_cell_bc(blk) bind(C, name="biocfd_tlm_solid_cell_bc")
      import :: c_int
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_adj_solid_cell_bc(blk) bind(C, name="biocfd_adj_solid_cell_bc")
      import :: c_int
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_tlm_forcing(blk, routine, kin, kin_dot, rec) bind(C, name="biocfd_tlm_forcing")
      import :: c_int, c_double
      integer(c_int), value :: blk, routine, rec
      real(c_double) :: kin(*), kin_dot(*)
    end function

    integer(c_int) function biocfd_adj_forcing(blk, routine, kin, rec) bind(C, name="biocfd_adj_forcing")
      import :: c_int, c_double
      integer(c_int), value :: blk, routine, rec
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_tlm_poisson(blk, n_iter, dt, uc) bind(C, name="biocfd_tlm_poisson")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt, uc
    end function

    integer(c_int) function biocfd_adj_poisson(blk, n_iter, dt, uc) bind(C, name="biocfd_adj_poisson")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt, uc
    end function

    integer(c_int) function biocfd_stress_sums(blk, mu_f, rho_f, kin, out) bind(C, name="biocfd_stress_sums")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: mu_f, rho_f
      real(c_double) :: kin(*), out(*)
    end function

    integer(c_int) function biocfd_tlm_stress(blk, mu_f, rho_f, kin, kin_dot, geom, out) bind(C, &
        name="biocfd_tlm_stress")
      import :: c_int, c_double
      integer(c_int), value :: blk, geom
      real(c_double), value :: mu_f, rho_f
      real(c_double) :: kin(*), kin_dot(*), out(*)
    end function

    integer(c_int) function biocfd_adj_stress(blk, mu_f, rho_f, kin, sbar, geom) bind(C, name="biocfd_adj_stress")
      import :: c_int, c_double
      integer(c_int), value :: blk, geom
      real(c_double), value :: mu_f, rho_f
      real(c_double) :: kin(*), sbar(*)
    end function

    integer(c_int) function biocfd_tlm_step_start(blk, n_iter, dt, re, uc, kin, kin_dot, rec) bind(C, &
        name="biocfd_tlm_step_start")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk, rec
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt, re, uc
      real(c_double) :: kin(*), kin_dot(*)
    end function

    integer(c_int) function biocfd_tlm_step_end(blk, kin, kin_dot, rec) bind(C, name="biocfd_tlm_step_end")
      import :: c_int, c_double
      integer(c_int), value :: blk, rec
      real(c_double) :: kin(*), kin_dot(*)
    end function

    integer(c_int) function biocfd_adj_step_end(blk, kin, rec) bind(C, name="biocfd_adj_step_end")
      import :: c_int, c_double
      integer(c_int), value :: blk, rec
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_adj_step_start(blk, n_iter, dt, re, uc, kin, rec) bind(C, &
        name="biocfd_adj_step_start")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk, rec
      integer(c_int64_t), value :: n_iter
      real(c_double), value :: dt, re, uc
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_ad_record(blk, slot) bind(C, name="biocfd_ad_record")
      import :: c_int
      integer(c_int), value :: blk, slot
    end function

    integer(c_int) function biocfd_ad_geometry_tangent(blk, inor, xdot, ydot, zdot) bind(C, &
        name="biocfd_ad_geometry_tangent")
      import :: c_int, c_double
      integer(c_int), value :: blk
      real(c_double), value :: inor
      real(c_double) :: xdot(*), ydot(*), zdot(*)
    end function

    integer(c_int) function biocfd_ad_geometry_adjoint(blk, inor, xbar, ybar, zbar, clear) bind(C, &
        name="biocfd_ad_geometry_adjoint")
      import :: c_int, c_double
      integer(c_int), value :: blk, clear
      real(c_double), value :: inor
      real(c_double) :: xbar(*), ybar(*), zbar(*)
    end function

    integer(c_int) function biocfd_snapshot_save(blk, slot) bind(C, name="biocfd_snapshot_save")
      import :: c_int
      integer(c_int), value :: blk, slot
    end function

    integer(c_int) function biocfd_snapshot_restore(blk, slot) bind(C, name="biocfd_snapshot_restore")
      import :: c_int
      integer(c_int), value :: blk, slot
    end function

    integer(c_int) function biocfd_snapshot_free(blk, slot) bind(C, name="biocfd_snapshot_free")
      import :: c_int
      integer(c_int), value :: blk, slot
    end function

    integer(c_int) function biocfd_coarse_update_adjoint(itf, comp, lam_coarse, lam_fine) bind(C, &
        name="biocfd_coarse_update_adjoint")
      import :: c_int, c_double
      integer(c_int), value :: itf, comp
      real(c_double) :: lam_coarse(*), lam_fine(*)
    end function

    integer(c_int) function biocfd_fine_update_bd_adjoint(itf, comp, lam_fine, lam_coarse) bind(C, &
        name="biocfd_fine_update_bd_adjoint")
      import :: c_int, c_double
      integer(c_int), value :: itf, comp
      real(c_double) :: lam_fine(*), lam_coarse(*)
    end function

    integer(c_int) function biocfd_compute_surface_norm(blk, xnode1, ynode1, znode1, inor) bind(C, &
        name="biocfd_compute_surface_norm")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: xnode1(*)
      real(c_double) :: ynode1(*)
      real(c_double) :: znode1(*)
      integer(c_int64_t), value :: inor
    end function

    integer(c_int) function biocfd_tagging(blk) bind(C, name="biocfd_tagging")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_selective_retagging(blk, blk_mv_tag) bind(C, &
        name="biocfd_selective_retagging")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      integer(c_int64_t), value :: blk_mv_tag
    end function

    integer(c_int) function biocfd_cell_count_solid(blk) bind(C, name="biocfd_cell_count_solid")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_cell_count_solid_coarse(blk) bind(C, name="biocfd_cell_count_solid_coarse")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_fine_block_cell(coarse_blk) bind(C, name="biocfd_fine_block_cell")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: coarse_blk
    end function

    integer(c_int) function biocfd_compute_norm_distance(blk) bind(C, name="biocfd_compute_norm_distance")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_find_ts_cells(blk) bind(C, name="biocfd_find_ts_cells")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
    end function

    integer(c_int) function biocfd_velocity_forcing1(blk, kin) bind(C, name="biocfd_velocity_forcing1")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_pressure_forcing1(blk, kin) bind(C, name="biocfd_pressure_forcing1")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_velocity_forcing_field(blk, kin) bind(C, &
        name="biocfd_velocity_forcing_field")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_pressure_forcing_field(blk, kin) bind(C, &
        name="biocfd_pressure_forcing_field")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_velocity_forcing_ghost(blk, kin) bind(C, &
        name="biocfd_velocity_forcing_ghost")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_pressure_forcing_ghost(blk, kin) bind(C, &
        name="biocfd_pressure_forcing_ghost")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double) :: kin(*)
    end function

    integer(c_int) function biocfd_stress_calculation(blk, mu_f, rho_f, kin, out) bind(C, &
        name="biocfd_stress_calculation")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: mu_f
      real(c_double), value :: rho_f
      real(c_double) :: kin(*)
      real(c_double) :: out(*)
    end function

    integer(c_int) function biocfd_stress_calculation_noalloc(blk, mu_f, rho_f, kin, out) bind(C, &
        name="biocfd_stress_calculation_noalloc")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: blk
      real(c_double), value :: mu_f
      real(c_double), value :: rho_f
      real(c_double) :: kin(*)
      real(c_double) :: out(*)
    end function

    integer(c_int) function biocfd_interface_create(itf, a_blk, b_blk, a_msh, b_msh, &
        extent) bind(C, name="biocfd_interface_create")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: itf
      integer(c_int), value :: a_blk
      integer(c_int), value :: b_blk
      integer(c_int64_t), value :: a_msh
      integer(c_int64_t), value :: b_msh
      real(c_double) :: extent(*)
    end function

    integer(c_int) function biocfd_coarse_update(which) bind(C, name="biocfd_coarse_update")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: which
    end function

    integer(c_int) function biocfd_fine_update_bd(which) bind(C, name="biocfd_fine_update_bd")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: which
    end function

    integer(c_int) function biocfd_fine_update_bd_one(itf, which) bind(C, name="biocfd_fine_update_bd_one")
      import :: c_int
      integer(c_int), value :: itf
      integer(c_int), value :: which
    end function

    integer(c_int) function biocfd_interface_table(itf, comp, axis, host, counter, cap) bind(C, &
        name="biocfd_interface_table")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int), value :: itf
      integer(c_int), value :: comp
      integer(c_int), value :: axis
      integer(c_int64_t) :: host(*)
      integer(c_int64_t) :: counter(*)
      integer(c_int64_t), value :: cap
    end function

    integer(c_int) function biocfd_shift_surface_nodes_initial(n_nodes, xnode, ynode, znode, &
        node_id, aoa, piv_pt, ita, dt, dxmin, totime, state, xnode1, ynode1, znode1) bind(C, &
        name="biocfd_shift_surface_nodes_initial")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: n_nodes
      real(c_double) :: xnode(*)
      real(c_double) :: ynode(*)
      real(c_double) :: znode(*)
      integer(c_int64_t) :: node_id(*)
      real(c_double), value :: aoa
      real(c_double), value :: piv_pt
      integer(c_int64_t), value :: ita
      real(c_double), value :: dt
      real(c_double), value :: dxmin
      real(c_double), value :: totime
      real(c_double) :: state(*)
      real(c_double) :: xnode1(*)
      real(c_double) :: ynode1(*)
      real(c_double) :: znode1(*)
    end function

    integer(c_int) function biocfd_compute_surface_variables(n_nodes, xnode, ynode, znode, node_id, &
        phase_angle, piv_pt, ita, dt, dxmin, freq, totime, state, xnode1, ynode1, znode1) bind(C, &
        name="biocfd_compute_surface_variables")
      import :: c_int, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: n_nodes
      real(c_double) :: xnode(*)
      real(c_double) :: ynode(*)
      real(c_double) :: znode(*)
      integer(c_int64_t) :: node_id(*)
      real(c_double), value :: phase_angle
      real(c_double), value :: piv_pt
      integer(c_int64_t), value :: ita
      real(c_double), value :: dt
      real(c_double), value :: dxmin
      real(c_double), value :: freq
      real(c_double), value :: totime
      real(c_double) :: state(*)
      real(c_double) :: xnode1(*)
      real(c_double) :: ynode1(*)
      real(c_double) :: znode1(*)
    end function

    function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function

    integer(c_int) function biocfd_sor_batching(on) bind(C, name="biocfd_sor_batching")
      import :: c_int
      integer(c_int), value :: on
    end function
    integer(c_int) function biocfd_sor_graphs(on) bind(C, name="biocfd_sor_graphs")
      import :: c_int
      integer(c_int), value :: on
    end function
    integer(c_int) function biocfd_interp_batching(on) bind(C, name="biocfd_interp_batching")
      import :: c_int
      integer(c_int), value :: on
    end function
    integer(c_int) function biocfd_sor_graph_stats(captures, launches) bind(C, name="biocfd_sor_graph_stats")
      import :: c_int, c_int64_t
      integer(c_int64_t) :: captures, launches
    end function
    integer(c_int) function biocfd_momentum_config(variant, pow2_products) &
        bind(C, name="biocfd_momentum_config")
      import :: c_int
      integer(c_int), value :: variant, pow2_products
    end function
    integer(c_int) function biocfd_momentum_precision(fp32) bind(C, name="biocfd_momentum_precision")
      import :: c_int
      integer(c_int), value :: fp32
    end function
    integer(c_int) function biocfd_momentum_divisions(reciprocal_rows) bind(C, name="biocfd_momentum_divisions")
      import :: c_int
      integer(c_int), value :: reciprocal_rows
    end function

    integer(c_int) function biocfd_ts_pass_config(parallel) bind(C, name="biocfd_ts_pass_config")
      import :: c_int
      integer(c_int), value :: parallel
    end function

    ! ---- block move (Search.F90:21-47, :415-432, :916-1461; FineInterp.f90:10-281)
    integer(c_int) function biocfd_find_dist_node(n_nodes, xnode, znode, node_id, markers) bind(C, &
        name="biocfd_find_dist_node")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n_nodes
      real(c_double) :: xnode(*)
      real(c_double) :: znode(*)
      integer(c_int64_t) :: node_id(*)
      integer(c_int64_t) :: markers(3)
    end function

    integer(c_int) function biocfd_block_move_check(itf, xnode1, ynode1, markers, state, moved, &
        coarse_flcnt_check) bind(C, name="biocfd_block_move_check")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: itf
      real(c_double) :: xnode1(*)
      real(c_double) :: ynode1(*)
      integer(c_int64_t) :: markers(3)
      real(c_double) :: state(*)
      integer(c_int64_t) :: moved
      integer(c_int64_t) :: coarse_flcnt_check
    end function

    integer(c_int) function biocfd_change_block_coords(blk, move_check) bind(C, name="biocfd_change_block_coords")
      import :: c_int, c_int64_t
      integer(c_int), value :: blk
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_change_block_coords_interfaces(itf, move_check) bind(C, &
        name="biocfd_change_block_coords_interfaces")
      import :: c_int, c_int64_t
      integer(c_int), value :: itf
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_change_block_interface(itf, move_check) bind(C, &
        name="biocfd_change_block_interface")
      import :: c_int, c_int64_t
      integer(c_int), value :: itf
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_cell_count_solid_coarse_mv(coarse_blk, coarse_flcnt_check) bind(C, &
        name="biocfd_cell_count_solid_coarse_mv")
      import :: c_int, c_int64_t
      integer(c_int), value :: coarse_blk
      integer(c_int64_t) :: coarse_flcnt_check
    end function

    integer(c_int) function biocfd_fine_update_mv(itf) bind(C, name="biocfd_fine_update_mv")
      import :: c_int
      integer(c_int), value :: itf
    end function

    integer(c_int) function biocfd_tagging_move(blk, move_check) bind(C, name="biocfd_tagging_move")
      import :: c_int, c_int64_t
      integer(c_int), value :: blk
      integer(c_int64_t), value :: move_check
    end function

    integer(c_int) function biocfd_grid_download(blk, which, host, n) bind(C, name="biocfd_grid_download")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int64_t), value :: n
    end function

    ! ---- output sync (WriteOutput_corner1.F90:80-139)
    integer(c_int) function biocfd_output_fields(blk, u1, v1, w1, p, cell, cell_n, cell_pr) bind(C, &
        name="biocfd_output_fields")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: blk
      real(c_double) :: u1(*)
      real(c_double) :: v1(*)
      real(c_double) :: w1(*)
      real(c_double) :: p(*)
      integer(c_int64_t) :: cell(*)
      integer(c_int64_t) :: cell_n(*)
      integer(c_int64_t) :: cell_pr(*)
    end function
  end interface

contains

  ! the message of the last failed call
  function biocfd_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, q
    p = biocfd_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do q = 1, n
      msg(q:q) = chars(q)
    end do
  end function

  ! stop with the library's message when rc is non-zero
  subroutine biocfd_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(a,a,a,i0,a,a)') 'biocfd: ', what, ' failed (', rc, '): ', biocfd_error_message()
      error stop 1
    end if
  end subroutine

end module biocfd_hip
