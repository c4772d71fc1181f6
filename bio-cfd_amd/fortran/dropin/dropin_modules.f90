! Drop-in replacements of the reference's hot-path modules: the same module
! names, public procedure names and argument lists as the reference's
! (app/main.F90:4-27 `use`s them), over the reference's own global::block /
! global::intfr (modGlobal.f90:9-10), each procedure running its work on the
! GPU through the C ABI (module biocfd_hip).  Linked with the unchanged
! main.F90 and the reference's set-up and I/O modules, they make the GPU path
! a drop-in; host/device synchronisation is described in biocfd_dropin.f90.
! Reference file:line of each replaced procedure is given at its definition.

! biocfd_gpu_helpers (gpu_helpers.F90:17-40): one GPU, selected once
module biocfd_gpu_helpers
  use biocfd_dropin, only: dropin_register
  implicit none
  private
  public :: set_gpu
contains
  subroutine set_gpu()
    call dropin_register()
  end subroutine
end module biocfd_gpu_helpers

! biocfd_boundary_conditions (BoundaryConditions.f90)
module biocfd_boundary_conditions
  use, intrinsic :: iso_fortran_env, only: dp => real64
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, biocfd_velocity_bc, biocfd_solid_cell_bc, biocfd_solid_cell_bc_move
  use biocfd_dropin, only: dropin_register, block_no
  implicit none
  private
  public :: velocityBC, solidCellBC, solidCellBC_move
contains
  ! BoundaryConditions.f90:11-74
  subroutine velocityBC(blk, deltat, uc)
    type(Block_t), intent(inout) :: blk
    real(dp), intent(in) :: deltat, uc
    call dropin_register()
    call biocfd_check(biocfd_velocity_bc(block_no(blk), deltat, uc), 'velocityBC')
  end subroutine
  ! BoundaryConditions.f90:76-92
  subroutine solidCellBC(blk)
    type(Block_t), intent(inout) :: blk
    call dropin_register()
    call biocfd_check(biocfd_solid_cell_bc(block_no(blk)), 'solidCellBC')
  end subroutine
  ! BoundaryConditions.f90:94-115
  subroutine solidCellBC_move(blk)
    type(Block_t), intent(inout) :: blk
    call dropin_register()
    call biocfd_check(biocfd_solid_cell_bc_move(block_no(blk), blk%move_check), 'solidCellBC_move')
  end subroutine
end module biocfd_boundary_conditions

! biocfd_coarse_update (CoarseInterp.f90:16-437): fine -> coarse restriction
! over every interface
module biocfd_coarse_update
  use, intrinsic :: iso_c_binding, only: c_int
  use biocfd_hip, only: biocfd_check, biocfd_coarse_update_c => biocfd_coarse_update
  use biocfd_dropin, only: dropin_register
  implicit none
  private
  public :: coarseUpdate_newv, coarseUpdate_pc, coarseUpdate
contains
  subroutine coarseUpdate()          ! :16-204 (p, u, v, w)
    call dropin_register()
    call biocfd_check(biocfd_coarse_update_c(0_c_int), 'coarseUpdate')
  end subroutine
  subroutine coarseUpdate_newv()     ! :206-346 (ut, vt, wt)
    call dropin_register()
    call biocfd_check(biocfd_coarse_update_c(1_c_int), 'coarseUpdate_newv')
  end subroutine
  subroutine coarseUpdate_pc()       ! :347-437 (pc, pco)
    call dropin_register()
    call biocfd_check(biocfd_coarse_update_c(2_c_int), 'coarseUpdate_pc')
  end subroutine
end module biocfd_coarse_update

! biocfd_fine_interp_bound (FineInterp_bound.f90:17-544): coarse -> fine
! boundary prolongation
module biocfd_fine_interp_bound
  use, intrinsic :: iso_fortran_env, only: int64
  use, intrinsic :: iso_c_binding, only: c_int
  use biocfd_hip, only: biocfd_check, biocfd_fine_update_bd, biocfd_fine_update_bd_one
  use biocfd_dropin, only: dropin_register
  implicit none
  private
  public :: fineUpdate_pc_bd, fineUpdate_newv_bd, fineUpdate_bd_mv
contains
  subroutine fineUpdate_pc_bd()      ! :17-107
    call dropin_register()
    call biocfd_check(biocfd_fine_update_bd(2_c_int), 'fineUpdate_pc_bd')
  end subroutine
  subroutine fineUpdate_newv_bd()    ! :109-309
    call dropin_register()
    call biocfd_check(biocfd_fine_update_bd(1_c_int), 'fineUpdate_newv_bd')
  end subroutine
  subroutine fineUpdate_bd_mv(g)     ! :311-544, interface g
    integer(int64), intent(in) :: g
    call dropin_register()
    call biocfd_check(biocfd_fine_update_bd_one(int(g, c_int), 0_c_int), 'fineUpdate_bd_mv')
  end subroutine
end module biocfd_fine_interp_bound

! biocfd_fine_interp (FineInterp.f90:10-281): coarse -> fine volume fill
module biocfd_fine_interp
  use, intrinsic :: iso_fortran_env, only: int64
  use, intrinsic :: iso_c_binding, only: c_int
  use biocfd_hip, only: biocfd_check, biocfd_fine_update_mv
  use biocfd_dropin, only: dropin_register
  implicit none
  private
  public :: fineUpdate_mv
contains
  subroutine fineUpdate_mv(g)
    integer(int64), intent(in) :: g
    call dropin_register()
    call biocfd_check(biocfd_fine_update_mv(int(g, c_int)), 'fineUpdate_mv')
  end subroutine
end module biocfd_fine_interp

! biocfd_coefficient_matrix (CoefficientMatrix.f90:10-71)
module biocfd_coefficient_matrix
  use, intrinsic :: iso_fortran_env, only: int64
  use, intrinsic :: iso_c_binding, only: c_int
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, biocfd_coefficient_matrix_c => biocfd_coefficient_matrix, &
                        biocfd_coef_download
  use biocfd_dropin, only: dropin_register, block_no, dropin_fields_up
  implicit none
  private
  public :: coefficientMatrix
contains
  ! The driver calls it once per block after initialConditions /
  ! lastConditions (main.F90:100-113): the host's initial fields go to the
  ! device here.  Acx, Acy, Acz come back to the host arrays.
  subroutine coefficientMatrix(blk, is_coarse)
    type(Block_t), intent(inout) :: blk
    logical, intent(in) :: is_coarse
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call dropin_fields_up(g, blk)
    call biocfd_check(biocfd_coefficient_matrix_c(g, merge(1_c_int, 0_c_int, is_coarse)), 'coefficientMatrix')
    call biocfd_check(biocfd_coef_download(g, 0_c_int, blk%Acx, blk%nx), 'coef_download')
    call biocfd_check(biocfd_coef_download(g, 1_c_int, blk%Acy, blk%ny), 'coef_download')
    call biocfd_check(biocfd_coef_download(g, 2_c_int, blk%Acz, blk%nz), 'coef_download')
  end subroutine
end module biocfd_coefficient_matrix

! biocfd_navier_stokes (NavierStokes.f90)
module biocfd_navier_stokes
  use, intrinsic :: iso_fortran_env, only: int64, dp => real64
  use, intrinsic :: iso_c_binding, only: c_int
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, biocfd_non_uni_coeff, biocfd_ns_momentum, biocfd_table_download
  use biocfd_dropin, only: dropin_register, block_no, dropin_set_stepping
  implicit none
  private
  public :: non_uni_coeff, nsMomentum2order
contains
  ! NavierStokes.f90:13-250: the 18 upwind tables, computed by the library
  ! and copied to the host arrays the reference allocates here (:32-40)
  subroutine non_uni_coeff(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_non_uni_coeff(g), 'non_uni_coeff')
    call table(0, blk%nx, blk%ca_uu, blk%ck_uu)
    call table(1, blk%ny, blk%ca_vv, blk%ck_vv)
    call table(2, blk%nz, blk%ca_ww, blk%ck_ww)
    call table(3, blk%nx, blk%ca_uv, blk%ck_uv)
    call table(4, blk%nx, blk%ca_uw, blk%ck_uw)
    call table(5, blk%ny, blk%ca_vu, blk%ck_vu)
    call table(6, blk%ny, blk%ca_vw, blk%ck_vw)
    call table(7, blk%nz, blk%ca_wu, blk%ck_wu)
    call table(8, blk%nz, blk%ca_wv, blk%ck_wv)
  contains
    subroutine table(t, n, ca, ck)
      integer, intent(in) :: t
      integer(int64), intent(in) :: n
      real(dp), allocatable, intent(inout) :: ca(:, :), ck(:, :)
      if (allocated(ca)) deallocate(ca)
      if (allocated(ck)) deallocate(ck)
      allocate(ca(6, n + 2), ck(6, n + 2))
      call biocfd_check(biocfd_table_download(g, int(t, c_int), 0_c_int, ca, n + 2), 'table_download')
      call biocfd_check(biocfd_table_download(g, int(t, c_int), 1_c_int, ck, n + 2), 'table_download')
    end subroutine
  end subroutine

  ! NavierStokes.f90:252-912
  subroutine nsMomentum2order(blk, deltat, re)
    type(Block_t), intent(inout) :: blk
    real(dp), intent(in) :: deltat, re
    call dropin_register()
    call dropin_set_stepping()
    call biocfd_check(biocfd_ns_momentum(block_no(blk), deltat, re), 'nsMomentum2order')
  end subroutine
end module biocfd_navier_stokes

! biocfd_forcing (Forcing.f90): the kinematic scalars of the Block_t go with
! each call
module biocfd_forcing
  use, intrinsic :: iso_fortran_env, only: dp => real64
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, BK_N, biocfd_velocity_forcing1, biocfd_pressure_forcing1, &
                        biocfd_velocity_forcing_field, biocfd_pressure_forcing_field, &
                        biocfd_velocity_forcing_ghost, biocfd_pressure_forcing_ghost
  use biocfd_dropin, only: dropin_register, block_no, kin_of
  implicit none
  private
  public :: pressureForcing1, pressureForcingField, pressureForcingGhost
  public :: velocityForcing1, velocityForcingField, velocityForcingGhost
contains
  subroutine pressureForcing1(blk)          ! :14-90
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_pressure_forcing1(block_no(blk), k), 'pressureForcing1')
  end subroutine
  subroutine velocityForcing1(blk)          ! :92-360
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_velocity_forcing1(block_no(blk), k), 'velocityForcing1')
  end subroutine
  subroutine pressureForcingGhost(blk)      ! :362-448
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_pressure_forcing_ghost(block_no(blk), k), 'pressureForcingGhost')
  end subroutine
  subroutine velocityForcingGhost(blk)      ! :450-734
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_velocity_forcing_ghost(block_no(blk), k), 'velocityForcingGhost')
  end subroutine
  subroutine pressureForcingField(blk)      ! :737-811
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_pressure_forcing_field(block_no(blk), k), 'pressureForcingField')
  end subroutine
  subroutine velocityForcingField(blk)      ! :813-1075
    type(Block_t), intent(inout) :: blk
    real(dp) :: k(0:BK_N - 1)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_velocity_forcing_field(block_no(blk), k), 'velocityForcingField')
  end subroutine
end module biocfd_forcing

! biocfd_pcor_vcor (PcorVcor.F90)
module biocfd_pcor_vcor
  use, intrinsic :: iso_fortran_env, only: int64, dp => real64
  use, intrinsic :: iso_c_binding, only: c_int
  use global, only: block
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, biocfd_poisson_solver, biocfd_sor_state, biocfd_update_velocity_newv
  use biocfd_dropin, only: dropin_register, block_no
  implicit none
  private
  public :: poissonSolver, updateVelocity_newv
contains
  ! PcorVcor.F90:22-220 over blocks start..finish (step 1: one process
  ! drives the GPU); nIterPcor, derr2 and derrStdSt come back to the
  ! Block_t, and the iteration log and report of :162-184 are written as
  ! the reference writes them
  subroutine poissonSolver(pcItaMax, epsi, ita, deltat, totime, uc, totalTime, start, finish, step)
    integer(int64), intent(in) :: pcItaMax, ita
    real(dp), intent(in) :: epsi, deltat, totime, uc
    real(dp), intent(inout) :: totalTime
    integer, intent(in) :: start, finish, step
    integer :: g, u
    integer(int64) :: it(1), max_it
    real(dp) :: d2(1), ds(1), max_d2, max_ds
    call dropin_register()
    if (step /= 1) error stop 'biocfd drop-in: poissonSolver runs the blocks start..finish of one process'
    call biocfd_check(biocfd_poisson_solver(start, finish, pcItaMax, epsi, deltat, uc), 'poissonSolver')
    max_it = 0
    max_d2 = 0.0_dp
    max_ds = 0.0_dp
    do g = start, finish
      call biocfd_check(biocfd_sor_state(g, it, d2, ds), 'sor_state')
      block(g)%nIterPcor = it(1)
      block(g)%derr2 = d2(1)
      block(g)%derrStdSt = ds(1)
      if (block(g)%derr2 > max_d2) max_d2 = block(g)%derr2
      if (block(g)%derrStdSt > max_ds) max_ds = block(g)%derrStdSt
      if (block(g)%nIterPcor > max_it) max_it = block(g)%nIterPcor
      totalTime = totime + totalTime
    end do
    open(newunit=u, file='sphere_iter.dat', position='append', status='unknown')
    write(u, '(" ",I8, 2I10, 2F6.2,F14.9)') ita, block(1)%nIterPcor, block(2)%nIterPcor, block(1)%omega, &
                                            block(2)%omega, 0.0_dp
    write(*, '(" ",I8, I10, 4E15.6)') ita, max_it, max_d2, max_ds, totalTime
    close(u)
  end subroutine

  ! PcorVcor.F90:379-397
  subroutine updateVelocity_newv(blk)
    type(Block_t), intent(inout) :: blk
    call dropin_register()
    call biocfd_check(biocfd_update_velocity_newv(block_no(blk), blk%move_check), 'updateVelocity_newv')
  end subroutine
end module biocfd_pcor_vcor

! biocfd_search (Search.F90): kinematics are host arithmetic on the
! Block_t's node arrays and scalars; tagging, lists, distances and block
! moves run on the GPU
module biocfd_search
  use, intrinsic :: iso_fortran_env, only: int64, dp => real64
  use, intrinsic :: iso_c_binding, only: c_int, c_int64_t
  use global, only: block, intfr
  use biocfd_block_type, only: Block_t
  use biocfd_interface_type, only: Interface_t
  use biocfd_hip
  use biocfd_dropin
  implicit none
  private
  public :: findDistnode, shiftSurfaceNodesInitial, computeSurfaceNorm
  public :: tagging_th, tagging_th_move, block_move_check, cellCount_solid
  public :: cellCount_solid_coarse, cellCount_solid_coarse_mv, change_block_coords
  public :: change_block_interface, computeNormDistance, computeSurfaceVariables, findTScells
  public :: fine_block_cell, selectiveRetagging_th, change_block_coords_interfaces
contains
  ! :21-47: the marker nodes mk, mkx1, mkx2 among the ID-51 nodes (0 where
  ! none qualifies; block_move_check needs all three)
  subroutine findDistnode(blk)
    type(Block_t), intent(inout) :: blk
    integer(int64) :: mk(3)
    call dropin_register()
    call biocfd_check(biocfd_find_dist_node(blk%ibNodes, blk%xnode, blk%znode, blk%ibNodeId, mk), &
                      'findDistnode')
    blk%mk = mk(1)
    blk%mkx1 = mk(2)
    blk%mkx2 = mk(3)
  end subroutine

  ! :49-124
  subroutine shiftSurfaceNodesInitial(blk, aoa, piv_pt, ita, deltat, dxmin, totime)
    type(Block_t), intent(inout) :: blk
    real(dp), intent(in) :: aoa, piv_pt, deltat, dxmin, totime
    integer(int64), intent(in) :: ita
    real(dp) :: st(0:BKS_N - 1)
    call dropin_register()
    if (.not. allocated(blk%xnode1)) allocate(blk%xnode1(blk%ibNodes), blk%ynode1(blk%ibNodes), &
                                              blk%znode1(blk%ibNodes))
    call state_pack(blk, st)
    call biocfd_check(biocfd_shift_surface_nodes_initial(blk%ibNodes, blk%xnode, blk%ynode, blk%znode, &
                      blk%ibNodeId, aoa, piv_pt, ita, deltat, dxmin, totime, st, blk%xnode1, blk%ynode1, &
                      blk%znode1), 'shiftSurfaceNodesInitial')
    call state_unpack(st, blk)
  end subroutine

  ! :126-208, with the per-step index file of :177-183
  subroutine computeSurfaceVariables(blk, g, phase_angle, piv_pt, ita, deltat, dxmin, freq, re, totime)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: g, ita
    real(dp), intent(in) :: phase_angle, piv_pt, deltat, dxmin, freq, re, totime
    real(dp) :: st(0:BKS_N - 1)
    character(len=150) :: fname
    integer :: u
    call dropin_register()
    call state_pack(blk, st)
    call biocfd_check(biocfd_compute_surface_variables(blk%ibNodes, blk%xnode, blk%ynode, blk%znode, &
                      blk%ibNodeId, phase_angle, piv_pt, ita, deltat, dxmin, freq, totime, st, blk%xnode1, &
                      blk%ynode1, blk%znode1), 'computeSurfaceVariables')
    call state_unpack(st, blk)
    write(fname, '("d",I4.4,"_index.",F8.2,".",i3.1,".dat")') blk%fineg, re, g
    open(newunit=u, file=fname, position='append', status='unknown')
    write(u, '(7F15.8)') totime, blk%nxty_cent, blk%inity_cent, blk%ymove, blk%nxtx_cent, blk%xmove
    close(u)
  end subroutine

  ! :210-258: centroids, unit normals and element lengths, on the device and
  ! back into the Block_t's element arrays
  subroutine computeSurfaceNorm(blk, inor)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: inor
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_compute_surface_norm(g, blk%xnode1, blk%ynode1, blk%znode1, inor), &
                      'computeSurfaceNorm')
    call elem(BE_XCENT, blk%xcent)
    call elem(BE_YCENT, blk%ycent)
    call elem(BE_ZCENT, blk%zcent)
    call elem(BE_COSA, blk%cosAlpha)
    call elem(BE_COSB, blk%cosBeta)
    call elem(BE_COSG, blk%cosGamma)
    call elem(BE_ELEN, blk%element_length)
  contains
    subroutine elem(which, a)
      integer(c_int), intent(in) :: which
      real(dp), allocatable, intent(inout) :: a(:)
      if (.not. allocated(a)) return
      call biocfd_check(biocfd_elem_download(g, which, a, size(a, kind=int64)), 'elem_download')
    end subroutine
  end subroutine

  ! :388-413: tagging of the search box (tagging_th_core, :260-386), the
  ! tag and count files of :395-411
  subroutine tagging_th(blk, blk_no)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: blk_no
    integer :: g, u
    integer(int64) :: i, j, k
    character(len=120) :: fname
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_tagging(g), 'tagging_th')
    call mirror_tags(g, blk)
    blk%solidCellCount = count(blk%cell(2:blk%nx + 1, 2:blk%ny + 1, 2:blk%nz + 1) == 1, kind=int64)
    blk%fluidCellCount = count(blk%cell(2:blk%nx + 1, 2:blk%ny + 1, 2:blk%nz + 1) == 0, kind=int64)
    blk%ibCellCount = count(blk%cell(2:blk%nx + 1, 2:blk%ny + 1, 2:blk%nz + 1) == 2, kind=int64)
    write(fname, '("butter_f.",i3.3,".dat")') blk_no
    open(newunit=u, file=fname, status='unknown')
    do k = 1, blk%nz + 2
      do j = 1, blk%ny + 2
        do i = 1, blk%nx + 2
          write(u, *) blk%cell(i, j, k), blk%nodeIdTag(i, j, k)
        end do
      end do
    end do
    close(u)
    write(fname, '("butter_cellcount_f.",i3.3,".dat")') blk_no
    open(newunit=u, file=fname, form='formatted')
    write(u, *) blk%solidCellCount, blk%fluidCellCount, blk%ibCellCount
    close(u)
  end subroutine

  ! :415-432
  subroutine tagging_th_move(blk, id)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: id
    call dropin_register()
    if (int(id) /= block_no(blk)) error stop 'biocfd drop-in: tagging_th_move block number mismatch'
    call biocfd_check(biocfd_tagging_move(block_no(blk), blk%move_check), 'tagging_th_move')
  end subroutine

  ! :520-651
  subroutine selectiveRetagging_th(blk)
    type(Block_t), intent(inout) :: blk
    call dropin_register()
    call biocfd_check(biocfd_selective_retagging(block_no(blk), blk%blk_mv_tag), 'selectiveRetagging_th')
  end subroutine

  ! :653-741: counts, then the fluid, red, black, intercepted and solid lists
  ! (allocated as the reference allocates them; contents during the set-up)
  subroutine cellCount_solid(blk, blk_no)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: blk_no
    integer :: g
    call dropin_register()
    g = block_no(blk)
    if (int(blk_no) /= g) error stop 'biocfd drop-in: cellCount_solid block number mismatch'
    call biocfd_check(biocfd_cell_count_solid(g), 'cellCount_solid')
    call mirror_counts(g, blk)
    call alloc_ib_lists(blk)
    call alloc_fluid_lists(blk)
    if (dropin_host_lists()) call fill_lists(g, blk)
    if (.not. dropin_stepping()) call mirror_tags(g, blk)
  end subroutine

  ! :1463-1532
  subroutine cellCount_solid_coarse(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_cell_count_solid_coarse(g), 'cellCount_solid_coarse')
    call mirror_counts(g, blk)
    call alloc_fluid_lists(blk)
    if (dropin_host_lists()) call fill_lists(g, blk)
  end subroutine

  ! :916-995: block 1's lists rebuilt when the flag is 1 (then cleared)
  subroutine cellCount_solid_coarse_mv(coarse_flcnt_check)
    integer(int64), intent(inout) :: coarse_flcnt_check
    logical :: rebuilt
    call dropin_register()
    rebuilt = coarse_flcnt_check == 1
    call biocfd_check(biocfd_cell_count_solid_coarse_mv(1_c_int, coarse_flcnt_check), 'cellCount_solid_coarse_mv')
    if (rebuilt) then
      call mirror_counts(1, block(1))
      call alloc_fluid_lists(block(1))
      if (dropin_host_lists()) call fill_lists(1, block(1))
    end if
  end subroutine

  ! :863-914 on block 1 for every interface
  subroutine fine_block_cell()
    call dropin_register()
    call biocfd_check(biocfd_fine_block_cell(1_c_int), 'fine_block_cell')
    if (.not. dropin_stepping()) call mirror_tags(1, block(1))
  end subroutine

  ! :743-861: the nearest elements and normal distances of the IB cells
  subroutine computeNormDistance(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    integer(int64) :: n
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_compute_norm_distance(g), 'computeNormDistance')
    call mirror_counts(g, blk)
    n = blk%ibCellCount
    if (allocated(blk%nelp)) deallocate(blk%nelp, blk%nelu1, blk%nelu2, blk%nelv1, blk%nelv2, blk%nelw1, &
                                        blk%nelw2, blk%pNormDis, blk%u1NormDis, blk%u2NormDis, blk%v1NormDis, &
                                        blk%v2NormDis, blk%w1NormDis, blk%w2NormDis)
    allocate(blk%pNormDis(n), blk%nelp(n), blk%nelu1(n), blk%nelu2(n), blk%nelv1(n), blk%nelv2(n), &
             blk%nelw1(n), blk%nelw2(n), blk%u1NormDis(n), blk%u2NormDis(n), blk%v1NormDis(n), &
             blk%v2NormDis(n), blk%w1NormDis(n), blk%w2NormDis(n))
    if (.not. dropin_host_lists()) return
    call nel(BN_P, blk%nelp)
    call nel(BN_U1, blk%nelu1)
    call nel(BN_U2, blk%nelu2)
    call nel(BN_V1, blk%nelv1)
    call nel(BN_V2, blk%nelv2)
    call nel(BN_W1, blk%nelw1)
    call nel(BN_W2, blk%nelw2)
    call nd(BN_P, blk%pNormDis)
    call nd(BN_U1, blk%u1NormDis)
    call nd(BN_U2, blk%u2NormDis)
    call nd(BN_V1, blk%v1NormDis)
    call nd(BN_V2, blk%v2NormDis)
    call nd(BN_W1, blk%w1NormDis)
    call nd(BN_W2, blk%w2NormDis)
  contains
    subroutine nel(which, a)
      integer(c_int), intent(in) :: which
      integer(int64), intent(inout) :: a(:)
      call biocfd_check(biocfd_nel_download(g, which, a, n), 'nel_download')
    end subroutine
    subroutine nd(which, a)
      integer(c_int), intent(in) :: which
      real(dp), intent(inout) :: a(:)
      call biocfd_check(biocfd_normdis_download(g, which, a, n), 'normdis_download')
    end subroutine
  end subroutine

  ! :434-518: TS cells, their list, index_ts and the 14 ghosts
  subroutine findTScells(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_find_ts_cells(g), 'findTScells')
    call mirror_counts(g, blk)
    call alloc_ts_arrays(blk)
    if (.not. dropin_host_lists()) return
    call biocfd_check(biocfd_list_download(g, BL_TS, blk%TSIndexPtr, blk%TSCellCount), 'list_download')
    call biocfd_check(biocfd_index_ts_download(g, blk%index_ts, blk%TSCellCount), 'index_ts_download')
    call gh(BGH_P, blk%p_ghost)
    call gh(BGH_PT, blk%pt_ghost)
    call gh(BGH_U1, blk%u1_ghost)
    call gh(BGH_U1T, blk%u1t_ghost)
    call gh(BGH_U2, blk%u2_ghost)
    call gh(BGH_U2T, blk%u2t_ghost)
    call gh(BGH_V1, blk%v1_ghost)
    call gh(BGH_V1T, blk%v1t_ghost)
    call gh(BGH_V2, blk%v2_ghost)
    call gh(BGH_V2T, blk%v2t_ghost)
    call gh(BGH_W1, blk%w1_ghost)
    call gh(BGH_W1T, blk%w1t_ghost)
    call gh(BGH_W2, blk%w2_ghost)
    call gh(BGH_W2T, blk%w2t_ghost)
    if (.not. dropin_stepping()) call mirror_tags(g, blk)
  contains
    subroutine gh(which, a)
      integer(c_int), intent(in) :: which
      real(dp), intent(inout) :: a(:)
      call biocfd_check(biocfd_ghost_download(g, which, a, blk%TSCellCount), 'ghost_download')
    end subroutine
  end subroutine

  ! :997-1203 for every interface, in order: a fine block within the move
  ! margins of its interface moves (move_check = blk_mv_tag = 1).  The
  ! check reads the marker nodes findDistnode sets only from ID-51 nodes,
  ! so it runs only where every fine block has them.
  subroutine block_move_check(coarse_flcnt_check)
    integer(int64), intent(inout) :: coarse_flcnt_check
    integer :: q, b
    integer(int64) :: mk(3), moved
    real(dp) :: st(0:BKS_N - 1)
    call dropin_register()
    if (.not. allocated(intfr)) return
    do q = 1, size(intfr)
      b = intfr(q)%b_blk
      if (block(b)%mk <= 0 .or. block(b)%mkx1 <= 0 .or. block(b)%mkx2 <= 0) return
    end do
    do q = 1, size(intfr)
      b = intfr(q)%b_blk
      mk = [block(b)%mk, block(b)%mkx1, block(b)%mkx2]
      call state_pack(block(b), st)
      moved = 0
      call biocfd_check(biocfd_block_move_check(int(q, c_int), block(b)%xnode1, block(b)%ynode1, mk, st, &
                        moved, coarse_flcnt_check), 'block_move_check')
      call state_unpack(st, block(b))
      if (moved == 1) then
        block(b)%move_check = 1
        block(b)%blk_mv_tag = 1
      end if
    end do
  end subroutine

  ! :1208-1264: the block's grids shifted by its last move (also on the host)
  subroutine change_block_coords(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_change_block_coords(g, blk%move_check), 'change_block_coords')
    if (blk%move_check == 1) call download_grids(g, blk)
  end subroutine

  ! :1269-1370
  subroutine change_block_coords_interfaces(local_intfr, blk)
    type(Interface_t), intent(in) :: local_intfr
    type(Block_t), intent(inout) :: blk
    integer :: q
    call dropin_register()
    q = interface_no(local_intfr)
    if (block_no(blk) /= local_intfr%b_blk) error stop 'biocfd drop-in: change_block_coords_interfaces block mismatch'
    call biocfd_check(biocfd_change_block_coords_interfaces(int(q, c_int), blk%move_check), &
                      'change_block_coords_interfaces')
  end subroutine

  ! :1372-1461
  subroutine change_block_interface(local_intfr, blk_a, blk_b)
    type(Interface_t), intent(inout) :: local_intfr
    type(Block_t), intent(in) :: blk_a, blk_b
    integer :: q
    call dropin_register()
    q = interface_no(local_intfr)
    if (block_no(blk_a) /= local_intfr%a_blk .or. block_no(blk_b) /= local_intfr%b_blk) &
      error stop 'biocfd drop-in: change_block_interface block mismatch'
    call biocfd_check(biocfd_change_block_interface(int(q, c_int), blk_b%move_check), 'change_block_interface')
  end subroutine
end module biocfd_search

! biocfd_stress_calculation (StressCalc.f90:14-208): not called by
! main.F90; the four coefficients go to the Block_t
module biocfd_stress_calculation
  use, intrinsic :: iso_fortran_env, only: dp => real64
  use biocfd_block_type, only: Block_t
  use biocfd_hip, only: biocfd_check, BK_N, biocfd_stress_calculation_c => biocfd_stress_calculation
  use biocfd_dropin, only: dropin_register, block_no, kin_of
  implicit none
  private
  public :: stress_calculation
contains
  subroutine stress_calculation(blk, mu_f, rho_f)
    type(Block_t), intent(inout) :: blk
    real(dp), intent(in) :: mu_f, rho_f
    real(dp) :: k(0:BK_N - 1), out(4)
    call dropin_register()
    k = kin_of(blk)
    call biocfd_check(biocfd_stress_calculation_c(block_no(blk), mu_f, rho_f, k, out), 'stress_calculation')
    blk%viscous_drag_coefficient = out(1)
    blk%pressure_drag_coefficient = out(2)
    blk%viscous_lift_coefficient = out(3)
    blk%pressure_lift_coefficient = out(4)
  end subroutine
end module biocfd_stress_calculation
