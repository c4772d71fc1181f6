! State shared by the drop-in replacement modules (dropin_modules.f90): the
! reference driver (app/main.F90, unchanged) keeps its global::block(:) and
! global::intfr(:) (modGlobal.f90:9-10) on the host, and the replacement
! procedures run the hot path on the GPU through the C ABI (module
! biocfd_hip).  This module
!
!  * registers every block, its grids and immersed body, and every interface
!    with the library at the first call of any replacement procedure (after
!    readInput, readBlockInterface, interfaceDetail, allocateArrays and
!    readSurfaceMeshGmsh, main.F90:43-58);
!  * maps a Block_t / Interface_t argument to its block / interface number
!    (its position in global::block / global::intfr);
!  * packs and unpacks the kinematic scalars of a Block_t for the library's
!    host kinematics (state[35]) and forcing routines (kin[11]);
!  * keeps the host copies the reference's own out-of-scope code reads up to
!    date: during the set-up (before the first nsMomentum2order) the tags and
!    index lists after every routine that builds them (initialConditions
!    reads fluidIndexPtr, main.F90:100-104, InitialConditions.f90:24-33), the
!    initial fields go to the device at the first coefficientMatrix of a
!    block (after initialConditions/lastConditions, main.F90:100-113); in the
!    time loop the index lists the driver deallocates every step
!    (main.F90:240-250) are re-allocated with the device's counts and filled
!    only when BIOCFD_DROPIN_HOST_LISTS is set (no host code of the driver
!    reads them), and the fields and tags the output writes are downloaded
!    by the writer wrapper (dropin_output.f90) when a file is written.
!
! Block numbers must be 1..size(block) in one process (start = 1, step = 1):
! the library runs one GPU.
module biocfd_dropin
  use, intrinsic :: iso_fortran_env, only: int64, dp => real64
  use, intrinsic :: iso_c_binding, only: c_int, c_int64_t
  use global, only: block, intfr
  use biocfd_block_type, only: Block_t
  use biocfd_interface_type, only: Interface_t
  use biocfd_hip
  implicit none
  private

  public :: dropin_register, block_no, interface_no, state_pack, state_unpack, kin_of
  public :: dropin_stepping, dropin_set_stepping, dropin_fields_up, mirror_tags, mirror_counts
  public :: fill_lists, alloc_fluid_lists, alloc_ib_lists, alloc_ts_arrays, download_grids, dropin_host_lists
  public :: c_i

  logical, save :: registered = .false.
  logical, save :: stepping = .false.
  logical, save :: host_lists = .false.
  logical, allocatable, save :: fields_up(:)

contains

  integer(c_int) function c_i(v)
    integer(int64), intent(in) :: v
    c_i = int(v, c_int)
  end function

  ! the position of blk in global::block
  integer function block_no(blk)
    type(Block_t), intent(in) :: blk
    integer :: g
    do g = 1, size(block)
      if (loc(blk) == loc(block(g))) then
        block_no = g
        return
      end if
    end do
    error stop 'biocfd drop-in: the Block_t passed is not an element of global::block'
  end function

  ! the position of itf in global::intfr
  integer function interface_no(itf)
    type(Interface_t), intent(in) :: itf
    integer :: q
    do q = 1, size(intfr)
      if (loc(itf) == loc(intfr(q))) then
        interface_no = q
        return
      end if
    end do
    error stop 'biocfd drop-in: the Interface_t passed is not an element of global::intfr'
  end function

  logical function dropin_stepping()
    dropin_stepping = stepping
  end function

  ! the time loop has begun (nsMomentum2order, main.F90:138)
  subroutine dropin_set_stepping()
    if (.not. stepping) then
      !$omp critical (biocfd_dropin_state)
      stepping = .true.
      !$omp end critical (biocfd_dropin_state)
    end if
  end subroutine

  logical function dropin_host_lists()
    dropin_host_lists = host_lists .or. .not. stepping
  end function

  ! every block, grid, body and interface of global::block/intfr, once
  subroutine dropin_register()
    integer :: g, q, n
    character(len=16) :: env
    real(dp) :: ext(6)
    if (registered) return
    !$omp critical (biocfd_dropin_state)
    if (.not. registered) then
      call biocfd_check(biocfd_init(0_c_int), 'init')
      call get_environment_variable('BIOCFD_DROPIN_HOST_LISTS', env, n)
      host_lists = n > 0 .and. env(1:1) /= '0'
      allocate(fields_up(size(block)))
      fields_up = .false.
      do g = 1, size(block)
        call register_block(g, block(g))
      end do
      if (allocated(intfr)) then
        do q = 1, size(intfr)
          ext = [intfr(q)%xintf_start, intfr(q)%xintf_end, intfr(q)%yintf_start, intfr(q)%yintf_end, &
                 intfr(q)%zintf_start, intfr(q)%zintf_end]
          call biocfd_check(biocfd_interface_create(int(q, c_int), int(intfr(q)%a_blk, c_int), &
                            int(intfr(q)%b_blk, c_int), intfr(q)%a_msh, intfr(q)%b_msh, ext), &
                            'interface_create')
        end do
      end if
      do g = 2, size(block)
        call register_body(g, block(g))
      end do
      registered = .true.
    end if
    !$omp end critical (biocfd_dropin_state)
  end subroutine

  subroutine register_block(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(in) :: b
    if (.not. allocated(b%x1)) error stop 'biocfd drop-in: a block has no grid (readInput first)'
    call biocfd_check(biocfd_block_create(g, b%nx, b%ny, b%nz, b%omega, b%cintp, b%dx, b%dy, b%dz), &
                      'block_create')
    call up(BG_X1, b%x1)
    call up(BG_Y1, b%y1)
    call up(BG_Z1, b%z1)
    call up(BG_DX, b%deltax)
    call up(BG_DY, b%deltay)
    call up(BG_DZ, b%deltaz)
    call up(BG_XU, b%xu)
    call up(BG_YU, b%yu)
    call up(BG_ZU, b%zu)
    call up(BG_XV, b%xv)
    call up(BG_YV, b%yv)
    call up(BG_ZV, b%zv)
    call up(BG_XW, b%xw)
    call up(BG_YW, b%yw)
    call up(BG_ZW, b%zw)
    call up(BG_XP, b%xp)
    call up(BG_YP, b%yp)
    call up(BG_ZP, b%zp)
  contains
    subroutine up(which, a)
      integer(c_int), intent(in) :: which
      real(dp), intent(in) :: a(:)
      call biocfd_check(biocfd_grid_upload(g, which, a, size(a, kind=int64)), 'grid_upload')
    end subroutine
  end subroutine

  ! the triangulated body of block g (readSurfaceMeshGmsh, main.F90:57)
  subroutine register_body(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(in) :: b
    integer(int64) :: box(6)
    if (.not. allocated(b%ibElP1)) return
    box = [b%i_startSearch, b%i_endSearch, b%j_startSearch, b%j_endSearch, b%k_startSearch, b%k_endSearch]
    call biocfd_check(biocfd_body_create(g, b%ibNodes, b%ibElems, b%ibElP1, b%ibElP2, b%ibElP3, b%ibSurfId, &
                                         box), 'body_create')
  end subroutine

  ! the host's initial fields and cell_pr to the device, once per block
  ! (after initialConditions / lastConditions)
  subroutine dropin_fields_up(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(in) :: b
    if (fields_up(g)) return
    call biocfd_check(biocfd_field_upload(g, BF_U, b%u), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_V, b%v), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_W, b%w), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_UT, b%ut), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_VT, b%vt), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_WT, b%wt), 'field_upload')
    call biocfd_check(biocfd_field_upload(g, BF_P, b%p), 'field_upload')
    call biocfd_check(biocfd_tag_upload(g, BT_CELL_PR, b%cell_pr), 'tag_upload')
    fields_up(g) = .true.
  end subroutine

  ! Block_t scalars <-> the library's kinematic state (include/biocfd_hip.h,
  ! biocfd_shift_surface_nodes_initial)
  subroutine state_pack(b, st)
    type(Block_t), intent(in) :: b
    real(dp), intent(out) :: st(0:BKS_N - 1)
    st = [b%xshift, b%yshift, b%zshift, b%a0, b%yamp, b%bfreq, b%xpth1, b%xpth2, b%ypth2, b%xchg, b%ychg, &
          b%yt, b%ydot, b%yddot, b%xt, b%xdot, b%thetaDot, b%thetaDDot, b%thetaDot1, b%thetaDDot1, &
          b%thetaDot2, b%thetaDDot2, b%alphaDot, b%piv_x, b%piv_y, b%piv_z, b%xmove, b%ymove, b%zmove, &
          b%inity_cent, b%nxty_cent, b%initx_cent, b%nxtx_cent, b%initz_cent, b%nxtz_cent]
  end subroutine

  subroutine state_unpack(st, b)
    real(dp), intent(in) :: st(0:BKS_N - 1)
    type(Block_t), intent(inout) :: b
    b%xshift = st(0); b%yshift = st(1); b%zshift = st(2); b%a0 = st(3); b%yamp = st(4); b%bfreq = st(5)
    b%xpth1 = st(6); b%xpth2 = st(7); b%ypth2 = st(8); b%xchg = st(9); b%ychg = st(10); b%yt = st(11)
    b%ydot = st(12); b%yddot = st(13); b%xt = st(14); b%xdot = st(15); b%thetaDot = st(16)
    b%thetaDDot = st(17); b%thetaDot1 = st(18); b%thetaDDot1 = st(19); b%thetaDot2 = st(20)
    b%thetaDDot2 = st(21); b%alphaDot = st(22); b%piv_x = st(23); b%piv_y = st(24); b%piv_z = st(25)
    b%xmove = st(26); b%ymove = st(27); b%zmove = st(28); b%inity_cent = st(29); b%nxty_cent = st(30)
    b%initx_cent = st(31); b%nxtx_cent = st(32); b%initz_cent = st(33); b%nxtz_cent = st(34)
  end subroutine

  ! the kinematic scalars the forcing routines read (enum biocfd_kin)
  function kin_of(b) result(k)
    type(Block_t), intent(in) :: b
    real(dp) :: k(0:BK_N - 1)
    k = [b%xdot, b%ydot, b%yddot, b%thetaDot1, b%thetaDDot1, b%thetaDot2, b%thetaDDot2, b%alphaDot, &
         b%piv_x, b%piv_y, b%piv_z]
  end function

  ! host tags of block g from the device (those allocated)
  subroutine mirror_tags(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(inout) :: b
    if (allocated(b%cell)) call biocfd_check(biocfd_tag_download(g, BT_CELL, b%cell), 'tag_download')
    if (allocated(b%cell2)) call biocfd_check(biocfd_tag_download(g, BT_CELL2, b%cell2), 'tag_download')
    if (allocated(b%nodeIdTag)) call biocfd_check(biocfd_tag_download(g, BT_NODEIDTAG, b%nodeIdTag), &
                                                  'tag_download')
    if (allocated(b%cell_n)) call biocfd_check(biocfd_tag_download(g, BT_CELL_N, b%cell_n), 'tag_download')
    if (allocated(b%cell_pr)) call biocfd_check(biocfd_tag_download(g, BT_CELL_PR, b%cell_pr), 'tag_download')
  end subroutine

  ! the device's cell counts into the Block_t (Blocks.f90:42-43)
  subroutine mirror_counts(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(inout) :: b
    integer(int64) :: c(0:BC_N - 1)
    call biocfd_check(biocfd_counts(g, c), 'counts')
    b%fluidCellCount = c(BC_FLUID)
    b%solidCellCount = c(BC_SOLID)
    b%ibCellCount = c(BC_IB)
    b%redCellCount = c(BC_RED)
    b%blackCellCount = c(BC_BLACK)
    b%TSCellCount = c(BC_TS)
  end subroutine

  ! fluid, red and black lists (count,3), re-allocated with the counts
  subroutine alloc_fluid_lists(b)
    type(Block_t), intent(inout) :: b
    if (allocated(b%fluidIndexPtr)) deallocate(b%fluidIndexPtr)
    if (allocated(b%redCellIndexPtr)) deallocate(b%redCellIndexPtr)
    if (allocated(b%blackCellIndexPtr)) deallocate(b%blackCellIndexPtr)
    allocate(b%fluidIndexPtr(b%fluidCellCount, 3), b%redCellIndexPtr(b%redCellCount, 3), &
             b%blackCellIndexPtr(b%blackCellCount, 3))
  end subroutine

  ! intercepted and solid lists (cellCount_solid, Search.F90:666-668)
  subroutine alloc_ib_lists(b)
    type(Block_t), intent(inout) :: b
    if (allocated(b%interceptedIndexPtr)) deallocate(b%interceptedIndexPtr)
    if (allocated(b%solidIndexPtr)) deallocate(b%solidIndexPtr)
    allocate(b%interceptedIndexPtr(b%ibCellCount, 3), b%solidIndexPtr(b%solidCellCount, 3))
  end subroutine

  ! TS list, index_ts and the 14 ghosts (findTScells, Search.F90:464-498)
  subroutine alloc_ts_arrays(b)
    type(Block_t), intent(inout) :: b
    integer(int64) :: t
    t = b%TSCellCount
    if (allocated(b%TSIndexPtr)) deallocate(b%TSIndexPtr)
    if (allocated(b%index_ts)) deallocate(b%index_ts)
    if (allocated(b%p_ghost)) deallocate(b%p_ghost, b%pt_ghost, b%u1_ghost, b%u1t_ghost, b%u2_ghost, &
                                         b%u2t_ghost, b%v1_ghost, b%v1t_ghost, b%v2_ghost, b%v2t_ghost, &
                                         b%w1_ghost, b%w1t_ghost, b%w2_ghost, b%w2t_ghost)
    allocate(b%TSIndexPtr(t, 3), b%index_ts(t))
    allocate(b%p_ghost(t), b%pt_ghost(t), b%u1_ghost(t), b%u1t_ghost(t), b%u2_ghost(t), b%u2t_ghost(t), &
             b%v1_ghost(t), b%v1t_ghost(t), b%v2_ghost(t), b%v2t_ghost(t), b%w1_ghost(t), b%w1t_ghost(t), &
             b%w2_ghost(t), b%w2t_ghost(t))
  end subroutine

  ! the contents of the allocated lists of block g from the device
  subroutine fill_lists(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(inout) :: b
    if (allocated(b%fluidIndexPtr)) call get(BL_FLUID, b%fluidIndexPtr)
    if (allocated(b%redCellIndexPtr)) call get(BL_RED, b%redCellIndexPtr)
    if (allocated(b%blackCellIndexPtr)) call get(BL_BLACK, b%blackCellIndexPtr)
    if (allocated(b%interceptedIndexPtr)) call get(BL_INTERCEPTED, b%interceptedIndexPtr)
    if (allocated(b%solidIndexPtr)) call get(BL_SOLID, b%solidIndexPtr)
    if (allocated(b%TSIndexPtr)) call get(BL_TS, b%TSIndexPtr)
  contains
    subroutine get(which, a)
      integer(c_int), intent(in) :: which
      integer(int64), intent(inout) :: a(:, :)
      call biocfd_check(biocfd_list_download(g, which, a, size(a, 1, kind=int64)), 'list_download')
    end subroutine
  end subroutine

  ! the 18 grid arrays of block g from the device (after a block move)
  subroutine download_grids(g, b)
    integer, intent(in) :: g
    type(Block_t), intent(inout) :: b
    call dn(BG_X1, b%x1)
    call dn(BG_Y1, b%y1)
    call dn(BG_Z1, b%z1)
    call dn(BG_DX, b%deltax)
    call dn(BG_DY, b%deltay)
    call dn(BG_DZ, b%deltaz)
    call dn(BG_XU, b%xu)
    call dn(BG_YU, b%yu)
    call dn(BG_ZU, b%zu)
    call dn(BG_XV, b%xv)
    call dn(BG_YV, b%yv)
    call dn(BG_ZV, b%zv)
    call dn(BG_XW, b%xw)
    call dn(BG_YW, b%yw)
    call dn(BG_ZW, b%zw)
    call dn(BG_XP, b%xp)
    call dn(BG_YP, b%yp)
    call dn(BG_ZP, b%zp)
  contains
    subroutine dn(which, a)
      integer(c_int), intent(in) :: which
      real(dp), intent(inout) :: a(:)
      call biocfd_check(biocfd_grid_download(g, which, a, size(a, kind=int64)), 'grid_download')
    end subroutine
  end subroutine

end module biocfd_dropin
