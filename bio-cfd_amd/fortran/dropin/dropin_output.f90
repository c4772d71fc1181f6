! biocfd_write_output_corner1 as a sync-then-write wrapper: the fields live
! on the GPU, so before the reference's own writers (WriteOutput_corner1.F90,
! compiled unchanged under the module name biocfd_ref_write_output) write a
! file, the arrays that file holds are copied to the Block_t.  The copies are
! made only on the steps where the writer writes: the field file after the
! set-up, at steps 1, 2 and every 200th (:90), the checkpoint every 500th
! step (:125); the body plot (:142-160) reads only host node arrays.
module biocfd_write_output_corner1
  use, intrinsic :: iso_fortran_env, only: int64, dp => real64
  use biocfd_block_type, only: Block_t
  use biocfd_ref_write_output, only: ref_write_output_ascii => write_output_ascii, ref_body_plot => body_plot, &
                                     ref_writeresult => writeresult
  use biocfd_hip, only: biocfd_check, biocfd_field_download, biocfd_tag_download, BF_U, BF_V, BF_W, BF_P, &
                        BT_CELL, BT_CELL_N, BT_CELL_PR
  use biocfd_dropin, only: dropin_register, block_no
  implicit none
  private
  public :: write_output_ascii, body_plot, writeResult
contains
  subroutine fields_down(blk)
    type(Block_t), intent(inout) :: blk
    integer :: g
    call dropin_register()
    g = block_no(blk)
    call biocfd_check(biocfd_field_download(g, BF_U, blk%u), 'field_download')
    call biocfd_check(biocfd_field_download(g, BF_V, blk%v), 'field_download')
    call biocfd_check(biocfd_field_download(g, BF_W, blk%w), 'field_download')
    call biocfd_check(biocfd_field_download(g, BF_P, blk%p), 'field_download')
  end subroutine

  subroutine write_output_ascii(blk, blk_no, char_f, ita, re, totime)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: blk_no, ita
    character(len=3), intent(in) :: char_f
    real(dp), intent(in) :: re, totime
    integer :: g
    if (mod(ita, 200_int64) == 0 .or. ita <= 2) then
      call fields_down(blk)
      g = block_no(blk)
      call biocfd_check(biocfd_tag_download(g, BT_CELL, blk%cell), 'tag_download')
      call biocfd_check(biocfd_tag_download(g, BT_CELL_N, blk%cell_n), 'tag_download')
      call biocfd_check(biocfd_tag_download(g, BT_CELL_PR, blk%cell_pr), 'tag_download')
    end if
    call ref_write_output_ascii(blk, blk_no, char_f, ita, re, totime)
  end subroutine

  subroutine writeResult(blk, blk_no, char_f, ita, ita1, re, totime)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: blk_no, ita, ita1
    character(len=3), intent(in) :: char_f
    real(dp), intent(in) :: re, totime
    if (mod(ita, 500_int64) == 0) call fields_down(blk)
    call ref_writeresult(blk, blk_no, char_f, ita, ita1, re, totime)
  end subroutine

  subroutine body_plot(blk, blk_id, ita)
    type(Block_t), intent(inout) :: blk
    integer(int64), intent(in) :: blk_id, ita
    call ref_body_plot(blk, blk_id, ita)
  end subroutine
end module biocfd_write_output_corner1
