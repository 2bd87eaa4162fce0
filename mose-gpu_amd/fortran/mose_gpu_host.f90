!> The Fortran host side of the MI355X path: the sequencer calls MOSE's step makes, with the reference's error
!> behaviour (a failing state or thermo guard stops the run with the cell that failed, as State_Check_blk /
!> thermo_guard_report do, Lib_Newstate.f90:200-246, Mod_Thermo_Cache.f90:323-335).
module mose_gpu_host
  use iso_c_binding
  use mose_gpu_c
  implicit none

contains

  !> n explicit steps of one case on the device (Explicit_Step, Mod_Explicit.f90:52-397, time accurate, without
  !> chemistry).  On a failed state the run stops with the first bad cell (block, i, j, k).
  subroutine MOSE_GPU_Explicit_Step(ctx, case_id, nsteps, rep)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: case_id, nsteps
    type(mose_report), intent(out) :: rep
    integer(c_int) :: rc
    rc = mose_case_step(ctx, case_id, nsteps, rep)
    if (rc == MOSE_ERR_STATE .or. rc == MOSE_ERR_THERMO) then
      write(*,'(a,i0,a,4(1x,i0))') ' [ERROR] state check: ', rep%bad_cells, ' bad cells; first (block i j k):', &
        rep%first_bad_block + 1, rep%first_bad_i, rep%first_bad_j, rep%first_bad_k
      write(*,'(2a)') ' ', mose_error_text(ctx)
      error stop 'MOSE_GPU_Explicit_Step: bad state'
    else if (rc /= MOSE_OK) then
      write(*,'(2a)') ' [ERROR] ', mose_error_text(ctx)
      error stop 'MOSE_GPU_Explicit_Step'
    end if
  end subroutine MOSE_GPU_Explicit_Step

  !> Stops with the library's message when a call failed.
  subroutine mose_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= MOSE_OK) then
      if (c_associated(ctx)) then
        write(*,'(4a)') ' [ERROR] ', what, ': ', mose_error_text(ctx)
      else
        write(*,'(3a,i0)') ' [ERROR] ', what, ': code ', rc
      end if
      error stop 'mose_check'
    end if
  end subroutine mose_check
end module mose_gpu_host
