!> The Fortran host side of the MI355X path: the sequencer calls MOSE's step makes, with the reference's error
!> behaviour (a failing state or thermo guard stops the run with the cell that failed, as State_Check_blk /
!> thermo_guard_report do, Lib_Newstate.f90:200-246, Mod_Thermo_Cache.f90:323-335).
module mose_gpu_host
  use iso_c_binding
  use mose_gpu_c
  implicit none

contains

  !> n explicit steps of one case on the device (Explicit_Step, Mod_Explicit.f90:52-397, time accurate; Strang-split
  !> with chemistry once MOSE_GPU_Chemistry_Attach has coupled a mechanism).  On a failed state the run stops with the
  !> first bad cell (block, i, j, k).  Chemistry cells whose integration ended with idid < 0 are reported and the run
  !> goes on, as Chemistry_Newstate_blk writes such a cell back and carries on (Lib_Chemistry.f90:62-70).
  subroutine MOSE_GPU_Explicit_Step(ctx, case_id, nsteps, rep)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: case_id, nsteps
    type(mose_report), intent(out) :: rep
    integer(c_int) :: rc
    rc = mose_case_step(ctx, case_id, nsteps, rep)
    if (rc == MOSE_ERR_ODE) then
      write(*,'(2a)') ' [WARNING] chemistry: ', mose_error_text(ctx)
      rc = MOSE_OK
    end if
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

  !> Chemistry_Newstate_blk (src/lib/physics/Lib_Chemistry.f90:33-76) on a batch of ncell cells resident in the
  !> chemistry object chem_id: state(ncell, ns + 1) holds the species densities and, in column ns + 1, the pressure
  !> (the chemistry columns of P_flat).  Every cell: T = p / (rho R), the integrator (MOSE_ODE_RODAS or
  !> MOSE_ODE_RADAU5, raw tolerances per component) over [0, dt], floors, p = T rho R; hr(c) = |T_new - T_old| / T_old.
  !> Like the reference's loop, a cell whose integration fails (idid < 0) is written back as the integrator left it;
  !> the count of such cells is returned in nfailed and reported once.
  subroutine MOSE_GPU_Chemistry_Newstate(ctx, chem_id, ncell, ns, state, dt, method, rtol, atol, hr, nfailed)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: chem_id, ns, method
    integer(c_int64_t), intent(in) :: ncell
    real(c_double), intent(inout) :: state(ncell, ns + 1)
    real(c_double), intent(in) :: dt, rtol(ns + 1), atol(ns + 1)
    real(c_double), intent(out) :: hr(ncell)
    integer(c_int64_t), intent(out) :: nfailed
    type(mose_ode_opts) :: opts
    type(mose_chem_report) :: rep
    integer(c_int32_t), allocatable :: idid(:), nstep(:)
    integer(c_int) :: rc
    call mose_check(ctx, mose_chem_upload(ctx, chem_id, ncell, state), 'mose_chem_upload')
    opts%method = method
    opts%reserved = 0
    opts%rtol = 0.0d0
    opts%atol = 0.0d0
    opts%rtol(1:ns + 1) = rtol
    opts%atol(1:ns + 1) = atol
    rc = mose_chem_advance(ctx, chem_id, dt, opts, rep)
    if (rc /= MOSE_OK .and. rc /= MOSE_ERR_ODE) call mose_check(ctx, rc, 'mose_chem_advance')
    nfailed = rep%failed
    if (nfailed > 0) write(*,'(a,i0,a)') ' [WARNING] chemistry: ', nfailed, ' cells ended with idid < 0'
    allocate(idid(ncell), nstep(ncell))
    call mose_check(ctx, mose_chem_download(ctx, chem_id, state, hr, idid, nstep), 'mose_chem_download')
  end subroutine MOSE_GPU_Chemistry_Newstate

  !> Couples the mechanism of chemistry object chem_id to a case: its steps become Strang-split steps (half transport,
  !> Chemistry_Newstate_blk over dt on the resident state, half transport; set_StrangSplitting, Lib_Strang.f90:10-22)
  !> with the integrator `method` and raw tolerances per component.
  subroutine MOSE_GPU_Chemistry_Attach(ctx, case_id, chem_id, ns, method, rtol, atol)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: case_id, chem_id, ns, method
    real(c_double), intent(in) :: rtol(ns + 1), atol(ns + 1)
    type(mose_ode_opts) :: opts
    opts%method = method
    opts%reserved = 0
    opts%rtol = 0.0d0
    opts%atol = 0.0d0
    opts%rtol(1:ns + 1) = rtol
    opts%atol(1:ns + 1) = atol
    call mose_check(ctx, mose_case_chem_attach(ctx, case_id, chem_id, opts), 'mose_case_chem_attach')
  end subroutine MOSE_GPU_Chemistry_Attach

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
