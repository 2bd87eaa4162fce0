!> ISO_C_BINDING interface to the MI355X residual-path library (include/mose_gpu.h).
!>
!> The derived types are the C structs of the header, field for field; the interfaces are the header's entry points.
!> A MOSE Fortran host calls these in place of the module procedures of Explicit_Step
!> (src/lib/numerics/time/explicit/Mod_Explicit.f90:155-294): Fill_Ghost_Cell, Refresh_Thermo_Cache, Fluxes,
!> BC_Fluxes, RANS_Source_Terms and RK_Newstate become one mose_case_step per step batch, Compute_dt becomes
!> mose_case_dt, Compute_Residual_blk becomes mose_case_residual_norm and field_state_hash mose_case_state_hash.
!> Chemistry_Newstate_blk (src/lib/physics/Lib_Chemistry.f90:33-76) becomes mose_chem_upload / mose_chem_advance /
!> mose_chem_download on a batch of cells.
module mose_gpu_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MOSE_OK = 0, MOSE_ERR_ARG = -1, MOSE_ERR_HIP = -2, MOSE_ERR_NOMEM = -3, &
                               MOSE_ERR_MESH = -4, MOSE_ERR_STATE = -5, MOSE_ERR_THERMO = -6, MOSE_ERR_NODEV = -7, &
                               MOSE_ERR_ODE = -8
  integer(c_int), parameter :: MOSE_BC_NONE = 0, MOSE_BC_CONNECTION = 1, MOSE_BC_SYMMETRY = 3, MOSE_BC_FARFIELD = 4, &
                               MOSE_BC_WALL_ADIABATIC = 5, MOSE_BC_WALL_ISOTHERMAL = 6, MOSE_BC_EXTRAPOLATION = 11
  integer(c_int), parameter :: MOSE_LIM_MINMOD = 1, MOSE_LIM_VANLEER = 2, MOSE_LIM_VANALBADA = 3, MOSE_LIM_MC = 4, &
                               MOSE_LIM_SUPERBEE = 5, MOSE_LIM_FIRST = 6
  integer(c_int), parameter :: MOSE_RS_HLLC = 1, MOSE_RS_HLLE = 2, MOSE_RS_SLAU = 3, MOSE_RS_AUSMP = 4, MOSE_RS_LMROE = 12
  integer(c_int), parameter :: MOSE_PATH_FUSED = 0, MOSE_PATH_PLAIN = 1, MOSE_PATH_FAST = 2
  integer(c_int), parameter :: MOSE_INT_CONS = 0, MOSE_INT_PRIM = 1
  integer(c_int), parameter :: MOSE_NTAB = 10001, MOSE_GC = 2, MOSE_MAX_SPECIES = 16

  type, bind(C) :: mose_bc_record
    integer(c_int32_t) :: b, i, j, k, f, typ
    integer(c_int32_t) :: bs, is, js, ks, fs
    integer(c_int32_t) :: d11, d12, d21, d22
    integer(c_int32_t) :: periodic, ff, pad
    real(c_double) :: tw
    real(c_double) :: tvec(3)
  end type mose_bc_record

  type, bind(C) :: mose_farfield
    real(c_double) :: mach, t0, p0, pamb, alpha, beta, rel_fac, reserved
    real(c_double) :: ci(16)
    real(c_double) :: rans(2)
  end type mose_farfield

  type, bind(C) :: mose_scheme
    integer(c_int32_t) :: nsc, nrans, limiter, riemann, viscous, n_rk, ndir, point_implicit, thermo_guard
    integer(c_int32_t) :: integration
    integer(c_int32_t) :: irs, irs_njacobi
    integer(c_int32_t) :: reserved(4)
    real(c_double) :: mco, cfl, vnn, prandtl, sc, sct, prt
    real(c_double) :: irs_beta
    real(c_double) :: reserved_d(7)
  end type mose_scheme

  type, bind(C) :: mose_tables
    integer(c_int32_t) :: nsc, pad
    type(c_ptr) :: wm, cp, h, mu, k
  end type mose_tables

  type, bind(C) :: mose_report
    real(c_double) :: dt, time
    integer(c_int64_t) :: steps
    integer(c_int32_t) :: bad_cells, first_bad_block, first_bad_i, first_bad_j, first_bad_k, status
  end type mose_report

  type, bind(C) :: mose_tuning
    integer(c_int32_t) :: path, ti, tj, tk, timing, no_graph
    integer(c_int32_t) :: reserved(2)
  end type mose_tuning

  ! stiff chemistry (Chemistry_Newstate_blk, src/lib/physics/Lib_Chemistry.f90:33-76)
  integer(c_int), parameter :: MOSE_ODE_RODAS = 1, MOSE_ODE_RADAU5 = 2

  type, bind(C) :: mose_mechanism
    integer(c_int32_t) :: ns, n_arrh, n_troe, n_lind, tmin, tmax
    type(c_ptr) :: wm, cp, h, ni1, ni2, eps, kf, kb, troe, lind
  end type mose_mechanism

  type, bind(C) :: mose_ode_opts
    integer(c_int32_t) :: method
    integer(c_int32_t) :: reserved(3)
    real(c_double) :: rtol(MOSE_MAX_SPECIES + 1), atol(MOSE_MAX_SPECIES + 1)
  end type mose_ode_opts

  type, bind(C) :: mose_chem_report
    integer(c_int64_t) :: cells, failed, steps_total, rejected_total, rhs_total
    integer(c_int32_t) :: steps_max, first_failed_idid
    integer(c_int64_t) :: first_failed_cell
    real(c_double) :: ms
  end type mose_chem_report

  interface
    integer(c_int) function mose_device_count(n) bind(C, name="mose_device_count")
      import :: c_int
      integer(c_int), intent(out) :: n
    end function
    integer(c_int) function mose_ctx_create(device, tables, ctx) bind(C, name="mose_ctx_create")
      import :: c_int, c_ptr, mose_tables
      integer(c_int), value :: device
      type(mose_tables), intent(in) :: tables
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mose_ctx_destroy(ctx) bind(C, name="mose_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mose_last_error(ctx) bind(C, name="mose_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mose_metrics_build(nb, dims, nodes, nrec, rec, ndir, dl, m, vol, yn, fm1, fm2, fm3) &
        bind(C, name="mose_metrics_build")
      import :: c_int, c_double, mose_bc_record
      integer(c_int), value :: nb, nrec, ndir
      integer(c_int), intent(in) :: dims(*)
      real(c_double), intent(in) :: nodes(*)
      type(mose_bc_record), intent(in) :: rec(*)
      real(c_double), intent(out) :: dl(*), m(*), vol(*), yn(*), fm1(*), fm2(*), fm3(*)
    end function
    integer(c_int) function mose_records_tangent(nb, dims, nodes, nrec, rec) bind(C, name="mose_records_tangent")
      import :: c_int, c_double, mose_bc_record
      integer(c_int), value :: nb, nrec
      integer(c_int), intent(in) :: dims(*)
      real(c_double), intent(in) :: nodes(*)
      type(mose_bc_record), intent(inout) :: rec(*)
    end function
    integer(c_int) function mose_case_create(ctx, scheme, nb, dims, dl, m, vol, yn, fm1, fm2, fm3, nrec, rec, nff, ff, &
                                             case_id) bind(C, name="mose_case_create")
      import :: c_int, c_ptr, c_double, mose_scheme, mose_bc_record, mose_farfield
      type(c_ptr), value :: ctx
      type(mose_scheme), intent(in) :: scheme
      integer(c_int), value :: nb, nrec, nff
      integer(c_int), intent(in) :: dims(*)
      real(c_double), intent(in) :: dl(*), m(*), vol(*), yn(*), fm1(*), fm2(*), fm3(*)
      type(mose_bc_record), intent(in) :: rec(*)
      type(mose_farfield), intent(in) :: ff(*)
      integer(c_int), intent(out) :: case_id
    end function
    integer(c_int) function mose_case_destroy(ctx, case_id) bind(C, name="mose_case_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
    end function
    integer(c_int) function mose_case_bytes(scheme, nb, dims, nrec, bytes) bind(C, name="mose_case_bytes")
      import :: c_int, c_int64_t, mose_scheme
      type(mose_scheme), intent(in) :: scheme
      integer(c_int), value :: nb, nrec
      integer(c_int), intent(in) :: dims(*)
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int) function mose_state_upload(ctx, case_id, P) bind(C, name="mose_state_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(in) :: P(*)
    end function
    integer(c_int) function mose_state_download(ctx, case_id, P) bind(C, name="mose_state_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(out) :: P(*)
    end function
    integer(c_int) function mose_case_step(ctx, case_id, nsteps, rep) bind(C, name="mose_case_step")
      import :: c_int, c_ptr, mose_report
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id, nsteps
      type(mose_report), intent(out) :: rep
    end function
    integer(c_int) function mose_cases_step(ctx, ids, ncases, nsteps, reps) bind(C, name="mose_cases_step")
      import :: c_int, c_ptr, mose_report
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: ids(*)
      integer(c_int), value :: ncases, nsteps
      type(mose_report), intent(out) :: reps(*)
    end function
    integer(c_int) function mose_case_residual(ctx, case_id, R) bind(C, name="mose_case_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(out) :: R(*)
    end function
    integer(c_int) function mose_case_dt(ctx, case_id, dt) bind(C, name="mose_case_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function mose_case_residual_norm(ctx, case_id, out) bind(C, name="mose_case_residual_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mose_case_state_hash(ctx, case_id, h_lo, h_hi) bind(C, name="mose_case_state_hash")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      integer(c_int64_t), intent(out) :: h_lo, h_hi
    end function
    integer(c_int) function mose_case_tune(ctx, case_id, tuning, chosen) bind(C, name="mose_case_tune")
      import :: c_int, c_ptr, mose_tuning
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      type(mose_tuning), intent(in) :: tuning
      type(mose_tuning), intent(out) :: chosen
    end function
    integer(c_int) function mose_chem_create(ctx, mech, ncell_max, chem_id) bind(C, name="mose_chem_create")
      import :: c_int, c_ptr, c_int64_t, mose_mechanism
      type(c_ptr), value :: ctx
      type(mose_mechanism), intent(in) :: mech
      integer(c_int64_t), value :: ncell_max
      integer(c_int), intent(out) :: chem_id
    end function
    integer(c_int) function mose_chem_destroy(ctx, chem_id) bind(C, name="mose_chem_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: chem_id
    end function
    integer(c_int) function mose_chem_upload(ctx, chem_id, ncell, state) bind(C, name="mose_chem_upload")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: chem_id
      integer(c_int64_t), value :: ncell
      real(c_double), intent(in) :: state(*)
    end function
    integer(c_int) function mose_chem_download(ctx, chem_id, state, hr, idid, nstep) bind(C, name="mose_chem_download")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: chem_id
      real(c_double), intent(out) :: state(*), hr(*)
      integer(c_int32_t), intent(out) :: idid(*), nstep(*)
    end function
    integer(c_int) function mose_chem_advance(ctx, chem_id, dt, opts, report) bind(C, name="mose_chem_advance")
      import :: c_int, c_ptr, c_double, mose_ode_opts, mose_chem_report
      type(c_ptr), value :: ctx
      integer(c_int), value :: chem_id
      real(c_double), value :: dt
      type(mose_ode_opts), intent(in) :: opts
      type(mose_chem_report), intent(out) :: report
    end function
    ! chemistry on a case's resident state and the Strang-split step (Lib_Chemistry.f90:33-76, Mod_Explicit.f90:155-294)
    integer(c_int) function mose_case_chem_attach(ctx, case_id, chem_id, opts) bind(C, name="mose_case_chem_attach")
      import :: c_int, c_ptr, mose_ode_opts
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id, chem_id
      type(mose_ode_opts), intent(in) :: opts
    end function
    integer(c_int) function mose_case_chem_advance(ctx, case_id, dt, report) bind(C, name="mose_case_chem_advance")
      import :: c_int, c_ptr, c_double, mose_chem_report
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), value :: dt
      type(mose_chem_report), intent(out) :: report
    end function
    integer(c_int) function mose_case_chem_report(ctx, case_id, report) bind(C, name="mose_case_chem_report")
      import :: c_int, c_ptr, mose_chem_report
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      type(mose_chem_report), intent(out) :: report
    end function
    ! serving: asynchronous steps of a case, the stage kernel's timing, the device's memory
    integer(c_int) function mose_case_launch(ctx, case_id, nsteps) bind(C, name="mose_case_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id, nsteps
    end function
    integer(c_int) function mose_case_poll(ctx, case_id, done, rep) bind(C, name="mose_case_poll")
      import :: c_int, c_ptr, mose_report
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      integer(c_int), intent(out) :: done
      type(mose_report), intent(out) :: rep
    end function
    integer(c_int) function mose_case_timing(ctx, case_id, avg_ms, total_ms, launches) bind(C, name="mose_case_timing")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: case_id
      real(c_double), intent(out) :: avg_ms, total_ms
      integer(c_int64_t), intent(out) :: launches
    end function
    integer(c_int) function mose_ctx_info(ctx, total, freeb, cus, lds) bind(C, name="mose_ctx_info")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: total, freeb
      integer(c_int), intent(out) :: cus, lds
    end function
  end interface

contains

  !> The library's last error text as a Fortran string.
  function mose_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    type(c_ptr) :: s
    s = mose_last_error(ctx)
    n = int(c_strlen(s))
    call c_f_pointer(s, p, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function mose_error_text
end module mose_gpu_c
