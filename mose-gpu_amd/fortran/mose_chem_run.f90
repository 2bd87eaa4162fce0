!> A Fortran caller of the chemistry path: reads a mechanism and a batch of cells from raw binary files, integrates
!> them over dt through MOSE_GPU_Chemistry_Newstate, and writes the new state and hr.
!>
!>   mose_chem_run <mechanism.bin> <cells.bin> <out.bin> <dt> <method> <rtol> <atol>
!>
!> mechanism.bin: int32 ns, n_arrh, n_troe, n_lind, tmin, tmax, then float64 wm(ns), cp(rows, ns), h(rows, ns),
!>   ni1/ni2/eps((ns + 1), nr), kf(rows, n_arrh), kb(rows, n_arrh), troe(rows, 4, n_troe), lind(rows, 3, n_lind)
!>   (the C layouts of mose_mechanism, read as Fortran arrays with the fastest index first).
!> cells.bin: int64 ncell, then float64 state(ncell, ns + 1).  out.bin: float64 state(ncell, ns + 1), hr(ncell).
program mose_chem_run
  use iso_c_binding
  use mose_gpu_c
  use mose_gpu_host
  implicit none
  character(len=512) :: fmech, fcells, fout, arg
  integer(c_int32_t) :: hdr(6)
  integer(c_int) :: ns, na, nt, nl, rows, nr, method, chem_id
  integer(c_int64_t) :: ncell, nfailed
  real(c_double), allocatable, target :: wm(:), cp(:, :), h(:, :), ni1(:, :), ni2(:, :), eps(:, :), kf(:, :), kb(:, :)
  real(c_double), allocatable, target :: troe(:, :, :), lind(:, :, :), mu(:, :), kk(:, :)
  real(c_double), allocatable :: state(:, :), hr(:), rtol(:), atol(:)
  real(c_double) :: dt, rt, at
  type(mose_mechanism) :: mech
  type(mose_tables) :: tabs
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, fmech)
  call get_command_argument(2, fcells)
  call get_command_argument(3, fout)
  call get_command_argument(4, arg); read(arg, *) dt
  call get_command_argument(5, arg); read(arg, *) method
  call get_command_argument(6, arg); read(arg, *) rt
  call get_command_argument(7, arg); read(arg, *) at

  open(newunit=u, file=trim(fmech), access='stream', form='unformatted', status='old')
  read(u) hdr
  ns = hdr(1); na = hdr(2); nt = hdr(3); nl = hdr(4)
  rows = hdr(6) - hdr(5) + 1
  nr = na + nt + nl
  allocate(wm(ns), cp(rows, ns), h(rows, ns), ni1(ns + 1, nr), ni2(ns + 1, nr), eps(ns + 1, nr))
  allocate(kf(rows, max(na, 1)), kb(rows, max(na, 1)), troe(rows, 4, max(nt, 1)), lind(rows, 3, max(nl, 1)))
  read(u) wm, cp, h, ni1, ni2, eps
  if (na > 0) read(u) kf(:, 1:na), kb(:, 1:na)
  if (nt > 0) read(u) troe(:, :, 1:nt)
  if (nl > 0) read(u) lind(:, :, 1:nl)
  close(u)

  open(newunit=u, file=trim(fcells), access='stream', form='unformatted', status='old')
  read(u) ncell
  allocate(state(ncell, ns + 1), hr(ncell))
  read(u) state
  close(u)

  ! a context needs the thermo tables of the flow path; the chemistry object carries its own
  allocate(mu(rows, ns), kk(rows, ns))
  mu = 1.0d-5
  kk = 1.0d-2
  tabs%nsc = ns
  tabs%pad = 0
  tabs%wm = c_loc(wm); tabs%cp = c_loc(cp); tabs%h = c_loc(h); tabs%mu = c_loc(mu); tabs%k = c_loc(kk)
  call mose_check(c_null_ptr, mose_ctx_create(0, tabs, ctx), 'mose_ctx_create')

  mech%ns = ns; mech%n_arrh = na; mech%n_troe = nt; mech%n_lind = nl; mech%tmin = hdr(5); mech%tmax = hdr(6)
  mech%wm = c_loc(wm); mech%cp = c_loc(cp); mech%h = c_loc(h)
  mech%ni1 = c_loc(ni1); mech%ni2 = c_loc(ni2); mech%eps = c_loc(eps)
  mech%kf = c_loc(kf); mech%kb = c_loc(kb); mech%troe = c_loc(troe); mech%lind = c_loc(lind)
  call mose_check(ctx, mose_chem_create(ctx, mech, ncell, chem_id), 'mose_chem_create')

  allocate(rtol(ns + 1), atol(ns + 1))
  rtol = rt
  atol = at
  call MOSE_GPU_Chemistry_Newstate(ctx, chem_id, ncell, ns, state, dt, method, rtol, atol, hr, nfailed)
  write(*,'(a,i0,a,i0)') ' chemistry cells ', ncell, ', failed ', nfailed

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) state, hr
  close(u)
  call mose_check(ctx, mose_chem_destroy(ctx, chem_id), 'mose_chem_destroy')
  call mose_check(ctx, mose_ctx_destroy(ctx), 'mose_ctx_destroy')
end program mose_chem_run
