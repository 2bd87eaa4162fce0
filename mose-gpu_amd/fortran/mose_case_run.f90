!> A Fortran caller of the whole residual path: reads a case (scheme, blocks, geometry, boundary records, far-field
!> rows, thermo tables, initial state and, for a reacting case, the mechanism and the integrator's tolerances) from a
!> raw binary file, steps it through MOSE_GPU_Explicit_Step (Strang-split steps when a mechanism is given) and writes
!> the final state, the time and the step count.
!>
!>   mose_case_run <case.bin> <out.bin>
!>
!> case.bin (little-endian stream; the C layouts of include/mose_gpu.h read as Fortran arrays, fastest index first):
!>   int32 hdr(12) = version (1), nb, nrec, nff, nsteps, path, ti, tj, tk, chem (0/1), method, viscous tables (0/1)
!>   mose_scheme (raw), int32 dims(3, nb)
!>   float64 wm(nsc), cp(NTAB, nsc), h(NTAB, nsc), [mu(NTAB, nsc), k(NTAB, nsc)]
!>   int64 NC, NF1, NF2, NF3; float64 dl(NC, 3), m(NC, 9), vol(NC), yn(NC), fm1(NF1, 4), fm2(NF2, 4), fm3(NF3, 4)
!>   mose_bc_record(nrec), mose_farfield(nff) (raw), float64 P(NC, nprim)
!>   chemistry only: the mechanism as mose_chem_run reads it, then float64 rtol(nsc + 1), atol(nsc + 1)
!> out.bin: float64 P(NC, nprim), time; int64 steps.
program mose_case_run
  use iso_c_binding
  use mose_gpu_c
  use mose_gpu_host
  implicit none
  character(len=512) :: fcase, fout
  integer(c_int32_t) :: hdr(12), mh(6)
  integer(c_int) :: nb, nrec, nff, nsteps, nsc, nprim, case_id, chem_id, rows, na, nt, nl, nr
  integer(c_int64_t) :: NC, NF(3)
  integer(c_int), allocatable, target :: dims(:)
  type(mose_scheme) :: sc
  type(mose_tables) :: tabs
  type(mose_tuning) :: tun, chosen
  type(mose_report) :: rep
  type(mose_mechanism) :: mech
  type(mose_bc_record), allocatable, target :: rec(:)
  type(mose_farfield), allocatable, target :: ff(:)
  real(c_double), allocatable, target :: wm(:), cp(:, :), h(:, :), mu(:, :), kk(:, :)
  real(c_double), allocatable, target :: dl(:), m(:), vol(:), yn(:), fm1(:), fm2(:), fm3(:), P(:)
  real(c_double), allocatable, target :: cwm(:), ccp(:, :), ch(:, :), ni1(:, :), ni2(:, :), eps(:, :)
  real(c_double), allocatable, target :: kf(:, :), kb(:, :), troe(:, :, :), lind(:, :, :)
  real(c_double), allocatable :: rtol(:), atol(:)
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, fcase)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fcase), access='stream', form='unformatted', status='old')
  read(u) hdr
  if (hdr(1) /= 1) error stop 'mose_case_run: unknown case file version'
  nb = hdr(2); nrec = hdr(3); nff = hdr(4); nsteps = hdr(5)
  read(u) sc
  nsc = sc%nsc
  nprim = sc%nsc + 4 + sc%nrans
  allocate(dims(3 * nb))
  read(u) dims
  allocate(wm(nsc), cp(MOSE_NTAB, nsc), h(MOSE_NTAB, nsc), mu(MOSE_NTAB, nsc), kk(MOSE_NTAB, nsc))
  read(u) wm, cp, h
  if (hdr(12) /= 0) then
    read(u) mu, kk
  end if
  read(u) NC, NF
  allocate(dl(3 * NC), m(9 * NC), vol(NC), yn(NC), fm1(4 * NF(1)), fm2(4 * NF(2)), fm3(4 * NF(3)))
  read(u) dl, m, vol, yn, fm1, fm2, fm3
  allocate(rec(max(nrec, 1)), ff(max(nff, 1)))
  if (nrec > 0) read(u) rec(1:nrec)
  if (nff > 0) read(u) ff(1:nff)
  allocate(P(nprim * NC))
  read(u) P

  tabs%nsc = nsc
  tabs%pad = 0
  tabs%wm = c_loc(wm); tabs%cp = c_loc(cp); tabs%h = c_loc(h)
  if (hdr(12) /= 0) then
    tabs%mu = c_loc(mu); tabs%k = c_loc(kk)
  else
    tabs%mu = c_null_ptr; tabs%k = c_null_ptr
  end if
  call mose_check(c_null_ptr, mose_ctx_create(0, tabs, ctx), 'mose_ctx_create')
  call mose_check(ctx, mose_case_create(ctx, sc, nb, dims, dl, m, vol, yn, fm1, fm2, fm3, nrec, rec, nff, ff, &
                                        case_id), 'mose_case_create')
  call mose_check(ctx, mose_state_upload(ctx, case_id, P), 'mose_state_upload')
  tun%path = hdr(6); tun%ti = hdr(7); tun%tj = hdr(8); tun%tk = hdr(9)
  tun%timing = 0; tun%no_graph = 0; tun%reserved = 0
  call mose_check(ctx, mose_case_tune(ctx, case_id, tun, chosen), 'mose_case_tune')

  chem_id = -1
  if (hdr(10) /= 0) then
    read(u) mh
    na = mh(2); nt = mh(3); nl = mh(4)
    rows = mh(6) - mh(5) + 1
    nr = na + nt + nl
    allocate(cwm(mh(1)), ccp(rows, mh(1)), ch(rows, mh(1)), ni1(mh(1) + 1, nr), ni2(mh(1) + 1, nr), eps(mh(1) + 1, nr))
    allocate(kf(rows, max(na, 1)), kb(rows, max(na, 1)), troe(rows, 4, max(nt, 1)), lind(rows, 3, max(nl, 1)))
    read(u) cwm, ccp, ch, ni1, ni2, eps
    if (na > 0) read(u) kf(:, 1:na), kb(:, 1:na)
    if (nt > 0) read(u) troe(:, :, 1:nt)
    if (nl > 0) read(u) lind(:, :, 1:nl)
    allocate(rtol(nsc + 1), atol(nsc + 1))
    read(u) rtol, atol
    mech%ns = mh(1); mech%n_arrh = na; mech%n_troe = nt; mech%n_lind = nl; mech%tmin = mh(5); mech%tmax = mh(6)
    mech%wm = c_loc(cwm); mech%cp = c_loc(ccp); mech%h = c_loc(ch)
    mech%ni1 = c_loc(ni1); mech%ni2 = c_loc(ni2); mech%eps = c_loc(eps)
    mech%kf = c_loc(kf); mech%kb = c_loc(kb); mech%troe = c_loc(troe); mech%lind = c_loc(lind)
    call mose_check(ctx, mose_chem_create(ctx, mech, 1_c_int64_t, chem_id), 'mose_chem_create')
    call MOSE_GPU_Chemistry_Attach(ctx, case_id, chem_id, nsc, hdr(11), rtol, atol)
  end if
  close(u)

  call MOSE_GPU_Explicit_Step(ctx, case_id, nsteps, rep)
  call mose_check(ctx, mose_state_download(ctx, case_id, P), 'mose_state_download')
  write(*,'(a,i0,a,es23.15e3,a,4(1x,i0))') ' steps ', rep%steps, ' time ', rep%time, ' tile', chosen%ti, chosen%tj, &
    chosen%tk, chosen%path
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) P, rep%time, rep%steps
  close(u)
  if (chem_id >= 0) call mose_check(ctx, mose_chem_destroy(ctx, chem_id), 'mose_chem_destroy')
  call mose_check(ctx, mose_case_destroy(ctx, case_id), 'mose_case_destroy')
  call mose_check(ctx, mose_ctx_destroy(ctx), 'mose_ctx_destroy')
end program mose_case_run
