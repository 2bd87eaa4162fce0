!> A reference test case run by the Fortran host from its own files: input.ini and INPUT/ (phase.txt, thermo.dat,
!> transport.dat for a viscous case, ic.tec, bc.txt), read by mose_case_files; the geometry from the nodes
!> (mose_metrics_build, Setup_Metrics); the case created, uploaded and stepped through MOSE_GPU_Explicit_Step.
!>
!>   mose_case_dir <case_dir> <nsteps> <out.bin> [path [ti tj tk]]
!>   mose_case_dir <case_dir> <nsteps> <out.bin> dump
!>
!> nsteps < 0 takes [MOSE-Parameters] iter-threshold; a case without one runs while its time is below time-threshold,
!> one step at a time (the reference's time loop, do while time < timemax).  path: 0 = fused (the default), 1 = plain,
!> 2 = performance;
!> the tile 0 0 0 (the default) asks the autotuner.  out.bin (little-endian stream): float64 P(NC, nprim), time;
!> int64 steps.  "dump" writes what was read instead and touches no GPU:
!>   int32 nb, nsc, nrans, nrec, ndir; mose_scheme (raw); int32 dims(3, nb); float64 wm(nsc), cp(NTAB, nsc),
!>   h(NTAB, nsc), [mu, k for a viscous case]; int64 NC, NF(3); float64 dl(NC, 3), m(NC, 9), vol(NC),
!>   fm1(NF1, 4), fm2(NF2, 4), fm3(NF3, 4); mose_bc_record(nrec) (raw); int32 nff, mose_farfield(nff) (raw);
!>   float64 P(NC, nprim).
program mose_case_dir
  use iso_c_binding
  use mose_gpu_c
  use mose_gpu_host
  use mose_case_files
  implicit none
  character(len=1024) :: dir, fout, arg
  type(ini_t) :: ini
  type(mose_scheme) :: sc
  type(mose_tables) :: tabs
  type(mose_tuning) :: tun, chosen
  type(mose_report) :: rep
  type(mose_bc_record), allocatable, target :: rec(:)
  type(mose_farfield), allocatable, target :: ff(:)
  real(c_double), allocatable, target :: wm(:), cp(:, :), h(:, :), mu(:, :), kk(:, :), nodes(:), fields(:, :)
  real(c_double), allocatable, target :: dl(:), m(:), vol(:), yn(:), fm1(:), fm2(:), fm3(:), P(:, :)
  real(c_double) :: tmax
  integer, allocatable :: dims(:, :)
  integer(c_int), allocatable, target :: cdims(:)
  integer(c_int64_t), allocatable :: boff(:)
  integer(c_int64_t) :: NC, NF(3), row0, c0
  integer :: nb, nvcc, nrec, nsc, nprim, ndir, nsteps, b, i, j, k, s, u, ios, nargs, nff
  integer(c_int) :: case_id
  logical :: viscous, dump
  type(c_ptr) :: ctx

  nargs = command_argument_count()
  if (nargs < 3) then
    write(0, '(a)') 'usage: mose_case_dir <case_dir> <nsteps> <out.bin> [path [ti tj tk] | dump]'
    error stop 2
  end if
  call get_command_argument(1, dir)
  call get_command_argument(2, arg)
  read(arg, *, iostat=ios) nsteps
  if (ios /= 0) call case_stop('nsteps: not an integer')
  call get_command_argument(3, fout)
  dump = .false.
  tun%path = MOSE_PATH_FUSED; tun%ti = 0; tun%tj = 0; tun%tk = 0
  tun%timing = 0; tun%no_graph = 0; tun%reserved = 0
  if (nargs >= 4) then
    call get_command_argument(4, arg)
    if (trim(arg) == 'dump') then
      dump = .true.
    else
      read(arg, *, iostat=ios) tun%path
      if (ios /= 0) call case_stop('path: not an integer')
      if (nargs >= 7) then
        call get_command_argument(5, arg); read(arg, *) tun%ti
        call get_command_argument(6, arg); read(arg, *) tun%tj
        call get_command_argument(7, arg); read(arg, *) tun%tk
      end if
    end if
  end if

  ! ---- the files
  call ini_read(trim(dir) // '/input.ini', ini)
  viscous = trim(ini_get(ini, 'MOSE-Physics', 'equations', '')) == 'navier-stokes'
  call read_tables(trim(dir) // '/INPUT', viscous, nsc, wm, cp, h, mu, kk)
  call read_block_tec(trim(dir) // '/INPUT/ic.tec', nb, dims, nodes, fields, nvcc)
  ndir = mesh_ndir(dims(1, 1), dims(2, 1), dims(3, 1), nodes(1:3 * (dims(1, 1) + 1) * (dims(2, 1) + 1) * (dims(3, 1) + 1)))
  call scheme_from_ini(ini, ndir, nsc, sc)
  nprim = sc%nsc + 4 + sc%nrans
  if (nvcc /= nprim) then
    write(arg, '(a,i0,a,i0,a)') 'ic.tec: ', nvcc, ' cell variables, ', nprim, ' expected'
    call case_stop(trim(arg))
  end if
  call read_bc_open(trim(dir) // '/INPUT/bc.txt', rec, nrec)
  call sort_records(rec, nrec)
  tmax = 0.0_c_double
  if (nsteps < 0) then
    nsteps = int(num_or_zero(ini_get(ini, 'MOSE-Parameters', 'iter-threshold', '0')))
    if (nsteps == 0) tmax = num_or_zero(ini_get(ini, 'MOSE-Parameters', 'time-threshold', '0'))
  end if

  ! ---- the arena: the initial state of every block's interior, ghosts from the nearest interior cell
  allocate(boff(nb), cdims(3 * nb))
  call arena_sizes(nb, dims, NC, NF, boff)
  cdims = reshape(dims, [3 * nb])
  allocate(P(NC, nprim))
  P = 0.0_c_double
  c0 = 0
  do b = 1, nb
    do k = 1, dims(3, b)
      do j = 1, dims(2, b)
        do i = 1, dims(1, b)
          c0 = c0 + 1
          row0 = boff(b) + (i + MOSE_GC) + int(j + MOSE_GC - 1, c_int64_t) * (dims(1, b) + 2 * MOSE_GC) + &
                 int(k + MOSE_GC - 1, c_int64_t) * (dims(1, b) + 2 * MOSE_GC) * (dims(2, b) + 2 * MOSE_GC)
          do s = 1, nprim
            P(row0, s) = fields(c0, s)
          end do
        end do
      end do
    end do
  end do
  call fill_ghosts_nearest(nb, dims, boff, NC, nprim, P)

  ! ---- the geometry (Setup_Metrics)
  allocate(dl(3 * NC), m(9 * NC), vol(NC), yn(NC), fm1(4 * NF(1)), fm2(4 * NF(2)), fm3(4 * NF(3)))
  dl = 0.0_c_double; m = 0.0_c_double; vol = 0.0_c_double; yn = 0.0_c_double
  fm1 = 0.0_c_double; fm2 = 0.0_c_double; fm3 = 0.0_c_double
  if (mose_metrics_build(nb, cdims, nodes, nrec, rec, ndir, dl, m, vol, yn, fm1, fm2, fm3) /= MOSE_OK) &
    call case_stop('mose_metrics_build: a degenerate cell or a bad record')
  call inlet_rows(nb, dims, NF, rec, nrec, fm1, fm2, fm3, nsc, sc%nrans, ff, nff)

  if (dump) then
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) int([nb, nsc, sc%nrans, nrec, ndir], c_int32_t)
    write(u) sc
    write(u) cdims
    write(u) wm, cp, h
    if (viscous) write(u) mu, kk
    write(u) NC, NF
    write(u) dl, m, vol, fm1, fm2, fm3
    if (nrec > 0) write(u) rec(1:nrec)
    write(u) int(nff, c_int32_t)
    if (nff > 0) write(u) ff(1:nff)
    write(u) P
    close(u)
  else
    call run()
  end if

contains

  !> the run: context, case, upload, tune, steps, download, out.bin
  subroutine run()
    tabs%nsc = nsc
    tabs%pad = 0
    tabs%wm = c_loc(wm); tabs%cp = c_loc(cp); tabs%h = c_loc(h)
    if (viscous) then
      tabs%mu = c_loc(mu); tabs%k = c_loc(kk)
    else
      tabs%mu = c_null_ptr; tabs%k = c_null_ptr
    end if
    call mose_check(c_null_ptr, mose_ctx_create(0, tabs, ctx), 'mose_ctx_create')
    call mose_check(ctx, mose_case_create(ctx, sc, nb, cdims, dl, m, vol, yn, fm1, fm2, fm3, nrec, rec, nff, ff, &
                                          case_id), 'mose_case_create')
    call mose_check(ctx, mose_state_upload(ctx, case_id, P), 'mose_state_upload')
    call mose_check(ctx, mose_case_tune(ctx, case_id, tun, chosen), 'mose_case_tune')
    if (tmax > 0.0_c_double) then
      rep%time = 0.0_c_double
      do while (rep%time < tmax)
        call MOSE_GPU_Explicit_Step(ctx, case_id, 1, rep)
      end do
    else
      call MOSE_GPU_Explicit_Step(ctx, case_id, nsteps, rep)
    end if
    call mose_check(ctx, mose_state_download(ctx, case_id, P), 'mose_state_download')
    write(*, '(a,a,a,i0,a,es23.15e3,a,4(1x,i0))') ' case ', trim(dir), ' steps ', rep%steps, ' time ', rep%time, &
      ' tile', chosen%ti, chosen%tj, chosen%tk, chosen%path
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) P, rep%time, rep%steps
    close(u)
    call mose_check(ctx, mose_case_destroy(ctx, case_id), 'mose_case_destroy')
    call mose_check(ctx, mose_ctx_destroy(ctx), 'mose_ctx_destroy')
  end subroutine run

  real(c_double) function num_or_zero(s)
    character(len=*), intent(in) :: s
    read(s, *, iostat=ios) num_or_zero
    if (ios /= 0) num_or_zero = 0.0_c_double
  end function num_or_zero

end program mose_case_dir
