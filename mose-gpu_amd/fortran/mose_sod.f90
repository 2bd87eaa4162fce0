!> The 1D shock tube through the Fortran host and the C ABI: n x 1 x 1 cells on x in [0, 1], y in [0, 0.01],
!> z in [-5e-4, 5e-4], the constant-cp air table of test/1D/Sod79/INPUT (cp = 1004.5, h = cp T, wm = 28.970418),
!> (rho, u, p) = (1, 0, 1) | (0.125, 0, 0.1) at x = 0.5, minmod + HLLC, SSP-RK3, CFL 0.5, extrapolation (MOSE 400)
!> on every face.  Usage: mose_sod [n] [nsteps] [path: 0 fused | 1 plain].  Prints the state hash, the time and the
!> residual norm of the last step, and writes rho, u, p of the cells to mose_sod.out.
program mose_sod
  use iso_c_binding
  use mose_gpu_c
  use mose_gpu_host
  implicit none
  integer(c_int) :: n, nsteps, path, ncell, nrec, rc, cid, e, f, i, ndev
  integer(c_int), target :: dims(3)
  real(c_double), allocatable, target :: nodes(:), dl(:), m(:), vol(:), yn(:), fm1(:), fm2(:), fm3(:), P(:, :)
  real(c_double), allocatable, target :: cp(:), h(:), wm(:)
  type(mose_bc_record), allocatable :: rec(:)
  type(mose_farfield) :: ff(1)
  type(mose_tables) :: tab
  type(mose_scheme) :: sc
  type(mose_report) :: rep
  type(mose_tuning) :: tun, chosen
  type(c_ptr) :: ctx
  integer(c_int64_t) :: hlo, hhi
  real(c_double) :: rn(5), x
  character(len=32) :: arg
  integer :: pi, pj, NC, row, k, j

  n = 1000; nsteps = 200; path = MOSE_PATH_FUSED
  if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read(arg, *) n; end if
  if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read(arg, *) nsteps; end if
  if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read(arg, *) path; end if
  rc = mose_device_count(ndev)
  if (ndev < 1) error stop 'mose_sod: no GPU'

  dims = [n, 1, 1]
  pi = n + 2 * MOSE_GC; pj = 1 + 2 * MOSE_GC
  NC = pi * pj * (1 + 2 * MOSE_GC)
  ncell = n
  ! nodes, i fastest: (n+1) x 2 x 2 points
  allocate(nodes(3 * (n + 1) * 4))
  e = 0
  do k = 0, 1
    do j = 0, 1
      do i = 0, n
        nodes(e + 1) = real(i, c_double) / real(n, c_double)
        nodes(e + 2) = 0.01d0 * j
        nodes(e + 3) = -5d-4 + 1d-3 * k
        e = e + 3
      end do
    end do
  end do
  ! records: every boundary cell-face, grouped by face 1..6, type 11
  nrec = 2 + 4 * n
  allocate(rec(nrec))
  e = 0
  do f = 1, 6
    if (f <= 2) then
      e = e + 1
      rec(e) = blank_record(merge(1, n, f == 1), 1, 1, f)
    else
      do i = 1, n
        e = e + 1
        rec(e) = blank_record(i, 1, 1, f)
      end do
    end if
  end do
  allocate(dl(3 * NC), m(9 * NC), vol(NC), yn(NC), fm1(4 * (n + 1)), fm2(4 * 2 * n), fm3(4 * 2 * n))
  call mose_check(c_null_ptr, mose_metrics_build(1, dims, nodes, nrec, rec, 1, dl, m, vol, yn, fm1, fm2, fm3), &
                  'mose_metrics_build')

  allocate(cp(MOSE_NTAB), h(MOSE_NTAB), wm(1))
  do i = 0, MOSE_NTAB - 1
    cp(i + 1) = 1004.5d0
    h(i + 1) = 1004.5d0 * i
  end do
  wm(1) = 28.970418d0
  tab%nsc = 1; tab%pad = 0
  tab%wm = c_loc(wm); tab%cp = c_loc(cp); tab%h = c_loc(h); tab%mu = c_null_ptr; tab%k = c_null_ptr
  call mose_check(c_null_ptr, mose_ctx_create(0, tab, ctx), 'mose_ctx_create')

  sc%nsc = 1; sc%nrans = 0; sc%limiter = MOSE_LIM_MINMOD; sc%riemann = MOSE_RS_HLLC; sc%viscous = 0
  sc%n_rk = 3; sc%ndir = 1; sc%point_implicit = 0; sc%thermo_guard = 1; sc%reserved = 0
  sc%mco = 1d0; sc%cfl = 0.5d0; sc%vnn = 0.5d0; sc%prandtl = 0d0; sc%sc = 0.5d0; sc%sct = 0.9d0; sc%prt = 0.9d0
  sc%reserved_d = 0d0
  call mose_check(ctx, mose_case_create(ctx, sc, 1, dims, dl, m, vol, yn, fm1, fm2, fm3, nrec, rec, 0, ff, cid), &
                  'mose_case_create')

  ! the state, var-last: P(row, var); ghosts take the nearest interior cell
  allocate(P(NC, 5))
  P = 0d0
  do k = 1 - MOSE_GC, 1 + MOSE_GC
    do j = 1 - MOSE_GC, 1 + MOSE_GC
      do i = 1 - MOSE_GC, n + MOSE_GC
        x = (real(min(max(i, 1), n), c_double) - 0.5d0) / real(n, c_double)
        row = (i + MOSE_GC - 1) + (j + MOSE_GC - 1) * pi + (k + MOSE_GC - 1) * pi * pj + 1
        P(row, 1) = merge(1d0, 0.125d0, x < 0.5d0)
        P(row, 5) = merge(1d0, 0.1d0, x < 0.5d0)
      end do
    end do
  end do
  call mose_check(ctx, mose_state_upload(ctx, cid, P), 'mose_state_upload')
  tun%path = path; tun%ti = merge(256, 0, path == MOSE_PATH_FUSED); tun%tj = merge(1, 0, path == MOSE_PATH_FUSED)
  tun%tk = merge(1, 0, path == MOSE_PATH_FUSED); tun%timing = 0; tun%no_graph = 0; tun%reserved = 0
  call mose_check(ctx, mose_case_tune(ctx, cid, tun, chosen), 'mose_case_tune')

  call MOSE_GPU_Explicit_Step(ctx, cid, nsteps, rep)

  call mose_check(ctx, mose_case_state_hash(ctx, cid, hlo, hhi), 'mose_case_state_hash')
  call mose_check(ctx, mose_case_residual_norm(ctx, cid, rn), 'mose_case_residual_norm')
  call mose_check(ctx, mose_state_download(ctx, cid, P), 'mose_state_download')
  write(*,'(a,i0,1x,i0)') 'hash ', hlo, hhi
  write(*,'(a,es24.16)') 'time ', rep%time
  write(*,'(a,5es24.16)') 'resnorm ', rn
  open(unit=10, file='mose_sod.out', status='replace')
  do i = 1, n
    row = (i + MOSE_GC - 1) + (MOSE_GC) * pi + (MOSE_GC) * pi * pj + 1
    write(10, '(3es24.16)') P(row, 1), P(row, 2), P(row, 5)
  end do
  close(10)
  call mose_check(ctx, mose_case_destroy(ctx, cid), 'mose_case_destroy')
  rc = mose_ctx_destroy(ctx)

contains

  function blank_record(ii, jj, kk, ff_) result(r)
    integer, intent(in) :: ii, jj, kk, ff_
    type(mose_bc_record) :: r
    r%b = 0; r%i = ii; r%j = jj; r%k = kk; r%f = ff_; r%typ = MOSE_BC_EXTRAPOLATION
    r%bs = 0; r%is = 0; r%js = 0; r%ks = 0; r%fs = 0
    r%d11 = 1; r%d12 = 0; r%d21 = 0; r%d22 = 1; r%periodic = 0; r%ff = -1; r%pad = 0
    r%tw = 0d0; r%tvec = 0d0
  end function blank_record
end program mose_sod
