!> The reference's case format read by the Fortran host: input.ini (the MOSE_open dialect), INPUT/phase.txt,
!> INPUT/thermo.dat and INPUT/transport.dat (Tecplot POINT zones, one per species), INPUT/ic.tec (Tecplot BLOCK
!> zones, one per block: nodal x y z, cell-centred primitives) and INPUT/bc.txt (three-digit MOSE_open boundary
!> types), into the arrays of the C ABI (include/mose_gpu.h): the scheme, the thermo tables, the blocks' nodes and
!> initial state in the var-last arena, and the boundary records.
!>
!> Follows IO_Ini.f90:9-193 and IO_Ini_Open.f90 (keys), Load_ThermoTransport.f90:7-275 (tables on T = 1, 2, ... K),
!> IO_BC_Open.f90:132-399 (bc.txt records and their property lines) and Check_Mesh_Type (Lib_Metrics.f90:139-196).
!> What this path does not run (local time stepping, other solvers and boundary types, multicomponent diffusion,
!> axisymmetric meshes) stops with a message that names it.
module mose_case_files
  use iso_c_binding
  use mose_gpu_c
  implicit none
  private
  public :: ini_t, ini_read, ini_get, ini_has, case_stop, read_tables, read_block_tec, read_bc_open, mesh_ndir, &
            scheme_from_ini, fill_ghosts_nearest, sort_records, arena_sizes, inlet_rows

  integer, parameter :: LL = 16384
  !> IO_BC_Open.f90:40: the pamb marker of a 405 record, which BC_Inflow_Outflow (Lib_BC.f90:1574-1583) sends to the
  !> supersonic closure at any Mach and closes on outward flow
  real(c_double), parameter :: PAMB_FORCE_SUPERSONIC = -2.71_c_double

  !> a 405 record's property line (IO_BC_Open.f90:286-300): Mach, static T and p, the two angles (radians, or the
  !> token "normal": the face normal's angle, resolved after the metrics), rel_fac, then the mass fractions and the
  !> RANS values
  type :: inlet_t
    real(c_double) :: mach = 0, t = 0, p = 0, alpha = 0, beta = 0, rel = 0
    logical :: an = .false., bn = .false.
    integer :: nv = 0
    real(c_double) :: v(MOSE_MAX_SPECIES + 2) = 0
  end type inlet_t
  type(inlet_t), allocatable :: inlets(:)
  integer :: n_inlets = 0

  type :: ini_t
    integer :: n = 0
    character(len=64), allocatable :: sec(:), key(:)
    character(len=256), allocatable :: val(:)
  end type ini_t

contains

  subroutine case_stop(msg)
    character(len=*), intent(in) :: msg
    write(0, '(a)') 'mose case: ' // trim(msg)
    error stop 2
  end subroutine case_stop

  ! ---------------------------------------------------------------------------------------------- input.ini
  !> [section] / key = value lines; ';' and '#' start a comment (IO_Ini.f90:9-89)
  subroutine ini_read(path, ini)
    character(len=*), intent(in) :: path
    type(ini_t), intent(out) :: ini
    character(len=LL) :: line
    character(len=64) :: sec
    integer :: u, ios, p, cap
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call case_stop('cannot open ' // trim(path))
    cap = 64
    allocate(ini%sec(cap), ini%key(cap), ini%val(cap))
    sec = ''
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      p = scan(line, ';#')
      if (p > 0) line(p:) = ' '
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '[') then
        p = index(line, ']')
        if (p > 2) sec = adjustl(line(2:p - 1))
        cycle
      end if
      p = index(line, '=')
      if (p < 2 .or. len_trim(sec) == 0) cycle
      if (ini%n == cap) call grow()
      ini%n = ini%n + 1
      ini%sec(ini%n) = sec
      ini%key(ini%n) = adjustl(line(1:p - 1))
      ini%val(ini%n) = adjustl(line(p + 1:))
    end do
    close(u)
  contains
    subroutine grow()
      character(len=64), allocatable :: s2(:), k2(:)
      character(len=256), allocatable :: v2(:)
      allocate(s2(2 * cap), k2(2 * cap), v2(2 * cap))
      s2(1:cap) = ini%sec; k2(1:cap) = ini%key; v2(1:cap) = ini%val
      call move_alloc(s2, ini%sec); call move_alloc(k2, ini%key); call move_alloc(v2, ini%val)
      cap = 2 * cap
    end subroutine grow
  end subroutine ini_read

  logical function ini_has(ini, sec, key)
    type(ini_t), intent(in) :: ini
    character(len=*), intent(in) :: sec, key
    integer :: q
    ini_has = .false.
    do q = 1, ini%n
      if (trim(ini%sec(q)) == sec .and. trim(ini%key(q)) == key) ini_has = .true.
    end do
  end function ini_has

  !> the last value of sec/key, or def
  function ini_get(ini, sec, key, def) result(v)
    type(ini_t), intent(in) :: ini
    character(len=*), intent(in) :: sec, key, def
    character(len=256) :: v
    integer :: q
    v = def
    do q = 1, ini%n
      if (trim(ini%sec(q)) == sec .and. trim(ini%key(q)) == key) v = ini%val(q)
    end do
  end function ini_get

  logical function truth(v)
    character(len=*), intent(in) :: v
    character(len=16) :: w
    w = lower(adjustl(v))
    truth = any(trim(w) == [character(len=16) :: 'true', 't', '.true.', 'yes', '1'])
  end function truth

  function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: q
    t = s
    do q = 1, len(s)
      if (t(q:q) >= 'A' .and. t(q:q) <= 'Z') t(q:q) = achar(iachar(t(q:q)) + 32)
    end do
  end function lower

  real(c_double) function num(s, what)
    character(len=*), intent(in) :: s, what
    integer :: ios
    read(s, *, iostat=ios) num
    if (ios /= 0) call case_stop(what // ' = ' // trim(s) // ': not a number')
  end function num

  !> The scheme of the [MOSE-Physics] / [MOSE-Numerics] keys (IO_Ini_Open.f90), the subset this path runs
  subroutine scheme_from_ini(ini, ndir, nsc, sc)
    type(ini_t), intent(in) :: ini
    integer, intent(in) :: ndir, nsc
    type(mose_scheme), intent(out) :: sc
    character(len=256) :: eq, turb, ts, recon, lim, rs, iv
    logical :: viscous
    real(c_double) :: x
    integer :: ios
    sc%nsc = nsc; sc%nrans = 0; sc%reserved = 0; sc%reserved_d = 0.0_c_double; sc%integration = MOSE_INT_CONS
    sc%irs = 0; sc%irs_njacobi = 2; sc%irs_beta = 0.0_c_double
    sc%thermo_guard = 1; sc%point_implicit = 1; sc%ndir = ndir
    sc%sct = 0.9_c_double; sc%prt = 0.9_c_double; sc%vnn = 0.5_c_double; sc%prandtl = 0.0_c_double
    sc%sc = 0.5_c_double
    eq = ini_get(ini, 'MOSE-Physics', 'equations', '')
    if (trim(eq) /= 'euler' .and. trim(eq) /= 'navier-stokes') call case_stop('equations = ' // trim(eq) // &
                                                                             ': euler or navier-stokes')
    viscous = trim(eq) == 'navier-stokes'
    sc%viscous = merge(1, 0, viscous)
    turb = 'none'
    if (viscous) turb = ini_get(ini, 'MOSE-Physics', 'turbulence', 'none')
    if (trim(turb) /= 'none' .and. trim(turb) /= 'SA') call case_stop('turbulence = ' // trim(turb) // &
                                                                     ': only SA is on this path')
    if (trim(turb) == 'SA') then
      sc%nrans = 1
      sc%point_implicit = merge(1, 0, truth(ini_get(ini, 'MOSE-Turbulence', 'point-implicit', 'true')))
    end if
    if (.not. truth(ini_get(ini, 'MOSE-Numerics', 'time-accurate', 'false'))) &
      call case_stop('time-accurate = false (local time stepping) is not on this path')
    ts = ini_get(ini, 'MOSE-Numerics', 'time-scheme', '')
    select case (trim(ts))
    case ('euler'); sc%n_rk = 1
    case ('RK2'); sc%n_rk = 2
    case ('RK3'); sc%n_rk = 3
    case default; call case_stop('time-scheme = ' // trim(ts) // ': euler, RK2 or RK3')
    end select
    ! the state update's variables (Assign_Integration_Variables, Lib_Newstate.f90:43-60)
    iv = ini_get(ini, 'MOSE-Numerics', 'integration-variables', 'cons')
    select case (trim(iv))
    case ('cons'); sc%integration = MOSE_INT_CONS
    case ('prim'); sc%integration = MOSE_INT_PRIM
    case default; call case_stop('integration-variables = ' // trim(iv) // &
                                 ': cons or prim (the preconditioned update, prec, is not on this path)')
    end select
    ! implicit residual smoothing: on with irs = true or irs-beta > 0, beta = irs-beta (IO_Ini_Open.f90:138-146)
    read(ini_get(ini, 'MOSE-Numerics', 'irs-beta', '0.0'), *, iostat=ios) x
    if (ios /= 0) x = 0.0_c_double
    if (x > 0.0_c_double .or. truth(ini_get(ini, 'MOSE-Numerics', 'irs', 'false'))) then
      sc%irs = 1; sc%irs_beta = x
    end if
    recon = ini_get(ini, 'MOSE-Numerics', 'space-reconstruction', '')
    if (trim(recon) /= 'MUSCL' .and. trim(recon) /= 'first-order') &
      call case_stop('space-reconstruction = ' // trim(recon) // ': MUSCL or first-order')
    if (trim(recon) == 'first-order') then
      lim = 'first-order'
    else
      lim = lower(ini_get(ini, 'MOSE-Numerics', 'flux-limiter', ''))
    end if
    select case (trim(lim))
    case ('minmod'); sc%limiter = MOSE_LIM_MINMOD
    case ('vanleer'); sc%limiter = MOSE_LIM_VANLEER
    case ('vanalbada'); sc%limiter = MOSE_LIM_VANALBADA
    case ('mc'); sc%limiter = MOSE_LIM_MC
    case ('superbee'); sc%limiter = MOSE_LIM_SUPERBEE
    case ('first-order'); sc%limiter = MOSE_LIM_FIRST
    case default; call case_stop('flux-limiter = ' // trim(lim) // ': not a limiter of this path')
    end select
    rs = ini_get(ini, 'MOSE-Numerics', 'riemann-solver', 'HLLC')
    select case (trim(rs))
    case ('HLLC'); sc%riemann = MOSE_RS_HLLC
    case ('HLLE'); sc%riemann = MOSE_RS_HLLE
    case ('SLAU'); sc%riemann = MOSE_RS_SLAU
    case ('AUSM+'); sc%riemann = MOSE_RS_AUSMP
    case ('LMRoe'); sc%riemann = MOSE_RS_LMROE
    case default; call case_stop('riemann-solver = ' // trim(rs) // ': HLLC, HLLE, SLAU, AUSM+ or LMRoe on this path')
    end select
    sc%mco = num(ini_get(ini, 'MOSE-Numerics', 'riemann-options-Mco', '1.0'), 'riemann-options-Mco')
    if (.not. ini_has(ini, 'MOSE-Numerics', 'cfl')) call case_stop('[MOSE-Numerics] cfl is required')
    sc%cfl = num(ini_get(ini, 'MOSE-Numerics', 'cfl', ''), 'cfl')
    if (viscous) then
      if (.not. ini_has(ini, 'MOSE-Numerics', 'vnn')) &
        call case_stop('[MOSE-Numerics] vnn is required by this loader for a viscous case')
      if (.not. ini_has(ini, 'MOSE-Physics', 'schmidt')) &
        call case_stop('a viscous case needs [MOSE-Physics] schmidt > 0 (multicomponent diffusion is not on this path)')
      sc%sc = num(ini_get(ini, 'MOSE-Physics', 'schmidt', ''), 'schmidt')
      if (.not. (sc%sc > 0.0_c_double)) &
        call case_stop('a viscous case needs [MOSE-Physics] schmidt > 0 (multicomponent diffusion is not on this path)')
      sc%vnn = num(ini_get(ini, 'MOSE-Numerics', 'vnn', ''), 'vnn')
      sc%prandtl = num(ini_get(ini, 'MOSE-Physics', 'prandtl', '0'), 'prandtl')
    end if
  end subroutine scheme_from_ini

  ! ---------------------------------------------------------------------------------------------- tables
  !> the numbers of one line (',' separates like a blank); n = 0 when the line holds anything else
  subroutine line_numbers(line, x, n)
    character(len=*), intent(in) :: line
    real(c_double), intent(out) :: x(:)
    integer, intent(out) :: n
    character(len=len(line)) :: s
    integer :: p, q, ios
    s = line
    do p = 1, len_trim(s)
      if (s(p:p) == ',' .or. s(p:p) == achar(9)) s(p:p) = ' '
    end do
    n = 0
    p = 1
    do
      do while (p <= len_trim(s))
        if (s(p:p) /= ' ') exit
        p = p + 1
      end do
      if (p > len_trim(s)) exit
      q = index(s(p:), ' ') + p - 2
      if (q < p) q = len_trim(s)
      if (n == size(x)) exit
      read(s(p:q), *, iostat=ios) x(n + 1)
      if (ios /= 0) then
        n = 0
        return
      end if
      n = n + 1
      p = q + 1
    end do
  end subroutine line_numbers

  !> The zones of a Tecplot POINT file (one per species): rows of T and two columns ca, cb, as this library's rows
  !> T = 0..NTAB-1 (Load_ThermoTransport.f90): rows must run T = 1, 2, ... K; row 0 repeats row 1 (or 0 for an
  !> enthalpy, h0 = true); rows above the file's last repeat it.
  subroutine read_point_zones(path, nsc, ca, cb, h0, ta, tb)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nsc, ca, cb
    logical, intent(in) :: h0
    real(c_double), intent(out) :: ta(MOSE_NTAB, nsc), tb(MOSE_NTAB, nsc)
    character(len=LL) :: line
    real(c_double) :: x(16)
    integer :: u, ios, n, z, r
    character(len=16) :: head
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call case_stop('cannot open ' // trim(path))
    z = 0
    r = 0
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      head = adjustl(line)
      if (lower(head(1:4)) == 'zone') then
        if (z > 0) call finish_zone()
        z = z + 1
        if (z > nsc) call case_stop(trim(path) // ': more zones than species')
        r = 0
        cycle
      end if
      call line_numbers(line, x, n)
      if (n < max(ca, cb) .or. z == 0) cycle
      r = r + 1
      if (r > MOSE_NTAB - 1) call case_stop(trim(path) // ': more rows than this library''s tables (T <= 10000 K)')
      if (nint(x(1)) /= r .or. abs(x(1) - r) > 1e-9_c_double) &
        call case_stop(trim(path) // ': rows must run T = 1, 2, ... K')
      ta(r + 1, z) = x(ca)
      tb(r + 1, z) = x(cb)
    end do
    close(u)
    if (z > 0) call finish_zone()
    if (z /= nsc) call case_stop(trim(path) // ': fewer zones than species')
  contains
    subroutine finish_zone()
      if (r < 1) call case_stop(trim(path) // ': an empty zone')
      ta(r + 2:, z) = ta(r + 1, z)
      tb(r + 2:, z) = tb(r + 1, z)
      ta(1, z) = ta(2, z)
      tb(1, z) = merge(0.0_c_double, tb(2, z), h0)
    end subroutine finish_zone
  end subroutine read_point_zones

  !> phase.txt (a header line, then name and molecular weight per species), thermo.dat (cp, h) and, for a
  !> viscous case, transport.dat (mu, k)
  subroutine read_tables(inp, viscous, nsc, wm, cp, h, mu, kk)
    character(len=*), intent(in) :: inp
    logical, intent(in) :: viscous
    integer, intent(out) :: nsc
    real(c_double), allocatable, intent(out) :: wm(:), cp(:, :), h(:, :), mu(:, :), kk(:, :)
    character(len=LL) :: line
    character(len=64) :: name
    real(c_double) :: w(16)
    integer :: u, ios, n
    open(newunit=u, file=inp // '/phase.txt', status='old', action='read', iostat=ios)
    if (ios /= 0) call case_stop('cannot open ' // inp // '/phase.txt')
    read(u, '(a)', iostat=ios) line
    allocate(wm(MOSE_MAX_SPECIES))
    nsc = 0
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      read(line, *, iostat=ios) name
      if (ios /= 0) cycle
      n = index(line, trim(name)) + len_trim(name)
      call line_numbers(line(n:), w, n)
      if (n < 1) cycle
      nsc = nsc + 1
      if (nsc > MOSE_MAX_SPECIES) call case_stop('phase.txt: more species than this library holds')
      wm(nsc) = w(1)
    end do
    close(u)
    if (nsc < 1) call case_stop('phase.txt: no species')
    wm = wm(1:nsc)
    allocate(cp(MOSE_NTAB, nsc), h(MOSE_NTAB, nsc))
    call read_point_zones(inp // '/thermo.dat', nsc, 2, 3, .true., cp, h)
    if (viscous) then
      allocate(mu(MOSE_NTAB, nsc), kk(MOSE_NTAB, nsc))
      call read_point_zones(inp // '/transport.dat', nsc, 2, 3, .false., mu, kk)
    end if
  end subroutine read_tables

  ! ---------------------------------------------------------------------------------------------- ic.tec
  integer function zone_int(header, key)
    character(len=*), intent(in) :: header, key
    character(len=len(header)) :: up
    integer :: p, q, ios
    up = header
    do p = 1, len(up)
      if (up(p:p) >= 'a' .and. up(p:p) <= 'z') up(p:p) = achar(iachar(up(p:p)) - 32)
    end do
    zone_int = 1
    p = 0
    do
      q = index(up(p + 1:), key // '=')
      if (q == 0) return
      p = p + q
      if (p == 1) exit
      if (up(p - 1:p - 1) == ' ' .or. up(p - 1:p - 1) == ',') exit
    end do
    q = scan(up(p + len(key) + 1:), ', ') + p + len(key) - 1
    if (q < p + len(key) + 1) q = len_trim(up)
    read(up(p + len(key) + 1:q), *, iostat=ios) zone_int
    if (ios /= 0) call case_stop('ic.tec: bad ' // key // '= in a ZONE header')
  end function zone_int

  !> VARLOCATION=([a-b]=NODAL,[c-d]=CELLCENTERED): true for the cell-centred variables; nodal where absent
  subroutine varlocation(header, nvar, cc)
    character(len=*), intent(in) :: header
    integer, intent(in) :: nvar
    logical, intent(out) :: cc(nvar)
    character(len=len(header)) :: up
    integer :: p, e, q, r, a, b, ios, dash
    logical :: cell
    up = header
    do p = 1, len(up)
      if (up(p:p) >= 'a' .and. up(p:p) <= 'z') up(p:p) = achar(iachar(up(p:p)) - 32)
    end do
    cc = .false.
    p = index(up, 'VARLOCATION')
    if (p == 0) return
    p = p + index(up(p:), '(')
    e = p + index(up(p:), ')') - 2
    q = p
    do
      r = index(up(q:e), '[')
      if (r == 0) exit
      r = q + r
      q = r + index(up(r:e), ']') - 1
      cell = index(up(q:min(e, q + 16)), 'CELL') > 0
      dash = index(up(r:q - 1), '-')
      if (dash > 0) then
        read(up(r:r + dash - 2), *, iostat=ios) a
        read(up(r + dash:q - 1), *, iostat=ios) b
      else
        read(up(r:q - 1), *, iostat=ios) a
        b = a
      end if
      if (ios /= 0 .or. a < 1 .or. b > nvar) call case_stop('ic.tec: bad VARLOCATION')
      cc(a:b) = cell
    end do
  end subroutine varlocation

  !> The zones of ic.tec (Tecplot BLOCK, one zone per block): dims(3, nb) = cells per block, nodes (3 x nodes of
  !> every block, node (i, j, k) fastest in i, then j, then k: the layout of mose_metrics_build) and the cell-centred
  !> variables fields(ncell_total, nvar_cc), block after block, i fastest.
  subroutine read_block_tec(path, nb, dims, nodes, fields, nvcc)
    character(len=*), intent(in) :: path
    integer, intent(out) :: nb, nvcc
    integer, allocatable, intent(out) :: dims(:, :)
    real(c_double), allocatable, intent(out) :: nodes(:), fields(:, :)
    character(len=LL) :: line, header
    real(c_double), allocatable :: x(:), tmp(:)
    real(c_double) :: row(4096)
    integer :: u, ios, nvar, I, J, K, need, got, n, v, o, nn, nc, zn, q, nodes0, cell0
    logical, allocatable :: cc(:)
    integer, allocatable :: dz(:, :)
    real(c_double), allocatable :: nd(:), fl(:, :)
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call case_stop('cannot open ' // trim(path))
    nvar = 0
    nb = 0
    allocate(dz(3, 64), nd(0), fl(0, 0))
    nodes0 = 0
    cell0 = 0
    nvcc = -1
    read(u, '(a)', iostat=ios) line
    do while (ios == 0)
      header = adjustl(line)
      if (lower(header(1:9)) == 'variables') then
        nvar = count_names(header(index(header, '=') + 1:))
        read(u, '(a)', iostat=ios) line
      else if (lower(header(1:4)) == 'zone') then
        if (nvar < 4) call case_stop('ic.tec: a ZONE before VARIABLES')
        do
          read(u, '(a)', iostat=ios) line
          if (ios /= 0) call case_stop('ic.tec: a ZONE without values')
          call line_numbers(line, row, n)
          if (n > 0) exit
          header = trim(header) // ' ' // adjustl(line)
        end do
        I = zone_int(header, 'I'); J = zone_int(header, 'J'); K = zone_int(header, 'K')
        if (allocated(cc)) deallocate(cc)
        allocate(cc(nvar))
        call varlocation(header, nvar, cc)
        if (any(cc(1:3))) call case_stop('ic.tec: x, y, z must be nodal')
        if (nvcc < 0) nvcc = count(cc)
        if (count(cc) /= nvcc) call case_stop('ic.tec: zones with different cell-centred variables')
        nn = I * J * K
        nc = (I - 1) * max(J - 1, 1) * max(K - 1, 1)
        need = 0
        do v = 1, nvar
          need = need + merge(nc, nn, cc(v))
        end do
        allocate(x(need))
        got = 0
        do
          q = min(n, need - got)
          x(got + 1:got + q) = row(1:q)
          got = got + q
          if (got >= need) exit
          read(u, '(a)', iostat=ios) line
          if (ios /= 0) call case_stop('ic.tec: a zone ends early')
          call line_numbers(line, row, n)
        end do
        nb = nb + 1
        if (nb > size(dz, 2)) call case_stop('ic.tec: more than 64 blocks')
        dz(:, nb) = [I - 1, max(J - 1, 1), max(K - 1, 1)]
        ! nodes: x y z interleaved per node
        allocate(tmp(size(nd) + 3 * nn))
        tmp(1:size(nd)) = nd
        o = 0
        zn = 0
        do v = 1, nvar
          if (cc(v)) then
            o = o + nc
          else
            zn = zn + 1
            if (zn <= 3) tmp(size(nd) + zn:size(nd) + 3 * nn:3) = x(o + 1:o + nn)
            o = o + nn
          end if
        end do
        call move_alloc(tmp, nd)
        ! cell-centred variables, appended block after block
        block
          real(c_double), allocatable :: f2(:, :)
          allocate(f2(cell0 + nc, nvcc))
          if (cell0 > 0) f2(1:cell0, :) = fl(1:cell0, :)
          o = 0
          zn = 0
          do v = 1, nvar
            if (cc(v)) then
              zn = zn + 1
              f2(cell0 + 1:cell0 + nc, zn) = x(o + 1:o + nc)
              o = o + nc
            else
              o = o + nn
            end if
          end do
          call move_alloc(f2, fl)
        end block
        cell0 = cell0 + nc
        nodes0 = nodes0 + nn
        deallocate(x)
        read(u, '(a)', iostat=ios) line
      else
        read(u, '(a)', iostat=ios) line
      end if
    end do
    close(u)
    if (nb == 0) call case_stop('ic.tec: no zones')
    allocate(dims(3, nb))
    dims = dz(:, 1:nb)
    call move_alloc(nd, nodes)
    call move_alloc(fl, fields)
  contains
    integer function count_names(s)
      character(len=*), intent(in) :: s
      integer :: p
      logical :: inq, intok
      count_names = 0
      inq = .false.
      intok = .false.
      do p = 1, len_trim(s)
        if (s(p:p) == '"') then
          if (.not. inq) count_names = count_names + 1
          inq = .not. inq
          intok = .false.
        else if (.not. inq) then
          if (s(p:p) == ' ' .or. s(p:p) == ',') then
            intok = .false.
          else if (.not. intok) then
            count_names = count_names + 1
            intok = .true.
          end if
        end if
      end do
    end function count_names
  end subroutine read_block_tec

  !> Check_Mesh_Type (Lib_Metrics.f90:139-196) on block 1: 3 if n3 > 1, 1 if n2 = n3 = 1, else 2; a mesh that the
  !> reference takes as axisymmetric is refused.  nodes: block 1's nodes, (3, n1+1, n2+1, n3+1).
  integer function mesh_ndir(n1, n2, n3, nodes)
    integer, intent(in) :: n1, n2, n3
    real(c_double), intent(in) :: nodes(3, n1 + 1, n2 + 1, n3 + 1)
    real(c_double) :: t1, t2, r0, r1
    if (n3 > 1) then
      t1 = ang(nodes(:, 1, 2, n3 + 1)) - ang(nodes(:, 1, 2, 1))
      t2 = ang(nodes(:, 1, n2 + 1, n3 + 1)) - ang(nodes(:, 1, n2 + 1, 1))
      if ((t1 - t2) < 1e-5_c_double .and. t1 /= 0.0_c_double) call case_stop('axisymmetric 3D mesh: not on this path')
      mesh_ndir = 3
      return
    end if
    if (n2 == 1) then
      mesh_ndir = 1
      return
    end if
    t1 = ang(nodes(:, 1, 2, 2)) - ang(nodes(:, 1, 2, 1))
    t2 = ang(nodes(:, 1, n2 + 1, 2)) - ang(nodes(:, 1, n2 + 1, 1))
    r0 = hypot(nodes(2, 1, 2, 1), nodes(3, 1, 2, 1))
    r1 = hypot(nodes(2, 1, 2, 2), nodes(3, 1, 2, 2))
    if ((t1 - t2) < 1e-5_c_double .and. r0 > 1e-10_c_double .and. abs(r0 - r1) < 1e-5_c_double * r0 .and. &
        t1 /= 0.0_c_double) call case_stop('axisymmetric 2D mesh: not on this path')
    mesh_ndir = 2
  contains
    real(c_double) function ang(p)
      real(c_double), intent(in) :: p(3)
      ang = atan2(p(3), p(2))
    end function ang
  end function mesh_ndir

  ! ---------------------------------------------------------------------------------------------- bc.txt
  !> Records of a MOSE_open bc.txt in file order (IO_BC_Open.f90:132-399), 0-based block ids.  A line
  !> "b i j k f type" is followed by the type's property line: 101/201 "bs is js ks fs d11 d12 d21 d22",
  !> 301 "qw ks eps", 302 "Tw ks eps", 405 "mach T p alpha beta rel_fac ci(species) ci(rans)"; 0, 100, 300 and 400
  !> have none.  A 405 record's ff holds, until inlet_rows, the 1-based index of its parsed line.
  subroutine read_bc_open(path, rec, nrec)
    character(len=*), intent(in) :: path
    type(mose_bc_record), allocatable, intent(out) :: rec(:)
    integer, intent(out) :: nrec
    character(len=LL) :: line
    real(c_double) :: x(16)
    integer :: u, ios, n, typ, cap
    type(mose_bc_record), allocatable :: r2(:)
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call case_stop('cannot open ' // trim(path))
    cap = 256
    allocate(rec(cap))
    nrec = 0
    do
      call next_line(n)
      if (n < 0) exit
      if (n < 6) call case_stop('bc.txt: a record line with fewer than 6 numbers')
      if (nrec == cap) then
        allocate(r2(2 * cap))
        r2(1:cap) = rec
        call move_alloc(r2, rec)
        cap = 2 * cap
      end if
      nrec = nrec + 1
      associate (r => rec(nrec))
        r%b = nint(x(1)) - 1; r%i = nint(x(2)); r%j = nint(x(3)); r%k = nint(x(4)); r%f = nint(x(5))
        r%bs = 0; r%is = 0; r%js = 0; r%ks = 0; r%fs = 0
        r%d11 = 1; r%d12 = 0; r%d21 = 0; r%d22 = 1
        r%periodic = 0; r%ff = 0; r%pad = 0; r%tw = 0.0_c_double; r%tvec = 0.0_c_double
        typ = nint(x(6))
        select case (typ)
        case (0, 100); r%typ = MOSE_BC_NONE
        case (101, 201); r%typ = MOSE_BC_CONNECTION
        case (300); r%typ = MOSE_BC_SYMMETRY
        case (301); r%typ = MOSE_BC_WALL_ADIABATIC
        case (302); r%typ = MOSE_BC_WALL_ISOTHERMAL
        case (400); r%typ = MOSE_BC_EXTRAPOLATION
        case (405); r%typ = MOSE_BC_FARFIELD
        case default
          write(line, '(a,i0,a,i0,a,i0,a)') 'bc.txt: boundary type ', typ, ' (block ', r%b + 1, ', face ', r%f, &
                                            ') is not on this path'
          call case_stop(trim(line))
        end select
        if (typ == 101 .or. typ == 201) then
          call next_line(n)
          if (n < 9) call case_stop('bc.txt: a connection record without its donor line')
          r%bs = nint(x(1)) - 1; r%is = nint(x(2)); r%js = nint(x(3)); r%ks = nint(x(4)); r%fs = nint(x(5))
          r%d11 = nint(x(6)); r%d12 = nint(x(7)); r%d21 = nint(x(8)); r%d22 = nint(x(9))
          if (typ == 201) r%periodic = 1      ! ghost nodes translated by one period (metrics)
        else if (typ == 301 .or. typ == 302) then
          call next_line(n)
          if (n < 3) call case_stop('bc.txt: a wall record without its property line')
          if (x(2) /= 0.0_c_double .or. x(3) /= 0.0_c_double) &
            call case_stop('bc.txt: wall roughness / emissivity extras are not on this path')
          if (typ == 301 .and. x(1) /= 0.0_c_double) &
            call case_stop('bc.txt: a 301 wall with a nonzero heat flux is not on this path')
          if (typ == 302) r%tw = x(1)
        else if (typ == 405) then
          call next_line(n)
          if (n < 0) call case_stop('bc.txt: a 405 record without its property line')
          call parse_inlet(line)
          r%ff = n_inlets
        end if
      end associate
    end do
    close(u)
  contains
    subroutine next_line(n)
      integer, intent(out) :: n
      do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) then
          n = -1
          return
        end if
        if (len_trim(line) == 0) cycle
        call line_numbers(line, x, n)
        return
      end do
    end subroutine next_line
  end subroutine read_bc_open

  subroutine parse_inlet(line)
    character(len=*), intent(in) :: line
    character(len=64) :: w(MOSE_MAX_SPECIES + 8)
    character(len=len(line)) :: s
    type(inlet_t), allocatable :: g(:)
    integer :: nw, p, q, ios, e
    s = line
    do p = 1, len_trim(s)
      if (s(p:p) == ',' .or. s(p:p) == achar(9)) s(p:p) = ' '
    end do
    nw = 0
    p = 1
    do
      do while (p <= len_trim(s))
        if (s(p:p) /= ' ') exit
        p = p + 1
      end do
      if (p > len_trim(s) .or. nw == size(w)) exit
      q = index(s(p:), ' ') + p - 2
      if (q < p) q = len_trim(s)
      nw = nw + 1
      w(nw) = s(p:q)
      p = q + 1
    end do
    if (nw < 6) call case_stop('bc.txt: a 405 property line with fewer than 6 values')
    if (.not. allocated(inlets)) allocate(inlets(64))
    if (n_inlets == size(inlets)) then
      allocate(g(2 * size(inlets)))
      g(1:n_inlets) = inlets(1:n_inlets)
      call move_alloc(g, inlets)
    end if
    n_inlets = n_inlets + 1
    associate (t => inlets(n_inlets))
      t%mach = value(w(1)); t%t = value(w(2)); t%p = value(w(3))
      t%an = index(w(4), 'normal') > 0
      t%bn = index(w(5), 'normal') > 0
      t%alpha = 0.0_c_double; t%beta = 0.0_c_double
      if (.not. t%an) t%alpha = value(w(4))
      if (.not. t%bn) t%beta = value(w(5))
      t%rel = value(w(6))
      t%nv = nw - 6
      do e = 1, t%nv
        t%v(e) = value(w(6 + e))
      end do
    end associate
  contains
    real(c_double) function value(tok)
      character(len=*), intent(in) :: tok
      read(tok, *, iostat=ios) value
      if (ios /= 0) call case_stop('bc.txt: unreadable 405 value "' // trim(tok) // '"')
    end function value
  end subroutine parse_inlet

  !> The far-field table of the 405 records, one row per distinct set of values (bit patterns), and each record's
  !> 0-based row in ff.  A row: the Mach number, the static T and p (BC_Supersonic_Inflow takes p0 and T0 as static
  !> values), pamb = PAMB_FORCE_SUPERSONIC, rel_fac, the mass fractions and the RANS values.  An angle "normal" is
  !> the face normal's, alpha = atan(n2 / (n1 + 1e-20)), beta = atan(n3 / (n1 + 1e-20)) (Resolve_Open_BC_Direction,
  !> IO_BC_Open.f90:411-451, at the face of Face_Index, Lib_BC.f90:46-74).
  subroutine inlet_rows(nb, dims, NF, rec, nrec, fm1, fm2, fm3, nsc, nrans, ff, nff)
    integer, intent(in) :: nb, dims(3, nb), nrec, nsc, nrans
    integer(c_int64_t), intent(in) :: NF(3)
    type(mose_bc_record), intent(inout) :: rec(:)
    real(c_double), intent(in) :: fm1(:), fm2(:), fm3(:)
    type(mose_farfield), allocatable, intent(out) :: ff(:)
    integer, intent(out) :: nff
    integer(c_int64_t) :: off(3, nb), f
    integer :: b, e, q, dr, md, n1, n2, n3
    real(c_double) :: nx, ny, nz
    type(mose_farfield) :: row
    allocate(ff(max(n_inlets, 1)))
    nff = 0
    off(:, 1) = 0
    do b = 2, nb
      off(1, b) = off(1, b - 1) + int(dims(1, b - 1) + 1, c_int64_t) * dims(2, b - 1) * dims(3, b - 1)
      off(2, b) = off(2, b - 1) + int(dims(1, b - 1), c_int64_t) * (dims(2, b - 1) + 1) * dims(3, b - 1)
      off(3, b) = off(3, b - 1) + int(dims(1, b - 1), c_int64_t) * dims(2, b - 1) * (dims(3, b - 1) + 1)
    end do
    do e = 1, nrec
      if (rec(e)%typ /= MOSE_BC_FARFIELD) cycle
      associate (r => rec(e), t => inlets(rec(e)%ff))
        if (t%nv < nsc + nrans) call case_stop('bc.txt: a 405 record without its mass fractions')
        b = r%b + 1
        n1 = dims(1, b); n2 = dims(2, b); n3 = dims(3, b)
        dr = (r%f + 1) / 2
        md = mod(r%f, 2)
        if (dr == 1) then
          f = off(1, b) + (r%i - md) + int(n1 + 1, c_int64_t) * ((r%j - 1) + int(n2, c_int64_t) * (r%k - 1))
          nx = fm1(f + 1); ny = fm1(NF(1) + f + 1); nz = fm1(2 * NF(1) + f + 1)
        else if (dr == 2) then
          f = off(2, b) + (r%i - 1) + int(n1, c_int64_t) * ((r%j - md) + int(n2 + 1, c_int64_t) * (r%k - 1))
          nx = fm2(f + 1); ny = fm2(NF(2) + f + 1); nz = fm2(2 * NF(2) + f + 1)
        else
          f = off(3, b) + (r%i - 1) + int(n1, c_int64_t) * ((r%j - 1) + int(n2, c_int64_t) * (r%k - md))
          nx = fm3(f + 1); ny = fm3(NF(3) + f + 1); nz = fm3(2 * NF(3) + f + 1)
        end if
        row%mach = t%mach; row%t0 = t%t; row%p0 = t%p; row%pamb = PAMB_FORCE_SUPERSONIC
        row%alpha = t%alpha; row%beta = t%beta
        if (t%an) row%alpha = atan(ny / (nx + 1e-20_c_double))
        if (t%bn) row%beta = atan(nz / (nx + 1e-20_c_double))
        row%rel_fac = t%rel; row%reserved = 0.0_c_double
        row%ci = 0.0_c_double; row%rans = 0.0_c_double
        row%ci(1:nsc) = t%v(1:nsc)
        if (nrans > 0) row%rans(1:nrans) = t%v(nsc + 1:nsc + nrans)
        q = 0
        do b = 1, nff
          if (all(transfer(ff(b), [0_c_int64_t]) == transfer(row, [0_c_int64_t]))) then
            q = b
            exit
          end if
        end do
        if (q == 0) then
          nff = nff + 1
          ff(nff) = row
          q = nff
        end if
        r%ff = q - 1
      end associate
    end do
  end subroutine inlet_rows

  !> grouped by block, then by face, file order inside a group (a stable sort)
  subroutine sort_records(rec, nrec)
    type(mose_bc_record), intent(inout) :: rec(:)
    integer, intent(in) :: nrec
    type(mose_bc_record) :: t
    integer :: p, q
    do p = 2, nrec
      t = rec(p)
      q = p - 1
      do while (q >= 1)
        if (rec(q)%b < t%b .or. (rec(q)%b == t%b .and. rec(q)%f <= t%f)) exit
        rec(q + 1) = rec(q)
        q = q - 1
      end do
      rec(q + 1) = t
    end do
  end subroutine sort_records

  ! ---------------------------------------------------------------------------------------------- the arena
  !> the arena sizes of the blocks (rows of the padded cells, face counts) and each block's first row
  subroutine arena_sizes(nb, dims, NC, NF, boff)
    integer, intent(in) :: nb, dims(3, nb)
    integer(c_int64_t), intent(out) :: NC, NF(3)
    integer(c_int64_t), intent(out) :: boff(nb)
    integer :: b
    NC = 0; NF = 0
    do b = 1, nb
      boff(b) = NC
      NC = NC + int(dims(1, b) + 2 * MOSE_GC, c_int64_t) * (dims(2, b) + 2 * MOSE_GC) * (dims(3, b) + 2 * MOSE_GC)
      NF(1) = NF(1) + int(dims(1, b) + 1, c_int64_t) * dims(2, b) * dims(3, b)
      NF(2) = NF(2) + int(dims(1, b), c_int64_t) * (dims(2, b) + 1) * dims(3, b)
      NF(3) = NF(3) + int(dims(1, b), c_int64_t) * dims(2, b) * (dims(3, b) + 1)
    end do
  end subroutine arena_sizes

  !> ghost rows (never read before the first ghost fill) set to the nearest interior cell: the k ghosts first, then
  !> j, then i, each over the whole padded plane.  P(NC, nprim), block b at rows boff+1 .. of (pi, pj, pk).
  subroutine fill_ghosts_nearest(nb, dims, boff, NC, nprim, P)
    integer, intent(in) :: nb, dims(3, nb), nprim
    integer(c_int64_t), intent(in) :: boff(nb), NC
    real(c_double), intent(inout) :: P(NC, nprim)
    integer :: b, s, g, pi, pj, pk, n1, n2, n3
    do b = 1, nb
      n1 = dims(1, b); n2 = dims(2, b); n3 = dims(3, b)
      pi = n1 + 2 * MOSE_GC; pj = n2 + 2 * MOSE_GC; pk = n3 + 2 * MOSE_GC
      do s = 1, nprim
        call fill(P(boff(b) + 1:boff(b) + int(pi, c_int64_t) * pj * pk, s))
      end do
    end do
  contains
    subroutine fill(v)
      real(c_double), intent(inout) :: v(pi, pj, pk)
      do g = 1, MOSE_GC
        v(:, :, g) = v(:, :, MOSE_GC + 1)
        v(:, :, MOSE_GC + n3 + g) = v(:, :, MOSE_GC + n3)
      end do
      do g = 1, MOSE_GC
        v(:, g, :) = v(:, MOSE_GC + 1, :)
        v(:, MOSE_GC + n2 + g, :) = v(:, MOSE_GC + n2, :)
      end do
      do g = 1, MOSE_GC
        v(g, :, :) = v(MOSE_GC + 1, :, :)
        v(MOSE_GC + n1 + g, :, :) = v(MOSE_GC + n1, :, :)
      end do
    end subroutine fill
  end subroutine fill_ghosts_nearest

end module mose_case_files
