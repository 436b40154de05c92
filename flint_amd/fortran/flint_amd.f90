! Fortran drop-in for FLINT's per-cell thermochemistry calls, over the C ABI of libflint_amd
! (include/flint_amd.h) through iso_c_binding.
!
! Module flint_amd_c declares the C entry points.  Module flint_amd keeps one mechanism in module
! state, as FLINT does, and offers the FLINT procedure names with FLINT's argument lists:
!
!   read_idealgas_thermo(folder)            Load_ThermoTransport.f90:7-89
!   read_chemistry(folder, mech_name)       Load_Chemistry.f90:7-164
!   Assign_Mechanism(mad_world)             Lib_Chemistry_wdot.f90:24-92
!   chemistry_source(roi, temp, omegadot)   Lib_Chemistry_wdot.f90:6-20 (the procedure pointer)
!   rhs_native(nz, time, Z, F)              Lib_Chemistry_rhs.f90:15-52
!   CEA_initialize_global()                 Lib_CEA_setup.f90:6-84
!   CEA_solve(p_in, T_in, rhoi_in, T_out, rhoi_out)   Lib_CEA_solver.f90:6-17
!   free_chemistry_data()
!
! plus batched forms that hand a whole block of cells to the GPU in one call (the per-cell forms
! pay one host-device round trip per call):
!
!   chemistry_source_batch(rho, T, wdot)          rho(n, ns), T(n), wdot(n, ns)
!   rhs_native_batch(Z, F [, J])                  Z(n, ns+1), F(n, ns+1), J(n, ns+1, ns+1)
!   reactor_advance_batch(X, t_end, rtol, atol, constant_pressure, p, max_steps, status)
!   CEA_solve_batch(p, T, rho, T_out, rho_out, status)
!
! Batched arrays are cell-fastest, which is the C ABI's structure-of-arrays layout with ld = n.
! J(i, c, r) = dF_r/dZ_c of cell i.  Errors: the loaders return a nonzero ios (and print the
! library's message); the other procedures stop with the message, as FLINT's loaders do.
module flint_amd_c
  use iso_c_binding
  implicit none
  interface
    function flint_last_error() bind(C, name="flint_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function
    function flint_mech_load(folder, prefix, out) bind(C, name="flint_mech_load") result(rc)
      import :: c_char, c_ptr, c_int
      character(kind=c_char), intent(in) :: folder(*), prefix(*)
      type(c_ptr), intent(out) :: out
      integer(c_int) :: rc
    end function
    subroutine flint_mech_destroy(m) bind(C, name="flint_mech_destroy")
      import :: c_ptr
      type(c_ptr), value :: m
    end subroutine
    function flint_mech_query(m, ns, Tk_min, Tk_max, Th_min, Th_max) bind(C, name="flint_mech_query") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: m
      integer(c_int), intent(out) :: ns, Tk_min, Tk_max, Th_min, Th_max
      integer(c_int) :: rc
    end function
    function flint_source_host(m, n, rho, T, wdot) bind(C, name="flint_source_host") result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      real(c_double), intent(in) :: rho(*), T(*)
      real(c_double), intent(out) :: wdot(*)
      integer(c_int) :: rc
    end function
    function flint_rhs_host(m, n, Z, F, J) bind(C, name="flint_rhs_host") result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      real(c_double), intent(in) :: Z(*)
      real(c_double), intent(out) :: F(*)
      type(c_ptr), value :: J
      integer(c_int) :: rc
    end function
    function flint_reactor_advance_host(m, n, mode, X, p, t_end, rtol, atol, max_steps, status) &
        bind(C, name="flint_reactor_advance_host") result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      integer(c_int), value :: mode, max_steps
      real(c_double), intent(inout) :: X(*)
      type(c_ptr), value :: p, status
      real(c_double), value :: t_end, rtol, atol
      integer(c_int) :: rc
    end function
    function flint_cea_solve_host(m, n, hp, p, T, rho, T_out, Y_out, status) &
        bind(C, name="flint_cea_solve_host") result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      integer(c_int), value :: hp
      type(c_ptr), value :: p, status
      real(c_double), intent(in) :: T(*), rho(*)
      real(c_double), intent(out) :: T_out(*), Y_out(*)
      integer(c_int) :: rc
    end function
    function flint_cons2prim_host(m, n, nd, cons, T, prim, T_out, status) bind(C, name="flint_cons2prim_host") &
        result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      integer(c_int), value :: nd
      real(c_double), intent(in) :: cons(*), T(*)
      real(c_double), intent(out) :: prim(*)
      type(c_ptr), value :: T_out, status
      integer(c_int) :: rc
    end function
    function flint_prim2cons_host(m, n, nd, prim, cons) bind(C, name="flint_prim2cons_host") result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      integer(c_int), value :: nd
      real(c_double), intent(in) :: prim(*)
      real(c_double), intent(out) :: cons(*)
      integer(c_int) :: rc
    end function
    function flint_transport_host(m, n, rhoi, rho, T, mu, lam, status) bind(C, name="flint_transport_host") &
        result(rc)
      import :: c_ptr, c_long, c_double, c_int
      type(c_ptr), value :: m
      integer(c_long), value :: n
      real(c_double), intent(in) :: rhoi(*), rho(*), T(*)
      real(c_double), intent(out) :: mu(*), lam(*)
      type(c_ptr), value :: status
      integer(c_int) :: rc
    end function
  end interface
end module flint_amd_c

module flint_amd
  use iso_c_binding
  use flint_amd_c
  implicit none
  private

  !> number of species of the loaded mechanism (FLINT_Lib_Thermodynamic's ns)
  integer, public :: ns = 0
  !> equilibrium at fixed internal energy and volume (.true., CEA_initialize_global's default,
  !> Lib_CEA_setup.f90:25) or at fixed enthalpy and pressure (.false.)
  logical, public :: Vol = .true.
  !> mechanism name from chemistry-info.txt
  character(len=64), public :: loaded_mech_name = ''
  integer, public :: Tmin = 0, Tmax = 0

  type(c_ptr), save :: mech = c_null_ptr
  character(len=:), allocatable, save :: mech_folder

  public :: read_idealgas_thermo, read_chemistry, Assign_Mechanism, chemistry_source, rhs_native, &
            CEA_initialize_global, CEA_solve, free_chemistry_data, chemistry_source_batch, &
            rhs_native_batch, reactor_advance_batch, CEA_solve_batch, flint_error_message, cons2prim, &
            prim2cons, cons2prim_batch, prim2cons_batch, co_k_mi_lam_Wilke, f_laminarViscosity, &
            transport_batch

contains

  !> text of the library's last error
  function flint_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: k
    p = flint_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, s, [4096])
    k = 0
    do while (k < 4096)
      if (s(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: msg)
    msg = transfer(s(1:k), msg)
  end function

  subroutine fail(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(A,A,I0,A,A)') trim(where), ': flint error ', rc, ': ', flint_error_message()
      stop 1
    end if
  end subroutine

  subroutine need_mech(where)
    character(len=*), intent(in) :: where
    if (.not. c_associated(mech)) then
      write(*, '(A,A)') trim(where), ': no mechanism loaded (call read_idealgas_thermo first)'
      stop 1
    end if
  end subroutine

  !> Loads the native input folder (default INPUT, as FLINT).  The compiled kernels need the
  !> chemistry files of the same folder, so this reads the whole mechanism; read_chemistry on the
  !> same folder then only reports its name.
  function read_idealgas_thermo(folder) result(ios)
    character(len=*), intent(in), optional :: folder
    integer :: ios
    character(len=:), allocatable :: f
    integer(c_int) :: rc, n, a, b, c, d
    if (present(folder)) then
      f = trim(folder)
    else
      f = 'INPUT'
    end if
    call free_chemistry_data()
    rc = flint_mech_load(f // c_null_char, c_null_char, mech)
    if (rc /= 0) then
      write(*, '(A,A)') '[ERROR] ', flint_error_message()
      mech = c_null_ptr
      ios = int(rc)
      return
    end if
    rc = flint_mech_query(mech, n, a, b, c, d)
    ns = n
    Tmin = c
    Tmax = d
    mech_folder = f
    ios = read_name(f, loaded_mech_name)
  end function

  !> FLINT's read_chemistry: loads the folder if read_idealgas_thermo has not, and returns the
  !> mechanism name written in chemistry-info.txt.
  function read_chemistry(folder, mech_name) result(ios)
    character(len=*), intent(in), optional :: folder
    character(len=*), intent(out), optional :: mech_name
    integer :: ios
    character(len=:), allocatable :: f
    if (present(folder)) then
      f = trim(folder)
    else
      f = 'INPUT'
    end if
    ios = 0
    if (.not. c_associated(mech)) then
      ios = read_idealgas_thermo(f)
    else if (mech_folder /= f) then
      ios = read_idealgas_thermo(f)
    end if
    if (present(mech_name)) mech_name = loaded_mech_name
  end function

  function read_name(folder, name) result(ios)
    character(len=*), intent(in) :: folder
    character(len=*), intent(out) :: name
    integer :: ios, u
    name = ''
    open(newunit=u, file=folder // '/chemistry-info.txt', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(u, *, iostat=ios) name
    close(u)
  end function

  !> The loaded folder already selected its compiled kernels; this checks the name.
  subroutine Assign_Mechanism(mad_world)
    character(len=*), intent(in) :: mad_world
    call need_mech('Assign_Mechanism')
    if (trim(mad_world) /= trim(loaded_mech_name)) then
      write(*, '(A,A,A,A)') 'Assign_Mechanism: ', trim(mad_world), ' is not the loaded mechanism ', &
                            trim(loaded_mech_name)
      stop 1
    end if
  end subroutine

  !> Production rates [kg/m^3/s] of one cell.  Like FLINT's WD routine (WD.f90:18-21), negative
  !> densities are set to zero in roi for WD; the generated mechanisms leave roi as it is.
  subroutine chemistry_source(roi, temp, omegadot)
    real(8), intent(inout) :: roi(ns)
    real(8), intent(in) :: temp
    real(8), intent(out) :: omegadot(ns)
    real(8) :: t(1)
    call need_mech('chemistry_source')
    if (trim(loaded_mech_name) == 'WD') roi = max(roi, 0d0)
    t(1) = temp
    call fail(flint_source_host(mech, 1_c_long, roi, t, omegadot), 'chemistry_source')
  end subroutine

  !> Constant-volume reactor right-hand side of one cell, Z = (rho_1..rho_ns, T); F = -1 outside
  !> the thermo table.
  subroutine rhs_native(nz, time, Z, F)
    integer, intent(in) :: nz
    real(8), intent(in) :: time
    real(8), intent(in) :: Z(nz)
    real(8), intent(out) :: F(nz)
    call need_mech('rhs_native')
    call fail(flint_rhs_host(mech, 1_c_long, Z, F, c_null_ptr), 'rhs_native')
  end subroutine

  !> The equilibrium data were set up when the folder was loaded; this resets Vol to FLINT's
  !> default.
  subroutine CEA_initialize_global()
    call need_mech('CEA_initialize_global')
    Vol = .true.
  end subroutine

  !> Equilibrium of one state: with Vol (default) at the internal energy of rhoi_in at T_in and the
  !> volume 1e5/sum(rhoi_in); otherwise at p_in [bar] and the enthalpy of rhoi_in at T_in.
  !> rhoi_out = n_k W_k; T_out and rhoi_out are NaN when the solver fails.
  subroutine CEA_solve(p_in, T_in, rhoi_in, T_out, rhoi_out)
    real(8), dimension(:), intent(in) :: rhoi_in
    real(8), intent(in) :: p_in, T_in
    real(8), intent(out) :: T_out
    real(8), dimension(:), intent(out) :: rhoi_out
    real(c_double), target :: p(1)
    real(c_double) :: t(1), to(1), r(ns), y(ns)
    integer(c_int), target :: st(1)
    integer(c_int) :: hp
    call need_mech('CEA_solve')
    p(1) = p_in
    t(1) = T_in
    r = rhoi_in(1:ns)
    hp = merge(0, 1, Vol)
    call fail(flint_cea_solve_host(mech, 1_c_long, hp, c_loc(p), t, r, to, y, c_loc(st)), 'CEA_solve')
    T_out = to(1)
    rhoi_out(1:ns) = y
  end subroutine

  subroutine free_chemistry_data()
    if (c_associated(mech)) call flint_mech_destroy(mech)
    mech = c_null_ptr
    ns = 0
    loaded_mech_name = ''
  end subroutine

  !> Primitive variables (rho_1..rho_ns, u_1..u_nd, p) of one cell from its conservative variables
  !> (rho_1..rho_ns, rho u_1..rho u_nd, rho E), Newton on T from T (Lib_ThermoTransport.f90:130);
  !> 0 <= nd <= 3.  p is NaN when T leaves the thermo table.  Not pure, unlike the reference's.
  function cons2prim(cons, T) result(prim)
    real(8), intent(in) :: cons(:)
    real(8), intent(in) :: T
    real(8) :: prim(size(cons))
    real(c_double) :: c(size(cons)), q(size(cons)), tt(1)
    call need_mech('cons2prim')
    c = cons
    tt(1) = T
    call fail(flint_cons2prim_host(mech, 1_c_long, int(size(cons) - ns - 1, c_int), c, tt, q, c_null_ptr, &
                                   c_null_ptr), 'cons2prim')
    prim = q
  end function

  !> Mixture viscosity milam [Pa s] and conductivity klam [W/(m K)] of one cell by Wilke's rule
  !> (co_k_mi_lam_Wilke, Lib_ThermoTransport.f90:407); the tables come from transport.dat in the
  !> folder read_chemistry loaded.  NaN when T leaves the transport table.
  subroutine co_k_mi_lam_Wilke(rhoi, rho, T, milam, klam)
    real(8), intent(in) :: rhoi(:), rho, T
    real(8), intent(out) :: milam, klam
    real(c_double) :: r(size(rhoi)), a(1), b(1), m(1), k(1)
    call need_mech('co_k_mi_lam_Wilke')
    r = rhoi
    a(1) = rho
    b(1) = T
    call fail(flint_transport_host(mech, 1_c_long, r, a, b, m, k, c_null_ptr), 'co_k_mi_lam_Wilke')
    milam = m(1)
    klam = k(1)
  end subroutine

  !> Mixture viscosity at T = p / (Rtot rho) (f_laminarViscosity, Lib_ThermoTransport.f90:365).
  function f_laminarViscosity(rhoi, p, rho, Rtot) result(mil)
    real(8), intent(in) :: rhoi(:), p, rho, Rtot
    real(8) :: mil, klam
    call co_k_mi_lam_Wilke(rhoi, rho, p / (Rtot * rho), mil, klam)
  end function

  !> co_k_mi_lam_Wilke over n cells in one call: rhoi(n, ns), rho(n), T(n) -> milam(n), klam(n).
  subroutine transport_batch(rhoi, rho, T, milam, klam, status)
    real(8), intent(in) :: rhoi(:, :), rho(:), T(:)
    real(8), intent(out) :: milam(:), klam(:)
    integer, intent(out), optional, target :: status(:)
    integer(c_int), allocatable, target :: st(:)
    call need_mech('transport_batch')
    allocate(st(size(T)))
    call fail(flint_transport_host(mech, int(size(T), c_long), rhoi, rho, T, milam, klam, c_loc(st)), &
              'transport_batch')
    if (present(status)) status = st
  end subroutine

  !> Conservative variables of one cell from its primitive ones (Lib_ThermoTransport.f90:168).
  function prim2cons(prim) result(cons)
    real(8), intent(in) :: prim(:)
    real(8) :: cons(size(prim))
    real(c_double) :: q(size(prim)), c(size(prim))
    call need_mech('prim2cons')
    q = prim
    call fail(flint_prim2cons_host(mech, 1_c_long, int(size(prim) - ns - 1, c_int), q, c), 'prim2cons')
    cons = c
  end function

  !> cons2prim over a block of cells: cons(n, np), T(n) -> prim(n, np), with the converged T in
  !> T_out(n) and status(n) (0 ok, 2 T outside the thermo table) if present.
  subroutine cons2prim_batch(cons, T, prim, T_out, status)
    real(8), intent(in) :: cons(:, :), T(:)
    real(8), intent(out) :: prim(:, :)
    real(8), intent(out), optional :: T_out(:)
    integer, intent(out), optional :: status(:)
    real(c_double), allocatable :: c(:, :), q(:, :), tt(:)
    real(c_double), allocatable, target :: to(:)
    integer(c_int), allocatable, target :: st(:)
    integer(c_long) :: n
    integer :: npr
    call need_mech('cons2prim_batch')
    n = size(T, kind=c_long)
    npr = size(cons, 2)
    c = cons(1:n, 1:npr)
    tt = T
    allocate(q(n, npr), to(n), st(n))
    call fail(flint_cons2prim_host(mech, n, int(npr - ns - 1, c_int), c, tt, q, c_loc(to), c_loc(st)), &
              'cons2prim_batch')
    prim(1:n, 1:npr) = q
    if (present(T_out)) T_out(1:n) = to
    if (present(status)) status(1:n) = st
  end subroutine

  !> prim2cons over a block of cells: prim(n, np) -> cons(n, np).
  subroutine prim2cons_batch(prim, cons)
    real(8), intent(in) :: prim(:, :)
    real(8), intent(out) :: cons(:, :)
    real(c_double), allocatable :: q(:, :), c(:, :)
    integer(c_long) :: n
    integer :: npr
    call need_mech('prim2cons_batch')
    n = size(prim, 1, kind=c_long)
    npr = size(prim, 2)
    q = prim
    allocate(c(n, npr))
    call fail(flint_prim2cons_host(mech, n, int(npr - ns - 1, c_int), q, c), 'prim2cons_batch')
    cons(1:n, 1:npr) = c
  end subroutine

  !> chemistry_source over a block of cells: rho(n, ns), T(n) -> wdot(n, ns).
  subroutine chemistry_source_batch(rho, T, wdot)
    real(8), intent(in) :: rho(:, :), T(:)
    real(8), intent(out) :: wdot(:, :)
    real(c_double), allocatable :: r(:, :), tt(:), w(:, :)
    integer(c_long) :: n
    call need_mech('chemistry_source_batch')
    n = size(T, kind=c_long)
    r = rho(1:n, 1:ns)
    tt = T
    allocate(w(n, ns))
    call fail(flint_source_host(mech, n, r, tt, w), 'chemistry_source_batch')
    wdot(1:n, 1:ns) = w
  end subroutine

  !> rhs_native over a block of cells: Z(n, ns+1) -> F(n, ns+1) and, if present, the analytic
  !> Jacobian J(n, ns+1, ns+1) with J(i, c, r) = dF_r/dZ_c.
  subroutine rhs_native_batch(Z, F, J)
    real(8), intent(in) :: Z(:, :)
    real(8), intent(out) :: F(:, :)
    real(8), intent(out), optional, target :: J(:, :, :)
    real(c_double), allocatable :: zz(:, :), ff(:, :)
    real(c_double), allocatable, target :: jj(:, :, :)
    integer(c_long) :: n
    integer :: nz
    call need_mech('rhs_native_batch')
    n = size(Z, 1, kind=c_long)
    nz = ns + 1
    zz = Z(:, 1:nz)
    allocate(ff(n, nz))
    if (present(J)) then
      allocate(jj(n, nz, nz))
      call fail(flint_rhs_host(mech, n, zz, ff, c_loc(jj)), 'rhs_native_batch')
      J(1:n, 1:nz, 1:nz) = jj
    else
      call fail(flint_rhs_host(mech, n, zz, ff, c_null_ptr), 'rhs_native_batch')
    end if
    F(1:n, 1:nz) = ff
  end subroutine

  !> Advances a block of 0D reactors from 0 to t_end in place with the batched BDF integrator:
  !> X(n, ns+1) = (rho_k, T) at constant volume, or (Y_k, T) at constant pressure p(n) [Pa].
  !> status(n) (optional): 0 ok, 1 step underflow, 2 max_steps reached, 3 T outside the tables.
  subroutine reactor_advance_batch(X, t_end, rtol, atol, constant_pressure, p, max_steps, status)
    real(8), intent(inout) :: X(:, :)
    real(8), intent(in) :: t_end, rtol, atol
    logical, intent(in), optional :: constant_pressure
    real(8), intent(in), optional :: p(:)
    integer, intent(in), optional :: max_steps
    integer, intent(out), optional :: status(:)
    real(c_double), allocatable :: xx(:, :)
    real(c_double), allocatable, target :: pp(:)
    integer(c_int), allocatable, target :: st(:)
    integer(c_long) :: n
    integer(c_int) :: mode, ms
    type(c_ptr) :: pptr
    call need_mech('reactor_advance_batch')
    n = size(X, 1, kind=c_long)
    mode = 0
    if (present(constant_pressure)) then
      if (constant_pressure) mode = 1
    end if
    ms = 0
    if (present(max_steps)) ms = max_steps
    pptr = c_null_ptr
    if (mode == 1) then
      if (.not. present(p)) then
        write(*, '(A)') 'reactor_advance_batch: constant pressure needs p'
        stop 1
      end if
      pp = p(1:n)
      pptr = c_loc(pp)
    end if
    xx = X(:, 1:ns + 1)
    allocate(st(n))
    call fail(flint_reactor_advance_host(mech, n, mode, xx, pptr, t_end, rtol, atol, ms, c_loc(st)), &
              'reactor_advance_batch')
    X(:, 1:ns + 1) = xx
    if (present(status)) status(1:n) = st
  end subroutine

  !> CEA_solve over a block of states (Vol selects UV or HP as in CEA_solve): p(n) [bar], T(n),
  !> rho(n, ns) -> T_out(n), rho_out(n, ns) = n_k W_k; status(n) (optional): 0 converged, 1 no
  !> convergence, 2 singular, 3 T outside the thermo table.
  subroutine CEA_solve_batch(p, T, rho, T_out, rho_out, status)
    real(8), intent(in) :: p(:), T(:), rho(:, :)
    real(8), intent(out) :: T_out(:), rho_out(:, :)
    integer, intent(out), optional :: status(:)
    real(c_double), allocatable, target :: pp(:)
    real(c_double), allocatable :: tt(:), r(:, :), to(:), y(:, :)
    integer(c_int), allocatable, target :: st(:)
    integer(c_long) :: n
    call need_mech('CEA_solve_batch')
    n = size(T, kind=c_long)
    pp = p(1:n)
    tt = T
    r = rho(1:n, 1:ns)
    allocate(to(n), y(n, ns), st(n))
    call fail(flint_cea_solve_host(mech, n, merge(0, 1, Vol), c_loc(pp), tt, r, to, y, c_loc(st)), &
              'CEA_solve_batch')
    T_out(1:n) = to
    rho_out(1:n, 1:ns) = y
    if (present(status)) status(1:n) = st
  end subroutine

end module flint_amd
