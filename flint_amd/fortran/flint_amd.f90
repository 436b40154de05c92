! Fortran drop-in for FLINT's per-cell thermochemistry calls, over the C ABI of libflint_amd
! (include/flint_amd.h) through iso_c_binding.
!
! The modules carry FLINT's own names, so that a FLINT caller's `use` lines, calls and
! deallocations compile unchanged (src/test/Fortran/test-CEA.f90:1-5, test-wdot.f90:2-6):
!
!   FLINT_Lib_Thermodynamic     FLINT_phase_prefix, ns, ne, Tmin, Tmax, Runiv, species_names,
!                               elements_names, species_composition, wm_tab, Ri_tab, cp_tab, h_tab,
!                               s_tab, dcpi_tab (Lib_ThermoTransport.f90:1-25); the table accessors
!                               f_tabT, cpi, H0i, S0i, f_cp, f_Rtot, f_molecularWeight, mass2molar,
!                               molar2mass; cons2prim, prim2cons, co_k_mi_lam_Wilke and
!                               f_laminarViscosity on the GPU
!   FLINT_Load_ThermoTransport  read_idealgas_thermo, read_idealgas_composition,
!                               read_idealgas_transport (Load_ThermoTransport.f90:7-175)
!   FLINT_Lib_Chemistry_data    nrc, free_chemistry_data (Lib_Chemistry_data.f90:1-189)
!   FLINT_Load_Chemistry        read_chemistry (Load_Chemistry.f90:7-164)
!   FLINT_Lib_Chemistry_wdot    the procedure pointer chemistry_source, Assign_Mechanism, general
!                               (Lib_Chemistry_wdot.f90:4-92)
!   FLINT_Lib_Chemistry_rhs     rhs_native (Lib_Chemistry_rhs.f90:15-52)
!   FLINT_CEA_data              Hp, Tp, Vol, Rr (Lib_CEA_data.f90:62-65)
!   FLINT_CEA_setup             el_weight, CEA_initialize_global (Lib_CEA_setup.f90:1-84)
!   FLINT_CEA_solver            CEA_solve (Lib_CEA_solver.f90:6-17)
!
! Module flint_amd adds batched forms that hand a block of cells to the GPU in one call (a
! per-cell call pays one host-device round trip):
!
!   chemistry_source_batch(rho, T, wdot)          rho(n, ns), T(n), wdot(n, ns)
!   rhs_native_batch(Z, F [, J])                  Z(n, ns+1), F(n, ns+1), J(n, ns+1, ns+1)
!   reactor_advance_batch(X, t_end, rtol, atol, constant_pressure, p, max_steps, status)
!   CEA_solve_batch(p, T, rho, T_out, rho_out, status)
!   cons2prim_batch, prim2cons_batch, transport_batch
!
! Batched arrays are cell-fastest, which is the C ABI's structure-of-arrays layout with ld = n.
! J(i, c, r) = dF_r/dZ_c of cell i.  Differences from the reference:
!   * read_idealgas_thermo loads the whole native folder (thermo and chemistry: the GPU kernels
!     need both) into one library handle; read_chemistry on the same folder then reports the name;
!   * the procedures backed by the GPU (cons2prim, prim2cons, co_k_mi_lam_Wilke,
!     f_laminarViscosity) are not pure;
!   * free_chemistry_data releases the kinetics: chemistry_source fails until the next load, while
!     the thermodynamic module arrays stay with the caller, as in FLINT;
!   * errors of the library stop the program with its message, as FLINT's loaders do; the loaders
!     themselves return a nonzero ios.
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
    function flint_mech_thermo(m, W, cp, h, s, dcpi) bind(C, name="flint_mech_thermo") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: W(*), cp(*), h(*), s(*), dcpi(*)
      integer(c_int) :: rc
    end function
    function flint_mech_names(m, kind, buf, len, count) bind(C, name="flint_mech_names") result(rc)
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: m
      integer(c_int), value :: kind, len
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), intent(out) :: count
      integer(c_int) :: rc
    end function
    function flint_mech_composition(m, comp, el_weight) bind(C, name="flint_mech_composition") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: comp(*), el_weight(*)
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
    function flint_source_general_host(m, n, rho, T, wdot) bind(C, name="flint_source_general_host") result(rc)
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

!> The library handle of the loaded folder and the error helpers shared by the FLINT modules.
module flint_amd_state
  use iso_c_binding
  use flint_amd_c
  implicit none
  type(c_ptr), save :: flint_handle = c_null_ptr
  !> mechanism name from chemistry-info.txt, and whether its kinetics are loaded
  character(len=64), save :: loaded_mech_name = ''
  logical, save :: chemistry_loaded = .false.
  character(len=:), allocatable, save :: loaded_folder

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

  subroutine flint_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(A,A,I0,A,A)') trim(where), ': flint error ', rc, ': ', flint_error_message()
      stop 1
    end if
  end subroutine

  subroutine need_mech(where)
    character(len=*), intent(in) :: where
    if (.not. c_associated(flint_handle)) then
      write(*, '(A,A)') trim(where), ': no mechanism loaded (call read_idealgas_thermo first)'
      stop 1
    end if
  end subroutine
end module flint_amd_state

module FLINT_Lib_Thermodynamic
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  implicit none

  !> file-name prefix of the native input files (phase.txt, thermo.dat, ...)
  character(len=128) :: FLINT_phase_prefix = ''
  integer :: ns = 0                 ! species
  integer :: ne = 0                 ! atomic elements
  integer :: Tmin = 0, Tmax = 0     ! temperature range of the thermo tables
  real(kind=8), parameter :: Runiv = 8314.51d0
  integer, parameter :: s_str_len = 20
  character(len=s_str_len), allocatable :: species_names(:)
  character(len=s_str_len), allocatable :: elements_names(:)
  real(kind=8), dimension(:, :), allocatable :: species_composition
  real(kind=8), dimension(:), allocatable :: wm_tab, Ri_tab
  real(kind=8), dimension(:, :), allocatable :: cp_tab, dcpi_tab, h_tab, s_tab
  real(kind=8), dimension(:, :), allocatable :: mi_tab, k_tab

contains

  !> Value of a per-mass table tab(Tmin:Tmax, ns) for species sp at T: the 1 K bin int(T) and the
  !> linear interpolation a + (b - a)(T - int(T)).
  pure function f_tabT(T, sp, tab) result(v)
    real(8), intent(in) :: T, tab(Tmin:Tmax, ns)
    integer, intent(in) :: sp
    real(8) :: v
    integer :: i
    i = int(T)
    v = tab(i, sp) + (tab(i + 1, sp) - tab(i, sp)) * (T - i)
  end function

  pure function cpi(i, T) result(v)
    integer, intent(in) :: i
    real(8), intent(in) :: T
    real(8) :: v
    v = f_tabT(T, i, cp_tab)
  end function

  pure function H0i(i, T) result(v)
    integer, intent(in) :: i
    real(8), intent(in) :: T
    real(8) :: v
    v = f_tabT(T, i, h_tab)
  end function

  pure function S0i(i, T) result(v)
    integer, intent(in) :: i
    real(8), intent(in) :: T
    real(8) :: v
    v = f_tabT(T, i, s_tab)
  end function

  !> mixture c_p [J/(kg K)] of the partial densities rhoi at T, rho = sum(rhoi)
  pure function f_cp(rhoi, T, rho) result(v)
    real(8), intent(in) :: rhoi(ns), T, rho
    real(8) :: v
    integer :: k
    v = 0d0
    do k = 1, ns
      v = v + rhoi(k) / rho * cpi(k, T)
    end do
  end function

  !> mixture gas constant [J/(kg K)]
  pure function f_Rtot(rhoi) result(v)
    real(8), intent(in) :: rhoi(ns)
    real(8) :: v, rho
    integer :: k
    rho = sum(rhoi)
    v = 0d0
    do k = 1, ns
      v = v + Ri_tab(k) * rhoi(k) / rho
    end do
  end function

  !> mixture molecular weight [kg/kmol]
  pure function f_molecularWeight(rhoi) result(v)
    real(8), intent(in) :: rhoi(ns)
    real(8) :: v, rho
    integer :: k
    rho = sum(rhoi)
    v = 0d0
    do k = 1, ns
      v = v + rhoi(k) / rho / wm_tab(k)
    end do
    v = 1d0 / v
  end function

  pure function mass2molar(Y_species) result(x_species)
    real(8), intent(in) :: Y_species(ns)
    real(8) :: x_species(ns)
    x_species = Y_species / wm_tab
    x_species = x_species / sum(x_species)
  end function

  pure function molar2mass(x_species) result(y_species)
    real(8), intent(in) :: x_species(ns)
    real(8) :: y_species(ns)
    y_species = x_species * wm_tab
    y_species = y_species / sum(y_species)
  end function

  !> Primitive variables (rho_1..rho_ns, u_1..u_nd, p) of one cell from its conservative variables
  !> (rho_1..rho_ns, rho u_1..rho u_nd, rho E), Newton on T from T (Lib_ThermoTransport.f90:130);
  !> 0 <= nd <= 3.  p is NaN when T leaves the thermo table.
  function cons2prim(cons, T) result(prim)
    real(8), intent(in) :: cons(:)
    real(8), intent(in) :: T
    real(8) :: prim(size(cons))
    real(c_double) :: c(size(cons)), q(size(cons)), tt(1)
    call need_mech('cons2prim')
    c = cons
    tt(1) = T
    call flint_check(flint_cons2prim_host(flint_handle, 1_c_long, int(size(cons) - ns - 1, c_int), c, tt, q, &
                                          c_null_ptr, c_null_ptr), 'cons2prim')
    prim = q
  end function

  !> Conservative variables of one cell from its primitive ones (Lib_ThermoTransport.f90:168).
  function prim2cons(prim) result(cons)
    real(8), intent(in) :: prim(:)
    real(8) :: cons(size(prim))
    real(c_double) :: q(size(prim)), c(size(prim))
    call need_mech('prim2cons')
    q = prim
    call flint_check(flint_prim2cons_host(flint_handle, 1_c_long, int(size(prim) - ns - 1, c_int), q, c), &
                     'prim2cons')
    cons = c
  end function

  !> Mixture viscosity milam [Pa s] and conductivity klam [W/(m K)] of one cell by Wilke's rule
  !> (co_k_mi_lam_Wilke, Lib_ThermoTransport.f90:407); the tables come from transport.dat in the
  !> loaded folder.  NaN when T leaves the transport table.
  subroutine co_k_mi_lam_Wilke(rhoi, rho, T, milam, klam)
    real(8), intent(in) :: rhoi(:), rho, T
    real(8), intent(out) :: milam, klam
    real(c_double) :: r(ns), a(1), b(1), m(1), k(1)
    call need_mech('co_k_mi_lam_Wilke')
    if (size(rhoi) /= ns) then
      write(*, '(A,I0,A,I0)') 'co_k_mi_lam_Wilke: rhoi has ', size(rhoi), ' entries, the mechanism ', ns
      stop 1
    end if
    r = rhoi
    a(1) = rho
    b(1) = T
    call flint_check(flint_transport_host(flint_handle, 1_c_long, r, a, b, m, k, c_null_ptr), 'co_k_mi_lam_Wilke')
    milam = m(1)
    klam = k(1)
  end subroutine

  !> Mixture viscosity at T = p / (Rtot rho) (f_laminarViscosity, Lib_ThermoTransport.f90:365).
  function f_laminarViscosity(rhoi, p, rho, Rtot) result(mil)
    real(8), intent(in) :: rhoi(:), p, rho, Rtot
    real(8) :: mil, klam
    call co_k_mi_lam_Wilke(rhoi, rho, p / (Rtot * rho), mil, klam)
  end function
end module FLINT_Lib_Thermodynamic

module FLINT_Load_ThermoTransport
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Thermodynamic
  implicit none

contains

  !> Loads the native input folder (default INPUT; FLINT_phase_prefix prefixes its file names)
  !> and fills FLINT_Lib_Thermodynamic's arrays (ns, wm_tab, Ri_tab, species_names, cp_tab,
  !> h_tab, s_tab, dcpi_tab over Tmin:Tmax, and the composition), which the caller owns.  A
  !> previously loaded folder is replaced.  ios = 0 on success.
  function read_idealgas_thermo(folder) result(ios)
    character(len=*), intent(in), optional :: folder
    integer :: ios
    character(len=:), allocatable :: f
    integer(c_int) :: rc, n, a, b, c, d, cnt
    real(c_double), allocatable :: W(:), cp(:, :), h(:, :), s(:, :), dc(:, :)
    character(kind=c_char), allocatable :: buf(:)
    integer :: k
    if (present(folder)) then
      f = trim(folder)
    else
      f = 'INPUT'
    end if
    if (c_associated(flint_handle)) call flint_mech_destroy(flint_handle)
    flint_handle = c_null_ptr
    chemistry_loaded = .false.
    rc = flint_mech_load(f // c_null_char, trim(FLINT_phase_prefix) // c_null_char, flint_handle)
    if (rc /= 0) then
      write(*, '(A,A)') '[ERROR] ', flint_error_message()
      flint_handle = c_null_ptr
      ios = int(rc)
      return
    end if
    chemistry_loaded = .true.
    loaded_folder = f
    rc = flint_mech_query(flint_handle, n, a, b, c, d)
    ns = n
    Tmin = c
    Tmax = d
    allocate(W(ns), cp(Tmin:Tmax, ns), h(Tmin:Tmax, ns), s(Tmin:Tmax, ns), dc(Tmin:Tmax, ns))
    call flint_check(flint_mech_thermo(flint_handle, W, cp, h, s, dc), 'read_idealgas_thermo')
    if (allocated(wm_tab)) deallocate(wm_tab)
    if (allocated(Ri_tab)) deallocate(Ri_tab)
    if (allocated(species_names)) deallocate(species_names)
    if (allocated(cp_tab)) deallocate(cp_tab)
    if (allocated(h_tab)) deallocate(h_tab)
    if (allocated(s_tab)) deallocate(s_tab)
    if (allocated(dcpi_tab)) deallocate(dcpi_tab)
    allocate(wm_tab(1:ns), Ri_tab(1:ns), species_names(1:ns))
    wm_tab = W
    Ri_tab = Runiv / wm_tab
    allocate(cp_tab(Tmin:Tmax, 1:ns), h_tab(Tmin:Tmax, 1:ns), s_tab(Tmin:Tmax, 1:ns), dcpi_tab(Tmin:Tmax, 1:ns))
    cp_tab = cp
    h_tab = h
    s_tab = s
    dcpi_tab = dc
    allocate(buf(s_str_len * ns))
    call flint_check(flint_mech_names(flint_handle, 0_c_int, buf, int(s_str_len, c_int), cnt), &
                     'read_idealgas_thermo')
    do k = 1, ns
      species_names(k) = transfer(buf((k - 1) * s_str_len + 1:k * s_str_len), species_names(k))
    end do
    open(newunit=k, file=f // '/chemistry-info.txt', status='old', action='read', iostat=ios)
    if (ios == 0) then
      read(k, *, iostat=ios) loaded_mech_name
      close(k)
    end if
    ios = read_idealgas_composition(f)
  end function

  !> elements_names(1:ne) and species_composition(1:ne, 1:ns) of the loaded folder
  !> (composition.txt); ne = 0 and nothing allocated when the folder has none.
  function read_idealgas_composition(folder) result(ios)
    character(len=*), intent(in), optional :: folder
    integer :: ios
    integer(c_int) :: cnt
    character(kind=c_char), allocatable :: buf(:)
    real(c_double), allocatable :: comp(:, :), elw(:)
    integer :: e
    ios = 0
    if (present(folder)) then
      if (.not. allocated(loaded_folder)) then
        ios = read_idealgas_thermo(folder)
        return
      else if (trim(folder) /= loaded_folder) then
        ios = read_idealgas_thermo(folder)
        return
      end if
    end if
    call need_mech('read_idealgas_composition')
    allocate(buf(1))
    call flint_check(flint_mech_names(flint_handle, 1_c_int, buf, 1_c_int, cnt), 'read_idealgas_composition')
    ne = cnt
    if (ne == 0) then
      write(*, *) '[WARNING] Composition data not found'
      return
    end if
    deallocate(buf)
    allocate(buf(s_str_len * ne), comp(ne, ns), elw(ne))
    call flint_check(flint_mech_names(flint_handle, 1_c_int, buf, int(s_str_len, c_int), cnt), &
                     'read_idealgas_composition')
    call flint_check(flint_mech_composition(flint_handle, comp, elw), 'read_idealgas_composition')
    if (allocated(elements_names)) deallocate(elements_names)
    if (allocated(species_composition)) deallocate(species_composition)
    allocate(elements_names(1:ne), species_composition(1:ne, 1:ns))
    do e = 1, ne
      elements_names(e) = transfer(buf((e - 1) * s_str_len + 1:e * s_str_len), elements_names(e))
    end do
    species_composition = comp
  end function

  !> Mixture transport comes from transport.dat of the loaded folder, which read_idealgas_thermo
  !> has read with the rest; the per-species tables stay inside the library (mi_tab and k_tab are
  !> not filled).  ios = 0 when the folder has transport.dat.
  function read_idealgas_transport(folder) result(ios)
    character(len=*), intent(in), optional :: folder
    integer :: ios
    logical :: there
    character(len=:), allocatable :: f
    if (present(folder)) then
      f = trim(folder)
    else
      f = 'INPUT'
    end if
    inquire(file=f // '/' // trim(FLINT_phase_prefix) // 'transport.dat', exist=there)
    ios = merge(0, 1, there)
  end function
end module FLINT_Load_ThermoTransport

module FLINT_Lib_Chemistry_data
  use flint_amd_state
  implicit none
  integer :: nrc = 0  ! reactions of the loaded mechanism

contains

  !> Releases the kinetics of the loaded mechanism: chemistry_source stops with a message until
  !> the next read_chemistry / read_idealgas_thermo.  The thermodynamic data stay.
  subroutine free_chemistry_data()
    chemistry_loaded = .false.
    nrc = 0
  end subroutine
end module FLINT_Lib_Chemistry_data

module FLINT_Load_Chemistry
  use flint_amd_state
  use FLINT_Lib_Thermodynamic, only: ns, FLINT_phase_prefix
  use FLINT_Load_ThermoTransport, only: read_idealgas_thermo
  use FLINT_Lib_Chemistry_data
  implicit none

contains

  !> The kinetics of the native folder (default INPUT): loads it if read_idealgas_thermo has not
  !> (or after free_chemistry_data), and returns the mechanism name of chemistry-info.txt.
  function read_chemistry(folder, mech_name) result(ios)
    character(len=*), intent(in), optional :: folder
    character(len=*), intent(out), optional :: mech_name
    integer :: ios, u
    character(len=:), allocatable :: f
    character(len=256) :: line
    if (present(folder)) then
      f = trim(folder)
    else
      f = 'INPUT'
    end if
    ios = 0
    if (.not. c_associated(flint_handle) .or. .not. chemistry_loaded) then
      ios = read_idealgas_thermo(f)
    else if (loaded_folder /= f) then
      ios = read_idealgas_thermo(f)
    end if
    if (ios /= 0) return
    open(newunit=u, file=f // '/' // trim(FLINT_phase_prefix) // 'chemistry-info.txt', status='old', &
         action='read', iostat=ios)
    if (ios == 0) then
      read(u, '(A)', iostat=ios)
      read(u, '(A)', iostat=ios)
      read(u, '(A)', iostat=ios) line
      if (index(line, '=') > 0) read(line(index(line, '=') + 1:), *, iostat=ios) nrc
      close(u)
    end if
    ios = 0
    if (present(mech_name)) mech_name = loaded_mech_name
  end function
end module FLINT_Load_Chemistry

module FLINT_Lib_Chemistry_wdot
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Thermodynamic, only: ns
  implicit none

  !> the source-term routine of the assigned mechanism
  procedure(chemsource_if), pointer, public :: chemistry_source => null()

  abstract interface
    subroutine chemsource_if(roi, temp, omegadot)
      use FLINT_Lib_Thermodynamic, only: ns
      real(8), intent(inout) :: roi(ns)
      real(8), intent(in) :: temp
      real(8), intent(out) :: omegadot(ns)
    end subroutine chemsource_if
  end interface

contains

  !> Points chemistry_source at the loaded mechanism's own routine when mad_world names it (its
  !> GPU kernels, or the general routine for a mechanism compiled without them), otherwise warns
  !> and falls back to the data-driven general routine, as FLINT does for an unknown name.
  subroutine Assign_Mechanism(mad_world)
    character(len=*), intent(in) :: mad_world
    call need_mech('Assign_Mechanism')
    if (trim(mad_world) == trim(loaded_mech_name)) then
      chemistry_source => explicit_source
    else
      write(*, '(1X,A)') "[WARNING] Explicit procedure for " // trim(mad_world) // &
                         " not found, defaulting to the general procedure"
      chemistry_source => general
    end if
  end subroutine

  !> Production rates [kg/m^3/s] of one cell by the loaded mechanism's routine.  Like FLINT's WD
  !> routine (WD.f90:18-21) negative densities are set to zero in roi for WD; the generated
  !> mechanisms leave roi as it is.
  subroutine explicit_source(roi, temp, omegadot)
    real(8), intent(inout) :: roi(ns)
    real(8), intent(in) :: temp
    real(8), intent(out) :: omegadot(ns)
    real(8) :: t(1)
    call need_kinetics('chemistry_source')
    if (trim(loaded_mech_name) == 'WD') roi = max(roi, 0d0)
    t(1) = temp
    call flint_check(flint_source_host(flint_handle, 1_c_long, roi, t, omegadot), 'chemistry_source')
  end subroutine

  !> Production rates of one cell by the data-driven general routine (Lib_Chemistry_wdot.f90:96),
  !> from the Reaction definition block of chemistry-info.txt.
  subroutine general(roi, temp, omegadot)
    real(8), intent(inout) :: roi(ns)
    real(8), intent(in) :: temp
    real(8), intent(out) :: omegadot(ns)
    real(8) :: t(1)
    call need_kinetics('general')
    t(1) = temp
    call flint_check(flint_source_general_host(flint_handle, 1_c_long, roi, t, omegadot), 'general')
  end subroutine

  subroutine need_kinetics(where)
    character(len=*), intent(in) :: where
    call need_mech(where)
    if (.not. chemistry_loaded) then
      write(*, '(A,A)') trim(where), ': the kinetics were released (free_chemistry_data)'
      stop 1
    end if
  end subroutine
end module FLINT_Lib_Chemistry_wdot

module FLINT_Lib_Chemistry_rhs
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Chemistry_wdot
  implicit none
  logical :: reactive = .true.

contains

  !> Constant-volume reactor right-hand side of one cell, Z = (rho_1..rho_ns, T); F = -1 outside
  !> the thermo table (Lib_Chemistry_rhs.f90:15-52).
  subroutine rhs_native(nz, time, Z, F)
    integer, intent(in) :: nz
    real(8), intent(in) :: time
    real(8), intent(in) :: Z(nz)
    real(8), intent(out) :: F(nz)
    call need_mech('rhs_native')
    call flint_check(flint_rhs_host(flint_handle, 1_c_long, Z, F, c_null_ptr), 'rhs_native')
  end subroutine
end module FLINT_Lib_Chemistry_rhs

module FLINT_CEA_data
  implicit none
  !> equilibrium at fixed internal energy and volume (Vol, CEA_initialize_global's default,
  !> Lib_CEA_setup.f90:25) or at fixed enthalpy and pressure (Vol = .false.)
  logical :: Hp = .true., Tp = .false., Vol = .true.
  real(8), parameter :: Rr = 8314.51d0
end module FLINT_CEA_data

module FLINT_CEA_setup
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Thermodynamic, only: ne, ns
  implicit none
  real(kind=8), dimension(:), allocatable :: el_weight

contains

  !> Resets the equilibrium flags to FLINT's defaults (Hp, Vol) and sets el_weight(1:ne), the
  !> atomic weights of the elements; the equilibrium data were set up when the folder was loaded.
  subroutine CEA_initialize_global()
    use FLINT_CEA_data
    real(c_double), allocatable :: comp(:, :), elw(:)
    call need_mech('CEA_initialize_global')
    Hp = .true.
    Tp = .false.
    Vol = .true.
    if (allocated(el_weight)) deallocate(el_weight)
    allocate(comp(max(ne, 1), ns), elw(max(ne, 1)))
    call flint_check(flint_mech_composition(flint_handle, comp, elw), 'CEA_initialize_global')
    allocate(el_weight(1:ne))
    el_weight = elw(1:ne)
  end subroutine
end module FLINT_CEA_setup

module FLINT_CEA_solver
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Thermodynamic, only: ns
  implicit none

contains

  !> Equilibrium of one state: with Vol (default) at the internal energy of rhoi_in at T_in and the
  !> volume 1e5/sum(rhoi_in); otherwise at p_in [bar] and the enthalpy of rhoi_in at T_in.
  !> rhoi_out = n_k W_k; T_out and rhoi_out are NaN when the solver fails.
  subroutine CEA_solve(p_in, T_in, rhoi_in, T_out, rhoi_out)
    use FLINT_CEA_data, only: Vol
    real(8), dimension(:), intent(in) :: rhoi_in
    real(8), intent(in) :: p_in, T_in
    real(8), intent(out) :: T_out
    real(8), dimension(:), intent(out) :: rhoi_out
    real(c_double), target :: p(1)
    real(c_double) :: t(1), to(1), r(ns), y(ns)
    integer(c_int), target :: st(1)
    call need_mech('CEA_solve')
    p(1) = p_in
    t(1) = T_in
    r = rhoi_in(1:ns)
    call flint_check(flint_cea_solve_host(flint_handle, 1_c_long, merge(0, 1, Vol), c_loc(p), t, r, to, y, &
                                          c_loc(st)), 'CEA_solve')
    T_out = to(1)
    rhoi_out(1:ns) = y
  end subroutine
end module FLINT_CEA_solver

!> Batched forms of the FLINT calls: one GPU call for a block of cells.
module flint_amd
  use iso_c_binding
  use flint_amd_c
  use flint_amd_state
  use FLINT_Lib_Thermodynamic, only: ns
  use FLINT_CEA_data, only: Vol
  implicit none
  private
  public :: chemistry_source_batch, rhs_native_batch, reactor_advance_batch, CEA_solve_batch, cons2prim_batch, &
            prim2cons_batch, transport_batch, flint_error_message

contains

  subroutine need_rows(where, n, arrays_ok)
    character(len=*), intent(in) :: where
    integer(c_long), intent(in) :: n
    logical, intent(in) :: arrays_ok
    call need_mech(where)
    if (.not. arrays_ok) then
      write(*, '(A,A,I0,A)') trim(where), ': array shapes do not match ', n, ' cells of the loaded mechanism'
      stop 1
    end if
  end subroutine

  !> co_k_mi_lam_Wilke over n cells in one call: rhoi(n, ns), rho(n), T(n) -> milam(n), klam(n).
  subroutine transport_batch(rhoi, rho, T, milam, klam, status)
    real(8), intent(in) :: rhoi(:, :), rho(:), T(:)
    real(8), intent(out) :: milam(:), klam(:)
    integer, intent(out), optional :: status(:)
    integer(c_int), allocatable, target :: st(:)
    real(c_double), allocatable :: r(:, :)
    integer(c_long) :: n
    n = size(T, kind=c_long)
    call need_rows('transport_batch', n, size(rhoi, 1) == n .and. size(rhoi, 2) == ns .and. size(rho) == n .and. &
                   size(milam) >= n .and. size(klam) >= n)
    r = rhoi
    allocate(st(n))
    call flint_check(flint_transport_host(flint_handle, n, r, rho, T, milam, klam, c_loc(st)), 'transport_batch')
    if (present(status)) status(1:n) = st
  end subroutine

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
    n = size(T, kind=c_long)
    npr = size(cons, 2)
    call need_rows('cons2prim_batch', n, size(cons, 1) >= n .and. size(prim, 1) >= n .and. size(prim, 2) >= npr)
    c = cons(1:n, 1:npr)
    tt = T
    allocate(q(n, npr), to(n), st(n))
    call flint_check(flint_cons2prim_host(flint_handle, n, int(npr - ns - 1, c_int), c, tt, q, c_loc(to), &
                                          c_loc(st)), 'cons2prim_batch')
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
    n = size(prim, 1, kind=c_long)
    npr = size(prim, 2)
    call need_rows('prim2cons_batch', n, size(cons, 1) >= n .and. size(cons, 2) >= npr)
    q = prim
    allocate(c(n, npr))
    call flint_check(flint_prim2cons_host(flint_handle, n, int(npr - ns - 1, c_int), q, c), 'prim2cons_batch')
    cons(1:n, 1:npr) = c
  end subroutine

  !> chemistry_source over a block of cells: rho(n, ns), T(n) -> wdot(n, ns).
  subroutine chemistry_source_batch(rho, T, wdot)
    real(8), intent(in) :: rho(:, :), T(:)
    real(8), intent(out) :: wdot(:, :)
    real(c_double), allocatable :: r(:, :), tt(:), w(:, :)
    integer(c_long) :: n
    n = size(T, kind=c_long)
    call need_rows('chemistry_source_batch', n, size(rho, 1) >= n .and. size(rho, 2) >= ns .and. &
                   size(wdot, 1) >= n .and. size(wdot, 2) >= ns)
    r = rho(1:n, 1:ns)
    tt = T
    allocate(w(n, ns))
    call flint_check(flint_source_host(flint_handle, n, r, tt, w), 'chemistry_source_batch')
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
    n = size(Z, 1, kind=c_long)
    nz = ns + 1
    call need_rows('rhs_native_batch', n, size(Z, 2) >= nz .and. size(F, 1) >= n .and. size(F, 2) >= nz)
    zz = Z(:, 1:nz)
    allocate(ff(n, nz))
    if (present(J)) then
      allocate(jj(n, nz, nz))
      call flint_check(flint_rhs_host(flint_handle, n, zz, ff, c_loc(jj)), 'rhs_native_batch')
      J(1:n, 1:nz, 1:nz) = jj
    else
      call flint_check(flint_rhs_host(flint_handle, n, zz, ff, c_null_ptr), 'rhs_native_batch')
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
    n = size(X, 1, kind=c_long)
    call need_rows('reactor_advance_batch', n, size(X, 2) >= ns + 1)
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
    call flint_check(flint_reactor_advance_host(flint_handle, n, mode, xx, pptr, t_end, rtol, atol, ms, c_loc(st)), &
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
    n = size(T, kind=c_long)
    call need_rows('CEA_solve_batch', n, size(p) >= n .and. size(rho, 1) >= n .and. size(rho, 2) >= ns .and. &
                   size(T_out) >= n .and. size(rho_out, 1) >= n .and. size(rho_out, 2) >= ns)
    pp = p(1:n)
    tt = T
    r = rho(1:n, 1:ns)
    allocate(to(n), y(n, ns), st(n))
    call flint_check(flint_cea_solve_host(flint_handle, n, merge(0, 1, Vol), c_loc(pp), tt, r, to, y, c_loc(st)), &
                     'CEA_solve_batch')
    T_out(1:n) = to
    rho_out(1:n, 1:ns) = y
    if (present(status)) status(1:n) = st
  end subroutine

end module flint_amd
