! Harness around the FLINT reference modules -- TEST INFRASTRUCTURE ONLY.
!
! oracle/Makefile compiles the reference's own src/lib modules (in place under the reference
! checkout) together with this file into oracle/_ref/libflint_ref.so.  This file only fills the
! reference's public module arrays (what FLINT's loaders would do, Load_ThermoTransport.f90:30-126,
! Load_Chemistry.f90:93-162) and calls the reference routines unchanged through bind(C) entry
! points, so tests can compare the oracle and the GPU path with the reference arithmetic itself.
module ref_harness
  use iso_c_binding
  use FLINT_Lib_Thermodynamic
  use FLINT_Lib_Chemistry_data
  use FLINT_Lib_Chemistry_wdot
  implicit none
contains

  subroutine ref_setup(ns_in, W, nTk, Tk_min, na, kf, kb, nt, troe, nl, lind, nTh, Th_min, cp, h, s, &
                       ne_in, comp, names, elnames, mech, mechlen) bind(C, name='ref_setup')
    integer(c_int), value :: ns_in, nTk, Tk_min, na, nt, nl, nTh, Th_min, ne_in, mechlen
    real(c_double), intent(in) :: W(ns_in)
    real(c_double), intent(in) :: kf(na, nTk), kb(na, nTk)
    real(c_double), intent(in) :: troe(4, max(nt,1), nTk), lind(3, max(nl,1), nTk)
    real(c_double), intent(in) :: cp(ns_in, nTh), h(ns_in, nTh), s(ns_in, nTh)
    real(c_double), intent(in) :: comp(ne_in, ns_in)
    character(kind=c_char), intent(in) :: names(20*ns_in), elnames(20*max(ne_in,1)), mech(mechlen)
    character(len=64) :: mname
    integer :: i, k, Ti2

    call free_chemistry_data()
    if (allocated(wm_tab)) deallocate(wm_tab, Ri_tab, species_names)
    if (allocated(h_tab)) deallocate(h_tab, cp_tab, s_tab, dcpi_tab)
    if (allocated(elements_names)) deallocate(elements_names, species_composition)

    ns = ns_in
    allocate(wm_tab(ns), Ri_tab(ns), species_names(ns))
    wm_tab = W
    Ri_tab = Runiv/wm_tab
    do k = 1, ns
      species_names(k) = ''
      do i = 1, 20
        if (names(20*(k-1)+i) /= ' ') species_names(k)(i:i) = names(20*(k-1)+i)
      enddo
    enddo
    ne = ne_in
    if (ne > 0) then
      allocate(elements_names(ne), species_composition(ne, ns))
      species_composition = comp
      do k = 1, ne
        elements_names(k) = ''
        do i = 1, 20
          if (elnames(20*(k-1)+i) /= ' ') elements_names(k)(i:i) = elnames(20*(k-1)+i)
        enddo
      enddo
    endif

    ! thermo tables, rows Th_min .. Th_min+nTh-1
    Tmin = Th_min
    Tmax = Th_min + nTh - 1
    allocate(h_tab(Tmin:Tmax, ns), cp_tab(Tmin:Tmax, ns), s_tab(Tmin:Tmax, ns), dcpi_tab(Tmin:Tmax, ns))
    do i = Tmin, Tmax
      cp_tab(i, :) = cp(:, i - Tmin + 1)
      h_tab(i, :) = h(:, i - Tmin + 1)
      s_tab(i, :) = s(:, i - Tmin + 1)
    enddo
    dcpi_tab = 0.d0

    ! kinetics tables, rows Tk_min .. Tk_min+nTk-1
    Ti2 = Tk_min + nTk - 1
    nrc_arrh = na
    nrc_troe = nt
    nrc_lindemann = nl
    allocate(kf_tab(Tk_min:Ti2, na), kb_tab(Tk_min:Ti2, na))
    do i = Tk_min, Ti2
      kf_tab(i, :) = kf(:, i - Tk_min + 1)
      kb_tab(i, :) = kb(:, i - Tk_min + 1)
    enddo
    if (nt > 0) then
      allocate(kinf_troe_tab(Tk_min:Ti2, nt), k0_troe_tab(Tk_min:Ti2, nt), kc_troe_tab(Tk_min:Ti2, nt), &
               Fcent_tab(Tk_min:Ti2, nt))
      do i = Tk_min, Ti2
        kinf_troe_tab(i, :) = troe(1, 1:nt, i - Tk_min + 1)
        k0_troe_tab(i, :) = troe(2, 1:nt, i - Tk_min + 1)
        kc_troe_tab(i, :) = troe(3, 1:nt, i - Tk_min + 1)
        Fcent_tab(i, :) = troe(4, 1:nt, i - Tk_min + 1)
      enddo
    endif
    if (nl > 0) then
      allocate(kinf_lind_tab(Tk_min:Ti2, nl), k0_lind_tab(Tk_min:Ti2, nl), kc_lind_tab(Tk_min:Ti2, nl))
      do i = Tk_min, Ti2
        kinf_lind_tab(i, :) = lind(1, 1:nl, i - Tk_min + 1)
        k0_lind_tab(i, :) = lind(2, 1:nl, i - Tk_min + 1)
        kc_lind_tab(i, :) = lind(3, 1:nl, i - Tk_min + 1)
      enddo
    endif

    mname = ''
    do i = 1, mechlen
      mname(i:i) = mech(i)
    enddo
    call Assign_Mechanism(trim(mname))
  end subroutine ref_setup

  ! chemistry_source over n cells, SoA with leading dimension ld
  subroutine ref_wdot(n, ld, rho, T, wdot) bind(C, name='ref_wdot')
    integer(c_long), value :: n, ld
    real(c_double), intent(in) :: rho(ld, *), T(*)
    real(c_double), intent(out) :: wdot(ld, *)
    real(8) :: roi(ns), om(ns)
    integer(c_long) :: i
    do i = 1, n
      roi = rho(i, 1:ns)
      call chemistry_source(roi, T(i), om)
      wdot(i, 1:ns) = om
    enddo
  end subroutine ref_wdot

  ! CEA equilibrium (Lib_CEA_setup.f90 / Lib_CEA_solver.f90) over n states, UV (hp=0) or HP (hp=1)
  subroutine ref_cea(n, ld, p, T, rho, hp, Tout, rhoout) bind(C, name='ref_cea')
    use FLINT_CEA_data, only: Vol
    use FLINT_CEA_setup
    use FLINT_CEA_solver
    integer(c_long), value :: n, ld
    integer(c_int), value :: hp
    real(c_double), intent(in) :: p(*), T(*), rho(ld, *)
    real(c_double), intent(out) :: Tout(*), rhoout(ld, *)
    real(8) :: ri(ns), ro(ns), to
    integer(c_long) :: i
    if (.not. allocated(el_weight)) call CEA_initialize_global()
    if (hp /= 0) then
      Vol = .false.
    else
      Vol = .true.
    endif
    do i = 1, n
      ri = rho(i, 1:ns)
      to = transfer(-1_8, 1.d0)
      ro = 0.d0
      call CEA_solve(p(i), T(i), ri, to, ro)
      Tout(i) = to
      rhoout(i, 1:ns) = ro
    enddo
  end subroutine ref_cea

  ! the tables of the data-driven general routine (Load_Chemistry.f90:55-89): per reaction of each
  ! kind (Arrhenius and three-body, then Troe, then Lindemann, each in file order) ns+1 values of
  ! ni1, ni2 and epsch, the last row being the third body M
  subroutine ref_general_setup(na, nt, nl, ni1, ni2, eps) bind(C, name='ref_general_setup')
    integer(c_int), value :: na, nt, nl
    real(c_double), intent(in) :: ni1(ns + 1, *), ni2(ns + 1, *), eps(ns + 1, *)
    if (allocated(ni1_arrh_tab)) deallocate(ni1_arrh_tab, ni2_arrh_tab, epsch_arrh_tab)
    if (allocated(ni1_troe_tab)) deallocate(ni1_troe_tab, ni2_troe_tab, epsch_troe_tab)
    if (allocated(ni1_lind_tab)) deallocate(ni1_lind_tab, ni2_lind_tab, epsch_lind_tab)
    allocate(ni1_arrh_tab(ns + 1, na), ni2_arrh_tab(ns + 1, na), epsch_arrh_tab(ns + 1, na))
    ni1_arrh_tab = ni1(:, 1:na)
    ni2_arrh_tab = ni2(:, 1:na)
    epsch_arrh_tab = eps(:, 1:na)
    if (nt > 0) then
      allocate(ni1_troe_tab(ns + 1, nt), ni2_troe_tab(ns + 1, nt), epsch_troe_tab(ns + 1, nt))
      ni1_troe_tab = ni1(:, na + 1:na + nt)
      ni2_troe_tab = ni2(:, na + 1:na + nt)
      epsch_troe_tab = eps(:, na + 1:na + nt)
    endif
    if (nl > 0) then
      allocate(ni1_lind_tab(ns + 1, nl), ni2_lind_tab(ns + 1, nl), epsch_lind_tab(ns + 1, nl))
      ni1_lind_tab = ni1(:, na + nt + 1:na + nt + nl)
      ni2_lind_tab = ni2(:, na + nt + 1:na + nt + nl)
      epsch_lind_tab = eps(:, na + nt + 1:na + nt + nl)
    endif
  end subroutine ref_general_setup

  ! the general routine (Lib_Chemistry_wdot.f90:96-221) over n cells
  subroutine ref_general(n, ld, rho, T, wdot) bind(C, name='ref_general')
    integer(c_long), value :: n, ld
    real(c_double), intent(in) :: rho(ld, *), T(*)
    real(c_double), intent(out) :: wdot(ld, *)
    real(8) :: roi(ns), om(ns)
    integer(c_long) :: i
    do i = 1, n
      roi = rho(i, 1:ns)
      call general(roi, T(i), om)
      wdot(i, 1:ns) = om
    enddo
  end subroutine ref_general

  ! cons2prim (Lib_ThermoTransport.f90:130-165) over n cells: cons(ld, np) = (rho_1..rho_ns,
  ! rho u_1..rho u_nd, rho E) -> prim(ld, np) = (rho_1..rho_ns, u_1..u_nd, p), T the Newton start
  subroutine ref_cons2prim(n, ld, np, cons, T, prim) bind(C, name='ref_cons2prim')
    integer(c_long), value :: n, ld
    integer(c_int), value :: np
    real(c_double), intent(in) :: cons(ld, *), T(*)
    real(c_double), intent(out) :: prim(ld, *)
    real(8) :: c(np), q(np)
    integer(c_long) :: i
    do i = 1, n
      c = cons(i, 1:np)
      q = cons2prim(c, T(i))
      prim(i, 1:np) = q
    enddo
  end subroutine ref_cons2prim

  ! prim2cons (Lib_ThermoTransport.f90:168-185) over n cells, the inverse map
  subroutine ref_prim2cons(n, ld, np, prim, cons) bind(C, name='ref_prim2cons')
    integer(c_long), value :: n, ld
    integer(c_int), value :: np
    real(c_double), intent(in) :: prim(ld, *)
    real(c_double), intent(out) :: cons(ld, *)
    real(8) :: c(np), q(np)
    integer(c_long) :: i
    do i = 1, n
      q = prim(i, 1:np)
      c = prim2cons(q)
      cons(i, 1:np) = c
    enddo
  end subroutine ref_prim2cons

  ! species transport tables (what read_idealgas_transport fills, Load_ThermoTransport.f90:129-174)
  ! on the thermo rows Tmin..Tmax set by ref_setup: mu(nT, ns), k(nT, ns)
  subroutine ref_transport_setup(nT, mu, k) bind(C, name='ref_transport_setup')
    integer(c_int), value :: nT
    real(c_double), intent(in) :: mu(nT, *), k(nT, *)
    if (allocated(mi_tab)) deallocate(mi_tab, k_tab)
    allocate(mi_tab(Tmin:Tmax, ns), k_tab(Tmin:Tmax, ns))
    mi_tab(Tmin:Tmax, 1:ns) = mu(1:nT, 1:ns)
    k_tab(Tmin:Tmax, 1:ns) = k(1:nT, 1:ns)
  end subroutine ref_transport_setup

  ! co_k_mi_lam_Wilke (Lib_ThermoTransport.f90:407-421) and f_laminarViscosity (:365-404) per cell
  subroutine ref_wilke(n, ld, rhoi, rho, T, p, Rtot, mu, lam, mul) bind(C, name='ref_wilke')
    integer(c_long), value :: n, ld
    real(c_double), intent(in) :: rhoi(ld, *), rho(*), T(*), p(*), Rtot(*)
    real(c_double), intent(out) :: mu(*), lam(*), mul(*)
    real(8) :: r(ns), a, b
    integer(c_long) :: i
    do i = 1, n
      r = rhoi(i, 1:ns)
      call co_k_mi_lam_Wilke(r, rho(i), T(i), a, b)
      mu(i) = a
      lam(i) = b
      mul(i) = f_laminarViscosity(r, p(i), rho(i), Rtot(i))
    enddo
  end subroutine ref_wilke

end module ref_harness
