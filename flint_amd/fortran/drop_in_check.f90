! A FLINT caller, written against FLINT's own module names and call sequence: the use lines of
! src/test/Fortran/test-CEA.f90:1-5 and test-wdot.f90:2-6, read_idealgas_thermo / read_chemistry /
! Assign_Mechanism / Chemistry_Source as test-wdot.f90:27-51 calls them, CEA_initialize_global /
! CEA_solve as test-CEA.f90:19-32, and the deallocations of test-CEA.f90:46-50.  It writes every
! result at full precision, so that the same calls made through the C ABI can be compared.
!
!   drop_in_check <native input folder> <states file> <results file>
!
! states file:  n, then n lines "T rho_1 .. rho_ns"          (chemistry_source, rhs_native)
!               m, then m lines "T p_bar Y_1 .. Y_ns"        (CEA_solve, UV and HP)
!               k, then k lines "p_Pa Y_1 .. Y_ns T"         (constant-pressure reactors)
!               t_end rtol atol                              (reactor integration)
!               q nd, then q lines "T0 rho_1 .. rho_ns u_1 .. u_nd p"  (prim2cons, cons2prim, and
!               co_k_mi_lam_Wilke at (rho_k, sum rho_k, T0) when the folder has transport.dat)
! results file: blocks "# name rows cols" followed by the rows; "# species_names ns" followed by
!               the names.
program drop_in_check
  use FLINT_Lib_Thermodynamic
  use FLINT_Load_ThermoTransport
  use FLINT_CEA_setup
  use FLINT_CEA_solver
  use FLINT_Load_chemistry
  use FLINT_Lib_Chemistry_data
  use FLINT_Lib_Chemistry_wdot
  use FLINT_Lib_Chemistry_rhs
  use FLINT_CEA_data, only: Vol
  use flint_amd
  implicit none
  character(len=512) :: folder, fin, fout
  character(len=32) :: mech_name
  integer :: err, n, m, k, i, u, v, nz
  real(8), allocatable :: T(:), rho(:, :), w1(:, :), w2(:, :), w3(:, :), F1(:, :), F2(:, :), J(:, :, :)
  real(8), allocatable :: Tc(:), pc(:), Yc(:, :), To(:), Yo(:, :), Tb(:), Yb(:, :)
  real(8), allocatable :: pr(:), X(:, :), row(:), Rm(:, :)
  integer, allocatable :: st(:)
  real(8) :: t_end, rtol, atol
  integer :: q, nd, npr
  real(8), allocatable :: T0(:), P0(:, :), C1(:, :), C2(:, :), Q1(:, :), Q2(:, :), Tq(:)
  integer, allocatable :: sq(:)
  real(8), allocatable :: mu1(:), k1(:), mu2(:), k2(:), rq(:)
  logical :: has_tr

  call get_command_argument(1, folder)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  err = read_idealgas_thermo(trim(folder))
  if (err /= 0) stop 2
  err = read_chemistry( folder=trim(folder), mech_name=mech_name )
  if (err /= 0) stop 3
  call Assign_Mechanism(mech_name)
  nz = ns + 1

  open(newunit=u, file=trim(fin), status='old', action='read')
  read(u, *) n
  allocate(T(n), rho(n, ns), row(nz + 1))
  do i = 1, n
    read(u, *) row(1:nz)
    T(i) = row(1)
    rho(i, :) = row(2:nz)
  end do
  read(u, *) m
  allocate(Tc(m), pc(m), Yc(m, ns))
  do i = 1, m
    read(u, *) row(1:ns + 2)
    Tc(i) = row(1)
    pc(i) = row(2)
    Yc(i, :) = row(3:ns + 2)
  end do
  read(u, *) k
  allocate(pr(k), X(k, nz))
  do i = 1, k
    read(u, *) row(1:nz + 1)
    pr(i) = row(1)
    X(i, :) = row(2:nz + 1)
  end do
  read(u, *) t_end, rtol, atol
  read(u, *) q, nd
  npr = ns + nd + 1
  deallocate(row)
  allocate(row(max(nz + 1, npr + 1)), T0(q), P0(q, npr))
  do i = 1, q
    read(u, *) row(1:npr + 1)
    T0(i) = row(1)
    P0(i, :) = row(2:npr + 1)
  end do
  close(u)

  open(newunit=v, file=trim(fout), status='replace', action='write')

  ! the module arrays FLINT's loaders fill, as callers read them
  write(v, '(A,1X,A,1X,I0)') '#', 'species_names', ns
  do i = 1, ns
    write(v, '(A)') trim(species_names(i))
  end do
  call put(v, 'wm_tab', reshape([wm_tab, Ri_tab], [ns, 2]))
  call put(v, 'thermo_rows', reshape([cp_tab(Tmin, :), cp_tab(1000, :), cp_tab(Tmax, :), h_tab(1000, :), &
                                      s_tab(1000, :), dcpi_tab(1000, :)], [ns, 6]))
  call put(v, 'composition', transpose(species_composition))
  allocate(Rm(n, 3))
  do i = 1, n
    Rm(i, 1) = f_Rtot(rho(i, :))
    Rm(i, 2) = f_cp(rho(i, :), T(i), sum(rho(i, :)))
    Rm(i, 3) = H0i(1, T(i))
  end do
  call put(v, 'f_Rtot_f_cp_H0i', Rm)

  ! per-cell calls, as a CFD code makes them
  allocate(w1(n, ns), F1(n, nz))
  do i = 1, n
    row(1:ns) = rho(i, :)
    call Chemistry_Source(row(1:ns), T(i), w1(i, :))
    row(1:ns) = rho(i, :)
    row(nz) = T(i)
    call rhs_native(nz, 0d0, row(1:nz), F1(i, :))
  end do
  call put(v, 'chemistry_source', w1)
  call put(v, 'rhs_native', F1)

  ! an unknown mechanism name: FLINT warns and falls back to the general routine
  call Assign_Mechanism('nemo')
  allocate(w3(n, ns))
  do i = 1, n
    row(1:ns) = rho(i, :)
    call chemistry_source(row(1:ns), T(i), w3(i, :))
  end do
  call put(v, 'general', w3)
  call Assign_Mechanism(mech_name)

  ! the same cells in one call each
  allocate(w2(n, ns), F2(n, nz), J(n, nz, nz))
  call chemistry_source_batch(rho, T, w2)
  call put(v, 'chemistry_source_batch', w2)
  deallocate(F2)
  allocate(F2(n, nz))
  call rhs_native_batch(reshape([rho, T], [n, nz]), F2, J)
  call put(v, 'rhs_native_batch', F2)
  call put(v, 'rhs_native_batch_J', reshape(J, [n, nz * nz]))

  ! equilibrium, one state at a time (UV, FLINT's default) and batched (HP)
  call CEA_initialize_global()
  call put(v, 'el_weight', reshape(el_weight, [ne, 1]))
  allocate(To(m), Yo(m, ns), Tb(m), Yb(m, ns), st(max(m, k)))
  do i = 1, m
    call CEA_solve(pc(i), Tc(i), Yc(i, :), To(i), Yo(i, :))
  end do
  call put(v, 'CEA_solve_uv', reshape([To, Yo], [m, nz]))
  Vol = .false.
  call CEA_solve_batch(pc, Tc, Yc, Tb, Yb, st(1:m))
  call put(v, 'CEA_solve_batch_hp', reshape([Tb, Yb, dble(st(1:m))], [m, nz + 1]))

  ! constant-pressure reactors
  call reactor_advance_batch(X, t_end, rtol, atol, constant_pressure=.true., p=pr, status=st(1:k))
  call put(v, 'reactor_advance_batch', reshape([X, dble(st(1:k))], [k, nz + 1]))

  ! conservative <-> primitive variables, per cell and batched
  allocate(C1(q, npr), Q1(q, npr), C2(q, npr), Q2(q, npr), Tq(q), sq(q))
  do i = 1, q
    C1(i, :) = prim2cons(P0(i, :))
    Q1(i, :) = cons2prim(C1(i, :), T0(i))
  end do
  call put(v, 'prim2cons', C1)
  call put(v, 'cons2prim', Q1)
  call prim2cons_batch(P0, C2)
  call cons2prim_batch(C2, T0, Q2, Tq, sq)
  call put(v, 'prim2cons_batch', C2)
  call put(v, 'cons2prim_batch', reshape([Q2, Tq, dble(sq)], [q, npr + 2]))

  ! mixture transport, per cell and batched
  inquire(file=trim(folder)//'/transport.dat', exist=has_tr)
  if (has_tr) then
    allocate(mu1(q), k1(q), mu2(q), k2(q), rq(q))
    do i = 1, q
      rq(i) = sum(P0(i, 1:ns))
      call co_k_mi_lam_Wilke(P0(i, 1:ns), rq(i), T0(i), mu1(i), k1(i))
    end do
    call put(v, 'co_k_mi_lam_Wilke', reshape([mu1, k1], [q, 2]))
    call transport_batch(P0(:, 1:ns), rq, T0, mu2, k2, sq)
    call put(v, 'transport_batch', reshape([mu2, k2, dble(sq)], [q, 3]))
  end if
  close(v)

  ! the caller owns the thermodynamic arrays (test-CEA.f90:46-50, test-wdot.f90:93-95)
  deallocate(wm_tab); deallocate(Ri_tab)
  deallocate(h_tab); deallocate(cp_tab); deallocate(dcpi_tab); deallocate(s_tab)
  deallocate(species_names); deallocate(elements_names)
  deallocate(species_composition); deallocate(el_weight)
  call free_chemistry_data()

contains

  subroutine put(unit, name, a)
    integer, intent(in) :: unit
    character(len=*), intent(in) :: name
    real(8), intent(in) :: a(:, :)
    integer :: r
    write(unit, '(A,1X,A,1X,I0,1X,I0)') '#', name, size(a, 1), size(a, 2)
    do r = 1, size(a, 1)
      write(unit, '(*(ES25.17E3,1X))') a(r, :)
    end do
  end subroutine

end program drop_in_check
