! Fortran host of the MI355X NEVPT2 hot path: reads input.nevpt, the integral
! bundle named by .BUNDLE (Cholesky vectors, one-electron integrals, reference
! energies and, without .FILE04, the packed 4-RDM and for several roots the
! transition RDMs), and with .FILE04 the reference's file04 CI vectors, whose
! RDMs are then built on the GPU; runs one device-resident evaluation and prints
! the reference's per-class table, the QD effective Hamiltonians, the summary
! (qdnevpt.F:1004-1230) and, with .NATORB, the first-order density matrices and
! natural occupation numbers (qdnevpt.F:1240-1467).
program nevpt2_amd_main
  use iso_c_binding
  use nevpt2_hip
  implicit none
  type(nv2_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int64_t) :: hdr(8), ncore, nact, nvirt, metat, nele, nchovec, norb, npair, nwords
  real(c_double), allocatable :: e0(:), oneint(:), chol(:), ro4(:), tab(:), ms(:)
  real(c_double), allocatable :: t3(:), t2(:), t1(:), hsc(:), hpc(:), hh(:), esc(:), epc(:)
  real(c_double), allocatable :: rho(:), occ(:)
  integer(c_int64_t) :: npr
  logical :: qd, from_ci
  character(len=8) :: magic
  integer(8) :: f_norb, f_nocc, f_metat, f_ncf
  integer(2), allocatable :: f_ne(:), f_trou(:), f_part(:)
  integer(8), allocatable :: f_nd(:)
  real(8), allocatable :: f_c(:, :), f_e(:)
  real(c_double), allocatable, target :: ci(:)
  integer(c_int64_t) :: nalpha, nbeta
  integer :: u, ios, s, k, i
  integer(c_int) :: rc
  real(8) :: sc_tot, pc_tot, nsc_tot, npc_tot
  character(len=6), parameter :: lab(8) = [character(len=6) :: '(0)  ', '(+1) ', '(-1) ', '(+2) ', '(-2) ', &
                                           "(+1)'", "(-1)'", "(0)' "]

  call read_input('input.nevpt', cfg)
  if (cfg%bundle == ' ') stop 'ERROR: .BUNDLE (integral bundle) is mandatory input'
  from_ci = cfg%file04 /= ' '
  open(newunit=u, file=trim(cfg%bundle), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) stop 'ERROR: cannot open the input bundle'
  read(u) magic
  if (magic /= 'NV2BUNDL') stop 'ERROR: not an nevpt2_amd input bundle'
  read(u) hdr
  ncore = hdr(2); nact = hdr(3); nvirt = hdr(4); metat = hdr(5); nele = hdr(6); nchovec = hdr(7)
  if (cfg%nr_states > 0 .and. cfg%nr_states /= metat) stop 'ERROR: .STATES does not match the input bundle'
  if (cfg%nr_active_electrons > 0 .and. cfg%nr_active_electrons /= nele) stop 'ERROR: .NACTEL does not match'
  norb = ncore + nact + nvirt
  npair = norb * (norb + 1) / 2
  nwords = nact * (nact + 1) * (nact + 2) * (nact + 3) / 24
  allocate(e0(metat), oneint(npair), chol(nchovec * npair))
  read(u) e0, oneint, chol
  if (.not. from_ci) then
    if (hdr(1) == 3) stop 'ERROR: the bundle holds integrals only: give the CI vectors with .FILE04'
    allocate(ro4(nwords * nact**4 * metat))
    read(u) ro4
  end if
  qd = metat > 1 .and. (hdr(1) == 2 .or. from_ci)
  if (metat > 1 .and. .not. qd) stop 'ERROR: several roots need the transition RDMs (bundle version 2)'
  if (qd .and. .not. from_ci) then
    npr = metat * (metat - 1) / 2
    allocate(t3(npr * nact**6), t2(npr * nact**4), t1(npr * nact**2))
    read(u) t3, t2, t1
  end if
  close(u)

  ctx = nv2_ctx_create(ncore, nact, nvirt, metat, nele, nchovec)
  if (.not. c_associated(ctx)) then
    write(*, '(a)') trim(last_error())
    stop 'ERROR: device context'
  end if
  rc = nv2_ctx_set_integrals(ctx, chol, oneint)
  if (from_ci) then
    ! file04 CI vectors -> string-ordered arrays -> RDMs (and transition RDMs) on the GPU; E0 from file04
    call read_file04(trim(cfg%file04), f_norb, f_nocc, f_metat, f_ncf, f_ne, f_nd, f_trou, f_part, f_c, f_e)
    if (f_norb /= norb .or. f_metat /= metat) stop 'ERROR: file04 does not match the integral bundle'
    if (rc == 0) rc = nv2_ci_from_file04(norb, ncore, nact, f_nocc, metat, f_ncf, f_ne, f_nd, f_trou, f_part, &
                                         f_c, c_null_ptr, nalpha, nbeta)
    if (rc == 0) then
      if (nalpha + nbeta /= nele) stop 'ERROR: file04 electron count does not match .NACTEL / the bundle'
      allocate(ci(metat * nstr(nact, nalpha) * nstr(nact, nbeta)))
      rc = nv2_ci_from_file04(norb, ncore, nact, f_nocc, metat, f_ncf, f_ne, f_nd, f_trou, f_part, f_c, &
                              c_loc(ci), nalpha, nbeta)
    end if
    if (rc == 0) rc = nv2_ctx_set_ci(ctx, nalpha, nbeta, ci)
    e0 = f_e
  else
    if (rc == 0) rc = nv2_ctx_set_ro4(ctx, ro4)
    if (rc == 0 .and. qd) rc = nv2_ctx_set_trdm(ctx, t3, t2, t1)
  end if
  ! .NO4RDM: no A-D matrices and no (+1)'/(-1)' classes (koopro4QD.F90:357, qdnevpt.F:949-991)
  if (cfg%no_4rdm_terms) then
    write(6, '(a)') ' WARNING: .NO4RDM has been specified. V(+1p) and'
    write(6, '(a)') ' V(-1p) terms will not be calculated.'
  end if
  if (cfg%prec /= 0) then
    write(6, '(a)') ' WARNING: .MIXED runs the 4-RDM contractions in FP32. On classes with near-zero'
    write(6, '(a)') ' denominators (large class norms) the energies can differ from the FP64 path by'
    write(6, '(a)') ' far more than 1 microhartree; compare with a run without .MIXED.'
  end if
  if (rc == 0 .and. cfg%no_4rdm_terms) rc = nv2_ctx_set_option(ctx, 'no4rdm' // c_null_char, 1_c_int64_t)
  if (rc == 0 .and. cfg%natorb) rc = nv2_ctx_set_option(ctx, 'natorb' // c_null_char, 1_c_int64_t)
  if (rc == 0 .and. cfg%nr_frozen > 0) then
    write(6, '(/a/(5x,18i4))') '     Frozen orbitals:', cfg%frozen(1:cfg%nr_frozen)
    rc = nv2_ctx_set_frozen(ctx, int(cfg%nr_frozen, c_int64_t), cfg%frozen)
  end if
  allocate(tab(32 * metat), ms(NV2_NPHASE))
  if (rc == 0) rc = nv2_ctx_run(ctx, int(cfg%prec, c_int32_t), real(cfg%thr_lindep, c_double), tab)
  if (rc /= 0) then
    write(*, '(a)') trim(last_error())
    call nv2_ctx_destroy(ctx)
    stop 'ERROR: device evaluation failed'
  end if
  rc = nv2_ctx_timings(ctx, ms)
  allocate(hsc(metat * metat), hpc(metat * metat), hh(metat * metat), esc(metat), epc(metat))
  rc = nv2_ctx_get(ctx, 'Heff-SC' // c_null_char, hsc, metat * metat)
  if (rc == 0) rc = nv2_ctx_get(ctx, 'Heff-PC' // c_null_char, hpc, metat * metat)
  if (rc == 0 .and. cfg%natorb) then
    allocate(rho(norb * norb * metat), occ(norb * metat))
    rc = nv2_ctx_get(ctx, 'rho1' // c_null_char, rho, norb * norb * metat)
    if (rc == 0) rc = nv2_ctx_get(ctx, 'natocc' // c_null_char, occ, norb * metat)
  end if
  call nv2_ctx_destroy(ctx)
  if (rc /= 0) stop 'ERROR: effective Hamiltonian'

  if (qd) then
    write(6, '(//a//)') ' QD-NEVPT2: FINAL RESULTS'
  else
    write(6, '(//a//)') ' state-specific NEVPT2: FINAL RESULTS'
  end if
  write(6, '(a)') ' Second order contribution of each class to the energies'
  if (qd) write(6, '(a)') ' (to the diagonal elements of Heff)'
  do s = 1, int(metat)
    write(6, '(/a,i4)') ' state number:', s
    write(6, '(a)') ' -----------------'
    write(6, '(a)') ' class        norm(sc)           energy(sc)           norm(pc)           energy(pc)'
    nsc_tot = 0; sc_tot = 0; npc_tot = 0; pc_tot = 0
    do k = 1, 8
      i = 32 * (s - 1) + 4 * (k - 1)
      write(6, '(1x,a,4f20.15)') lab(k)(1:5), tab(i + 1:i + 4)
      nsc_tot = nsc_tot + tab(i + 1); sc_tot = sc_tot + tab(i + 2)
      npc_tot = npc_tot + tab(i + 3); pc_tot = pc_tot + tab(i + 4)
    end do
    write(6, '(1x,85a1)') ('-', i = 1, 85)
    write(6, '(1x,a,4f20.15)') 'Total', nsc_tot, sc_tot, npc_tot, pc_tot
  end do
  ! energies: diagonal sums (state-specific) or eigenvalues of the hermitised Heff (QD)
  do s = 1, int(metat)
    esc(s) = e0(s) + hsc(s + metat * (s - 1))
    epc(s) = e0(s) + hpc(s + metat * (s - 1))
  end do
  if (qd) then
    call print_heff('SC', hsc)
    rc = nv2_qd_heff(metat, hsc, e0, hh, esc)
    if (rc /= 0) stop 'ERROR: SC effective Hamiltonian'
    call print_heff_h('SC', hh)
    if (.not. cfg%no_pc) then
      call print_heff('PC', hpc)
      rc = nv2_qd_heff(metat, hpc, e0, hh, epc)
      if (rc /= 0) stop 'ERROR: PC effective Hamiltonian'
      call print_heff_h('PC', hh)
    end if
  end if
  write(6, '(/)')
  if (qd) then
    write(6, '(a)') '                 Summary of the QD-NEVPT2 energies'
  else
    write(6, '(a)') '                   Summary of the NEVPT2 energies'
  end if
  write(6, '(1x,72("-"))')
  write(6, '(a/)') ' state   diagonalization     strongly-contracted     partially-contracted'
  do s = 1, int(metat)
    if (cfg%no_pc) then
      write(6, '(a2,i4,2f19.7,9x,a)') ' @', s, e0(s), esc(s), ' not requested'
    else
      write(6, '(a2,i4,f18.7,6x,f18.7,7x,f18.7)') ' @', s, e0(s), esc(s), epc(s)
    end if
  end do
  write(6, '(1x,72("-"))')
  if (cfg%natorb) call print_natorb()
  write(6, '(a,6f10.3)') ' device phase times (ms): ', ms

contains

  ! .NATORB output (qdnevpt.F:1297-1460): each state's density matrix to first order, then per state its
  ! eigenvalues as dsyev returns them and their sums over the core, active and virtual counts after the
  ! descending-modulus order of order_eigvalvec (one irrep: the reference's C1 path)
  subroutine print_natorb()
    real(c_double) :: w(norb), t, tot_loc, tot_sum, sum_c, sum_a, sum_v
    integer :: s, i, j, m
    do s = 1, int(metat)
      write(6, '(a,i12)') ' Density matrix 0   for state ', s
      call matout(rho(1 + norb * norb * (s - 1):norb * norb * s), norb)
    end do
    do s = 1, int(metat)
      tot_sum = 0
      write(6, '(/a,i3)') ' Natural orbital occupation numbers for state ', s
      write(6, '(a)') ' !NOTE! Approximate numbers --> no i-i/v-v terms!'
      write(6, '(a/)') ' ------------------------------------------------'
      write(6, '(a,i3,a/)') ' irrep ', 1, ': orbital occupations'
      w = occ(1 + norb * (s - 1):norb * s)
      write(6, '(6f12.8)') w
      do i = 1, int(norb) - 1          ! order_eigvalvec: the largest modulus first, later ones win ties
        m = i
        t = w(i)
        do j = i + 1, int(norb)
          if (abs(w(j)) >= abs(t)) then
            t = w(j)
            m = j
          end if
        end do
        if (m /= i) then
          w(m) = w(i)
          w(i) = t
        end if
      end do
      sum_c = sum(w(1:ncore))
      sum_a = sum(w(ncore + 1:ncore + nact))
      sum_v = sum(w(ncore + nact + 1:norb))
      tot_loc = sum(w)
      tot_sum = tot_sum + tot_loc
      write(6, '(/a)') ' ------------------------------------------------'
      write(6, '(a,f15.6/)') ' Total sum of occupation numbers =', tot_loc
      write(6, '(a,f15.6)') ' = core    occupation numbers    :', sum_c
      write(6, '(a,f15.6)') ' + active  occupation numbers    :', sum_a
      write(6, '(a,f15.6/)') ' + virtual occupation numbers    :', sum_v
      write(6, '(a,f15.6/)') ' Final sum of occupation numbers =', tot_sum
    end do
  end subroutine

  ! number of strings of k electrons in n orbitals
  integer(c_int64_t) function nstr(n, k)
    integer(c_int64_t), intent(in) :: n, k
    integer(c_int64_t) :: i
    nstr = 1
    do i = 1, k
      nstr = nstr * (n - k + i) / i
    end do
  end function

  ! matout layout (qdnevpt.F:1639-1659): five columns per block
  subroutine matout(a, m)
    real(c_double), intent(in) :: a(:)
    integer(c_int64_t), intent(in) :: m
    integer :: i0, i1, r, c
    do i0 = 1, int(m), 5
      i1 = min(i0 + 4, int(m))
      write(6, '(7x,5(8x,i2,7x))') (c, c = i0, i1)
      do r = 1, int(m)
        write(6, '(1x,i2,4x,5(f17.10))') r, (a(r + m * (c - 1)), c = i0, i1)
      end do
    end do
  end subroutine

  subroutine print_heff(key, h)
    character(len=2), intent(in) :: key
    real(c_double), intent(in) :: h(:)
    real(c_double), allocatable :: z(:)
    integer :: s
    write(6, '(/a)') ' *******************************************'
    write(6, '(a)') ' * Second order Effective Hamiltonian (' // key // ') *'
    write(6, '(a/)') ' *******************************************'
    call matout(h, metat)
    z = h
    do s = 1, int(metat)
      z(s + metat * (s - 1)) = z(s + metat * (s - 1)) + e0(s)
    end do
    write(6, '(a)') ' **************************************************'
    write(6, '(a)') ' * Zero + second order Effective Hamiltonian (' // key // ') *'
    write(6, '(a/)') ' **************************************************'
    call matout(z, metat)
  end subroutine

  subroutine print_heff_h(key, h)
    character(len=2), intent(in) :: key
    real(c_double), intent(in) :: h(:)
    write(6, '(a)') '  Heff~ (Hermitian) --> ' // key
    call matout(h, metat)
  end subroutine
end program nevpt2_amd_main
