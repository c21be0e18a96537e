! Fortran host of the MI355X NEVPT2 hot path: reads input.nevpt, the input
! bundle named by .FILE04 (packed 4-RDM, Cholesky vectors, one-electron
! integrals, reference energies), runs one device-resident evaluation and prints
! the reference's per-class table and summary (qdnevpt.F:1004-1230).
program nevpt2_amd_main
  use iso_c_binding
  use nevpt2_hip
  implicit none
  type(nv2_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int64_t) :: hdr(8), ncore, nact, nvirt, metat, nele, nchovec, norb, npair, nwords
  real(c_double), allocatable :: e0(:), oneint(:), chol(:), ro4(:), tab(:), ms(:)
  character(len=8) :: magic
  integer :: u, ios, s, k, i
  integer(c_int) :: rc
  real(8) :: sc_tot, pc_tot, nsc_tot, npc_tot
  character(len=6), parameter :: lab(8) = [character(len=6) :: '(0)  ', '(+1) ', '(-1) ', '(+2) ', '(-2) ', &
                                           "(+1)'", "(-1)'", "(0)' "]

  call read_input('input.nevpt', cfg)
  if (cfg%file04 == ' ') stop 'ERROR: .FILE04 (input bundle) is mandatory input'
  open(newunit=u, file=trim(cfg%file04), access='stream', form='unformatted', status='old', iostat=ios)
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
  allocate(e0(metat), oneint(npair), chol(nchovec * npair), ro4(nwords * nact**4 * metat))
  read(u) e0, oneint, chol, ro4
  close(u)

  ctx = nv2_ctx_create(ncore, nact, nvirt, metat, nele, nchovec)
  if (.not. c_associated(ctx)) then
    write(*, '(a)') trim(last_error())
    stop 'ERROR: device context'
  end if
  rc = nv2_ctx_set_integrals(ctx, chol, oneint)
  if (rc == 0) rc = nv2_ctx_set_ro4(ctx, ro4)
  allocate(tab(32 * metat), ms(NV2_NPHASE))
  if (rc == 0) rc = nv2_ctx_run(ctx, int(cfg%prec, c_int32_t), real(cfg%thr_lindep, c_double), tab)
  if (rc /= 0) then
    write(*, '(a)') trim(last_error())
    call nv2_ctx_destroy(ctx)
    stop 'ERROR: device evaluation failed'
  end if
  rc = nv2_ctx_timings(ctx, ms)
  call nv2_ctx_destroy(ctx)

  write(6, '(//a//)') ' state-specific NEVPT2: FINAL RESULTS'
  write(6, '(a)') ' Second order contribution of each class to the energies'
  do s = 1, int(metat)
    write(6, '(/a,i4)') ' state number:', s
    write(6, '(a)') ' -----------------'
    write(6, '(a)') ' class        norm(sc)           energy(sc)           norm(pc)           energy(pc)'
    nsc_tot = 0; sc_tot = 0; npc_tot = 0; pc_tot = 0
    do k = 1, 8
      i = 32 * (s - 1) + 4 * (k - 1)
      if (cfg%no_4rdm_terms .and. (k == 6 .or. k == 7)) tab(i + 1:i + 4) = 0
      write(6, '(1x,a,4f20.15)') lab(k)(1:5), tab(i + 1:i + 4)
      nsc_tot = nsc_tot + tab(i + 1); sc_tot = sc_tot + tab(i + 2)
      npc_tot = npc_tot + tab(i + 3); pc_tot = pc_tot + tab(i + 4)
    end do
    write(6, '(1x,85a1)') ('-', i = 1, 85)
    write(6, '(1x,a,4f20.15)') 'Total', nsc_tot, sc_tot, npc_tot, pc_tot
  end do
  write(6, '(/)')
  write(6, '(a)') '                   Summary of the NEVPT2 energies'
  write(6, '(1x,72("-"))')
  write(6, '(a/)') ' state   diagonalization     strongly-contracted     partially-contracted'
  do s = 1, int(metat)
    sc_tot = e0(s); pc_tot = e0(s)
    do k = 1, 8
      i = 32 * (s - 1) + 4 * (k - 1)
      sc_tot = sc_tot + tab(i + 2); pc_tot = pc_tot + tab(i + 4)
    end do
    if (cfg%no_pc) then
      write(6, '(a2,i4,2f19.7,9x,a)') ' @', s, e0(s), sc_tot, ' not requested'
    else
      write(6, '(a2,i4,f18.7,6x,f18.7,7x,f18.7)') ' @', s, e0(s), sc_tot, pc_tot
    end if
  end do
  write(6, '(1x,72("-"))')
  write(6, '(a,6f10.3)') ' device phase times (ms): ', ms
end program nevpt2_amd_main
