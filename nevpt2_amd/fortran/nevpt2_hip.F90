! ISO_C_BINDING interface to libnevpt2_amd.so (declarations in include/nevpt2_amd.h)
! plus the input.nevpt reader of the Fortran host.
!
! The interface block is what a Fortran caller of the reference binds instead of
! koopro4QD_driver/qdnevpt (koopro4QD.F90:25, qdnevpt.F:15): dimensions are
! int64 (the reference builds with -fdefault-integer-8) and arrays keep the
! reference's column-major layouts, state index first.
module nevpt2_hip
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: NV2_PREC_FP64 = 0, NV2_PREC_FP32 = 1
  integer, parameter :: NV2_NPHASE = 6

  interface
    function nv2_last_error(buf, len) bind(C, name="nv2_last_error") result(rc)
      import :: c_int, c_char, c_int64_t
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int64_t), value :: len
      integer(c_int) :: rc
    end function
    function nv2_ctx_create(ncore, nact, nvirt, metat, nele, nchovec) bind(C, name="nv2_ctx_create") result(ctx)
      import :: c_ptr, c_int64_t
      integer(c_int64_t), value :: ncore, nact, nvirt, metat, nele, nchovec
      type(c_ptr) :: ctx
    end function
    function nv2_ctx_set_integrals(ctx, cholesky, oneint) bind(C, name="nv2_ctx_set_integrals") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: cholesky(*), oneint(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_set_ro4(ctx, ro4) bind(C, name="nv2_ctx_set_ro4") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ro4(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_run(ctx, prec, thr_lindep, out) bind(C, name="nv2_ctx_run") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: prec
      real(c_double), value :: thr_lindep
      real(c_double), intent(out) :: out(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_set_option(ctx, name, value) bind(C, name="nv2_ctx_set_option") result(rc)
      import :: c_ptr, c_int, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t), value :: value
      integer(c_int) :: rc
    end function
    function nv2_ctx_set_frozen(ctx, nfrozen, orbitals) bind(C, name="nv2_ctx_set_frozen") result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nfrozen
      integer(c_int64_t), intent(in) :: orbitals(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_set_ci(ctx, nalpha, nbeta, ci) bind(C, name="nv2_ctx_set_ci") result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nalpha, nbeta
      real(c_double), intent(in) :: ci(*)
      integer(c_int) :: rc
    end function
    function nv2_ci_from_file04(norb, ncore, nact, nocc, metat, ncf, ne, nd, trou, part, c, ci, nalpha, nbeta) &
        bind(C, name="nv2_ci_from_file04") result(rc)
      import :: c_ptr, c_int, c_int16_t, c_int64_t, c_double
      integer(c_int64_t), value :: norb, ncore, nact, nocc, metat, ncf
      integer(c_int16_t), intent(in) :: ne(*), trou(*), part(*)
      integer(c_int64_t), intent(in) :: nd(*)
      real(c_double), intent(in) :: c(*)
      type(c_ptr), value :: ci
      integer(c_int64_t), intent(out) :: nalpha, nbeta
      integer(c_int) :: rc
    end function
    function nv2_ctx_timings(ctx, ms) bind(C, name="nv2_ctx_timings") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_set_trdm(ctx, trdm3, trdm2, trdm1) bind(C, name="nv2_ctx_set_trdm") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: trdm3(*), trdm2(*), trdm1(*)
      integer(c_int) :: rc
    end function
    function nv2_ctx_get(ctx, name, host, count) bind(C, name="nv2_ctx_get") result(rc)
      import :: c_ptr, c_int, c_char, c_double, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: host(*)
      integer(c_int64_t), value :: count
      integer(c_int) :: rc
    end function
    function nv2_qd_heff(metat, heff2, e0, heff, energies) bind(C, name="nv2_qd_heff") result(rc)
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: metat
      real(c_double), intent(in) :: heff2(*), e0(*)
      real(c_double), intent(out) :: heff(*), energies(*)
      integer(c_int) :: rc
    end function
    subroutine nv2_ctx_destroy(ctx) bind(C, name="nv2_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface

  ! ------------------------------------------------------------------ input.nevpt (nevpt2_cfg.F90 defaults)
  type nv2_config
    real(8) :: thr = 1.0d-10
    real(8) :: thr_lindep = 1.0d-6
    integer :: nr_states = -1
    integer :: nr_active_electrons = -1
    logical :: do_cholesky = .false.
    logical :: no_4rdm_terms = .false.
    logical :: no_pc = .false.
    logical :: zorder = .true.
    logical :: natorb = .false.                 ! .NATORB: first-order density and natural occupations
    integer :: prec = 0
    character(len=256) :: file04 = ' '          ! the reference's file04 (CI vectors, koopro4QD.F90:147-176)
    character(len=256) :: bundle = ' '          ! this host's integral bundle (nevpt2_amd/bundle.py)
    integer :: nr_frozen = 0
    integer(8), allocatable :: frozen(:)        ! .FROZEN orbital numbers (qdnevpt.F:733-745)
  end type

contains

  function upcase(s) result(u)
    character(*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i, c
    u = s
    do i = 1, len(s)
      c = ichar(s(i:i))
      if (c >= ichar('a') .and. c <= ichar('z')) u(i:i) = char(c - 32)
    end do
  end function

  ! Sections *MODEL, *KOOPRO, *QDNEVP; 7-character, case-insensitive keywords;
  ! a keyword's value sits on the next line; input ends at *END OF
  ! (parse_input.F90:1-67, input_reader_sections.F90:28-161).  Unknown keywords stop.
  subroutine read_input(fname, cfg)
    character(*), intent(in) :: fname
    type(nv2_config), intent(inout) :: cfg
    character(len=7) :: word, section
    character(len=256) :: line
    integer :: u, ios
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) stop 'ERROR: cannot open input.nevpt'
    section = '       '
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      word = upcase(line(1:7))
      if (word == '       ') cycle
      if (word(1:1) == '!' .or. word(1:1) == '#') cycle
      if (word(1:1) == '*') then
        section = word
        if (word == '*END OF' .or. word == '**END O') exit
        cycle
      end if
      select case (section)
      case ('*MODEL ')
        select case (word)
        case ('.CHOLES'); cfg%do_cholesky = .true.
        case ('.NO4RDM'); cfg%no_4rdm_terms = .true.
        case ('.DMRGPT')
        case ('.MIXED '); cfg%prec = 1
        case default; call bad_kw(word, section)
        end select
      case ('*KOOPRO')
        select case (word)
        case ('.THRESH'); read(u, *) cfg%thr
        case ('.SPIN  '); read(u, *)
        case ('.STATES'); read(u, *) cfg%nr_states
        case ('.NACTEL'); read(u, *) cfg%nr_active_electrons
        case ('.NCORE ', '.NACTOR'); read(u, *)
        case ('.FILE04'); read(u, '(a)') cfg%file04
        case ('.FILE50'); read(u, *)
        case ('.NO ZOR'); cfg%zorder = .false.
        case default; call bad_kw(word, section)
        end select
      case ('*QDNEVP')
        select case (word)
        case ('.LINDEP'); read(u, *) cfg%thr_lindep
        case ('.STATES'); read(u, *) cfg%nr_states
        case ('.FILE04'); read(u, '(a)') cfg%file04
        case ('.BUNDLE'); read(u, '(a)') cfg%bundle     ! integrals (+ RDMs without .FILE04): host input
        case ('.NO PC '); cfg%no_pc = .true.
        case ('.NATORB'); cfg%natorb = .true.
        case ('.FROZEN')                  ! count, then the orbital numbers (input_reader_sections.F90:110-114)
          read(u, *) cfg%nr_frozen
          if (allocated(cfg%frozen)) deallocate(cfg%frozen)
          allocate(cfg%frozen(max(cfg%nr_frozen, 1)))
          cfg%frozen = 0
          if (cfg%nr_frozen > 0) read(u, *) cfg%frozen(1:cfg%nr_frozen)
        case default; call bad_kw(word, section)
        end select
      case default
      end select
    end do
    close(u)
  end subroutine

  ! The file04 record (koopro4QD.F90:147-176): header, then the excitation lists, the CI coefficients
  ! and the state energies; the fields after E are not needed here.  NOCA/NOCB are equalised as the
  ! reference does (koopro4QD.F90:199-201).
  subroutine read_file04(fname, norb, nocc, metat, ncf, ne, nd, trou, part, c, e)
    character(*), intent(in) :: fname
    integer(8), intent(out) :: norb, nocc, metat, ncf
    integer(2), allocatable, intent(out) :: ne(:), trou(:), part(:)
    integer(8), allocatable, intent(out) :: nd(:)
    real(8), allocatable, intent(out) :: c(:, :), e(:)
    integer(8) :: noca, nocb, nexc, ncoef, j, k
    integer :: u, ios
    open(newunit=u, file=fname, form='unformatted', access='sequential', status='old', iostat=ios)
    if (ios /= 0) stop 'ERROR: cannot open file04'
    read(u) norb, noca, nocb, metat, ncf, nexc, ncoef
    rewind(u)
    allocate(ne(ncf), nd(ncf), trou(max(nexc, 1_8)), part(max(nexc, 1_8)), c(ncf, metat), e(metat))
    read(u) norb, noca, nocb, metat, ncf, nexc, ncoef, (ne(j), nd(j), j = 1, ncf), (trou(j), part(j), j = 1, nexc), &
        ((c(j, k), j = 1, ncf), k = 1, metat), (e(j), j = 1, metat)
    close(u)
    nocc = min(noca, nocb)
    write(6, '(a)') ' FILE04 informations:'
    write(6, '(a,i8)') ' NORB = ', norb
    write(6, '(a,i8)') ' NOCA = ', noca
    write(6, '(a,i8)') ' NOCB = ', nocb
    write(6, '(a,i8)') ' METAT= ', metat
    write(6, '(a,i8)') ' NCF  = ', ncf
  end subroutine

  subroutine bad_kw(word, section)
    character(*), intent(in) :: word, section
    write(*, '(a)') ' illegal keyword '//word//' in section '//section
    stop 'illegal keyword in input'
  end subroutine

  function last_error() result(msg)
    character(len=1024) :: msg
    character(kind=c_char) :: buf(1024)
    integer :: i, rc
    rc = nv2_last_error(buf, 1024_c_int64_t)
    msg = ' '
    do i = 1, 1024
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
  end function

end module nevpt2_hip
