! C-callable wrappers (own code) around the reference's accessor modules, which
! the recipe in this folder compiles unchanged from the reference tree
! (eexx_functions.F90, indices_norm.F90, ord_utils.F90, swap_util.F90).
! Test infrastructure only: used to pin oracle/eexx and oracle/c against the
! reference's own arithmetic.  Indices are 1-based, as in the reference.
module ref_eexx_capi
  use iso_c_binding
  use indices_norm
  use eexx_functions
  implicit none
contains
  subroutine ref_init(nact) bind(C, name="ref_init")
    integer(c_int64_t), value :: nact
    integer :: n
    n = int(nact)
    if (allocated(ind_norm%lindi)) call finalize_indices_norm()
    call initialize_indices_norm(n)
  end subroutine

  function ref_nwords() bind(C, name="ref_nwords") result(r)
    integer(c_int64_t) :: r
    r = nwords
  end function

  function ref_eeee(ix, d3, d2, d1, ro4, nact, istate, metat) bind(C, name="ref_eeee") result(r)
    integer(c_int64_t), intent(in) :: ix(8)
    integer(c_int64_t), value :: nact, istate, metat
    real(c_double), intent(in) :: d3(*), d2(*), d1(*), ro4(*)
    real(c_double) :: r
    r = eeee(int(ix(1)),int(ix(2)),int(ix(3)),int(ix(4)),int(ix(5)),int(ix(6)),int(ix(7)),int(ix(8)), &
             d3,d2,d1,ro4,int(nact),int(istate),int(metat))
  end function

  function ref_eeeet(ix, d3, d2, d1, ro4, nact, istate, metat) bind(C, name="ref_eeeet") result(r)
    integer(c_int64_t), intent(in) :: ix(8)
    integer(c_int64_t), value :: nact, istate, metat
    real(c_double), intent(in) :: d3(*), d2(*), d1(*), ro4(*)
    real(c_double) :: r
    r = eeeet(int(ix(1)),int(ix(2)),int(ix(3)),int(ix(4)),int(ix(5)),int(ix(6)),int(ix(7)),int(ix(8)), &
              d3,d2,d1,ro4,int(nact),int(istate),int(metat))
  end function

  function ref_eee2(which, ix, d3, d2, d1, nact, istate, metat) bind(C, name="ref_eee2") result(r)
    integer(c_int64_t), value :: which, nact, istate, metat
    integer(c_int64_t), intent(in) :: ix(6)
    real(c_double), intent(in) :: d3(*), d2(*), d1(*)
    real(c_double) :: r
    integer :: a,b,c,d,e,f
    a=int(ix(1)); b=int(ix(2)); c=int(ix(3)); d=int(ix(4)); e=int(ix(5)); f=int(ix(6))
    select case (which)
    case (0); r = eee2(a,b,c,d,e,f,d3,d2,d1,int(nact),int(istate),int(metat))
    case (1); r = eee2t(a,b,c,d,e,f,d3,d2,d1,int(nact),int(istate),int(metat))
    case (2); r = eee2tl(a,b,c,d,e,f,d3,d2,d1,int(nact),int(istate),int(metat))
    case (3); r = ro3t(a,b,c,d,e,f,d3,d2,d1,int(nact),int(istate),int(metat))
    case default; r = 0.0d0
    end select
  end function

  function ref_ee2(which, ix, d2, d1, nact, istate, metat) bind(C, name="ref_ee2") result(r)
    integer(c_int64_t), value :: which, nact, istate, metat
    integer(c_int64_t), intent(in) :: ix(4)
    real(c_double), intent(in) :: d2(*), d1(*)
    real(c_double) :: r
    integer :: a,b,c,d
    a=int(ix(1)); b=int(ix(2)); c=int(ix(3)); d=int(ix(4))
    select case (which)
    case (0); r = ee2(a,b,c,d,d2,d1,int(nact),int(istate),int(metat))
    case (1); r = ee2t(a,b,c,d,d2,d1,int(nact),int(istate),int(metat))
    case (2); r = ee2tr(a,b,c,d,d2,d1,int(nact),int(istate),int(metat))
    case (3); r = ro2t(a,b,c,d,d2,d1,int(nact),int(istate),int(metat))
    case default; r = 0.0d0
    end select
  end function
end module ref_eexx_capi
