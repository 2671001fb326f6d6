!
! Replacement for module Iterate (General_Routines/mod_iterate.f90): the
! mixing routines run on the GPU, the small linear solvers on the host, with
! the reference's names, arguments and history-shift semantics.
!
module Iterate
  use Global
  use adscf_hip
  implicit none

  ! mod_iterate.f90:13.  A driver whose Global declares mixing_dim itself
  ! (Main_Multiblock/mod_global.f90:21) is built with -DADSCF_GLOBAL_MIXING_DIM,
  ! which takes the history depth from Global instead of clashing with it.
#ifndef ADSCF_GLOBAL_MIXING_DIM
  integer, parameter :: mixing_dim = 2
#endif

contains

  ! mod_iterate.f90:27-38
  subroutine Mixing_Simple(field, dfield, fieldsize, lambda, iter)
    double precision, intent(in) :: lambda
    integer, intent(in)          :: fieldsize
    integer, intent(in)          :: iter
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: field
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: dfield
    call adscf_check(adscf_mixing_simple(adscf_context, field, dfield, int(fieldsize, c_long), &
                                         int(mixing_dim, c_int), lambda), 'Mixing_Simple')
  end subroutine Mixing_Simple

  ! mod_iterate.f90:46-84
  subroutine Mixing_Lambda(field, dfield, fieldsize, lammax, iter)
    double precision, intent(in) :: lammax
    integer, intent(in)          :: fieldsize
    integer, intent(in)          :: iter
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: field
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: dfield
    call adscf_check(adscf_mixing_lambda(adscf_context, field, dfield, int(fieldsize, c_long), &
                                         int(mixing_dim, c_int), lammax, int(iter, c_int)), 'Mixing_Lambda')
  end subroutine Mixing_Lambda

  ! mod_iterate.f90:92-150
  subroutine Mixing_Anderson(field, dfield, fieldsize, lam_mix, iter)
    double precision, intent(in) :: lam_mix
    integer, intent(in)          :: fieldsize
    integer, intent(in)          :: iter
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: field
    double precision, dimension(fieldsize,0:mixing_dim), intent(inout) :: dfield
    call adscf_check(adscf_mixing_anderson(adscf_context, field, dfield, int(fieldsize, c_long), &
                                           int(mixing_dim, c_int), lam_mix, int(iter, c_int)), 'Mixing_Anderson')
  end subroutine Mixing_Anderson

  ! mod_iterate.f90:160-195
  subroutine Solve_LinearSystem_Cholesky(A, B, X, d)
    integer, intent(in) :: d
    double precision, dimension(d,d), intent(in) :: A
    double precision, dimension(d),   intent(in) :: B
    double precision, dimension(d),  intent(out) :: X
    call adscf_check(adscf_solve_cholesky(A, B, X, int(d, c_int)), 'Solve_LinearSystem_Cholesky')
  end subroutine Solve_LinearSystem_Cholesky

  ! mod_iterate.f90:203-224
  subroutine Solve_LinearSystem_LU(A, B, X, n)
    integer, intent(in) :: n
    double precision, dimension(n,n), intent(in) :: A
    double precision, dimension(n),   intent(in) :: B
    double precision, dimension(n),  intent(out) :: X
    call adscf_check(adscf_solve_lu(A, B, X, int(n, c_int)), 'Solve_LinearSystem_LU')
  end subroutine Solve_LinearSystem_LU

  ! mod_iterate.f90:232-298
  subroutine ludcmp(a, n, indx, d)
    integer, intent(in)                            :: n
    integer, dimension(n), intent(out)             :: indx
    double precision, intent(out)                  :: d
    double precision, dimension(n,n), intent(inout) :: a
    integer(c_int) :: ix(n)
    call adscf_check(adscf_ludcmp(a, int(n, c_int), ix, d), 'ludcmp')
    indx = ix
  end subroutine ludcmp

  ! mod_iterate.f90:302-337
  subroutine lubksb(a, n, indx, b)
    integer, intent(in)                           :: n
    integer, dimension(n), intent(in)             :: indx
    double precision, dimension(n), intent(inout) :: b
    double precision, dimension(n,n), intent(in)  :: a
    integer(c_int) :: ix(n)
    ix = indx
    call adscf_check(adscf_lubksb(a, int(n, c_int), ix, b), 'lubksb')
  end subroutine lubksb

end module Iterate
