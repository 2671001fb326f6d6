!
! bind(C) entry points over the reference routines that build without FFTW
! (General_Routines/mod_mt.f90, General_Routines/mod_iterate.f90,
! Interaction_Modules/mod_FloryHuggins.f90 with Main_Diblock_Melt/mod_global.f90).
! TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libref.so by build_ref.sh
! and called by the oracle tests to pin the C restatement.
!
module ref_shim
  use iso_c_binding
  use Global
  use mtmod
  use Iterate
  use Interactions_FloryHuggins
  implicit none
contains
  subroutine ref_sgrnd(seed) bind(C, name='ref_sgrnd')
    integer(c_int), value :: seed
    call sgrnd(int(seed))
  end subroutine
  real(c_double) function ref_grnd() bind(C, name='ref_grnd')
    ref_grnd = grnd()
  end function
  integer(c_int) function ref_mixing_dim() bind(C, name='ref_mixing_dim')
    ref_mixing_dim = mixing_dim
  end function
  subroutine ref_grid(n) bind(C, name='ref_grid')
    integer(c_int), intent(out) :: n(4)
    n = [gridx, gridy, gridz, monomer_types]
  end subroutine
  subroutine ref_ludcmp(a, n, indx, d) bind(C, name='ref_ludcmp')
    integer(c_int), value :: n
    real(c_double), intent(inout) :: a(n, n)
    integer(c_int), intent(out) :: indx(n)
    real(c_double), intent(out) :: d
    integer :: ix(n)
    call ludcmp(a, int(n), ix, d)
    indx = ix
  end subroutine
  subroutine ref_lubksb(a, n, indx, b) bind(C, name='ref_lubksb')
    integer(c_int), value :: n
    real(c_double), intent(in) :: a(n, n)
    integer(c_int), intent(in) :: indx(n)
    real(c_double), intent(inout) :: b(n)
    integer :: ix(n)
    ix = indx
    call lubksb(a, int(n), ix, b)
  end subroutine
  subroutine ref_mixing_anderson(field, dfield, fieldsize, lam, iter) bind(C, name='ref_mixing_anderson')
    integer(c_int), value :: fieldsize, iter
    real(c_double), value :: lam
    real(c_double), intent(inout) :: field(fieldsize, 0:mixing_dim), dfield(fieldsize, 0:mixing_dim)
    call Mixing_Anderson(field, dfield, int(fieldsize), lam, int(iter))
  end subroutine
  subroutine ref_mixing_lambda(field, dfield, fieldsize, lam, iter) bind(C, name='ref_mixing_lambda')
    integer(c_int), value :: fieldsize, iter
    real(c_double), value :: lam
    real(c_double), intent(inout) :: field(fieldsize, 0:mixing_dim), dfield(fieldsize, 0:mixing_dim)
    call Mixing_Lambda(field, dfield, int(fieldsize), lam, int(iter))
  end subroutine
  subroutine ref_mixing_simple(field, dfield, fieldsize, lam, iter) bind(C, name='ref_mixing_simple')
    integer(c_int), value :: fieldsize, iter
    real(c_double), value :: lam
    real(c_double), intent(inout) :: field(fieldsize, 0:mixing_dim), dfield(fieldsize, 0:mixing_dim)
    call Mixing_Simple(field, dfield, int(fieldsize), lam, int(iter))
  end subroutine
  subroutine ref_set_box(lx, ly, lz) bind(C, name='ref_set_box')
    real(c_double), value :: lx, ly, lz
    sizex = lx
    sizey = ly
    sizez = lz
    volume = sizex*sizey*sizez
    dvol = volume/dble(gridx*gridy*gridz)
  end subroutine
  subroutine ref_initialize_interactions() bind(C, name='ref_initialize_interactions')
    call Initialize_Interactions
  end subroutine
  subroutine ref_mean_fields(density, field) bind(C, name='ref_mean_fields')
    real(c_double), intent(in) :: density(gridx, gridy, gridz, monomer_types)
    real(c_double), intent(out) :: field(gridx, gridy, gridz, monomer_types)
    call Calculate_Mean_Fields(density, field)
  end subroutine
  real(c_double) function ref_interaction_energy(density) bind(C, name='ref_interaction_energy')
    real(c_double), intent(in) :: density(gridx, gridy, gridz, monomer_types)
    ref_interaction_energy = Interaction_Energy(density)
  end function
end module ref_shim
