!
! Replacement for module Propagate (General_Routines/mod_propagate.f90).
!
module Propagate
  use Global
  implicit none

contains

  ! mod_propagate.f90:22-32
  subroutine Calculate_Expfield(field, expfield)
    use adscf_hip
    double precision, dimension(gridx, gridy, gridz, monomer_types), intent(in)  :: field
    double precision, dimension(gridx, gridy, gridz, monomer_types), intent(out) :: expfield
    call adscf_check(adscf_expfield(adscf_context, field, expfield), 'Calculate_Expfield')
  end subroutine Calculate_Expfield

  ! mod_propagate.f90:40-68
  subroutine Propagate_Step(prop, expfield)
    use adscf_hip
    double precision, dimension(gridx, gridy, gridz), intent(in)    :: expfield
    double precision, dimension(gridx, gridy, gridz), intent(inout) :: prop
    call adscf_check(adscf_propagate_step(adscf_context, prop, expfield), 'Propagate_Step')
  end subroutine Propagate_Step

end module Propagate
