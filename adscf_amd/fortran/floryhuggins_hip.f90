!
! Replacement for module Interactions_FloryHuggins
! (Interaction_Modules/mod_FloryHuggins.f90).  The parameters are read from
! input_interactions exactly as the reference reads them and handed to the
! library; fields, energy and the volume-fraction check run on the GPU.
!
module Interactions_FloryHuggins
  use Global
  use adscf_hip
  implicit none
  private

  integer, parameter     :: components = monomer_types

  double precision, save :: kappa
  double precision, save :: chi(components,components)
  double precision, save :: monomer_volume(components)

  public :: Initialize_Interactions
  public :: Calculate_Mean_Fields
  public :: Interaction_Energy
  public :: Check_Volume_Fraction

contains

  ! mod_FloryHuggins.f90:37-66
  subroutine Initialize_Interactions
    character*30 :: file_name
    integer      :: io_error
    integer      :: component, component2
    file_name = 'input_interactions'
    open(unit=20, file=file_name, status='old', action='read', iostat=io_error)
    if (io_error == 0) then
      read(20,*) kappa
      read(20,*)
      do component = 1, components
        chi(component,component) = 0.0d0
        do component2 = component+1, components
          read(20,*) chi(component,component2)
          chi(component2,component) = chi(component,component2)
        end do
      end do
      read(20,*)
      do component = 1, components
        read(20,*) monomer_volume(component)
      end do
    else
      print*, 'Error', io_error, ' while trying to open', file_name
    end if
    close(unit=20)
    call adscf_check(adscf_set_interactions(adscf_context, kappa, chi, monomer_volume), 'Initialize_Interactions')
  end subroutine Initialize_Interactions

  ! mod_FloryHuggins.f90:74-96
  double precision function Interaction_Energy(density)
    double precision, dimension(gridx,gridy,gridz,components), intent(in) :: density
    double precision :: e
    call adscf_check(adscf_interaction_energy(adscf_context, density, e), 'Interaction_Energy')
    Interaction_Energy = e
  end function Interaction_Energy

  ! mod_FloryHuggins.f90:104-129
  subroutine Calculate_Mean_Fields(density, field)
    double precision, dimension(gridx,gridy,gridz,components), intent(in)  :: density
    double precision, dimension(gridx,gridy,gridz,components), intent(out) :: field
    call adscf_check(adscf_mean_fields(adscf_context, density, field), 'Calculate_Mean_Fields')
  end subroutine Calculate_Mean_Fields

  ! mod_FloryHuggins.f90:139-160
  subroutine Check_Volume_Fraction(density)
    double precision, dimension(gridx,gridy,gridz,components), intent(in) :: density
    double precision :: average_volume_fraction
    call adscf_check(adscf_average_volume_fraction(adscf_context, density, average_volume_fraction), &
                     'Check_Volume_Fraction')
    write(30,*) 'Average total volume fraction is', average_volume_fraction
    print*, 'Average total volume fraction is', average_volume_fraction
    if (abs(average_volume_fraction - 1.0d0) .gt. 1.0d-3) then
      write(30,*) 'Caution: Differs noticeably from 1'
      print*, 'Caution: Differs noticeably from 1'
    end if
  end subroutine Check_Volume_Fraction

end module Interactions_FloryHuggins
