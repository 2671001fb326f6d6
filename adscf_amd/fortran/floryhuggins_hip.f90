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

  ! mod_FloryHuggins.f90:37-66.  input_interactions holds, one value per record: kappa, a
  ! separator record, chi(i,j) for i < j in row order (i outer), a separator record, and the
  ! monomer volumes.  The matrix is stored symmetric with a zero diagonal.
  subroutine Initialize_Interactions
    integer :: u, ios, i, j
    kappa = 0.0d0
    chi = 0.0d0
    monomer_volume = 0.0d0
    open(newunit=u, file='input_interactions', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      print*, 'Error', ios, ' while trying to open input_interactions'
    else
      read(u,*) kappa
      read(u,*)
      do i = 1, components - 1
        do j = i + 1, components
          read(u,*) chi(i,j)
        end do
      end do
      chi = chi + transpose(chi)
      read(u,*)
      do i = 1, components
        read(u,*) monomer_volume(i)
      end do
      close(u)
    end if
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
