!
! Replacements for Polymer_Modules/mod_homopolymer.f90 and mod_solvent.f90:
! identical derived types and Initialize_* procedures; Propagate_Homopolymer runs
! the propagator sweep, the contour integral and the ensemble normalisation on
! the GPU (adscf_propagate_homopolymer), Update_Solvent its pointwise update
! (adscf_update_solvent).
!
module Molecule_Homopolymer
  use Global
  use Propagate
  use adscf_hip
  implicit none
  private

  type homopolymer
    integer          :: label
    integer          :: monomer_type
    double precision :: length
    integer          :: ensemble       ! canonical:1, grand canonical:2
    double precision :: rho            ! total polymer concentration
    double precision :: mu             ! chemical potential of polymer
    double precision :: energy         ! free energy in SCF field
    integer          :: segments
    double precision :: Qsum
    double precision, dimension(gridx,gridy,gridz)    :: distribution
    double precision, dimension(gridx,gridy,gridz)    :: density
  end type homopolymer

  public :: homopolymer
  public :: Initialize_Homopolymer
  public :: Propagate_Homopolymer

contains

  ! mod_homopolymer.f90:41-80
  subroutine Initialize_Homopolymer(polymer, label, monomer_type, length, ensemble, rho, mu)
    integer, intent(in)           :: label, ensemble, monomer_type
    double precision, intent(in)  :: length, rho, mu
    type(homopolymer), intent(out):: polymer
    polymer%label        = label
    polymer%monomer_type = monomer_type
    polymer%length       = length
    polymer%segments     = sum(adscf_contour_segments([length], ds))
    polymer%ensemble     = ensemble
    polymer%rho          = rho
    polymer%mu           = mu
    call adscf_announce('Homopolymer', label, ensemble)
    polymer%density = adscf_start_density(rho * polymer%segments * ds, volume, dvol, gridx*gridy*gridz)
  end subroutine Initialize_Homopolymer

  ! mod_homopolymer.f90:90-141
  subroutine Propagate_Homopolymer(polymer, expfield)
    type(homopolymer), intent(inout) :: polymer
    double precision, dimension(gridx,gridy,gridz,monomer_types), intent(in) :: expfield
    integer(c_int) :: rc
    rc = adscf_propagate_homopolymer(adscf_context, int(polymer%monomer_type, c_int), int(polymer%segments, c_int), &
                                     int(polymer%ensemble, c_int), polymer%rho, polymer%mu, expfield, &
                                     polymer%distribution, polymer%density, polymer%Qsum, polymer%energy)
    if (rc == -3) then
      print*, 'Invalid ensemble; Cannot calculate Density'
      stop
    end if
    call adscf_check(rc, 'Propagate_Homopolymer')
  end subroutine Propagate_Homopolymer

end module Molecule_Homopolymer

module Molecule_Solvent
  use Global
  use Propagate
  use adscf_hip
  implicit none
  private

  type monomeric_solvent
    integer          :: label
    integer          :: monomer_type
    integer          :: ensemble       ! canonical:1, grand canonical:2
    double precision :: rho            ! total solvent concentration
    double precision :: mu             ! chemical potential of polymer
    double precision :: energy         ! free energy in SCF field
    double precision :: Qsum
    double precision, dimension(gridx,gridy,gridz)    :: distribution
    double precision, dimension(gridx,gridy,gridz)    :: density
  end type monomeric_solvent

  public :: monomeric_solvent
  public :: Initialize_Solvent
  public :: Update_Solvent

contains

  ! mod_solvent.f90:41-76
  subroutine Initialize_Solvent(solvent, label, monomer_type, ensemble, rho, mu)
    integer, intent(in)          :: label, ensemble, monomer_type
    double precision, intent(in) :: rho, mu
    type(monomeric_solvent), intent(out):: solvent
    solvent%label        = label
    solvent%monomer_type = monomer_type
    solvent%ensemble     = ensemble
    solvent%rho          = rho
    solvent%mu           = mu
    call adscf_announce('Solvent', label, ensemble)
    solvent%density = adscf_start_density(rho, volume, dvol, gridx*gridy*gridz)
  end subroutine Initialize_Solvent

  ! mod_solvent.f90:85-108
  subroutine Update_Solvent(solvent, field)
    type(monomeric_solvent), intent(inout)  :: solvent
    double precision, dimension(gridx,gridy,gridz), intent(in)  :: field
    integer(c_int) :: rc
    rc = adscf_update_solvent(adscf_context, int(solvent%ensemble, c_int), solvent%rho, solvent%mu, field, &
                              solvent%distribution, solvent%density, solvent%Qsum, solvent%energy)
    if (rc == -3) then
      print*, 'Invalid ensemble; Cannot calculate Density'
      stop
    end if
    call adscf_check(rc, 'Update_Solvent')
  end subroutine Update_Solvent

end module Molecule_Solvent
