!
! Replacements for Polymer_Modules/mod_diblock.f90, mod_triblock.f90 and
! mod_multiblock.f90: identical derived types and Initialize_* procedures;
! Propagate_* runs the forward and complementary sweeps, the contour integral
! and the ensemble normalisation on the GPU (adscf_propagate_chain).
!
module Molecule_Diblock_Copolymer
  use Global
  use Propagate
  use adscf_hip
  implicit none
  private

  integer, parameter  :: block_number = 2

  type diblock_copolymer
    integer                                   :: label
    integer, dimension(block_number)          :: monomer_type
    double precision, dimension(block_number) :: block_length
    integer           :: ensemble         ! canonical:1, grand canonical:2
    double precision  :: rho              ! total polymer concentration
    double precision  :: mu               ! chemical potential of polymer
    double precision  :: energy           ! free energy in SCF field
    integer, dimension(block_number)          :: block_segments
    integer                                   :: segments
    double precision                          :: Qsum
    double precision, dimension(gridx,gridy,gridz,block_number) :: distribution
    double precision, dimension(gridx,gridy,gridz,block_number) :: density
  end type diblock_copolymer

  public :: diblock_copolymer
  public :: Initialize_Diblock
  public :: Propagate_Diblock

contains

  ! mod_diblock.f90:46-88
  subroutine Initialize_Diblock(copolymer, label, monomer_type, block_length, ensemble, rho, mu)
    integer, intent(in)                                   :: label, ensemble
    double precision, intent(in)                          :: rho, mu
    double precision, dimension(block_number), intent(in) :: block_length
    integer, dimension(block_number), intent(in)          :: monomer_type
    type(diblock_copolymer), intent(out)                 :: copolymer
    integer :: b
    copolymer%label          = label
    copolymer%monomer_type   = monomer_type
    copolymer%block_length   = block_length
    copolymer%block_segments = adscf_contour_segments(block_length, ds)
    copolymer%segments       = sum(copolymer%block_segments)
    copolymer%ensemble       = ensemble
    copolymer%rho            = rho
    copolymer%mu             = mu
    call adscf_announce('Diblock Copolymer', label, ensemble)
    do b = 1, block_number
      copolymer%density(:,:,:,b) = adscf_start_density(rho * copolymer%block_segments(b) * ds, volume, dvol, &
                                                       gridx*gridy*gridz)
    end do
  end subroutine Initialize_Diblock

  ! mod_diblock.f90:100-174
  subroutine Propagate_Diblock(copolymer, expfield)
    type(diblock_copolymer), intent(inout) :: copolymer
    double precision, dimension(gridx,gridy,gridz,monomer_types), intent(in) :: expfield
    integer(c_int) :: rc, mt(block_number), bs(block_number)
    mt = copolymer%monomer_type
    bs = copolymer%block_segments
    rc = adscf_propagate_chain(adscf_context, int(block_number, c_int), mt, bs, int(copolymer%ensemble, c_int), &
                               copolymer%rho, copolymer%mu, expfield, copolymer%distribution, copolymer%density, &
                               copolymer%Qsum, copolymer%energy)
    if (rc == -3) then
      print*, 'Invalid ensemble; Cannot calculate Density'
      stop
    end if
    call adscf_check(rc, 'Propagate_Diblock')
  end subroutine Propagate_Diblock

end module Molecule_Diblock_Copolymer

module Molecule_Triblock_Copolymer
  use Global
  use Propagate
  use adscf_hip
  implicit none
  private

  integer, parameter  :: block_number = 3

  type triblock_copolymer
    integer                                   :: label
    integer, dimension(block_number)          :: monomer_type
    double precision, dimension(block_number) :: block_length
    integer           :: ensemble
    double precision  :: rho
    double precision  :: mu
    double precision  :: energy
    integer, dimension(block_number)          :: block_segments
    integer                                   :: segments
    double precision                          :: Qsum
    double precision, dimension(gridx,gridy,gridz,block_number) :: distribution
    double precision, dimension(gridx,gridy,gridz,block_number) :: density
  end type triblock_copolymer

  public :: triblock_copolymer
  public :: Initialize_Triblock
  public :: Propagate_Triblock

contains

  ! mod_triblock.f90:46-90
  subroutine Initialize_Triblock(copolymer, label, monomer_type, block_length, ensemble, rho, mu)
    integer, intent(in)                                   :: label, ensemble
    double precision, intent(in)                          :: rho, mu
    double precision, dimension(block_number), intent(in) :: block_length
    integer, dimension(block_number), intent(in)          :: monomer_type
    type(triblock_copolymer), intent(out)                 :: copolymer
    integer :: b
    copolymer%label          = label
    copolymer%monomer_type   = monomer_type
    copolymer%block_length   = block_length
    copolymer%block_segments = adscf_contour_segments(block_length, ds)
    copolymer%segments       = sum(copolymer%block_segments)
    copolymer%ensemble       = ensemble
    copolymer%rho            = rho
    copolymer%mu             = mu
    call adscf_announce('Triblock Copolymer', label, ensemble)
    do b = 1, block_number
      copolymer%density(:,:,:,b) = adscf_start_density(rho * copolymer%block_segments(b) * ds, volume, dvol, &
                                                       gridx*gridy*gridz)
    end do
  end subroutine Initialize_Triblock

  ! mod_triblock.f90:103-177
  subroutine Propagate_Triblock(copolymer, expfield)
    type(triblock_copolymer), intent(inout) :: copolymer
    double precision, dimension(gridx,gridy,gridz,monomer_types), intent(in) :: expfield
    integer(c_int) :: rc, mt(block_number), bs(block_number)
    mt = copolymer%monomer_type
    bs = copolymer%block_segments
    rc = adscf_propagate_chain(adscf_context, int(block_number, c_int), mt, bs, int(copolymer%ensemble, c_int), &
                               copolymer%rho, copolymer%mu, expfield, copolymer%distribution, copolymer%density, &
                               copolymer%Qsum, copolymer%energy)
    if (rc == -3) then
      print*, 'Invalid ensemble; Cannot calculate Density'
      stop
    end if
    call adscf_check(rc, 'Propagate_Triblock')
  end subroutine Propagate_Triblock

end module Molecule_Triblock_Copolymer

module Molecule_Multiblock_Copolymer
  use Global
  use Propagate
  use adscf_hip
  implicit none
  private

  type multiblock_copolymer
    integer                                    :: label
    integer                                    :: block_number
    integer, dimension(:),allocatable          :: monomer_type
    double precision, dimension(:),allocatable :: block_length
    integer           :: ensemble
    double precision  :: rho
    double precision  :: mu
    double precision  :: energy
    integer, dimension(:),allocatable   :: block_segments
    integer           :: segments
    double precision  :: Qsum
    double precision, dimension(:,:,:,:),allocatable :: distribution
    double precision, dimension(:,:,:,:),allocatable :: density
  end type multiblock_copolymer

  public :: multiblock_copolymer
  public :: Initialize_Multiblock
  public :: Propagate_Multiblock

contains

  ! mod_multiblock.f90:48-99
  subroutine Initialize_Multiblock(copolymer, label, nblocks, monomer_type, block_length, ensemble, rho, mu)
    integer, intent(in)                        :: label, nblocks, ensemble
    double precision, intent(in)               :: rho, mu
    double precision, dimension(:), intent(in) :: block_length
    integer, dimension(:), intent(in)          :: monomer_type
    type(multiblock_copolymer), intent(out)    :: copolymer
    integer :: b
    copolymer%label        = label
    copolymer%block_number = nblocks
    copolymer%ensemble     = ensemble
    copolymer%rho          = rho
    copolymer%mu           = mu
    ! allocation on assignment sizes the per-block arrays
    copolymer%monomer_type   = monomer_type(1:nblocks)
    copolymer%block_length   = block_length(1:nblocks)
    copolymer%block_segments = adscf_contour_segments(copolymer%block_length, ds)
    copolymer%segments       = sum(copolymer%block_segments)
    allocate(copolymer%distribution(gridx,gridy,gridz,nblocks), copolymer%density(gridx,gridy,gridz,nblocks))
    call adscf_announce('Multiblock Copolymer', label, ensemble)
    do b = 1, nblocks
      copolymer%density(:,:,:,b) = adscf_start_density(rho * copolymer%block_segments(b) * ds, volume, dvol, &
                                                       gridx*gridy*gridz)
    end do
  end subroutine Initialize_Multiblock

  ! mod_multiblock.f90:112-188
  subroutine Propagate_Multiblock(copolymer, expfield)
    type(multiblock_copolymer), intent(inout) :: copolymer
    double precision, dimension(gridx,gridy,gridz,monomer_types), intent(in) :: expfield
    integer(c_int) :: rc
    integer(c_int), allocatable :: mt(:), bs(:)
    mt = copolymer%monomer_type
    bs = copolymer%block_segments
    rc = adscf_propagate_chain(adscf_context, int(copolymer%block_number, c_int), mt, bs, &
                               int(copolymer%ensemble, c_int), copolymer%rho, copolymer%mu, expfield, &
                               copolymer%distribution, copolymer%density, copolymer%Qsum, copolymer%energy)
    if (rc == -3) then
      print*, 'Invalid ensemble; Cannot calculate Density'
      stop
    end if
    call adscf_check(rc, 'Propagate_Multiblock')
  end subroutine Propagate_Multiblock

end module Molecule_Multiblock_Copolymer
