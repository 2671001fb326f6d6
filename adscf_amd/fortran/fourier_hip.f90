!
! Replacement for module Fourier_fftw3 (Geometry_Modules/Bulk/mod_fourier.f90):
! same module name, procedures, dummy arguments and public arrays; the
! transforms run on the GPU through libadscf_hip.so.
!
module Fourier_fftw3
  use Global
  use adscf_hip
  implicit none
  private

  double complex, save, dimension(gridx/2+1, gridy, gridz)    :: laplace
  double complex, save, dimension(gridx/2+1, gridy, gridz)    :: explaplace
  double complex, save, dimension(gridx/2+1, gridy, gridz, 3) :: nabla

  public :: Initialize_Fourier
  public :: End_Fourier
  public :: Fourier2Real
  public :: Real2Fourier
  public :: Calculate_Volume_Parameters
  public :: Calculate_Laplace
  public :: nabla
  public :: laplace
  public :: explaplace

contains

  ! mod_fourier.f90:49-54 (the FFTW plans become the library context)
  subroutine Initialize_Fourier
    if (c_associated(adscf_context)) return
    call adscf_check(adscf_init(adscf_context, int(gridx, c_int), int(gridy, c_int), int(gridz, c_int), &
                                int(monomer_types, c_int), sizex, sizey, sizez, ds, ADSCF_FP64), &
                     'Initialize_Fourier')
  end subroutine Initialize_Fourier

  ! mod_fourier.f90:62-67
  subroutine End_Fourier
    integer(c_int) :: rc
    if (.not. c_associated(adscf_context)) return
    rc = adscf_finalize(adscf_context)
    adscf_context = c_null_ptr
  end subroutine End_Fourier

  ! mod_fourier.f90:75-86: unnormalised c2r
  subroutine Fourier2Real(prop, propfourier)
    double precision, dimension(gridx, gridy, gridz), intent(out)    :: prop
    double complex, dimension(gridx/2+1, gridy, gridz), intent(in)   :: propfourier
    call adscf_check(adscf_fourier2real(adscf_context, propfourier, prop), 'Fourier2Real')
  end subroutine Fourier2Real

  ! mod_fourier.f90:94-105: r2c divided by gridx*gridy*gridz
  subroutine Real2Fourier(prop, propfourier)
    double precision, dimension(gridx, gridy, gridz), intent(in)     :: prop
    double complex, dimension(gridx/2+1, gridy, gridz), intent(out)  :: propfourier
    call adscf_check(adscf_real2fourier(adscf_context, prop, propfourier), 'Real2Fourier')
  end subroutine Real2Fourier

  ! mod_fourier.f90:152-157
  subroutine Calculate_Volume_Parameters
    volume = sizex*sizey*sizez
    dvol   = volume/dble(gridx*gridy*gridz)
    if (c_associated(adscf_context)) &
      call adscf_check(adscf_set_box(adscf_context, sizex, sizey, sizez), 'Calculate_Volume_Parameters')
  end subroutine Calculate_Volume_Parameters

  ! mod_fourier.f90:165-198: host arrays as in the reference; the device keeps
  ! its own per-axis factors for the same box and ds
  subroutine Calculate_Laplace
    call adscf_check(adscf_set_box(adscf_context, sizex, sizey, sizez), 'Calculate_Laplace')
    call adscf_check(adscf_set_ds(adscf_context, ds), 'Calculate_Laplace')
    call adscf_check(adscf_laplace(adscf_context, laplace, explaplace, nabla), 'Calculate_Laplace')
  end subroutine Calculate_Laplace

end module Fourier_fftw3
