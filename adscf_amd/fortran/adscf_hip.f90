!
! ISO_C_BINDING interface to libadscf_hip.so (include/adscf.h) and the one
! context the replacement modules share.  The context is created by
! Initialize_Fourier and destroyed by End_Fourier, matching the lifetime of the
! FFTW plans it replaces (Geometry_Modules/Bulk/mod_fourier.f90:49-67).
!
module adscf_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ADSCF_OK = 0, ADSCF_FP64 = 0, ADSCF_FP32 = 1
  type(c_ptr), save :: adscf_context = c_null_ptr

  interface
    integer(c_int) function adscf_init(ctx, nx, ny, nz, nt, lx, ly, lz, ds, prec) bind(C, name='adscf_init')
      import :: c_ptr, c_int, c_double
      type(c_ptr) :: ctx
      integer(c_int), value :: nx, ny, nz, nt, prec
      real(c_double), value :: lx, ly, lz, ds
    end function
    integer(c_int) function adscf_finalize(ctx) bind(C, name='adscf_finalize')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function adscf_set_box(ctx, lx, ly, lz) bind(C, name='adscf_set_box')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lx, ly, lz
    end function
    integer(c_int) function adscf_set_ds(ctx, ds) bind(C, name='adscf_set_ds')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ds
    end function
    type(c_ptr) function adscf_last_error(ctx) bind(C, name='adscf_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function adscf_real2fourier(ctx, prop, pf) bind(C, name='adscf_real2fourier')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: prop(*)
      complex(c_double_complex), intent(out) :: pf(*)
    end function
    integer(c_int) function adscf_fourier2real(ctx, pf, prop) bind(C, name='adscf_fourier2real')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: pf(*)
      real(c_double), intent(out) :: prop(*)
    end function
    integer(c_int) function adscf_laplace(ctx, lap, elap, nab) bind(C, name='adscf_laplace')
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: lap(*), elap(*), nab(*)
    end function
    integer(c_int) function adscf_expfield(ctx, field, expfield) bind(C, name='adscf_expfield')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: field(*)
      real(c_double), intent(out) :: expfield(*)
    end function
    integer(c_int) function adscf_propagate_step(ctx, prop, e) bind(C, name='adscf_propagate_step')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: prop(*)
      real(c_double), intent(in) :: e(*)
    end function
    integer(c_int) function adscf_propagate_chain(ctx, nblocks, monomer_type, block_segments, ensemble, rho, mu, &
        expfield, distribution, density, qsum, energy) bind(C, name='adscf_propagate_chain')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nblocks, ensemble
      integer(c_int), intent(in) :: monomer_type(*), block_segments(*)
      real(c_double), intent(inout) :: rho, mu
      real(c_double), intent(in) :: expfield(*)
      real(c_double), intent(out) :: distribution(*), density(*)
      real(c_double), intent(out) :: qsum, energy
    end function
    integer(c_int) function adscf_propagate_homopolymer(ctx, monomer_type, segments, ensemble, rho, mu, &
        expfield, distribution, density, qsum, energy) bind(C, name='adscf_propagate_homopolymer')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: monomer_type, segments, ensemble
      real(c_double), intent(inout) :: rho, mu
      real(c_double), intent(in) :: expfield(*)
      real(c_double), intent(out) :: distribution(*), density(*)
      real(c_double), intent(out) :: qsum, energy
    end function
    integer(c_int) function adscf_update_solvent(ctx, ensemble, rho, mu, field, distribution, density, qsum, &
        energy) bind(C, name='adscf_update_solvent')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ensemble
      real(c_double), intent(inout) :: rho, mu
      real(c_double), intent(in) :: field(*)
      real(c_double), intent(out) :: distribution(*), density(*)
      real(c_double), intent(out) :: qsum, energy
    end function
    integer(c_int) function adscf_set_interactions(ctx, kappa, chi, vol) bind(C, name='adscf_set_interactions')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: kappa
      real(c_double), intent(in) :: chi(*), vol(*)
    end function
    integer(c_int) function adscf_mean_fields(ctx, density, field) bind(C, name='adscf_mean_fields')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
      real(c_double), intent(out) :: field(*)
    end function
    integer(c_int) function adscf_interaction_energy(ctx, density, energy) bind(C, name='adscf_interaction_energy')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
      real(c_double), intent(out) :: energy
    end function
    integer(c_int) function adscf_average_volume_fraction(ctx, density, average) &
        bind(C, name='adscf_average_volume_fraction')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
      real(c_double), intent(out) :: average
    end function
    integer(c_int) function adscf_mixing_simple(ctx, field, dfield, fieldsize, m, lambda) &
        bind(C, name='adscf_mixing_simple')
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: field(*)
      real(c_double), intent(in) :: dfield(*)
      integer(c_long), value :: fieldsize
      integer(c_int), value :: m
      real(c_double), value :: lambda
    end function
    integer(c_int) function adscf_mixing_lambda(ctx, field, dfield, fieldsize, m, lammax, iter) &
        bind(C, name='adscf_mixing_lambda')
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: field(*), dfield(*)
      integer(c_long), value :: fieldsize
      integer(c_int), value :: m, iter
      real(c_double), value :: lammax
    end function
    integer(c_int) function adscf_mixing_anderson(ctx, field, dfield, fieldsize, m, lam_mix, iter) &
        bind(C, name='adscf_mixing_anderson')
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: field(*), dfield(*)
      integer(c_long), value :: fieldsize
      integer(c_int), value :: m, iter
      real(c_double), value :: lam_mix
    end function
    integer(c_int) function adscf_ludcmp(a, n, indx, d) bind(C, name='adscf_ludcmp')
      import :: c_int, c_double
      real(c_double), intent(inout) :: a(*)
      integer(c_int), value :: n
      integer(c_int), intent(out) :: indx(*)
      real(c_double), intent(out) :: d
    end function
    integer(c_int) function adscf_lubksb(a, n, indx, b) bind(C, name='adscf_lubksb')
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*)
      integer(c_int), value :: n
      integer(c_int), intent(in) :: indx(*)
      real(c_double), intent(inout) :: b(*)
    end function
    integer(c_int) function adscf_solve_lu(a, b, x, n) bind(C, name='adscf_solve_lu')
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), b(*)
      real(c_double), intent(out) :: x(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function adscf_solve_cholesky(a, b, x, n) bind(C, name='adscf_solve_cholesky')
      import :: c_int, c_double
      real(c_double), intent(in) :: a(*), b(*)
      real(c_double), intent(out) :: x(*)
      integer(c_int), value :: n
    end function
    ! device-resident loops (no counterpart in the reference; see include/adscf.h)
    integer(c_int) function adscf_scf_setup(ctx, nblocks, monomer_type, block_segments, ensemble, rho, mu, &
                                            mixing_dim, lam_mix) bind(C, name='adscf_scf_setup')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nblocks, ensemble, mixing_dim
      integer(c_int), intent(in) :: monomer_type(*), block_segments(*)
      real(c_double), value :: rho, mu, lam_mix
    end function
    integer(c_int) function adscf_scf_set_field(ctx, field) bind(C, name='adscf_scf_set_field')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: field(*)
    end function
    ! 0 trapezoid (the reference's contour rule, the default), 1 composite Simpson per block
    integer(c_int) function adscf_set_contour_rule(ctx, rule) bind(C, name='adscf_set_contour_rule')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: rule
    end function
    integer(c_int) function adscf_scf_iterate(ctx, niter, tol, iters, accuracy, converged) &
        bind(C, name='adscf_scf_iterate')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: niter
      real(c_double), value :: tol
      integer(c_int), intent(out) :: iters, converged
      real(c_double), intent(out) :: accuracy
    end function
    integer(c_int) function adscf_scf_get(ctx, field, density) bind(C, name='adscf_scf_get')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: field(*), density(*)
    end function
    ! EPD_Timestep_Euler (method 0) / EPD_Timestep_RungeKutta (method 1) on the device field
    integer(c_int) function adscf_epd_field_step(ctx, timestep, method) bind(C, name='adscf_epd_field_step')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: timestep
      integer(c_int), value :: method
    end function
    integer(c_int) function adscf_field_derivative(ctx, dfield) bind(C, name='adscf_field_derivative')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dfield(*)
    end function
    integer(c_int) function adscf_conjugate_fields(ctx, density, thresh, maxit, measure, iterations, accuracy) &
        bind(C, name='adscf_conjugate_fields')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
      real(c_double), value :: thresh
      integer(c_int), value :: maxit, measure
      integer(c_int), intent(out) :: iterations
      real(c_double), intent(out) :: accuracy
    end function
    integer(c_int) function adscf_ddft_set_mobility(ctx, mobility_qq) bind(C, name='adscf_ddft_set_mobility')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: mobility_qq(*)
    end function
    integer(c_int) function adscf_ddft_set_density(ctx, density) bind(C, name='adscf_ddft_set_density')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: density(*)
    end function
    integer(c_int) function adscf_ddft_get_density(ctx, density) bind(C, name='adscf_ddft_get_density')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: density(*)
    end function
    integer(c_int) function adscf_ddft_step(ctx, timestep, qdmin, accuracy_scf, itermax_scf, gcfstat, iterations) &
        bind(C, name='adscf_ddft_step')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: timestep, accuracy_scf
      integer(c_int), value :: qdmin, itermax_scf
      integer(c_int), intent(out) :: gcfstat(2), iterations(2)
    end function
    integer(c_int) function adscf_ddft_energy(ctx, energy, mu, rho) bind(C, name='adscf_ddft_energy')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: energy, mu, rho
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! The reference reports errors with print* and stop; so does the binding.
  subroutine adscf_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=512) :: text
    if (rc == ADSCF_OK) return
    text = ''
    p = adscf_last_error(adscf_context)
    if (c_associated(p)) then
      n = int(min(c_strlen(p), 512_c_size_t))
      call c_f_pointer(p, msg, [n])
      do i = 1, n
        text(i:i) = msg(i)
      end do
    end if
    print*, 'adscf error in ', where, ' (status', rc, '): ', trim(text)
    stop 1
  end subroutine adscf_check

  ! The chain set-up shared by the Initialize_* replacements: the announcement line of a
  ! species, the contour segments of a block of the given length, and the uniform starting
  ! density of a species carrying `amount` (rho times its contour length) on npoints points.
  subroutine adscf_announce(species, label, ensemble)
    character(*), intent(in) :: species
    integer, intent(in)      :: label, ensemble
    select case (ensemble)
    case (1)
      print*, species, ' ', label, ' is initialized (canonical ensemble).'
    case (2)
      print*, species, ' ', label, ' is initialized (grand canonical ensemble).'
    case default
      print*, species, ' ', label, ': invalid ensemble', ensemble
    end select
  end subroutine adscf_announce

  ! (not pure or elemental: the drivers build with -fno-automatic, which saves every local)
  function adscf_contour_segments(lengths, step) result(segments)
    double precision, intent(in) :: lengths(:), step
    integer :: segments(size(lengths))
    segments = nint(lengths / step)
  end function adscf_contour_segments

  double precision function adscf_start_density(amount, box_volume, cell_volume, npoints)
    double precision, intent(in) :: amount, box_volume, cell_volume
    integer, intent(in)          :: npoints
    adscf_start_density = amount * box_volume / cell_volume / dble(npoints)
  end function adscf_start_density

end module adscf_hip
