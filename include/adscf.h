/*
 * adscf.h -- C-ABI of libadscf_hip.so, the MI355X (gfx950) implementation of
 * ADSCF's Bulk_3D_FFT pseudo-spectral SCFT propagator path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int status (ADSCF_OK == 0).  On failure adscf_last_error() describes the
 * problem; the Fortran replacement modules print it and stop, which is the
 * reference's own error convention (print*, then stop).
 *
 * Layout conventions follow the reference Fortran exactly:
 *   - a field is real(8)(nx,ny,nz) column-major (x fastest);
 *   - a field set is real(8)(nx,ny,nz,k) with k the slowest index
 *     (monomer type, or block);
 *   - a half spectrum is complex(8)(nx/2+1,ny,nz) stored as interleaved
 *     {re,im} doubles (layout-compatible with Fortran double complex);
 *   - monomer types are 1-based;
 *   - Anderson history arrays are real(8)(fieldsize, 0:mixing_dim).
 *
 * All host arrays are owned by the caller.  The context owns every device
 * buffer and one HIP stream; calls are synchronous (outputs are valid on
 * return) and not re-entrant for one context.
 *
 * Citations name the reference procedure each entry point replaces, relative
 * to the reference tree's Bulk_3D_FFT directory.
 */
#ifndef ADSCF_H
#define ADSCF_H

#ifdef __cplusplus
extern "C" {
#endif

#define ADSCF_OK 0
#define ADSCF_ERR_ARG (-1)      /* invalid argument or size */
#define ADSCF_ERR_HIP (-2)      /* HIP runtime error (device, allocation, launch) */
#define ADSCF_ERR_ENSEMBLE (-3) /* invalid ensemble: mod_diblock.f90:164-166 stops here */
#define ADSCF_ERR_STATE (-4)    /* call order violated (e.g. iterate before set_field) */

#define ADSCF_FP64 0 /* plain FP64 path: every array and FFT in double */
#define ADSCF_FP32 1 /* FP32-mixed: propagators/FFTs in float; Q, F, phi sums, Anderson in double */

typedef struct adscf_ctx adscf_ctx;
typedef struct adscf_batch adscf_batch;

/* ---- lifetime ---------------------------------------------------------- */

/* Initialize (General_Routines/sub_wrap.f90:8-40): Initialize_Fourier
 * (Geometry_Modules/Bulk/mod_fourier.f90:49-54), Calculate_Laplace (:165-198) and
 * Calculate_Volume_Parameters (:152-157) for a runtime grid (the reference fixes
 * gridx/gridy/gridz at compile time in module Global, e.g. Main_Diblock_Melt/mod_global.f90:8-19).
 * ntypes = monomer_types.  precision = ADSCF_FP64 or ADSCF_FP32.  At most 2^27 grid
 * points (512^3); larger grids return ADSCF_ERR_ARG. */
int adscf_init(adscf_ctx** ctx, int nx, int ny, int nz, int ntypes, double lx, double ly, double lz, double ds,
               int precision);
/* Finish / End_Fourier (sub_wrap.f90:48-59, mod_fourier.f90:62-67) */
int adscf_finalize(adscf_ctx* ctx);
/* Box or contour step changed: Calculate_Laplace + Calculate_Volume_Parameters again */
int adscf_set_box(adscf_ctx* ctx, double lx, double ly, double lz);
int adscf_set_ds(adscf_ctx* ctx, double ds);
int adscf_set_precision(adscf_ctx* ctx, int precision);
/* volume and dvol (mod_fourier.f90:154-155) */
int adscf_volume(const adscf_ctx* ctx, double* volume, double* dvol);
const char* adscf_last_error(const adscf_ctx* ctx);

/* ---- Fourier_fftw3 (mod_fourier.f90) ------------------------------------ */

/* Real2Fourier (mod_fourier.f90:94-105): r2c, divided by nx*ny*nz */
int adscf_real2fourier(adscf_ctx* ctx, const double* prop, double* propfourier);
/* Fourier2Real (mod_fourier.f90:75-86): unnormalised c2r */
int adscf_fourier2real(adscf_ctx* ctx, const double* propfourier, double* prop);
/* public arrays laplace, explaplace, nabla of Calculate_Laplace (mod_fourier.f90:26-28,165-198);
 * each output is complex (nx/2+1,ny,nz) [nabla: (nx/2+1,ny,nz,3)]; NULL skips it. */
int adscf_laplace(adscf_ctx* ctx, double* laplace, double* explaplace, double* nabla);

/* ---- Propagate (General_Routines/mod_propagate.f90) ---------------------- */

/* Calculate_Expfield (mod_propagate.f90:22-32): expfield = exp(-0.5*ds*field), (nx,ny,nz,ntypes) */
int adscf_expfield(adscf_ctx* ctx, const double* field, double* expfield);
/* Propagate_Step (mod_propagate.f90:40-68): prop <- E * c2r(K * r2c(E*prop)/Ng) */
int adscf_propagate_step(adscf_ctx* ctx, double* prop, const double* expfield_type);

/* ---- Polymer_Modules: Propagate_Diblock / _Triblock / _Multiblock -------- */

/* Propagate_Multiblock (Polymer_Modules/mod_multiblock.f90:112-188); the same algorithm as
 * Propagate_Diblock (mod_diblock.f90:100-174) and Propagate_Triblock (mod_triblock.f90:103-177).
 * monomer_type[nblocks] (1-based), block_segments[nblocks]; ensemble 1 canonical, 2 grand
 * canonical.  rho, mu are in/out as in the reference type.  expfield (nx,ny,nz,ntypes);
 * distribution and density (nx,ny,nz,nblocks); either may be NULL.  qsum, energy: out. */
int adscf_propagate_chain(adscf_ctx* ctx, int nblocks, const int* monomer_type, const int* block_segments,
                          int ensemble, double* rho, double* mu, const double* expfield, double* distribution,
                          double* density, double* qsum, double* energy);
/* No counterpart in the reference.  adscf_propagate_chain on DEVICE arrays with no device
 * allocation, for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): expfield (nx,ny,nz,ntypes) in, distribution and density
 * (nx,ny,nz,nblocks) out, scalars[4] = {Qsum, mu, rho, energy} out, all device pointers; rho
 * and mu are inputs (canonical: rho; grand canonical: mu).  Every work array lives in one
 * 128 MiB scratch array per GPU, held until the process ends: the history of the chain needs
 * 2 (Ns+1) nx ny nz P bytes (P = 8 in FP64, 4 in FP32-mixed) plus about nx ny nz (2P + 4P
 * in FP32-mixed) and 64 KiB, and a chain that needs more returns ADSCF_ERR_ARG.  The
 * context's transform tables are uploaded at adscf_init, adscf_set_box, adscf_set_ds and
 * adscf_set_precision, which may allocate; this call does not.  Results equal
 * adscf_propagate_chain's bit for bit.  All calls on one GPU share the scratch: calls that
 * overlap on one GPU (two contexts, two threads) race.  Returns after the results are
 * written (the context stream is synchronised). */
int adscf_propagate_chain_noalloc(adscf_ctx* ctx, int nblocks, const int* monomer_type, const int* block_segments,
                                  int ensemble, double rho, double mu, const double* expfield, double* distribution,
                                  double* density, double* scalars);

/* ---- Polymer_Modules: Propagate_Homopolymer, Update_Solvent ---------------- */

/* Propagate_Homopolymer (Polymer_Modules/mod_homopolymer.f90:90-141): one propagator of
 * `segments` steps in monomer_type (1-based); distribution and density (nx,ny,nz), either
 * may be NULL; rho, mu in/out; qsum, energy out; ensemble 1 or 2 (else ADSCF_ERR_ENSEMBLE). */
int adscf_propagate_homopolymer(adscf_ctx* ctx, int monomer_type, int segments, int ensemble, double* rho,
                                double* mu, const double* expfield, double* distribution, double* density,
                                double* qsum, double* energy);
/* Update_Solvent (Polymer_Modules/mod_solvent.f90:85-108): distribution = exp(-field) for one
 * field (nx,ny,nz); Qsum, density and energy per ensemble; rho, mu in/out. */
int adscf_update_solvent(adscf_ctx* ctx, int ensemble, double* rho, double* mu, const double* field,
                         double* distribution, double* density, double* qsum, double* energy);

/* ---- Interactions_FloryHuggins (Interaction_Modules/mod_FloryHuggins.f90) - */

/* The values Initialize_Interactions reads from input_interactions (:37-66):
 * kappa, chi (ntypes x ntypes column-major, symmetric) and monomer volumes (ntypes). */
int adscf_set_interactions(adscf_ctx* ctx, double kappa, const double* chi, const double* monomer_volume);
/* Calculate_Mean_Fields (:104-129): density (nx,ny,nz,ntypes) -> field */
int adscf_mean_fields(adscf_ctx* ctx, const double* density, double* field);
/* Interaction_Energy (:74-96) */
int adscf_interaction_energy(adscf_ctx* ctx, const double* density, double* energy);
/* Check_Volume_Fraction (:139-160): the average total volume fraction it prints */
int adscf_average_volume_fraction(adscf_ctx* ctx, const double* density, double* average);

/* ---- Iterate (General_Routines/mod_iterate.f90) ------------------------- */

/* field, dfield: (fieldsize, 0:mixing_dim), updated in place as the reference does. */
int adscf_mixing_simple(adscf_ctx* ctx, double* field, const double* dfield, long fieldsize, int mixing_dim,
                        double lambda);                                            /* :27-38 */
int adscf_mixing_lambda(adscf_ctx* ctx, double* field, double* dfield, long fieldsize, int mixing_dim,
                        double lammax, int iter);                                  /* :46-84 */
int adscf_mixing_anderson(adscf_ctx* ctx, double* field, double* dfield, long fieldsize, int mixing_dim,
                          double lam_mix, int iter);                               /* :92-150 */
/* Host routines with the reference's semantics (column-major n x n): */
int adscf_ludcmp(double* a, int n, int* indx, double* d);                          /* :232-298 */
int adscf_lubksb(const double* a, int n, const int* indx, double* b);              /* :302-337 */
int adscf_solve_lu(const double* a, const double* b, double* x, int n);            /* :203-224 */
int adscf_solve_cholesky(const double* a, const double* b, double* x, int n);      /* :160-195 */
/* The device wavefront LU that the Anderson step runs (ludcmp + lubksb, mod_iterate.f90:232-337)
   on one n x n system, 1 <= n <= 40, for checking it against adscf_ludcmp / adscf_lubksb:
   a column-major (unchanged), b overwritten by the solution; lu (n*n) receives the factors and
   indx (n) the 1-based pivots when not null, singular (when not null) 1 if a row of a is zero. */
int adscf_device_lu_solve(adscf_ctx* ctx, const double* a, int n, double* b, double* lu, int* indx, int* singular);

/* ---- device-resident SCF loop ------------------------------------------ */
/* The loop of Find_SCF_Saddle_Point (Main_Multiblock/main_multiblock.f90:648-700):
 * Calculate_Densities (514-540) -> Calculate_Mean_Fields -> dfield = new - field ->
 * accuracy = maxval(abs(dfield)) (773-783) -> exit if .not.(accuracy > tol) ->
 * Mix_Fields/Mixing_Anderson with a SAVE'd history (708-765), all kept in HBM.
 * Iterations are captured once in a HIP graph and replayed. */
int adscf_scf_setup(adscf_ctx* ctx, int nblocks, const int* monomer_type, const int* block_segments, int ensemble,
                    double rho, double mu, int mixing_dim, double lam_mix);
/* load field (nx,ny,nz,ntypes), clear the mixing history and the iteration count */
int adscf_scf_set_field(adscf_ctx* ctx, const double* field);
/* run up to niter iterations; stops early when .not.(accuracy > tol).  iters_total is the
 * reference's loop index so far, accuracy the last max|dfield|, converged 0/1.  After a call
 * that stopped at convergence, a call with a smaller tol continues the same loop (the mixing
 * step of the last evaluation, then further iterations) exactly as one call with that tol.
 * A call with another residual, measure or mixing type than the previous call (or a
 * conjugate-field solve in between) starts a new loop on the current field: the count
 * restarts at 1, as each loop of the reference does (Mixing_Lambda's first step,
 * Anderson's lam00 steps up to iteration m), and the older history entries are never read
 * (Anderson first solves at iteration m+1, when every entry is the new loop's). */
int adscf_scf_iterate(adscf_ctx* ctx, int niter, double tol, int* iters_total, double* accuracy, int* converged);
/* recompute densities of the current field (main_multiblock.f90:319) without mixing */
int adscf_scf_evaluate(adscf_ctx* ctx);
/* copy the current field and the species densities of the last evaluation; NULL skips */
int adscf_scf_get(adscf_ctx* ctx, double* field, double* density);
/* Calculate_Energy (main_multiblock.f90:548-557) for the last evaluation: total F, F per chain
 * (F / (rho*V)), Qsum, mu, rho.  Any pointer may be NULL. */
int adscf_scf_energy(adscf_ctx* ctx, double* F, double* F_per_chain, double* qsum, double* mu, double* rho);
/* max|dfield| of every iteration evaluated since adscf_scf_set_field (up to cap values);
 * n receives the number available */
int adscf_scf_residuals(adscf_ctx* ctx, double* out, int cap, int* n);
/* mean device time per SCF iteration of the last adscf_scf_iterate call, and of the
 * propagator sweep alone, in milliseconds (HIP events on the context stream) */
int adscf_scf_timing(adscf_ctx* ctx, double* ms_per_iter, double* ms_sweep);

/* Residual of the loop adscf_scf_iterate runs.  kind 0: dfield = mean fields - field
 * (Find_SCF_Saddle_Point, main_multiblock.f90:673-676); kind 1: dfield = densities - target
 * (the loop of Get_Conjugate_Fields, main_multiblock.f90:602-604), target (nx,ny,nz,ntypes)
 * uploaded when non-NULL.  measure 0: accuracy = maxval(abs(dfield)) (main_multiblock.f90:780);
 * 1: sum(abs(dfield))*dvol (Main_Blend/main_blend.f90:567); 2: sum((density-newdensity)**2),
 * the change of the densities since the previous evaluation (Main_Diblock_Melt/main_diblock.f90:387),
 * the first iteration comparing with adscf_scf_set_density; not available for a blend of
 * several homopolymers (adscf_scf_iterate returns ADSCF_ERR_ARG: the densities are summed
 * over the polymers in place).  Default: kind 0, measure 0. */
int adscf_scf_set_residual(adscf_ctx* ctx, int kind, int measure, const double* target);
/* the densities (nx,ny,nz,ntypes) the next evaluation's measure 2 compares with (the driver's
 * density before Calculate_Saddle_Point, main_diblock.f90:398-412); after adscf_scf_set_field */
int adscf_scf_set_density(adscf_ctx* ctx, const double* density);

/* Mixing of the loop adscf_scf_iterate runs (Mix_Fields' mixing_type, main_multiblock.f90:738-747):
 * 1 Mixing_Simple (General_Routines/mod_iterate.f90:27-38, step lam_mix), 2 Mixing_Lambda
 * (:46-84, lammax = lam_mix; Main_Diblock_Melt's and Main_Blend's saddle-point search),
 * 3 Mixing_Anderson (:92-150, the default).  The conjugate-field solves always use 3, as
 * both drivers do. */
int adscf_scf_set_mixing(adscf_ctx* ctx, int mixing_type);

/* Contour rule of every density evaluation of the context (chains, homopolymers, the SCF
 * loop, the adjoint).  0: the reference's trapezoid (Polymer_Modules/mod_multiblock.f90:157-168,
 * mod_diblock.f90:143-154), the default and the rule parity is stated for.  1: composite
 * Simpson per block (1/3-rule panels; an odd number of segments ends with one 3/8-rule panel
 * over its last three; blocks of fewer than two segments keep the trapezoid); no counterpart
 * in the reference.  The weights of both rules sum to the block length, so the canonical
 * mass is the same. */
int adscf_set_contour_rule(adscf_ctx* ctx, int rule);

/* ---- conjugate fields and DDFT (Main_Multiblock/main_multiblock.f90) ------------ */
/* Get_Conjugate_Fields (main_multiblock.f90:566-640; main_blend.f90:376-428): iterate the
 * current field (adscf_scf_set_field, or the field a previous call left) with the Anderson
 * settings of adscf_scf_setup until the densities it produces match `density` (NULL: the
 * device-resident DDFT density); exits when .not.(accuracy > thresh), where the reference
 * uses thresh = accuracy_scf*0.5 (main_multiblock.f90:609) or accuracy_scf (main_blend.f90:412).
 * iterations: the reference's loop index at exit (maxit+1 if the loop ran out). */
int adscf_conjugate_fields(adscf_ctx* ctx, const double* density, double thresh, int maxit, int measure,
                           int* iterations, double* accuracy);
/* Read_Mobilities' table part (main_multiblock.f90:486-504): mobility_table(nq,ntypes,ntypes)
 * column-major (row iq at q = iq*dq), interpolated at |q| of every half-spectrum point as the
 * reference does, times |q|^2, capped at wcutoff.  The result (nx/2+1,ny,nz,ntypes,ntypes) is
 * kept on the device for adscf_ddft_step and copied to mobility_qq when non-NULL. */
int adscf_mobility_qq(adscf_ctx* ctx, int nq, double dq, const double* table, double wcutoff, double* mobility_qq);
/* upload a mobility_qq array computed by the caller (the driver's own Read_Mobilities) */
int adscf_ddft_set_mobility(adscf_ctx* ctx, const double* mobility_qq);
/* the density of the time-stepping entry points (DDFT and EPD), (nx,ny,nz,ntypes), kept on
 * the device between steps */
int adscf_ddft_set_density(adscf_ctx* ctx, const double* density);
int adscf_ddft_get_density(adscf_ctx* ctx, double* density);
/* DDFT_Timestep_Euler (main_multiblock.f90:794-880) on the device-resident density and field:
 * conjugate fields (accuracy_scf, itermax_scf; the driver passes 1e-3 and 2000), mulocal =
 * mean fields - field, its spectrum times -mobility_qq summed over types, back to real space,
 * density += timestep*ddensity, the 1e-4 floor when qdmin, conjugate fields again, and, when
 * qdmin and that solve was slow or failed, the densities of the field.  gcfstat[2] and
 * iterations[2] (may be NULL) receive both solves' status (0 ok, 1 slow, 2 failed) and loop
 * index.  A solve that ends with accuracy >= 0.1 returns ADSCF_ERR_STATE (the reference stops). */
int adscf_ddft_step(adscf_ctx* ctx, double timestep, int qdmin, double accuracy_scf, int itermax_scf, int* gcfstat,
                    int* iterations);
/* Calculate_Energy (main_multiblock.f90:548-557) with the DDFT density and field; mu and rho
 * of the last density evaluation.  Any pointer may be NULL. */
int adscf_ddft_energy(adscf_ctx* ctx, double* F, double* mu, double* rho);

/* ---- homopolymer blend and EPD (Main_Blend/main_blend.f90) --------------------- */
/* The device-resident loop over a blend of npol homopolymers instead of one copolymer:
 * polymer i of monomer type monomer_type[i] (1-based) with segments[i] segments, canonical
 * at concentration rho[i] (Initialize_Homopolymer, Polymer_Modules/mod_homopolymer.f90:43-78;
 * main_blend.f90:93-100).  Calculate_Densities of the blend (main_blend.f90:321-345) then
 * replaces the chain's in adscf_scf_iterate/_evaluate/_get/_energy, adscf_conjugate_fields
 * and adscf_epd_step; Anderson history mixing_dim (2 in Main_Blend's Iterate), step lam_mix.
 * adscf_scf_setup returns the context to a single chain. */
int adscf_blend_setup(adscf_ctx* ctx, int npol, const int* monomer_type, const int* segments, const double* rho,
                      int mixing_dim, double lam_mix);
/* per polymer (arrays of npol, any may be NULL): Qsum, mu and energy of the last evaluation */
int adscf_blend_scalars(adscf_ctx* ctx, double* qsum, double* mu, double* energy);
/* Chain_Dynamics_Timestep_Euler (main_blend.f90:605-667) on the device-resident density
 * (adscf_ddft_set_density) and field: densities and mean fields of the field, the gradient of
 * mulocal along each axis through nabla, the density response to field - eps*gradient
 * (eps = 0.0001 as the driver's single-precision literal) and its divergence, density +=
 * timestep/eps times their sum, then Get_Conjugate_Fields of the new density with the
 * accuracy sum|dfield|*dvol and exit threshold accuracy_scf (the driver: 0.01 single
 * precision, itermax 100000).  iterations: the solve's loop index. */
int adscf_epd_step(adscf_ctx* ctx, double timestep, double accuracy_scf, int itermax, int* iterations);
/* Calculate_Field_Derivative (main_blend.f90:709-740) at the device-resident field (after
 * adscf_scf_set_field, on a blend or a single chain): dfield (nx,ny,nz,ntypes) =
 * -laplace * (mean fields of the densities - field) in Fourier space. */
int adscf_field_derivative(adscf_ctx* ctx, double* dfield);
/* EPD_Timestep_Euler (main_blend.f90:580-596; method 0: field += timestep*dfield) or
 * EPD_Timestep_RungeKutta (:678-701; method 1: the classical four-stage rule) on the
 * device-resident field, which stays in HBM; the new field then starts the SCF state
 * afresh, as adscf_scf_set_field does (read it with adscf_scf_get). */
int adscf_epd_field_step(adscf_ctx* ctx, double timestep, int method);

/* ---- schedule autotuner (no counterpart in the reference) --------------------- */
/* Times the captured propagator sweep of the configured SCF problem for every line-FFT
 * schedule of each axis (register-resident four-step shapes N1 x N2 with 1, 2 or 4
 * wavefronts per workgroup, and the LDS Stockham kernel) and keeps the fastest.  Needs
 * adscf_scf_setup + adscf_scf_set_field.  report (may be NULL) receives the timings. */
int adscf_autotune(adscf_ctx* ctx, int reps, char* report, int report_cap);
/* current schedule of an axis (0 x, 1 y, 2 z): n1 x n2 four-step register engine, n1 = 0 the
 * Stockham kernel, n2 = 0 the n1^3 cube engine (n = n1^3 points, n1^2 threads per line),
 * n2 < 0 the (-n2) n1^2 engine (n1^2 threads per line, -n2 points each: 128 = 8 x 4 x 4) */
int adscf_get_schedule(const adscf_ctx* ctx, int axis, int* n, int* n1, int* n2, int* waves);
/* The axis whose operator gets one line pass per contour step in a 3-D sweep; the operators
 * of the other two meet across consecutive steps (the order alternates a,b,c / c,b,a) and
 * share merged passes.  -1 restores the default (y).  adscf_autotune also chooses it. */
int adscf_set_middle_axis(adscf_ctx* ctx, int axis);
int adscf_get_middle_axis(const adscf_ctx* ctx, int* axis);
int adscf_set_schedule(adscf_ctx* ctx, int axis, int n1, int n2, int waves);
/* tiles per workgroup (1, 2 or 4; FP64 1 or 2) of the sweep passes along an axis with a
 * register schedule of 32 points or more: a workgroup issues the loads of its next tile
 * before it transforms and stores the current one (same results bit for bit; other settings
 * run one tile); adscf_autotune also chooses it */
int adscf_set_pass_tiles(adscf_ctx* ctx, int axis, int tiles);
int adscf_get_pass_tiles(const adscf_ctx* ctx, int axis, int* tiles);
/* xy-plane passes (no counterpart in the reference): the x and y operators of a contour step
 * applied by one workgroup per chain and z-plane, so that a 3-D sweep alternates plane passes
 * and z passes and costs one dependent pass per contour step instead of two.  mode -1 (the
 * default) uses them where the grid allows (nx == ny in {32, 48, 64}, nz > 1), 0 never, 1
 * requires them (ADSCF_ERR_ARG where the grid does not allow them).  n1, n2 select the
 * engine shape of the plane operators as in adscf_set_schedule (n1 = 0 keeps the current
 * one).  adscf_autotune also chooses them.  active reports whether the sweep of the
 * context's precision uses them. */
int adscf_set_plane_passes(adscf_ctx* ctx, int mode, int n1, int n2);
int adscf_get_plane_passes(const adscf_ctx* ctx, int* active, int* n1, int* n2);

/* ---- batched independent instances (no counterpart in the reference) ---------- */
/* Many independent two-type SCF problems (parameter sweeps) of uneven grid and contour
 * sizes, run together: each line pass of the propagator sweep is one launch over all
 * instances, and one HIP graph replay is one SCF iteration of every instance.
 * Per instance i: grid[3i..3i+2], box[3i..3i+2], ds[i], nblocks[i] blocks whose types and
 * segments follow in monomer_type/block_segments (concatenated), kappa[i], chi[4i..4i+3]
 * (2x2 column-major), vol[2i..2i+1].  Axis lengths must be 8, 16, 24, 32, 48, 64 or 128;
 * at most 2048 instances. */
int adscf_batch_create(adscf_batch** batch, int ninst, const int* grid, const double* box, const double* ds,
                       const int* nblocks, const int* monomer_type, const int* block_segments, const double* kappa,
                       const double* chi, const double* vol, int ensemble, double rho, int mixing_dim,
                       double lam_mix, int precision);
int adscf_batch_destroy(adscf_batch* batch);
const char* adscf_batch_last_error(const adscf_batch* batch);
/* fields of all instances concatenated, each (nx,ny,nz,2) column-major */
int adscf_batch_set_fields(adscf_batch* batch, const double* fields);
int adscf_batch_set_precision(adscf_batch* batch, int precision);
/* Sweep of a batched iteration: enable 1 runs every propagator pass of every instance in one
 * persistent launch (workgroups take tiles pass by pass and wait only for the previous pass of
 * their own instance; bounded waits, a wait that expires makes adscf_batch_iterate fail);
 * 0 (the default) runs one launch per pass and register shape.  The persistent sweep carries
 * the shapes 4x8, 8x4, 8x6, 6x8, 8x8 and cube4 (axis lengths 32, 48 and 64); active receives 1 when
 * the batch's schedules allow it, else the per-pass launches stay.  Results are identical. */
int adscf_batch_set_persistent(adscf_batch* batch, int enable, int* active);
/* xy-plane passes of the batched sweep (see adscf_set_plane_passes): -1 (the default) on every
 * instance whose grid allows them, 0 line passes only, 1 required on every instance
 * (ADSCF_ERR_ARG otherwise).  active (may be NULL) receives the number of instances whose
 * sweep uses them.  The persistent sweep carries line passes only. */
int adscf_batch_set_plane_passes(adscf_batch* batch, int mode, int* active);
/* register shape (as in adscf_set_schedule; the batched line passes run four wavefronts of
 * lines per workgroup) of every instance axis of length n, in place of the default or adscf_batch_autotune's choice; ADSCF_ERR_ARG when
 * the batched kernels have no such shape or no instance has an axis of that length. */
int adscf_batch_set_schedule(adscf_batch* batch, int n, int n1, int n2);
/* per-instance outputs (arrays of ninst); stops when every instance has converged */
int adscf_batch_iterate(adscf_batch* batch, int niter, double tol, int* iters, double* accuracy, int* converged);
int adscf_batch_energy(adscf_batch* batch, double* F_per_chain);
int adscf_batch_get(adscf_batch* batch, int instance, double* field, double* density);
int adscf_batch_timing(adscf_batch* batch, double* ms_per_iter, double* ms_sweep);
/* picks the register-resident shape per distinct axis length by timing the batched sweep */
int adscf_batch_autotune(adscf_batch* batch, int reps, char* report, int report_cap);

/* ---- comparator (no counterpart in the reference) ---------------------------- */
/* Mean time in microseconds of one single-chain contour step built like the reference's
 * Propagate_Step (E*q, r2c, K/Ng, c2r, E*q) with hipFFT plans on this context's grid and
 * precision.  For timing comparisons only; the library's own path does not use hipFFT. */
int adscf_hipfft_step_time(adscf_ctx* ctx, int reps, double* us_per_step);
/* Measured device bandwidth beside the 8 TB/s peak: a device-to-device copy of `bytes`
 * (copy_GBs counts read + write) and a streaming read of `bytes` (read_GBs), median of reps
 * timed with HIP events on the context stream; bytes >= 1 MiB.  Diagnostic only. */
int adscf_hbm_bandwidth(adscf_ctx* ctx, long bytes, int reps, double* copy_GBs, double* read_GBs);

/* ---- adjoint (no counterpart in the reference) ---------------------------- */
/* Reverse-mode derivative of one canonical density evaluation of the chain
 * (mod_multiblock.f90:112-188 with Calculate_Energy, main_multiblock.f90:548-557) for
 *   L = sum_r (phi_target_type(r) - phi_target(r))^2 + lam * F.
 * field (nx,ny,nz,ntypes), phi_target (nx,ny,nz), target_type 1-based.  Propagators are
 * stored only every checkpoint_every contour steps (0: about sqrt(Ns)) and recomputed
 * by segments.  Outputs: loss, free_energy (F), grad_field = dL/dw (nx,ny,nz,ntypes),
 * grad_box = dL/d(lx,ly,lz) (3), phi (nx,ny,nz,ntypes); every output may be NULL, and
 * with both gradients NULL only the forward part runs. */
int adscf_adjoint(adscf_ctx* ctx, int nblocks, const int* monomer_type, const int* block_segments, double rho,
                  const double* field, int target_type, const double* phi_target, double lam, int checkpoint_every,
                  double* loss, double* free_energy, double* grad_field, double* grad_box, double* phi);

#ifdef __cplusplus
}
#endif
#endif
