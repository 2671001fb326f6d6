/*
 * CPU FP64 restatement of ADSCF's Bulk_3D_FFT pseudo-spectral SCFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it, and only as the checker.  The product (libadscf_hip.so) never
 * links, loads or calls it, and has no CPU fallback.
 *
 * Conventions (same as the reference Fortran):
 *   - every field is real(8)(nx,ny,nz[,k]) in column-major order, x fastest;
 *   - a half spectrum is complex(8)(nx/2+1,ny,nz), stored as interleaved
 *     {re,im} doubles;
 *   - monomer types and blocks are 1-based in the arguments that carry them.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef ADSCF_ORACLE_H
#define ADSCF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;
  double lx, ly, lz; /* sizex, sizey, sizez */
  double ds;
} orc_grid;

/* ---- MT19937 as in General_Routines/mod_mt.f90 ---------------------- */
typedef struct {
  unsigned int mt[624];
  int mti;
} orc_mt;
void orc_mt_sgrnd(orc_mt* st, int seed);          /* mod_mt.f90:71-87 */
unsigned int orc_mt_next32(orc_mt* st);           /* tempered word, mod_mt.f90:109-140 */
double orc_mt_grnd(orc_mt* st);                   /* mod_mt.f90:142-146: y/(2^32-1) */

/* ---- Fourier module (Geometry_Modules/Bulk/mod_fourier.f90) ---------- */
/* Real2Fourier: forward r2c, divided by nx*ny*nz (mod_fourier.f90:94-105) */
void orc_real2fourier(const orc_grid* g, const double* prop, double* propfourier);
/* Fourier2Real: unnormalised c2r (mod_fourier.f90:75-86) */
void orc_fourier2real(const orc_grid* g, const double* propfourier, double* prop);
/* Calculate_Laplace (mod_fourier.f90:165-198); any output may be NULL.
 * laplace, explaplace: complex (nx/2+1,ny,nz); nabla: complex (nx/2+1,ny,nz,3) */
void orc_calculate_laplace(const orc_grid* g, double* laplace, double* explaplace, double* nabla);
/* Calculate_Volume_Parameters (mod_fourier.f90:152-157) */
void orc_volume_parameters(const orc_grid* g, double* volume, double* dvol);
/* Full complex DFT of one line (length n, stride 1), sign -1 forward / +1 backward, unnormalised. */
void orc_dft_line(int n, const double* in, double* out, int sign);

/* ---- Propagate module (General_Routines/mod_propagate.f90) ----------- */
void orc_calculate_expfield(const orc_grid* g, int ntypes, const double* field, double* expfield); /* :22-32 */
void orc_propagate_step(const orc_grid* g, double* prop, const double* expfield_type);            /* :40-68 */

/* ---- Chain propagation (Polymer_Modules/mod_multiblock.f90:112-188;
 *      identical algorithm in mod_diblock.f90:100-174, mod_triblock.f90:103-177) */
typedef struct {
  int nblocks;
  const int* monomer_type;   /* 1-based types, length nblocks */
  const int* block_segments; /* length nblocks */
  int ensemble;              /* 1 canonical, 2 grand canonical */
  double rho, mu;            /* inout as in the reference type */
  double Qsum, energy;       /* outputs */
} orc_chain;
/* expfield: (nx,ny,nz,ntypes); distribution, density: (nx,ny,nz,nblocks).
 * q_out / qdag_out, when non-NULL, receive all (Ns+1) slices of qq / qqdag.
 * Returns 0, or -1 for an invalid ensemble (the reference stops there). */
int orc_propagate_chain(const orc_grid* g, orc_chain* c, const double* expfield,
                        double* distribution, double* density, double* q_out, double* qdag_out);
/* block_segments = nint(block_length/ds) (mod_multiblock.f90:85, mod_diblock.f90:74) */
int orc_block_segments(double block_length, double ds);
/* Contour rule of orc_propagate_chain and orc_propagate_homopolymer (process-wide; tests
 * set it back to 0).  0: the reference's trapezoid (mod_multiblock.f90:157-168), the
 * default.  1: composite Simpson, an extra mode with no reference counterpart. */
void orc_set_contour_rule(int rule);
/* weight of point j (0..nb) of a block of nb segments under `rule` */
double orc_contour_weight(int rule, int j, int nb, double ds);

/* ---- Homopolymer and solvent (Polymer_Modules/mod_homopolymer.f90, mod_solvent.f90) */
/* Propagate_Homopolymer (mod_homopolymer.f90:90-141): one propagator qq(s), s = 0..segments,
 * of monomer_type (1-based); distribution = trapezoid sum of qq(s) qq(mono - s); the
 * ensemble normalisation of :127-138.  c->nblocks, monomer_type, block_segments give the
 * single type and segments.  Returns 0, or -1 for an invalid ensemble. */
int orc_propagate_homopolymer(const orc_grid* g, orc_chain* c, const double* expfield, double* distribution,
                              double* density);
/* Update_Solvent (mod_solvent.f90:85-108): distribution = exp(-field), Qsum = dvol sum,
 * density and energy per ensemble (grand canonical rho without the ds*segments factor). */
int orc_update_solvent(const orc_grid* g, orc_chain* c, const double* field, double* distribution,
                       double* density);

/* ---- Flory-Huggins interactions (Interaction_Modules/mod_FloryHuggins.f90) */
typedef struct {
  int ncomp;
  double kappa;
  const double* chi;    /* ncomp x ncomp, column-major, symmetric, zero diagonal */
  const double* volume; /* monomer volumes, length ncomp */
} orc_fh;
void orc_mean_fields(const orc_grid* g, const orc_fh* fh, const double* density, double* field); /* :104-129 */
double orc_interaction_energy(const orc_grid* g, const orc_fh* fh, const double* density);      /* :74-96 */
double orc_average_volume_fraction(const orc_grid* g, const orc_fh* fh, const double* density); /* :139-160 */

/* ---- Iterate module (General_Routines/mod_iterate.f90) -------------- */
/* field, dfield: (fieldsize, 0:m) column-major */
void orc_mixing_simple(double* field, const double* dfield, long fieldsize, int m, double lambda);       /* :27-38 */
void orc_mixing_lambda(double* field, double* dfield, long fieldsize, int m, double lammax, int iter);   /* :46-84 */
void orc_mixing_anderson(double* field, double* dfield, long fieldsize, int m, double lam_mix, int iter);/* :92-150 */
int orc_ludcmp(double* a, int n, int* indx, double* d);  /* :232-298, returns 1 if a row was all zero */
void orc_lubksb(const double* a, int n, const int* indx, double* b);                                      /* :302-337 */
void orc_solve_lu(const double* A, const double* B, double* X, int n);                                    /* :203-224 */
void orc_solve_cholesky(const double* A, const double* B, double* X, int n);                              /* :160-195 */

/* ---- SCF drivers ------------------------------------------------------ */
/* Calculate_Densities (main_multiblock.f90:514-540): species densities (nx,ny,nz,ntypes) */
int orc_species_densities(const orc_grid* g, orc_chain* c, int ntypes, const double* field,
                          double* density);
/* Calculate_Energy (main_multiblock.f90:548-557) with Interaction_Energy */
double orc_total_energy(const orc_grid* g, const orc_chain* c, const orc_fh* fh, int ntypes,
                        const double* density, const double* field);
/* Find_SCF_Saddle_Point (main_multiblock.f90:648-700) with Mix_Fields (708-765, SAVE history)
 * and Check_Accuracy = max|dfield| (773-783).  Anderson mixing, history m, step lam.
 * field: inout (nx,ny,nz,ntypes); density: out.  Stops after maxit iterations or when
 * .not.(accuracy > tol).  Returns the number of iterations performed (the reference's
 * loop index at exit; maxit+1 if not converged, as with a Fortran DO loop), writes the last
 * accuracy.  The caller's field holds the mixed field for the next iteration; density
 * holds the densities of the last evaluated field (main_multiblock.f90:696). */
int orc_scf_anderson(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* field,
                     double* density, int m, double lam, double tol, int maxit, double* accuracy,
                     double* residual_history);

/* Calculate_Saddle_Point of Main_Diblock_Melt (main_diblock.f90:398-451) with Mix_Fields
 * (:315-372) and Check_Accuracy (:380-390): per iteration the densities of the field, accuracy
 * sum((density-newdensity)**2) against the previous iteration's densities (density on entry
 * for the first), exit when .not.(accuracy > tol), else the mean fields, dfield = newfield -
 * field, mixing_type 1/2/3 with lam (Mixing_Lambda's lammax) and iter_status = iter, density =
 * newdensity.  On return density holds the last newdensity and field the last newfield
 * (:447-448), the mixed field in mixed_field (may be NULL).  Returns the loop index at exit. */
int orc_diblock_saddle(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* field, double* density,
                       int m, int mixing_type, double lam, double tol, int maxit, double* accuracy,
                       double* mixed_field);

/* ---- DDFT (Main_Multiblock/main_multiblock.f90) ------------------------ */
/* Read_Mobilities, table part (main_multiblock.f90:486-504): for every half-spectrum
 * point, |q| = sqrt(-laplace), iq = int(|q|/dq), the reference's interpolation of
 * table(nq, nt, nt) (column-major, row iq = 1..nq at q = iq*dq), times |q|^2, capped at
 * wcutoff.  mobility_qq: (nx/2+1, ny, nz, nt, nt) column-major. */
void orc_mobility_qq(const orc_grid* g, int nt, int nq, double dq, const double* table, double wcutoff,
                     double* mobility_qq);
/* Get_Conjugate_Fields (main_multiblock.f90:566-640; main_blend.f90:376-428): Anderson
 * mixing (history m, step lam; fresh history, iter_status = min(iter, m+1)) of the field
 * on dfield = newdensity - density.  measure 0: accuracy = maxval(abs(dfield))
 * (main_multiblock.f90:780); 1: sum(abs(dfield))*dvol (main_blend.f90:567).  Exit when
 * .not.(accuracy > thresh).  field: inout; density: the target, unchanged.  Returns the
 * loop index at exit (maxit+1 when the loop ran out) or -1 on an invalid ensemble. */
int orc_conjugate_fields(const orc_grid* g, orc_chain* c, int ntypes, const double* density, double* field,
                         int m, double lam, double thresh, int maxit, int measure, double* accuracy);
/* gcfstat of main_multiblock.f90:619-634 for accuracy_scf: 0 success, 1 slow, 2 failed,
 * 3 failed with accuracy >= 0.1 (the reference stops there). */
int orc_gcfstat(double accuracy, double accuracy_scf);
/* DDFT_Timestep_Euler (main_multiblock.f90:794-880) with the conjugate-field solves of
 * the reference (accuracy_scf, itermax_scf, Anderson with history m and lam = 0.5).
 * density, field: inout (nx,ny,nz,ntypes).  gcfstat: the two solves' statuses (may be
 * NULL); iterations: their loop indices (may be NULL).  Returns 0, -1 on an invalid
 * ensemble, or 3 when a solve ends with accuracy >= 0.1 (the reference stops). */
int orc_ddft_step(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* density, double* field,
                  const double* mobility_qq, double timestep, int qdmin, double accuracy_scf, int itermax_scf,
                  int m, int* gcfstat, int* iterations);

/* ---- Homopolymer blend and EPD (Main_Blend/main_blend.f90) -------------- */
/* A blend of npol homopolymers, polymer i of monomer type types[i] (1-based) with segs[i]
 * segments, canonical with concentration rho[i] (Initialize_Homopolymer,
 * mod_homopolymer.f90:43-78; main_blend.f90:93-100). */
typedef struct {
  int npol;
  const int* types;
  const int* segs;
  const double* rho;
  double mu[8], energy[8]; /* outputs of the last density evaluation */
} orc_blend;
/* Calculate_Densities of the blend (main_blend.f90:321-345): expfield of the field, every
 * polymer's Propagate_Homopolymer, species densities summed in polymer order. */
int orc_blend_densities(const orc_grid* g, orc_blend* b, int ntypes, const double* field, double* density);
/* Get_Conjugate_Fields of Main_Blend (main_blend.f90:376-428): Anderson (history m, lam)
 * on dfield = newdensity - density with accuracy sum(abs(dfield))*dvol, exit when
 * .not.(accuracy > thresh) (the driver's accuracy_scf = 0.01, a single-precision literal:
 * (double)0.01f).  Returns the loop index at
 * exit (maxit+1 when the loop ran out), -1 on an invalid ensemble. */
int orc_blend_conjugate_fields(const orc_grid* g, orc_blend* b, int ntypes, const double* density, double* field,
                               int m, double lam, double thresh, int maxit, double* accuracy);
/* Find_SCF_Saddle_Point of Main_Blend (main_blend.f90:437-486) with Mix_Fields
 * (:494-553): densities of the blend, mean fields, dfield = newfield - field, accuracy
 * sum(abs(dfield))*dvol (:567), exit when .not.(accuracy > thresh), else mixing_type 1
 * (Mixing_Simple), 2 (Mixing_Lambda, the driver's choice, lammax = lam) or 3
 * (Mixing_Anderson) with iter_status = iter.  field: inout; density: the densities of the
 * last evaluated field.  Returns the loop index at exit (maxit+1 when the loop ran out). */
int orc_blend_scf(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field, double* density,
                  int m, int mixing_type, double lam, double thresh, int maxit, double* accuracy);
/* Chain_Dynamics_Timestep_Euler (main_blend.f90:605-667): densities and mean fields of the
 * field, mulocal = newfield - field, its gradient along each axis through nabla, the density
 * response to field - eps*gradient (eps = (double)0.0001f, the driver's single-precision
 * literal) per axis, its divergence through nabla,
 * density += sum over axes * timestep/eps, then the conjugate fields of the new density
 * (history m, lam 0.5, accuracy_scf, itermax).  density, field: inout.  iterations: the
 * solve's loop index (may be NULL).  Returns 0 or -1. */
int orc_epd_step(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* density, double* field,
                 double timestep, int m, double accuracy_scf, int itermax, int* iterations);
/* Calculate_Field_Derivative (main_blend.f90:709-740): dfield = the Fourier product
 * -laplace * (newfield - field) with newfield the mean fields of the blend's densities at
 * field.  Returns 0 or -1. */
int orc_field_derivative(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, const double* field,
                         double* dfield);
/* EPD_Timestep_Euler (main_blend.f90:580-596) and EPD_Timestep_RungeKutta (:678-701) on
 * field (inout).  Return 0 or -1. */
int orc_epd_timestep_euler(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field,
                           double timestep);
int orc_epd_timestep_rk4(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field,
                         double timestep);

#ifdef __cplusplus
}
#endif
#endif
