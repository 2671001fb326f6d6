/*
 * CPU FP64 restatement of ADSCF's Bulk_3D_FFT SCFT path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the parity oracle.  It is written for obviousness, not speed:
 * the FFT is a recursive mixed-radix DFT with directly evaluated twiddles,
 * every array pass follows the order of the reference Fortran, and all
 * arithmetic is FP64.  Only tests, the smoke check and the benchmark's CPU
 * baseline may use it.  Reference citations are relative to the reference
 * tree's Bulk_3D_FFT directory.
 */
#include "scft_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ===================================================================== */
/* MT19937 (General_Routines/mod_mt.f90)                                  */
/* ===================================================================== */

/* sgrnd: mt(0)=seed; mt(i)=69069*mt(i-1) mod 2^32 (mod_mt.f90:71-87) */
void orc_mt_sgrnd(orc_mt* st, int seed) {
  st->mt[0] = (unsigned int)seed;
  for (int i = 1; i < 624; ++i) st->mt[i] = 69069u * st->mt[i - 1];
  st->mti = 624;
}

/* grnd body (mod_mt.f90:90-140): regenerate 624 words when exhausted, then temper. */
unsigned int orc_mt_next32(orc_mt* st) {
  const unsigned int MATA = 0x9908b0dfu, UMASK = 0x80000000u, LMASK = 0x7fffffffu;
  if (st->mti >= 624) {
    if (st->mti == 625) orc_mt_sgrnd(st, 4357); /* default seed, mod_mt.f90:50,117-121 */
    int kk;
    unsigned int y;
    for (kk = 0; kk < 624 - 397; ++kk) {
      y = (st->mt[kk] & UMASK) | (st->mt[kk + 1] & LMASK);
      st->mt[kk] = st->mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    }
    for (; kk < 623; ++kk) {
      y = (st->mt[kk] & UMASK) | (st->mt[kk + 1] & LMASK);
      st->mt[kk] = st->mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    }
    y = (st->mt[623] & UMASK) | (st->mt[0] & LMASK);
    st->mt[623] = st->mt[396] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
    st->mti = 0;
  }
  unsigned int y = st->mt[st->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* grnd = y / (2^32 - 1), y read as unsigned (mod_mt.f90:142-146) */
double orc_mt_grnd(orc_mt* st) { return (double)orc_mt_next32(st) / (4294967296.0 - 1.0); }

/* ===================================================================== */
/* FFT: recursive mixed-radix DFT (stands in for FFTW's r2c/c2r 3-D plans, */
/* mod_fourier.f90:51-52; FFTW computes the same unnormalised DFT sums)   */
/* ===================================================================== */

static int smallest_factor(int n) {
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_j in[j*is] * W^(j*k), W = exp(sign*2*pi*i/n).  tw[t] = exp(sign*2*pi*i*t/N), n | N */
static void dft_rec(int n, const cplx* in, int is, cplx* out, int N, const cplx* tw, cplx* scratch) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; ++r) dft_rec(m, in + (long)r * is, is * p, out + (long)r * m, N, tw, scratch + n);
  int step = N / n;
  for (int k = 0; k < n; ++k) {
    int k1 = k % m;
    cplx s = 0;
    for (int r = 0; r < p; ++r) s += out[(long)r * m + k1] * tw[(long)step * (((long)r * k) % n)];
    scratch[k] = s;
  }
  memcpy(out, scratch, sizeof(cplx) * (size_t)n);
}

static cplx* make_twiddles(int n, int sign) {
  cplx* tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  for (int t = 0; t < n; ++t) {
    double a = 6.283185307179586476925286766559 * (double)t / (double)n;
    tw[t] = cos(a) + sign * sin(a) * I;
  }
  return tw;
}

static int scratch_len(int n) { return 2 * n + 64; } /* sum over recursion levels n + n/p + ... < 2n */

void orc_dft_line(int n, const double* in, double* out, int sign) {
  cplx* tw = make_twiddles(n, sign);
  cplx* sc = (cplx*)malloc(sizeof(cplx) * (size_t)scratch_len(n));
  dft_rec(n, (const cplx*)in, 1, (cplx*)out, n, tw, sc);
  free(sc);
  free(tw);
}

/* Apply a length-n DFT to `count` lines of a complex array; line l starts at base(l), stride `st`. */
static void dft_lines(cplx* a, int n, long st, long count, const long* bases, int sign) {
  if (n == 1) return;
  cplx* tw = make_twiddles(n, sign);
  cplx* sc = (cplx*)malloc(sizeof(cplx) * (size_t)scratch_len(n));
  cplx* lin = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  cplx* lout = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  for (long l = 0; l < count; ++l) {
    cplx* p = a + bases[l];
    for (int j = 0; j < n; ++j) lin[j] = p[(long)j * st];
    dft_rec(n, lin, 1, lout, n, tw, sc);
    for (int j = 0; j < n; ++j) p[(long)j * st] = lout[j];
  }
  free(lout);
  free(lin);
  free(sc);
  free(tw);
}

/* y- and z-direction DFTs over a (nh, ny, nz) complex array */
static void dft_yz(const orc_grid* g, cplx* a, int nh, int sign) {
  int ny = g->ny, nz = g->nz;
  long cnt = (long)nh * nz;
  long* b = (long*)malloc(sizeof(long) * (size_t)(cnt > (long)nh * ny ? cnt : (long)nh * ny));
  long c = 0;
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nh; ++x) b[c++] = (long)z * nh * ny + x;
  dft_lines(a, ny, nh, c, b, sign);
  c = 0;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nh; ++x) b[c++] = (long)y * nh + x;
  dft_lines(a, nz, (long)nh * ny, c, b, sign);
  free(b);
}

/* Real2Fourier (mod_fourier.f90:94-105): r2c with sign -1 (dfftw_plan_dft_r2c_3d,
 * :51), then propfourier = ftout / (gridx*gridy*gridz) (:102). */
void orc_real2fourier(const orc_grid* g, const double* prop, double* propfourier) {
  int nx = g->nx, ny = g->ny, nz = g->nz, nh = nx / 2 + 1;
  cplx* out = (cplx*)propfourier;
  cplx* tw = make_twiddles(nx, -1);
  cplx* sc = (cplx*)malloc(sizeof(cplx) * (size_t)scratch_len(nx));
  cplx* lin = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
  cplx* lout = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      const double* r = prop + ((long)z * ny + y) * nx;
      for (int x = 0; x < nx; ++x) lin[x] = r[x];
      dft_rec(nx, lin, 1, lout, nx, tw, sc);
      cplx* o = out + ((long)z * ny + y) * nh;
      for (int k = 0; k < nh; ++k) o[k] = lout[k];
    }
  free(lout);
  free(lin);
  free(sc);
  free(tw);
  dft_yz(g, out, nh, -1);
  double inv = 1.0 / (double)(nx * ny * nz); /* integer product as in :102 */
  long nk = (long)nh * ny * nz;
  for (long i = 0; i < nk; ++i) out[i] = out[i] * inv;
}

/* Fourier2Real (mod_fourier.f90:75-86): unnormalised c2r, sign +1 (:52).  The full
 * x-line is rebuilt from the half spectrum by Hermitian symmetry X[nx-k] = conj(X[k]). */
void orc_fourier2real(const orc_grid* g, const double* propfourier, double* prop) {
  int nx = g->nx, ny = g->ny, nz = g->nz, nh = nx / 2 + 1;
  long nk = (long)nh * ny * nz;
  cplx* a = (cplx*)malloc(sizeof(cplx) * (size_t)nk);
  memcpy(a, propfourier, sizeof(cplx) * (size_t)nk);
  dft_yz(g, a, nh, +1);
  cplx* tw = make_twiddles(nx, +1);
  cplx* sc = (cplx*)malloc(sizeof(cplx) * (size_t)scratch_len(nx));
  cplx* lin = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
  cplx* lout = (cplx*)malloc(sizeof(cplx) * (size_t)nx);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      const cplx* h = a + ((long)z * ny + y) * nh;
      for (int k = 0; k < nh; ++k) lin[k] = h[k];
      for (int k = nh; k < nx; ++k) lin[k] = conj(h[nx - k]);
      dft_rec(nx, lin, 1, lout, nx, tw, sc);
      double* r = prop + ((long)z * ny + y) * nx;
      for (int x = 0; x < nx; ++x) r[x] = creal(lout[x]);
    }
  free(lout);
  free(lin);
  free(sc);
  free(tw);
  free(a);
}

/* Calculate_Laplace (mod_fourier.f90:165-198).  1-based loop indices x,y,z; the y and z
 * wave numbers fold to negative values only when the index exceeds n/2+1 (:181-186), so
 * the Nyquist index stays positive; x is never folded. */
void orc_calculate_laplace(const orc_grid* g, double* laplace, double* explaplace, double* nabla) {
  const double pi = 3.141592653589793; /* Global: pi = 3.141592653589793d0 */
  int nx = g->nx, ny = g->ny, nz = g->nz, nh = nx / 2 + 1;
  double dx = g->lx / (double)nx, dy = g->ly / (double)ny, dz = g->lz / (double)nz;
  long nk = (long)nh * ny * nz;
  for (int z = 1; z <= nz; ++z)
    for (int y = 1; y <= ny; ++y)
      for (int x = 1; x <= nh; ++x) {
        double qx = 2.0 * pi * (double)(x - 1) / g->lx;
        double qy = 2.0 * pi * (double)(y - 1) / g->ly;
        double qz = 2.0 * pi * (double)(z - 1) / g->lz;
        if (y > ny / 2 + 1) qy = 2.0 * pi * ((double)(y - 1) - (double)ny) / g->ly;
        if (z > nz / 2 + 1) qz = 2.0 * pi * ((double)(z - 1) - (double)nz) / g->lz;
        long i = ((long)(z - 1) * ny + (y - 1)) * nh + (x - 1);
        double k2 = qx * qx + qy * qy + qz * qz;
        if (laplace) {
          laplace[2 * i] = -k2;
          laplace[2 * i + 1] = 0.0;
        }
        if (explaplace) {
          explaplace[2 * i] = exp(-g->ds * k2);
          explaplace[2 * i + 1] = 0.0;
        }
        if (nabla) {
          nabla[2 * i] = 0.0;
          nabla[2 * i + 1] = sin(qx * dx) / dx;
          nabla[2 * (i + nk)] = 0.0;
          nabla[2 * (i + nk) + 1] = sin(qy * dy) / dy;
          nabla[2 * (i + 2 * nk)] = 0.0;
          nabla[2 * (i + 2 * nk) + 1] = sin(qz * dz) / dz;
        }
      }
}

/* Calculate_Volume_Parameters (mod_fourier.f90:152-157) */
void orc_volume_parameters(const orc_grid* g, double* volume, double* dvol) {
  double v = g->lx * g->ly * g->lz;
  *volume = v;
  *dvol = v / (double)(g->nx * g->ny * g->nz);
}

/* ===================================================================== */
/* Propagate module (General_Routines/mod_propagate.f90)                  */
/* ===================================================================== */

/* expfield = exp(-0.5*ds*field) (mod_propagate.f90:28) */
void orc_calculate_expfield(const orc_grid* g, int ntypes, const double* field, double* expfield) {
  long n = (long)g->nx * g->ny * g->nz * ntypes;
  for (long i = 0; i < n; ++i) expfield[i] = exp(-0.5 * g->ds * field[i]);
}

/* Propagate_Step (mod_propagate.f90:40-68) */
void orc_propagate_step(const orc_grid* g, double* prop, const double* expfield_type) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nk = (long)(g->nx / 2 + 1) * g->ny * g->nz;
  double* pf = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  double* el = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  orc_calculate_laplace(g, NULL, el, NULL);
  for (long i = 0; i < ng; ++i) prop[i] = expfield_type[i] * prop[i]; /* :56 */
  orc_real2fourier(g, prop, pf);                                        /* :58 */
  for (long i = 0; i < nk; ++i) {                                       /* :60 complex*complex */
    cplx a = pf[2 * i] + pf[2 * i + 1] * I, b = el[2 * i] + el[2 * i + 1] * I;
    cplx c = a * b;
    pf[2 * i] = creal(c);
    pf[2 * i + 1] = cimag(c);
  }
  orc_fourier2real(g, pf, prop);                                        /* :62 */
  for (long i = 0; i < ng; ++i) prop[i] = expfield_type[i] * prop[i]; /* :64 */
  free(el);
  free(pf);
}

/* ===================================================================== */
/* Contour rules                                                          */
/* ===================================================================== */

static int g_contour_rule = 0;
void orc_set_contour_rule(int rule) { g_contour_rule = rule; }

/* Rule 0 is the reference's trapezoid (mod_multiblock.f90:157-168): half weights at the two
 * ends of a block; a block of zero segments adds both halves at its one point.  Rule 1
 * restates the textbook composite Simpson rule panel by panel: panels [2p, 2p+2] with
 * weights ds/3 (1, 4, 1), and for an odd nb the last three intervals as one Simpson 3/8
 * panel, 3ds/8 (1, 3, 3, 1); a block of one segment keeps the trapezoid. */
double orc_contour_weight(int rule, int j, int nb, double ds) {
  if (nb == 0) return ds;
  if (rule == 0 || nb == 1) return (j == 0 || j == nb) ? 0.5 * ds : ds;
  const int npanel = (nb % 2 == 0) ? nb / 2 : (nb - 3) / 2;
  double w = 0.0;
  for (int p = 0; p < npanel; ++p) {
    const int a = 2 * p;
    if (j == a || j == a + 2) w += ds / 3.0;
    if (j == a + 1) w += 4.0 * ds / 3.0;
  }
  if (nb % 2 == 1) {
    const int a = 2 * npanel;
    if (j == a || j == a + 3) w += 3.0 * ds / 8.0;
    if (j == a + 1 || j == a + 2) w += 9.0 * ds / 8.0;
  }
  return w;
}

/* nint(block_length/ds): Fortran NINT rounds half away from zero (mod_multiblock.f90:85) */
/* mod_homopolymer.f90:90-141 */
int orc_propagate_homopolymer(const orc_grid* g, orc_chain* c, const double* expfield, double* distribution,
                              double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  const int mono = c->block_segments[0];
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  const double ds = g->ds;
  double* qq = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)(mono + 1));
  double* prop = (double*)malloc(sizeof(double) * (size_t)ng);
  if (!qq || !prop) {
    free(qq);
    free(prop);
    return -2;
  }
  /* :106-115 */
  for (long i = 0; i < ng; ++i) prop[i] = 1.0;
  memcpy(qq, prop, sizeof(double) * (size_t)ng);
  const double* E = expfield + (long)(c->monomer_type[0] - 1) * ng;
  for (int s = 1; s <= mono; ++s) {
    orc_propagate_step(g, prop, E);
    memcpy(qq + (long)s * ng, prop, sizeof(double) * (size_t)ng);
  }
  /* :117 */
  double sum = 0.0;
  for (long i = 0; i < ng; ++i) sum += qq[(long)mono * ng + i];
  c->Qsum = dvol * sum;
  /* :119-125 */
  if (g_contour_rule == 0) {
    for (long i = 0; i < ng; ++i) distribution[i] = 0.5 * ds * qq[i] * qq[(long)mono * ng + i];
    for (int s = 1; s <= mono - 1; ++s)
      for (long i = 0; i < ng; ++i)
        distribution[i] = distribution[i] + ds * qq[(long)s * ng + i] * qq[(long)(mono - s) * ng + i];
    for (long i = 0; i < ng; ++i) distribution[i] = distribution[i] + 0.5 * ds * qq[(long)mono * ng + i] * qq[i];
  } else {
    for (long i = 0; i < ng; ++i) distribution[i] = 0.0;
    for (int s = 0; s <= mono; ++s) {
      const double w = orc_contour_weight(g_contour_rule, s, mono, ds);
      for (long i = 0; i < ng; ++i)
        distribution[i] = distribution[i] + w * qq[(long)s * ng + i] * qq[(long)(mono - s) * ng + i];
    }
  }
  free(prop);
  free(qq);
  /* :127-138 */
  if (c->ensemble == 1) {
    for (long i = 0; i < ng; ++i) density[i] = c->rho * volume * distribution[i] / c->Qsum;
    c->mu = log(c->rho * volume / c->Qsum);
    c->energy = c->rho * volume * (c->mu - 1.0);
  } else if (c->ensemble == 2) {
    const double em = exp(c->mu);
    double tot = 0.0;
    for (long i = 0; i < ng; ++i) {
      density[i] = em * distribution[i];
      tot += density[i];
    }
    c->rho = tot * dvol / volume / (ds * mono);
    c->energy = -c->rho * volume;
  } else {
    return -1;
  }
  return 0;
}

/* mod_solvent.f90:85-108 */
int orc_update_solvent(const orc_grid* g, orc_chain* c, const double* field, double* distribution,
                       double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double sum = 0.0;
  for (long i = 0; i < ng; ++i) {
    distribution[i] = exp(-field[i]);
    sum += distribution[i];
  }
  c->Qsum = dvol * sum;
  if (c->ensemble == 1) {
    for (long i = 0; i < ng; ++i) density[i] = c->rho * volume * distribution[i] / c->Qsum;
    c->mu = log(c->rho * volume / c->Qsum);
    c->energy = c->rho * volume * (c->mu - 1.0);
  } else if (c->ensemble == 2) {
    const double em = exp(c->mu);
    double tot = 0.0;
    for (long i = 0; i < ng; ++i) {
      density[i] = em * distribution[i];
      tot += density[i];
    }
    c->rho = tot * dvol / volume;
    c->energy = -c->rho * volume;
  } else {
    return -1;
  }
  return 0;
}

int orc_block_segments(double block_length, double ds) { return (int)lround(block_length / ds); }

/* Propagate_Multiblock (mod_multiblock.f90:112-188) */
int orc_propagate_chain(const orc_grid* g, orc_chain* c, const double* expfield, double* distribution,
                        double* density, double* q_out, double* qdag_out) {
  long ng = (long)g->nx * g->ny * g->nz;
  int nb = c->nblocks;
  int mono = 0;
  for (int b = 0; b < nb; ++b) mono += c->block_segments[b];
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double ds = g->ds;
  double* qq = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)(mono + 1));
  double* qd = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)(mono + 1));
  double* prop = (double*)malloc(sizeof(double) * (size_t)ng);
  if (!qq || !qd || !prop) {
    free(qq); free(qd); free(prop);
    return -2;
  }
  /* forward sweep (:130-140) */
  int s0 = 0;
  for (long i = 0; i < ng; ++i) prop[i] = 1.0;
  memcpy(qq, prop, sizeof(double) * (size_t)ng);
  for (int b = 0; b < nb; ++b) {
    const double* E = expfield + (long)(c->monomer_type[b] - 1) * ng;
    for (int s = s0 + 1; s <= s0 + c->block_segments[b]; ++s) {
      orc_propagate_step(g, prop, E);
      memcpy(qq + (long)s * ng, prop, sizeof(double) * (size_t)ng);
    }
    s0 += c->block_segments[b];
  }
  /* complementary sweep, blocks reversed (:142-152) */
  s0 = 0;
  for (long i = 0; i < ng; ++i) prop[i] = 1.0;
  memcpy(qd, prop, sizeof(double) * (size_t)ng);
  for (int b = nb - 1; b >= 0; --b) {
    const double* E = expfield + (long)(c->monomer_type[b] - 1) * ng;
    for (int s = s0 + 1; s <= s0 + c->block_segments[b]; ++s) {
      orc_propagate_step(g, prop, E);
      memcpy(qd + (long)s * ng, prop, sizeof(double) * (size_t)ng);
    }
    s0 += c->block_segments[b];
  }
  /* Qsum = dvol * sum(qq(:,:,:,mono)) (:154-155) */
  double sum = 0.0;
  for (long i = 0; i < ng; ++i) sum += qq[(long)mono * ng + i];
  c->Qsum = dvol * sum;
  /* trapezoid contour integral per block (:157-168) */
  s0 = 0;
  for (int b = 0; b < nb && g_contour_rule != 0; ++b) { /* or another rule's weights */
    double* dist = distribution + (long)b * ng;
    const int n = c->block_segments[b];
    for (long i = 0; i < ng; ++i) dist[i] = 0.0;
    for (int j = 0; j <= n; ++j) {
      const double w = orc_contour_weight(g_contour_rule, j, n, ds);
      const double* a = qq + (long)(s0 + j) * ng;
      const double* d = qd + (long)(mono - s0 - j) * ng;
      for (long i = 0; i < ng; ++i) dist[i] = dist[i] + w * a[i] * d[i];
    }
    s0 += n;
  }
  s0 = 0;
  for (int b = 0; b < nb && g_contour_rule == 0; ++b) {
    double* dist = distribution + (long)b * ng;
    const double* a0 = qq + (long)s0 * ng;
    const double* d0 = qd + (long)(mono - s0) * ng;
    for (long i = 0; i < ng; ++i) dist[i] = 0.5 * ds * a0[i] * d0[i];
    for (int s = s0 + 1; s <= s0 + c->block_segments[b] - 1; ++s) {
      const double* a = qq + (long)s * ng;
      const double* d = qd + (long)(mono - s) * ng;
      for (long i = 0; i < ng; ++i) dist[i] = dist[i] + ds * a[i] * d[i];
    }
    s0 += c->block_segments[b];
    const double* a1 = qq + (long)s0 * ng;
    const double* d1 = qd + (long)(mono - s0) * ng;
    for (long i = 0; i < ng; ++i) dist[i] = dist[i] + 0.5 * ds * a1[i] * d1[i];
  }
  if (q_out) memcpy(q_out, qq, sizeof(double) * (size_t)ng * (size_t)(mono + 1));
  if (qdag_out) memcpy(qdag_out, qd, sizeof(double) * (size_t)ng * (size_t)(mono + 1));
  free(prop);
  free(qd);
  free(qq);
  /* ensemble normalisation (:173-184) */
  long nt = ng * nb;
  if (c->ensemble == 1) {
    for (long i = 0; i < nt; ++i) density[i] = c->rho * volume * distribution[i] / c->Qsum;
    c->mu = log(c->rho * volume / c->Qsum);
    c->energy = c->rho * volume * (c->mu - 1.0);
  } else if (c->ensemble == 2) {
    double em = exp(c->mu), tot = 0.0;
    for (long i = 0; i < nt; ++i) {
      density[i] = em * distribution[i];
      tot += density[i];
    }
    c->rho = tot * dvol / volume / (ds * (double)mono);
    c->energy = -c->rho * volume;
  } else {
    return -1; /* 'Invalid ensemble; Cannot calculate Density' + stop (:181-183) */
  }
  return 0;
}

/* ===================================================================== */
/* Flory-Huggins (Interaction_Modules/mod_FloryHuggins.f90)               */
/* ===================================================================== */

/* Calculate_Mean_Fields (:104-129) */
void orc_mean_fields(const orc_grid* g, const orc_fh* fh, const double* density, double* field) {
  long ng = (long)g->nx * g->ny * g->nz;
  int nc = fh->ncomp;
  for (long i = 0; i < ng; ++i) {
    double vf[16], tot = 0.0;
    for (int c = 0; c < nc; ++c) vf[c] = density[c * ng + i] * fh->volume[c]; /* :113-115 */
    for (int c = 0; c < nc; ++c) tot += vf[c];                                  /* :117 */
    for (int c = 0; c < nc; ++c) {
      double f = fh->kappa * (tot - 1.0);                                       /* :121 */
      for (int c2 = 0; c2 < nc; ++c2) f = f + fh->chi[c + c2 * nc] * vf[c2];   /* :122-125 */
      field[c * ng + i] = f * fh->volume[c];                                    /* :126 */
    }
  }
}

/* Interaction_Energy (:74-96) */
double orc_interaction_energy(const orc_grid* g, const orc_fh* fh, const double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  int nc = fh->ncomp;
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double s = 0.0;
  for (long i = 0; i < ng; ++i) {
    double tot = 0.0;
    for (int c = 0; c < nc; ++c) tot += density[c * ng + i] * fh->volume[c];
    s += (tot - 1.0) * (tot - 1.0);
  }
  double e = 0.5 * fh->kappa * dvol * s; /* :87 */
  for (int c = 0; c < nc; ++c)
    for (int c2 = c + 1; c2 < nc; ++c2) {
      double p = 0.0;
      for (long i = 0; i < ng; ++i) p += density[c * ng + i] * fh->volume[c] * density[c2 * ng + i] * fh->volume[c2];
      e = e + fh->chi[c + c2 * nc] * dvol * p; /* :91-92 */
    }
  return e;
}

/* Check_Volume_Fraction (:139-160): the average total volume fraction */
double orc_average_volume_fraction(const orc_grid* g, const orc_fh* fh, const double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  double s = 0.0;
  for (int c = 0; c < fh->ncomp; ++c)
    for (long i = 0; i < ng; ++i) s += density[c * ng + i] * fh->volume[c];
  return s / (double)ng;
}

/* ===================================================================== */
/* Iterate (General_Routines/mod_iterate.f90)                             */
/* ===================================================================== */

#define COL(a, j) ((a) + (long)(j) * fieldsize)

/* Mixing_Simple (:27-38) */
void orc_mixing_simple(double* field, const double* dfield, long fieldsize, int m, double lambda) {
  (void)m;
  for (long i = 0; i < fieldsize; ++i) field[i] = field[i] + lambda * dfield[i];
}

/* Mixing_Lambda (:46-84) */
void orc_mixing_lambda(double* field, double* dfield, long fieldsize, int m, double lammax, int iter) {
  const double lammin = 0.00001;
  double lambda;
  (void)m;
  if (iter == 1) {
    lambda = lammin;
  } else {
    double num = 0.0, nom = 0.0;
    for (long i = 0; i < fieldsize; ++i) {
      double a = COL(field, 0)[i] - COL(field, 1)[i];
      double b = COL(dfield, 0)[i] - COL(dfield, 1)[i];
      num += a * a;
      nom += b * b;
    }
    if (nom > 1.0e-50) {
      lambda = sqrt(num / nom);
      if (lambda < lammin) lambda = lammin;
      else if (lambda > lammax) lambda = lammax;
    } else {
      lambda = lammin;
    }
  }
  memcpy(COL(field, 1), COL(field, 0), sizeof(double) * (size_t)fieldsize);
  memcpy(COL(dfield, 1), COL(dfield, 0), sizeof(double) * (size_t)fieldsize);
  for (long i = 0; i < fieldsize; ++i) field[i] = field[i] + lambda * dfield[i];
}

/* Mixing_Anderson (:92-150) */
void orc_mixing_anderson(double* field, double* dfield, long fieldsize, int m, double lam_mix, int iter) {
  const double lam00 = 0.001; /* :94 */
  double* fmix = (double*)malloc(sizeof(double) * (size_t)fieldsize);
  double* dmix = (double*)malloc(sizeof(double) * (size_t)fieldsize);
  double lambda;
  if (iter > m) {
    double* ddiff = (double*)malloc(sizeof(double) * (size_t)fieldsize * (size_t)m);
    double* u = (double*)calloc((size_t)m * (size_t)m, sizeof(double));
    double* v = (double*)calloc((size_t)m, sizeof(double));
    double* x = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < m; ++i) /* :112-114 */
      for (long k = 0; k < fieldsize; ++k) ddiff[(long)i * fieldsize + k] = COL(dfield, 0)[k] - COL(dfield, i + 1)[k];
    for (int i = 0; i < m; ++i) { /* :116-121 */
      double s = 0.0;
      for (long k = 0; k < fieldsize; ++k) s += ddiff[(long)i * fieldsize + k] * COL(dfield, 0)[k];
      v[i] = s;
      for (int j = 0; j < m; ++j) {
        double t = 0.0;
        for (long k = 0; k < fieldsize; ++k) t += ddiff[(long)i * fieldsize + k] * ddiff[(long)j * fieldsize + k];
        u[i + (long)j * m] = t;
      }
    }
    orc_solve_lu(u, v, x, m); /* :123 */
    memcpy(fmix, COL(field, 0), sizeof(double) * (size_t)fieldsize);
    memcpy(dmix, COL(dfield, 0), sizeof(double) * (size_t)fieldsize);
    for (int i = 0; i < m; ++i) /* :127-130 */
      for (long k = 0; k < fieldsize; ++k) {
        fmix[k] = fmix[k] - x[i] * (COL(field, 0)[k] - COL(field, i + 1)[k]);
        dmix[k] = dmix[k] - x[i] * ddiff[(long)i * fieldsize + k];
      }
    lambda = lam_mix;
    free(x); free(v); free(u); free(ddiff);
  } else {
    memcpy(fmix, COL(field, 0), sizeof(double) * (size_t)fieldsize);
    memcpy(dmix, COL(dfield, 0), sizeof(double) * (size_t)fieldsize);
    lambda = lam00;
  }
  for (int i = m; i >= 1; --i) { /* :143-146 */
    memcpy(COL(field, i), COL(field, i - 1), sizeof(double) * (size_t)fieldsize);
    memcpy(COL(dfield, i), COL(dfield, i - 1), sizeof(double) * (size_t)fieldsize);
  }
  for (long k = 0; k < fieldsize; ++k) field[k] = fmix[k] + lambda * dmix[k]; /* :148 */
  free(dmix);
  free(fmix);
}

/* ludcmp: Crout LU with implicit-scaling partial pivoting (:232-298).  a is n x n column-major. */
int orc_ludcmp(double* a, int n, int* indx, double* d) {
#define A(i, j) a[(long)((i) - 1) + (long)((j) - 1) * n]
  const double tiny = 1.0e-20;
  int singular = 0, imax = 0;
  double* vv = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  *d = 1.0;
  for (int i = 1; i <= n; ++i) {
    double aamax = 0.0;
    for (int j = 1; j <= n; ++j)
      if (fabs(A(i, j)) > aamax) aamax = fabs(A(i, j));
    if (aamax == 0.0) singular = 1; /* 'singular matrix in ludcmp' (:251) */
    vv[i] = 1.0 / aamax;
  }
  for (int j = 1; j <= n; ++j) {
    for (int i = 1; i <= j - 1; ++i) {
      double s = A(i, j);
      for (int k = 1; k <= i - 1; ++k) s = s - A(i, k) * A(k, j);
      A(i, j) = s;
    }
    double aamax = 0.0;
    for (int i = j; i <= n; ++i) {
      double s = A(i, j);
      for (int k = 1; k <= j - 1; ++k) s = s - A(i, k) * A(k, j);
      A(i, j) = s;
      double dum = vv[i] * fabs(s);
      if (dum >= aamax) { /* :272 '.ge.' keeps the last maximum */
        imax = i;
        aamax = dum;
      }
    }
    if (j != imax) {
      for (int k = 1; k <= n; ++k) {
        double dum = A(imax, k);
        A(imax, k) = A(j, k);
        A(j, k) = dum;
      }
      *d = -*d;
      vv[imax] = vv[j];
    }
    indx[j - 1] = imax;
    if (A(j, j) == 0.0) A(j, j) = tiny; /* :287 */
    if (j != n) {
      double dum = 1.0 / A(j, j);
      for (int i = j + 1; i <= n; ++i) A(i, j) = A(i, j) * dum;
    }
  }
  free(vv);
  return singular;
#undef A
}

/* lubksb (:302-337) */
void orc_lubksb(const double* a, int n, const int* indx, double* b) {
#define A(i, j) a[(long)((i) - 1) + (long)((j) - 1) * n]
  int ii = 0;
  for (int i = 1; i <= n; ++i) {
    int ll = indx[i - 1];
    double s = b[ll - 1];
    b[ll - 1] = b[i - 1];
    if (ii != 0) {
      for (int j = ii; j <= i - 1; ++j) s = s - A(i, j) * b[j - 1];
    } else if (s != 0.0) {
      ii = i;
    }
    b[i - 1] = s;
  }
  for (int i = n; i >= 1; --i) {
    double s = b[i - 1];
    for (int j = i + 1; j <= n; ++j) s = s - A(i, j) * b[j - 1];
    b[i - 1] = s / A(i, i);
  }
#undef A
}

/* Solve_LinearSystem_LU (:203-224) */
void orc_solve_lu(const double* Am, const double* B, double* X, int n) {
  double* aa = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
  int* indx = (int*)malloc(sizeof(int) * (size_t)n);
  double d;
  memcpy(X, B, sizeof(double) * (size_t)n);
  memcpy(aa, Am, sizeof(double) * (size_t)n * (size_t)n);
  orc_ludcmp(aa, n, indx, &d);
  orc_lubksb(aa, n, indx, X);
  free(indx);
  free(aa);
}

/* Solve_LinearSystem_Cholesky (:160-195) */
void orc_solve_cholesky(const double* Am, const double* B, double* X, int n) {
#define AA(i, j) Am[(long)((i) - 1) + (long)((j) - 1) * n]
#define L(i, j) l[(long)((i) - 1) + (long)((j) - 1) * n]
  double* l = (double*)calloc((size_t)n * (size_t)n, sizeof(double));
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  L(1, 1) = sqrt(AA(1, 1));
  for (int i = 2; i <= n; ++i) {
    for (int j = 1; j <= i - 1; ++j) {
      double s = 0.0;
      for (int k = 1; k <= j - 1; ++k) s += L(i, k) * L(j, k);
      L(i, j) = (AA(i, j) - s) / L(j, j);
    }
    double s = 0.0;
    for (int k = 1; k <= i - 1; ++k) s += L(i, k) * L(i, k);
    L(i, i) = sqrt(AA(i, i) - s);
  }
  for (int i = 1; i <= n; ++i) {
    double s = 0.0;
    for (int k = 1; k <= i - 1; ++k) s += L(i, k) * y[k - 1];
    y[i - 1] = (B[i - 1] - s) / L(i, i);
  }
  for (int j = n; j >= 1; --j) {
    double s = 0.0;
    for (int k = j + 1; k <= n; ++k) s += L(k, j) * X[k - 1];
    X[j - 1] = (y[j - 1] - s) / L(j, j);
  }
  free(y);
  free(l);
#undef L
#undef AA
}

/* ===================================================================== */
/* SCF drivers                                                            */
/* ===================================================================== */

/* Calculate_Densities (main_multiblock.f90:514-540) */
int orc_species_densities(const orc_grid* g, orc_chain* c, int ntypes, const double* field, double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  double* E = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)ntypes);
  double* dist = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)c->nblocks);
  double* bden = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)c->nblocks);
  orc_calculate_expfield(g, ntypes, field, E);
  int rc = orc_propagate_chain(g, c, E, dist, bden, NULL, NULL);
  memset(density, 0, sizeof(double) * (size_t)ng * (size_t)ntypes);
  for (int b = 0; b < c->nblocks; ++b) {
    double* d = density + (long)(c->monomer_type[b] - 1) * ng;
    const double* s = bden + (long)b * ng;
    for (long i = 0; i < ng; ++i) d[i] = d[i] + s[i];
  }
  free(bden);
  free(dist);
  free(E);
  return rc;
}

/* Calculate_Energy (main_multiblock.f90:548-557) */
double orc_total_energy(const orc_grid* g, const orc_chain* c, const orc_fh* fh, int ntypes, const double* density,
                        const double* field) {
  long n = (long)g->nx * g->ny * g->nz * ntypes;
  double volume, dvol, s = 0.0;
  orc_volume_parameters(g, &volume, &dvol);
  for (long i = 0; i < n; ++i) s += density[i] * field[i];
  double chain_energy = c->energy - dvol * s;
  return chain_energy + orc_interaction_energy(g, fh, density);
}

/* Find_SCF_Saddle_Point (main_multiblock.f90:648-700) */
int orc_scf_anderson(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* field, double* density,
                     int m, double lam, double tol, int maxit, double* accuracy, double* residual_history) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  double* var = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));  /* SAVE'd, zero-initialised */
  double* dvar = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* newfield = (double*)malloc(sizeof(double) * (size_t)nvar);
  double acc = 0.0;
  int iter;
  for (iter = 1; iter <= maxit; ++iter) {
    if (orc_species_densities(g, c, ntypes, field, density) != 0) { iter = -1; break; } /* :673 */
    orc_mean_fields(g, fh, density, newfield);                                           /* :674 */
    acc = 0.0;
    for (long i = 0; i < nvar; ++i) {                                                    /* :675-676, 780 */
      double df = newfield[i] - field[i];
      dvar[i] = df;
      double a = fabs(df);
      if (a > acc || isnan(a)) acc = isnan(acc) ? acc : a;
    }
    if (residual_history) residual_history[iter - 1] = acc;
    if (!(acc > tol)) break;                                                             /* :681 */
    int iter_status = iter < 1000 ? iter : 1000;                                          /* :686 */
    memcpy(var, field, sizeof(double) * (size_t)nvar);                                   /* :725-736 */
    orc_mixing_anderson(var, dvar, nvar, m, lam, iter_status);                          /* :744 */
    memcpy(field, var, sizeof(double) * (size_t)nvar);                                   /* :750-761 */
  }
  *accuracy = acc;
  free(newfield);
  free(dvar);
  free(var);
  return iter;
}

/* ======================================================================
 * DDFT: Main_Multiblock/main_multiblock.f90
 * ====================================================================== */

/* main_multiblock.f90:486-504 (the table part of Read_Mobilities, 441-506) */
void orc_mobility_qq(const orc_grid* g, int nt, int nq, double dq, const double* table, double wcutoff,
                     double* mobility_qq) {
  int nh = g->nx / 2 + 1;
  long nk = (long)nh * g->ny * g->nz;
  double* lap = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  orc_calculate_laplace(g, lap, NULL, NULL);
  for (long i = 0; i < nk; ++i) {
    double qq = -lap[2 * i];  /* real(-laplace(x,y,z)) (:489) */
    double q = sqrt(qq);      /* :490 */
    int iq = (int)(q / dq);   /* int(): truncation (:491) */
    for (int t2 = 0; t2 < nt; ++t2)
      for (int t1 = 0; t1 < nt; ++t1) {
        const double* col = table + (long)(t1 + t2 * nt) * nq; /* mobility_table(:, t1, t2) */
        double mq;
        if (iq == 0)
          mq = col[0] * q / dq;                                 /* :493 */
        else if (iq >= nq)
          mq = col[nq - 1];                                     /* :495 */
        else                                                    /* :497-499, the reference's weights */
          mq = (col[iq - 1] * (q - (double)iq * dq) + col[iq] * ((double)(iq + 1) * dq - q)) / dq;
        double v = qq * mq;                                     /* :501 */
        mobility_qq[(long)(t1 + t2 * nt) * nk + i] = v < wcutoff ? v : wcutoff;
      }
  }
  free(lap);
}

/* main_multiblock.f90:566-640 (measure 0) and main_blend.f90:376-428 (measure 1) */
int orc_conjugate_fields(const orc_grid* g, orc_chain* c, int ntypes, const double* density, double* field, int m,
                         double lam, double thresh, int maxit, int measure, double* accuracy) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double* var = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* dvar = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* newdensity = (double*)malloc(sizeof(double) * (size_t)nvar);
  double acc = 0.0;
  int iter;
  for (iter = 1; iter <= maxit; ++iter) {
    if (orc_species_densities(g, c, ntypes, field, newdensity) != 0) { iter = -1; break; } /* :602 */
    acc = 0.0;
    for (long i = 0; i < nvar; ++i) {                                                       /* :603-604, 780 / main_blend.f90:567 */
      double df = newdensity[i] - density[i];
      dvar[i] = df;
      double a = fabs(df);
      if (measure == 0) {
        if (a > acc || isnan(a)) acc = isnan(acc) ? acc : a;
      } else {
        acc += a;
      }
    }
    if (measure != 0) acc = acc * dvol;
    if (!(acc > thresh)) break;                                                             /* :609 */
    int iter_status = iter < m + 1 ? iter : m + 1;                                          /* :614 */
    memcpy(var, field, sizeof(double) * (size_t)nvar);
    orc_mixing_anderson(var, dvar, nvar, m, lam, iter_status);                              /* :615 */
    memcpy(field, var, sizeof(double) * (size_t)nvar);
  }
  *accuracy = acc;
  free(newdensity);
  free(dvar);
  free(var);
  return iter;
}

/* main_multiblock.f90:619-634 */
int orc_gcfstat(double accuracy, double accuracy_scf) {
  if (!(accuracy < accuracy_scf * 0.5)) {
    if (!(accuracy < accuracy_scf)) return (!(accuracy < 0.1)) ? 3 : 2;
    return 1;
  }
  return 0;
}

/* main_multiblock.f90:794-880 */
int orc_ddft_step(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* density, double* field,
                  const double* mobility_qq, double timestep, int qdmin, double accuracy_scf, int itermax_scf, int m,
                  int* gcfstat, int* iterations) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  int nh = g->nx / 2 + 1;
  long nk = (long)nh * g->ny * g->nz;
  double acc;
  int it, st;
  it = orc_conjugate_fields(g, c, ntypes, density, field, m, 0.5, accuracy_scf * 0.5, itermax_scf, 0, &acc); /* :823 */
  if (it < 0) return -1;
  st = orc_gcfstat(acc, accuracy_scf);
  if (iterations) iterations[0] = it;
  if (gcfstat) gcfstat[0] = st;
  if (st == 3) return 3;
  double* newfield = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* mulocal = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* muq = (double*)malloc(sizeof(double) * 2 * (size_t)nk * (size_t)ntypes);
  double* ddq = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  double* dd = (double*)malloc(sizeof(double) * (size_t)nvar);
  orc_mean_fields(g, fh, density, newfield);                                 /* :824 */
  for (long i = 0; i < nvar; ++i) mulocal[i] = newfield[i] - field[i];       /* :825 */
  for (int t = 0; t < ntypes; ++t)                                           /* :834-839 */
    orc_real2fourier(g, mulocal + (long)t * ng, muq + 2 * (long)t * nk);
  for (int t1 = 0; t1 < ntypes; ++t1) {                                      /* :843-850 */
    for (long i = 0; i < 2 * nk; ++i) ddq[i] = 0.0;
    for (int t2 = 0; t2 < ntypes; ++t2) {
      const double* mob = mobility_qq + (long)(t2 + t1 * ntypes) * nk;       /* mobility_qq(:,:,:,type2,type1) */
      const double* mu = muq + 2 * (long)t2 * nk;
      for (long i = 0; i < nk; ++i) {
        ddq[2 * i] = ddq[2 * i] - mob[i] * mu[2 * i];
        ddq[2 * i + 1] = ddq[2 * i + 1] - mob[i] * mu[2 * i + 1];
      }
    }
    orc_fourier2real(g, ddq, dd + (long)t1 * ng);                            /* :854-858 */
  }
  for (long i = 0; i < nvar; ++i) density[i] = density[i] + timestep * dd[i]; /* :866 */
  if (qdmin) {                                                                /* :865-871 */
    const double dsmall = 1.0e-4;
    double dum = density[0];
    for (long i = 1; i < nvar; ++i) dum = density[i] < dum ? density[i] : dum;
    if (dum <= dsmall)
      for (long i = 0; i < nvar; ++i) density[i] = density[i] > dsmall ? density[i] : dsmall;
  }
  free(dd);
  free(ddq);
  free(muq);
  free(mulocal);
  free(newfield);
  it = orc_conjugate_fields(g, c, ntypes, density, field, m, 0.5, accuracy_scf * 0.5, itermax_scf, 0, &acc); /* :873 */
  if (it < 0) return -1;
  st = orc_gcfstat(acc, accuracy_scf);
  if (iterations) iterations[1] = it;
  if (gcfstat) gcfstat[1] = st;
  if (st == 3) return 3;
  if (qdmin && st >= 1)                                                       /* :874-876 */
    if (orc_species_densities(g, c, ntypes, field, density) != 0) return -1;
  return 0;
}

/* ======================================================================
 * Homopolymer blend and EPD: Main_Blend/main_blend.f90
 * ====================================================================== */

/* main_blend.f90:321-345 */
int orc_blend_densities(const orc_grid* g, orc_blend* b, int ntypes, const double* field, double* density) {
  long ng = (long)g->nx * g->ny * g->nz;
  double* E = (double*)malloc(sizeof(double) * (size_t)ng * (size_t)ntypes);
  double* dist = (double*)malloc(sizeof(double) * (size_t)ng);
  double* den = (double*)malloc(sizeof(double) * (size_t)ng);
  int rc = 0;
  orc_calculate_expfield(g, ntypes, field, E);                                  /* :333 */
  memset(density, 0, sizeof(double) * (size_t)ng * (size_t)ntypes);             /* :338 */
  for (int i = 0; i < b->npol && rc == 0; ++i) {                                /* :334-336 */
    orc_chain c = {1, &b->types[i], &b->segs[i], 1, b->rho[i], 0.0, 0.0, 0.0};
    rc = orc_propagate_homopolymer(g, &c, E, dist, den);
    b->mu[i] = c.mu;
    b->energy[i] = c.energy;
    double* d = density + (long)(b->types[i] - 1) * ng;                         /* :339-342 */
    for (long k = 0; k < ng; ++k) d[k] = d[k] + den[k];
  }
  free(den);
  free(dist);
  free(E);
  return rc;
}

/* main_blend.f90:376-428 with Mix_Fields (:490-554) and Check_Accuracy (:561-569) */
int orc_blend_conjugate_fields(const orc_grid* g, orc_blend* b, int ntypes, const double* density, double* field,
                               int m, double lam, double thresh, int maxit, double* accuracy) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double* var = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* dvar = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* newdensity = (double*)malloc(sizeof(double) * (size_t)nvar);
  double acc = 0.0;
  int iter;
  for (iter = 1; iter <= maxit; ++iter) {
    if (orc_blend_densities(g, b, ntypes, field, newdensity) != 0) { iter = -1; break; } /* :405 */
    acc = 0.0;
    for (long i = 0; i < nvar; ++i) {                                                   /* :406-407 */
      double df = newdensity[i] - density[i];
      dvar[i] = df;
      acc += fabs(df);
    }
    acc = acc * dvol;                                                                   /* :567 */
    if (!(acc > thresh)) break;                                                         /* :412 */
    memcpy(var, field, sizeof(double) * (size_t)nvar);
    orc_mixing_anderson(var, dvar, nvar, m, lam, iter);                                 /* :418, :532 */
    memcpy(field, var, sizeof(double) * (size_t)nvar);
  }
  *accuracy = acc;
  free(newdensity);
  free(dvar);
  free(var);
  return iter;
}

/* main_blend.f90:437-486 with Mix_Fields (:494-553) and Check_Accuracy (:561-569) */
int orc_blend_scf(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field, double* density,
                  int m, int mixing_type, double lam, double thresh, int maxit, double* accuracy) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  double volume, dvol;
  orc_volume_parameters(g, &volume, &dvol);
  double* var = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* dvar = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* newfield = (double*)malloc(sizeof(double) * (size_t)nvar);
  double acc = 0.0;
  int iter;
  for (iter = 1; iter <= maxit; ++iter) {
    if (orc_blend_densities(g, b, ntypes, field, density) != 0) { iter = -1; break; } /* :460 */
    orc_mean_fields(g, fh, density, newfield);                                        /* :461 */
    acc = 0.0;
    for (long i = 0; i < nvar; ++i) {                                                 /* :462-463 */
      double df = newfield[i] - field[i];
      dvar[i] = df;
      acc += fabs(df);
    }
    acc = acc * dvol;
    if (!(acc > thresh)) break;                                                       /* :468 */
    memcpy(var, field, sizeof(double) * (size_t)nvar);
    if (mixing_type == 1)                                                             /* :527-532 */
      orc_mixing_simple(var, dvar, nvar, m, lam);
    else if (mixing_type == 2)
      orc_mixing_lambda(var, dvar, nvar, m, lam, iter);
    else
      orc_mixing_anderson(var, dvar, nvar, m, lam, iter);
    memcpy(field, var, sizeof(double) * (size_t)nvar);
  }
  *accuracy = acc;
  free(newfield);
  free(dvar);
  free(var);
  return iter;
}

/* nabla_a * spectrum, nabla pure imaginary (mod_fourier.f90:189-191): (0 + i b)(x + i y) */
static void mul_nabla(long nk, const double* nab, const double* in, double* out) {
  for (long i = 0; i < nk; ++i) {
    const double bb = nab[2 * i + 1], x = in[2 * i], y = in[2 * i + 1];
    out[2 * i] = 0.0 * x - bb * y;
    out[2 * i + 1] = 0.0 * y + bb * x;
  }
}

/* main_blend.f90:605-667 */
int orc_epd_step(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* density, double* field,
                 double timestep, int m, double accuracy_scf, int itermax, int* iterations) {
  /* eps = 0.0001 (:630) assigns a default-kind (single precision) literal to a double:
     eps is the float nearest 1e-4, 9.99999974737875e-5 */
  const double eps = (double)0.0001f;
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  int nh = g->nx / 2 + 1;
  long nk = (long)nh * g->ny * g->nz;
  double* nabla = (double*)malloc(sizeof(double) * 2 * (size_t)nk * 3);
  double* old = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* newfield = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* mulocal = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* dfield = (double*)malloc(sizeof(double) * (size_t)nvar * 3);
  double* dden = (double*)malloc(sizeof(double) * (size_t)nvar * 3);
  double* newden = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* f0 = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  double* f1 = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)ng);
  int rc = 0;
  orc_calculate_laplace(g, NULL, NULL, nabla);
  if (orc_blend_densities(g, b, ntypes, field, old) != 0) rc = -1;               /* :632 */
  if (rc == 0) {
    orc_mean_fields(g, fh, old, newfield);                                      /* :633 */
    for (long i = 0; i < nvar; ++i) mulocal[i] = newfield[i] - field[i];        /* :634 */
    for (int t = 0; t < ntypes; ++t) {                                          /* :636-644 */
      orc_real2fourier(g, mulocal + (long)t * ng, f0);
      for (int a = 0; a < 3; ++a) {
        mul_nabla(nk, nabla + 2 * (long)a * nk, f0, f1);
        orc_fourier2real(g, f1, dfield + ((long)a * ntypes + t) * ng);
      }
    }
    for (int a = 0; a < 3 && rc == 0; ++a) {                                     /* :646-657 */
      const double* dfa = dfield + (long)a * ntypes * ng;
      for (long i = 0; i < nvar; ++i) newfield[i] = field[i] - eps * dfa[i];
      if (orc_blend_densities(g, b, ntypes, newfield, newden) != 0) { rc = -1; break; }
      for (int t = 0; t < ntypes; ++t) {
        double* dd = dden + ((long)a * ntypes + t) * ng;
        for (long i = 0; i < ng; ++i) tmp[i] = newden[(long)t * ng + i] - old[(long)t * ng + i];
        orc_real2fourier(g, tmp, f0);
        mul_nabla(nk, nabla + 2 * (long)a * nk, f0, f1);
        orc_fourier2real(g, f1, dd);
      }
    }
  }
  if (rc == 0) {
    for (long i = 0; i < nvar; ++i) {                                           /* :662 */
      double sum = dden[i];
      sum = sum + dden[nvar + i];
      sum = sum + dden[2 * nvar + i];
      density[i] = density[i] + sum * timestep / eps;
    }
    double acc;
    int it = orc_blend_conjugate_fields(g, b, ntypes, density, field, m, 0.5, accuracy_scf, itermax, &acc); /* :663 */
    if (iterations) *iterations = it;
    if (it < 0) rc = -1;
  }
  free(tmp);
  free(f1);
  free(f0);
  free(newden);
  free(dden);
  free(dfield);
  free(mulocal);
  free(newfield);
  free(old);
  free(nabla);
  return rc;
}

/* Calculate_Field_Derivative (main_blend.f90:709-740): densities and mean fields of the
 * field, dfield = newfield - field, then per type the spectral product
 * dummyfourier = - laplace*dummyfourier (:735; Fortran's unary minus applies to the
 * product), back to real space. */
int orc_field_derivative(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, const double* field,
                         double* dfield) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  long nk = (long)(g->nx / 2 + 1) * g->ny * g->nz;
  double* den = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* newfield = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* lap = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  double* spec = (double*)malloc(sizeof(double) * 2 * (size_t)nk);
  int rc = orc_blend_densities(g, b, ntypes, field, den); /* :727 */
  if (rc == 0) {
    orc_mean_fields(g, fh, den, newfield);                                  /* :728 */
    for (long i = 0; i < nvar; ++i) dfield[i] = newfield[i] - field[i];    /* :729 */
    orc_calculate_laplace(g, lap, NULL, NULL);
    for (int t = 0; t < ntypes; ++t) {                                      /* :731-737 */
      orc_real2fourier(g, dfield + (long)t * ng, spec);
      for (long k = 0; k < nk; ++k) {
        const cplx p = (lap[2 * k] + lap[2 * k + 1] * I) * (spec[2 * k] + spec[2 * k + 1] * I);
        spec[2 * k] = -creal(p);
        spec[2 * k + 1] = -cimag(p);
      }
      orc_fourier2real(g, spec, dfield + (long)t * ng);
    }
  }
  free(spec);
  free(lap);
  free(newfield);
  free(den);
  return rc == 0 ? 0 : -1;
}

/* EPD_Timestep_Euler (main_blend.f90:580-596): field = field + timestep*dfield */
int orc_epd_timestep_euler(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field,
                           double timestep) {
  long nvar = (long)g->nx * g->ny * g->nz * ntypes;
  double* df = (double*)malloc(sizeof(double) * (size_t)nvar);
  int rc = orc_field_derivative(g, b, fh, ntypes, field, df);
  if (rc == 0)
    for (long i = 0; i < nvar; ++i) field[i] = field[i] + timestep * df[i];
  free(df);
  return rc;
}

/* EPD_Timestep_RungeKutta (main_blend.f90:678-701): the classical four-stage rule,
 * field += timestep*(df1/6 + df2/3 + df3/3 + df4/6) with the stages at field,
 * field + timestep/2 df1, field + timestep/2 df2 and field + timestep df3 */
int orc_epd_timestep_rk4(const orc_grid* g, orc_blend* b, const orc_fh* fh, int ntypes, double* field,
                         double timestep) {
  long nvar = (long)g->nx * g->ny * g->nz * ntypes;
  double* df = (double*)malloc(sizeof(double) * (size_t)nvar * 4);
  double* stage = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* d1 = df;
  double* d2 = df + nvar;
  double* d3 = df + 2 * nvar;
  double* d4 = df + 3 * nvar;
  int rc = orc_field_derivative(g, b, fh, ntypes, field, d1);                        /* :690 */
  if (rc == 0) {
    for (long i = 0; i < nvar; ++i) stage[i] = field[i] + 0.5 * timestep * d1[i];   /* :691 */
    rc = orc_field_derivative(g, b, fh, ntypes, stage, d2);                         /* :692 */
  }
  if (rc == 0) {
    for (long i = 0; i < nvar; ++i) stage[i] = field[i] + 0.5 * timestep * d2[i];   /* :693 */
    rc = orc_field_derivative(g, b, fh, ntypes, stage, d3);                         /* :694 */
  }
  if (rc == 0) {
    for (long i = 0; i < nvar; ++i) stage[i] = field[i] + timestep * d3[i];         /* :695 */
    rc = orc_field_derivative(g, b, fh, ntypes, stage, d4);                         /* :696 */
  }
  if (rc == 0)
    for (long i = 0; i < nvar; ++i)                                                 /* :698 */
      field[i] = field[i] + timestep * (d1[i] / 6.0 + d2[i] / 3.0 + d3[i] / 3.0 + d4[i] / 6.0);
  free(stage);
  free(df);
  return rc;
}

/* ======================================================================
 * Main_Diblock_Melt/main_diblock.f90:398-451
 * ====================================================================== */
int orc_diblock_saddle(const orc_grid* g, orc_chain* c, const orc_fh* fh, int ntypes, double* field, double* density,
                       int m, int mixing_type, double lam, double tol, int maxit, double* accuracy,
                       double* mixed_field) {
  long ng = (long)g->nx * g->ny * g->nz;
  long nvar = ng * ntypes;
  double* var = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* dvar = (double*)calloc((size_t)nvar * (size_t)(m + 1), sizeof(double));
  double* newdensity = (double*)malloc(sizeof(double) * (size_t)nvar);
  double* newfield = (double*)calloc((size_t)nvar, sizeof(double));
  double acc = 0.0;
  int iter;
  for (iter = 1; iter <= maxit; ++iter) {
    if (orc_species_densities(g, c, ntypes, field, newdensity) != 0) { iter = -1; break; } /* :412 */
    acc = 0.0;                                                                             /* :414, :387 */
    for (long i = 0; i < nvar; ++i) {
      double d = density[i] - newdensity[i];
      acc += d * d;
    }
    if (!(acc > tol)) break;                                                               /* :421 */
    orc_mean_fields(g, fh, newdensity, newfield);                                         /* :425 */
    for (long i = 0; i < nvar; ++i) dvar[i] = newfield[i] - field[i];                      /* :426 */
    memcpy(var, field, sizeof(double) * (size_t)nvar);                                     /* :427 */
    if (mixing_type == 1)
      orc_mixing_simple(var, dvar, nvar, m, lam);
    else if (mixing_type == 2)
      orc_mixing_lambda(var, dvar, nvar, m, lam, iter);
    else
      orc_mixing_anderson(var, dvar, nvar, m, lam, iter);
    memcpy(field, var, sizeof(double) * (size_t)nvar);
    memcpy(density, newdensity, sizeof(double) * (size_t)nvar);                            /* :428 */
  }
  *accuracy = acc;
  if (mixed_field) memcpy(mixed_field, field, sizeof(double) * (size_t)nvar);
  if (iter > 0) {
    memcpy(density, newdensity, sizeof(double) * (size_t)nvar);                            /* :447 */
    memcpy(field, newfield, sizeof(double) * (size_t)nvar);                                /* :448 */
  }
  free(newfield);
  free(newdensity);
  free(dvar);
  free(var);
  return iter;
}
