// Pointwise and reduction kernels of one SCF iteration: expfield, rescale
// schedule, Q, the trapezoid contour integral fused with ensemble
// normalisation, Flory-Huggins mean fields and the field residual, the
// Anderson Gram matrix, the LU solve and the mixing update.
//
// Reductions are deterministic: every block writes its partial result to a
// fixed slot and one block combines the slots in a fixed order.
#pragma once

#include "kernels.hpp"
#include "lu.hpp"

namespace adscf {

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// max that propagates NaN (a NaN residual must end the loop as the reference's
// .not.(accuracy > tol) test does)
__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = nanmax(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// E = exp(-0.5*ds*w) (mod_propagate.f90:28) for nt fields, plus per-block
// partial sums of E^2 per type (they set the rescale schedule).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_expfield(const double* __restrict__ w, T* __restrict__ E, long ng, int nt,
                                                       double ds, double* __restrict__ e2part, const int* done) {
  if (done && *done) return;
  __shared__ double sh[kThreads];
  for (int t = 0; t < nt; ++t) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
      const double e = exp(-0.5 * ds * w[t * ng + i]);
      const T et = (T)e;
      E[t * ng + i] = et;
      acc += (double)et * (double)et;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0 && e2part) e2part[blockIdx.x * kMaxTypes + t] = acc;
  }
}

// partial sums of E^2 per type for an E supplied by the caller
template <typename T>
__global__ void __launch_bounds__(kThreads) k_e2(const T* __restrict__ E, long ng, int nt, double* __restrict__ e2part) {
  __shared__ double sh[kThreads];
  for (int t = 0; t < nt; ++t) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
      const double e = (double)E[t * ng + i];
      acc += e * e;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) e2part[blockIdx.x * kMaxTypes + t] = acc;
  }
}

// Power-of-two rescale schedule.  A step of type t multiplies q by about
// mean(E_t^2) on average; e_c(s) = -round(sum of log2 of those factors) keeps
// the stored q'(s) = q(s) * 2^e_c(s) near one.  Powers of two scale every
// rounding step exactly, so the FP64 path computes the same bits as an
// unscaled one; in FP32 the two packed chains stay balanced.
// One block of kThreads: the per-type means come from the E^2 partials, then
// every thread fills contour indices s in closed form (the log2 sum is
// piecewise linear in s within a block).
__global__ void __launch_bounds__(kThreads) k_schedule(ChainDesc cd, const double* __restrict__ e2part, int nblk,
                                                       long ng, int nt, int* __restrict__ e1, int* __restrict__ e2,
                                                       int* __restrict__ d1, int* __restrict__ d2, ScfState* st,
                                                       int rotate, int m, const int* done) {
  if (done && *done) return;
  __shared__ double sh[kThreads];
  __shared__ double lg[kMaxTypes];
  __shared__ double sig0[2][kMaxBlocks];  // log2 sum at the start of each block, chain order
  __shared__ int s0c[2][kMaxBlocks];      // first contour index of each block, chain order
  for (int t = 0; t < nt; ++t) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc += e2part[b * kMaxTypes + t];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) {
      const double mean = acc / (double)ng;
      lg[t] = (mean > 0.0 && mean < 1e300) ? log2(mean) : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int c = threadIdx.x;
    double sig = 0.0;
    int s = 0;
    for (int bb = 0; bb < cd.nblocks; ++bb) {
      const int b = c == 0 ? bb : cd.nblocks - 1 - bb;
      sig0[c][bb] = sig;
      s0c[c][bb] = s;
      sig += cd.seg[b] * lg[cd.type[b]];
      s += cd.seg[b];
    }
  }
  __syncthreads();
  auto expo = [&](int c, int s) -> int {
    if (s == 0) return 0;
    int bb = cd.nblocks - 1;
    while (bb > 0 && s0c[c][bb] >= s) --bb;
    const int b = c == 0 ? bb : cd.nblocks - 1 - bb;
    double r = -rint(sig0[c][bb] + (s - s0c[c][bb]) * lg[cd.type[b]]);
    if (r > 900.0) r = 900.0;
    if (r < -900.0) r = -900.0;
    return (int)r;
  };
  for (int s = threadIdx.x; s <= cd.ns; s += blockDim.x) {
    const int a1 = expo(0, s), a2 = expo(1, s);
    e1[s] = a1;
    e2[s] = a2;
    d1[s] = s == 0 ? 0 : a1 - expo(0, s - 1);
    d2[s] = s == 0 ? 0 : a2 - expo(1, s - 1);
  }
  if (rotate && st && threadIdx.x == 0) st->head = (st->head + m) % (m + 1);
}

// Streaming read of n doubles (16-byte loads, two in flight per thread and step) into
// per-block sums: the read bandwidth probe of adscf_hbm_bandwidth.
__global__ void __launch_bounds__(kThreads) k_read_sum(const double* __restrict__ x, long n, double* __restrict__ part) {
  __shared__ double sh[kThreads];
  const double2* x2 = reinterpret_cast<const double2*>(x);
  const long n2 = n / 2, stride = (long)gridDim.x * blockDim.x;
  double acc = 0.0;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + stride < n2; i += 2 * stride) {
    const double2 u = x2[i], v = x2[i + stride];
    acc += (u.x + u.y) + (v.x + v.y);
  }
  for (; i < n2; i += stride) acc += x2[i].x + x2[i].y;
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Update_Solvent (mod_solvent.f90:97-98): distribution = exp(-field), partial sums
__global__ void __launch_bounds__(kThreads) k_solvent_dist(const double* __restrict__ w, long ng,
                                                           double* __restrict__ dist, double* __restrict__ part) {
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    const double d = exp(-w[i]);
    dist[i] = d;
    acc += d;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// density = f * distribution (mod_solvent.f90:100-106), with partial sums of the density
__global__ void __launch_bounds__(kThreads) k_scale_sum(const double* __restrict__ dist, long ng, double f,
                                                        double* __restrict__ den, double* __restrict__ part) {
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    const double d = f * dist[i];
    den[i] = d;
    acc += d;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// partial sums of q'(Ns) for Qsum (mod_multiblock.f90:155)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_qsum(const T* __restrict__ q, long ng, double* __restrict__ part,
                                                   const int* done) {
  if (done && *done) return;
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x)
    acc += (double)q[i];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Qsum and the canonical normalisation (mod_multiblock.f90:155, 173-176);
// grand canonical only needs exp(mu) here (178).
__global__ void k_qfinal(const double* __restrict__ part, int nblk, const int* __restrict__ e1, int ns, double dvol,
                         double volume, int ensemble, double rho_in, double mu_in, ScfState* st, const int* done) {
  if (done && *done) return;
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc += part[b];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double q = dvol * acc * ldexp(1.0, -e1[ns]);
    st->qsum = q;
    if (ensemble == 1) {
      st->rho = rho_in;
      st->norm = rho_in * volume / q;
      st->mu = log(rho_in * volume / q);
      st->energy = rho_in * volume * (st->mu - 1.0);
    } else {
      st->mu = mu_in;
      st->norm = exp(mu_in);
    }
  }
}

enum ContourMode : int { CONTOUR_BLOCKS = 0, CONTOUR_SCF = 1 };
enum ContourRule : int { RULE_TRAPEZOID = 0, RULE_SIMPSON = 1 };

// Weight of contour point s0 + j of a block of nb segments (0 <= j <= nb).
// RULE_TRAPEZOID is the reference's rule (mod_multiblock.f90:157-168): ds/2 at both ends
// and ds inside; an empty block (nb = 0) takes both halves at its one point.
// RULE_SIMPSON is composite Simpson: 1/3-rule panels of two intervals (ds/3 [1 4 2 ... 4 1]);
// an odd nb >= 3 ends with one 3/8-rule panel over the last three intervals
// (3 ds/8 [1 3 3 1]), so the rule stays exact for cubics in s; nb < 2 keeps the trapezoid.
// Both rules' weights sum to nb ds (nb = 0: ds), so the canonical mass is unchanged.
__host__ __device__ inline double contour_weight(int rule, int j, int nb, double ds) {
  if (rule != RULE_SIMPSON || nb < 2) {
    if (nb == 0) return ds;
    return (j == 0 || j == nb) ? 0.5 * ds : ds;
  }
  const int n13 = (nb & 1) ? nb - 3 : nb;  // intervals under 1/3-rule panels
  double w = 0.0;
  if (n13 > 0 && j <= n13) w += (j == 0 || j == n13) ? ds / 3.0 : ((j & 1) ? 4.0 * ds / 3.0 : 2.0 * ds / 3.0);
  if (nb & 1) {
    const int k = j - n13;
    if (k >= 0) w += (k == 0 || k == 3) ? 3.0 * ds / 8.0 : 9.0 * ds / 8.0;
  }
  return w;
}

template <typename T>
struct ContourArgs {
  long ng;
  ChainDesc cd;
  FHDesc fh;
  double ds;
  const T* h1;        // q'(s) slices, (Ns+1) x ng
  const T* h2;        // q+'(s) slices
  const int* e1;      // rescale exponents of q' and q+'
  const int* e2;
  const ScfState* st; // norm
  int mode;
  int rule;           // ContourRule
  double* dist;       // CONTOUR_BLOCKS: (nb, ng) or null
  double* dens;       // CONTOUR_BLOCKS: (nb, ng) or null
  const double* w0;   // CONTOUR_SCF: field ring base; slot `head` is the current field (nt, ng)
  double* phi;        // CONTOUR_SCF: species densities (nt, ng)
  double* dring;      // CONTOUR_SCF: residual history ring (m+1, nt*ng) or null
  int m;
  const double* target;  // CONTOUR_SCF: null -> dfield = mean field - field; else dfield = phi - target
  int measure;        // 0: per-block max|dfield|; 1: sum|dfield|; 2: sum (phi_old - phi)^2
  const ScfState* rst;  // CONTOUR_SCF: state holding the ring head (null: st)
  int accumulate;     // CONTOUR_SCF: add this chain's densities to phi (blends: one chain per polymer)
  int no_resid;       // CONTOUR_SCF: densities only (a blend's chains before the last)
  double* rpart;      // per-block max (measure 0) or sum (measure 1) of |dfield|
  double* spart;      // per-block sum of densities (grand canonical rho)
  const int* done;
};

template <typename T, int VP>
struct alignas(sizeof(T) * VP) PackT {
  T v[VP];
};

// Trapezoid contour integral (mod_multiblock.f90:157-168) in the reference's
// summation order (or, with RULE_SIMPSON, the weights of contour_weight in contour
// order), then the ensemble normalisation (173-180).  In SCF mode the
// species densities (main_multiblock.f90:530-534) and the residual follow in the
// same pass: either the mean fields (mod_FloryHuggins.f90:104-129) and
// dfield = new - field (main_multiblock.f90:674-676), or, for the conjugate-field
// solve, dfield = newdensity - density (main_multiblock.f90:602-603); then
// max|dfield| (main_multiblock.f90:780) or sum|dfield| (main_blend.f90:567).
// Each thread owns VP consecutive points (16-byte loads; ng % VP == 0) and
// loads kUnroll contour slices ahead of the sum, which keeps the stream of
// slice reads deep enough to run at HBM rate.
constexpr int kContourUnroll = 8;

template <typename T, int VP>
__global__ void __launch_bounds__(kThreads) k_contour(ContourArgs<T> a) {
  if (a.done && *a.done) return;
  using P = PackT<T, VP>;
  constexpr int U = kContourUnroll;
  __shared__ double sh[kThreads];
  const long ng = a.ng;
  const int ns = a.cd.ns;
  const double ds = a.ds, hds = 0.5 * a.ds;
  const double norm = a.st->norm;
  double rmax = 0.0, dsum = 0.0;
  // nontemporal loads: each slice is read once, as a stream (the slices of a 64^3 FP32
  // sweep, ~400 MB, exceed what the caches could keep)
  typedef T vecT __attribute__((ext_vector_type(VP)));
  auto ld = [&](const T* base, int s, long i0) {
    const vecT x = __builtin_nontemporal_load(reinterpret_cast<const vecT*>(base + (long)s * ng + i0));
    P r;
#pragma unroll
    for (int v = 0; v < VP; ++v) r.v[v] = x[v];
    return r;
  };
  auto scl = [&](int s) { return ldexp(1.0, -(a.e1[s] + a.e2[ns - s])); };
  const long npk = ng / VP;
  for (long ip = blockIdx.x * (long)blockDim.x + threadIdx.x; ip < npk; ip += (long)gridDim.x * blockDim.x) {
    const long i0 = ip * VP;
    double phi[kMaxTypes][VP];
    for (int t = 0; t < a.fh.nt; ++t)
#pragma unroll
      for (int v = 0; v < VP; ++v) phi[t][v] = a.accumulate ? a.phi[(long)t * ng + i0 + v] : 0.0;
    for (int b = 0; b < a.cd.nblocks; ++b) {
      const int s0 = a.cd.start[b], nb = a.cd.seg[b];
      double acc[VP];
      if (a.rule == RULE_SIMPSON && nb >= 2) {
#pragma unroll
        for (int v = 0; v < VP; ++v) acc[v] = 0.0;
        int j = 0;
        for (; j + U <= nb + 1; j += U) {
          P x1[U], x2[U];
          double sc[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            x1[u] = ld(a.h1, s0 + j + u, i0);
            x2[u] = ld(a.h2, ns - s0 - j - u, i0);
            sc[u] = scl(s0 + j + u) * contour_weight(RULE_SIMPSON, j + u, nb, ds);
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VP; ++v) acc[v] = acc[v] + ((double)x1[u].v[v] * (double)x2[u].v[v]) * sc[u];
        }
        for (; j <= nb; ++j) {
          const P x1 = ld(a.h1, s0 + j, i0), x2 = ld(a.h2, ns - s0 - j, i0);
          const double sc = scl(s0 + j) * contour_weight(RULE_SIMPSON, j, nb, ds);
#pragma unroll
          for (int v = 0; v < VP; ++v) acc[v] = acc[v] + ((double)x1.v[v] * (double)x2.v[v]) * sc;
        }
      } else {
        {
          const P x1 = ld(a.h1, s0, i0), x2 = ld(a.h2, ns - s0, i0);
          const double sc = scl(s0);
#pragma unroll
          for (int v = 0; v < VP; ++v) acc[v] = ((hds * (double)x1.v[v]) * (double)x2.v[v]) * sc;
        }
        int s = s0 + 1;
        for (; s + U <= s0 + nb; s += U) {
          P x1[U], x2[U];
          double sc[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            x1[u] = ld(a.h1, s + u, i0);
            x2[u] = ld(a.h2, ns - s - u, i0);
            sc[u] = scl(s + u);
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < VP; ++v) acc[v] = acc[v] + ((ds * (double)x1[u].v[v]) * (double)x2[u].v[v]) * sc[u];
        }
        for (; s <= s0 + nb - 1; ++s) {
          const P x1 = ld(a.h1, s, i0), x2 = ld(a.h2, ns - s, i0);
          const double sc = scl(s);
#pragma unroll
          for (int v = 0; v < VP; ++v) acc[v] = acc[v] + ((ds * (double)x1.v[v]) * (double)x2.v[v]) * sc;
        }
        {
          const int se = s0 + nb;
          const P x1 = ld(a.h1, se, i0), x2 = ld(a.h2, ns - se, i0);
          const double sc = scl(se);
#pragma unroll
          for (int v = 0; v < VP; ++v) acc[v] = acc[v] + ((hds * (double)x1.v[v]) * (double)x2.v[v]) * sc;
        }
      }
#pragma unroll
      for (int v = 0; v < VP; ++v) {
        const long i = i0 + v;
        const double den = norm * acc[v];
        dsum += den;
        if (a.mode == CONTOUR_BLOCKS) {
          if (a.dist) a.dist[(long)b * ng + i] = acc[v];
          if (a.dens) a.dens[(long)b * ng + i] = den;
        } else {
          phi[a.cd.type[b]][v] += den;
        }
      }
    }
    if (a.mode == CONTOUR_SCF && a.no_resid) {
      for (int t = 0; t < a.fh.nt; ++t)
#pragma unroll
        for (int v = 0; v < VP; ++v) a.phi[(long)t * ng + i0 + v] = phi[t][v];
    } else if (a.mode == CONTOUR_SCF) {
      const int nt = a.fh.nt;
      const int head = (a.rst ? a.rst : a.st)->head;
      double* d0 = a.dring ? a.dring + (long)head * nt * ng : nullptr;
      const double* w0 = a.w0 + (long)head * nt * ng;  // current field: ring slot `head`
#pragma unroll
      for (int v = 0; v < VP; ++v) {
        const long i = i0 + v;
        if (a.measure == 2) {  // sum((density - newdensity)**2) against phi's previous contents
          for (int t = 0; t < nt; ++t) {
            const double dd = a.phi[(long)t * ng + i] - phi[t][v];
            rmax = rmax + dd * dd;
          }
        }
        for (int t = 0; t < nt; ++t) a.phi[(long)t * ng + i] = phi[t][v];
        if (a.target) {
          for (int t = 0; t < nt; ++t) {
            const double df = phi[t][v] - a.target[(long)t * ng + i];
            if (d0) d0[(long)t * ng + i] = df;
            if (a.measure != 2) rmax = a.measure ? rmax + fabs(df) : nanmax(rmax, fabs(df));
          }
          continue;
        }
        double vf[kMaxTypes], tot = 0.0;
        for (int t = 0; t < nt; ++t) vf[t] = phi[t][v] * a.fh.vol[t];
        for (int t = 0; t < nt; ++t) tot += vf[t];
        for (int t = 0; t < nt; ++t) {
          double f = a.fh.kappa * (tot - 1.0);
          for (int t2 = 0; t2 < nt; ++t2) f = f + a.fh.chi[t + t2 * nt] * vf[t2];
          f = f * a.fh.vol[t];
          const double df = f - w0[(long)t * ng + i];
          if (d0) d0[(long)t * ng + i] = df;
          if (a.measure != 2) rmax = a.measure ? rmax + fabs(df) : nanmax(rmax, fabs(df));
        }
      }
    }
  }
  rmax = a.measure ? block_sum(rmax, sh) : block_max(rmax, sh);
  dsum = block_sum(dsum, sh);
  if (threadIdx.x == 0) {
    if (a.rpart) a.rpart[blockIdx.x] = rmax;
    if (a.spart) a.spart[blockIdx.x] = dsum;
  }
}

// Anderson Gram sums (mod_iterate.f90:112-121) over the history ring:
// y_0 = D_0, y_i = D_0 - D_i (i = 1..m); every pair (i <= j) of
// <y_i, y_j> is accumulated, giving v_i = <y_i, y_0> and U_ij = <y_i, y_j>.
constexpr int kGramTile = 64;
constexpr int kGramBlocks = 1024;  // partial-sum slots of k_gram

// Partial Gram sums: block b covers tiles b, b + gridDim.x, ... and writes its
// np pair sums to part[p * gridDim.x + b] (pair-major, so k_gram_reduce reads
// them contiguously).  Pairs are register-blocked 3 x 3: a thread owns one block
// (bi <= bj) of the (m+1) x (m+1) pair matrix and one of R sub-ranges of the
// tile's elements, so each LDS value it reads feeds three products.
constexpr int kGramBlk = 3;

__host__ __device__ constexpr int gram_pair_blocks(int m) {
  return ((m + 1 + kGramBlk - 1) / kGramBlk) * ((m + 1 + kGramBlk - 1) / kGramBlk + 1) / 2;
}
// sub-ranges per tile: the largest power of two with R * blocks <= kThreads and R <= kGramTile
__host__ __device__ constexpr int gram_subranges(int m) {
  int r = 1;
  while (2 * r * gram_pair_blocks(m) <= kThreads && 2 * r <= kGramTile) r *= 2;
  return r;
}
// dynamic LDS of k_gram in doubles: the tile, then the per-thread block sums
__host__ __device__ constexpr int gram_lds_doubles(int m) {
  return (m + 1) * (kGramTile + 1) + gram_pair_blocks(m) * gram_subranges(m) * kGramBlk * kGramBlk;
}

__global__ void __launch_bounds__(kThreads) k_gram(const double* __restrict__ dring, long nvar, int m,
                                                   const ScfState* st, int force, double* __restrict__ part,
                                                   const int* done) {
  if (done && *done) return;
  if (!force) {
    const int it = st->iter + 1;
    if ((it < 1000 ? it : 1000) <= m) return;  // this iteration mixes without the history
  }
  extern __shared__ double ys[];  // (m+1) x (kGramTile+1), then block sums
  const int ld = kGramTile + 1;
  const int nv = m + 1;
  const int nb3 = (nv + kGramBlk - 1) / kGramBlk;
  const int npb = gram_pair_blocks(m);
  const int R = gram_subranges(m);
  const int sub = kGramTile / R;
  double* red = ys + nv * ld;
  const int head = st ? st->head : 0;
  const int tid = threadIdx.x;
  const bool worker = tid < npb * R;
  int bi = 0, bj = 0, r = 0;
  if (worker) {
    int rem = tid / R;
    r = tid % R;
    while (rem >= nb3 - bi) {
      rem -= nb3 - bi;
      ++bi;
    }
    bj = bi + rem;
  }
  double acc[kGramBlk * kGramBlk];
  for (int q = 0; q < kGramBlk * kGramBlk; ++q) acc[q] = 0.0;
  __shared__ const double* slot[kMaxHist + 1];  // history entry i = ring slot (head + i) mod (m + 1)
  for (int i = tid; i < nv; i += blockDim.x) slot[i] = dring + (long)((head + i) % nv) * nvar;
  __syncthreads();
  const long ntile = (nvar + kGramTile - 1) / kGramTile;
  const double* d0 = dring + (long)head * nvar;
  // The next tile's history values are loaded into registers before the current tile's
  // products, so each block keeps its loads in flight behind its arithmetic.
  constexpr int kPre = ((kMaxHist + 1) * kGramTile + kThreads - 1) / kThreads;
  double pre0[kPre], prei[kPre];
  auto fetch = [&](long tile) {
    const long k0 = tile * kGramTile;
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
      const int e = tid + q * kThreads;
      pre0[q] = prei[q] = 0.0;
      if (e < nv * kGramTile) {
        const int i = e / kGramTile, c = e % kGramTile;
        const long k = k0 + c;
        if (k < nvar) {
          pre0[q] = d0[k];
          if (i != 0) prei[q] = slot[i][k];
        }
      }
    }
  };
  if ((long)blockIdx.x < ntile) fetch(blockIdx.x);
  for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
      const int e = tid + q * kThreads;
      if (e < nv * kGramTile) {
        const int i = e / kGramTile, c = e % kGramTile;
        ys[i * ld + c] = i == 0 ? pre0[q] : pre0[q] - prei[q];
      }
    }
    __syncthreads();
    if (tile + gridDim.x < ntile) fetch(tile + gridDim.x);
    if (worker) {
      for (int c = r * sub; c < (r + 1) * sub; ++c) {
        double yi[kGramBlk], yj[kGramBlk];
#pragma unroll
        for (int u = 0; u < kGramBlk; ++u) {
          const int i = kGramBlk * bi + u, j = kGramBlk * bj + u;
          yi[u] = i < nv ? ys[i * ld + c] : 0.0;
          yj[u] = j < nv ? ys[j * ld + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kGramBlk; ++u)
#pragma unroll
          for (int v = 0; v < kGramBlk; ++v) acc[u * kGramBlk + v] += yi[u] * yj[v];
      }
    }
    __syncthreads();
  }
  if (worker)
    for (int q = 0; q < kGramBlk * kGramBlk; ++q) red[(long)tid * kGramBlk * kGramBlk + q] = acc[q];
  __syncthreads();
  // unique pairs (i <= j): the R sub-range sums of its block entry, in order
  const int np = nv * (nv + 1) / 2;
  for (int p = tid; p < np; p += blockDim.x) {
    int i = 0, rem = p;
    while (rem >= nv - i) {
      rem -= nv - i;
      ++i;
    }
    const int j = i + rem;
    const int pi = i / kGramBlk, pj = j / kGramBlk;
    const int pb = pi * nb3 - pi * (pi - 1) / 2 + (pj - pi);
    const int q = (i % kGramBlk) * kGramBlk + (j % kGramBlk);
    double s = 0.0;
    for (int rr = 0; rr < R; ++rr) s += red[(long)(pb * R + rr) * kGramBlk * kGramBlk + q];
    part[(long)p * gridDim.x + blockIdx.x] = s;
  }
}

// One block per pair: gsum[p] = sum over the k_gram blocks, fixed tree order.
__global__ void __launch_bounds__(kThreads) k_gram_reduce(const double* __restrict__ part, int nblk, int m,
                                                          const ScfState* st, int force, double* __restrict__ gsum,
                                                          const int* done) {
  if (done && *done) return;
  if (!force) {
    const int it = st->iter + 1;
    if ((it < 1000 ? it : 1000) <= m) return;
  }
  __shared__ double sh[kThreads];
  const int p = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(long)p * nblk + b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) gsum[p] = s;
}

// ludcmp + lubksb (mod_iterate.f90:232-337) by one wavefront, lanes = rows.
// Every matrix element sees the operations of lu_decompose / lu_backsubstitute
// (lu.hpp) in the same order: a left-looking sum s - a(i,1)a(1,j) - a(i,2)a(2,j)
// - ... becomes a sequence of right-looking updates over k = 1, 2, ..., which
// is the same sequence for that element, while all rows update in parallel.
// The backward substitution keeps the reference's increasing-j order per row,
// so it stays sequential in i with broadcast operands.  a: column-major m x m
// in LDS; b: right-hand side in LDS, overwritten by the solution; indx gets the
// 1-based pivots.  Returns 1 if a row of a is entirely zero.
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The loops over k, i and j run in unrolled groups of kLuBatch with no branch per
// step: the range test is part of each lane's predicate, and the reads of a(lane, k)
// for a group are issued together (rows beyond m read row 0 and columns are clamped;
// those values are never used).  Padding lanes take values that leave a step without
// effect: -1 in the pivot scan (never >= the running maximum) and 0 in the backward
// chain (s - 0 = s).  The pivot column stays in registers: the row interchange swaps
// the other columns in LDS and column j in registers, and lane j keeps indx(j).
constexpr int kLuBatch = 4;

__device__ int lu_solve_wave(double* a, double* b, int m, int* indx) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const bool row = lane < m;
  const int rl = row ? lane : 0;
  double vv = 0.0;
  int zero_row = 0;
  if (row) {
    double aamax = 0.0;
    for (int j = 0; j < m; ++j) {
      const double v = fabs(a[lane + j * m]);
      if (v > aamax) aamax = v;
    }
    zero_row = aamax == 0.0;
    vv = 1.0 / aamax;
  }
  const int singular = __ballot(zero_row) != 0;
  int imax = 0, piv_lane = 0;
  for (int j = 0; j < m; ++j) {
    double c = row ? a[lane + j * m] : 0.0;
    for (int k0 = 0; k0 < j; k0 += kLuBatch) {
      double aik[kLuBatch];
#pragma unroll
      for (int u = 0; u < kLuBatch; ++u) aik[u] = a[rl + min(k0 + u, j - 1) * m];
#pragma unroll
      for (int u = 0; u < kLuBatch; ++u) {
        const int k = k0 + u;
        const double akj = lane_bcast(c, k);  // final: lane k has applied every k' < k
        if (row && lane > k && k < j) c = c - aik[u] * akj;
      }
    }
    // pivot search of ludcmp (:263-270): the scan over i = j..n in order, '.ge.'
    const double dum_l = row ? vv * fabs(c) : -1.0;
    double aamax = 0.0;
    for (int i0 = j; i0 < m; i0 += kLuBatch) {
#pragma unroll
      for (int u = 0; u < kLuBatch; ++u) {
        const double dum = lane_bcast(dum_l, i0 + u);
        const bool ge = dum >= aamax;
        imax = ge ? i0 + u : imax;
        aamax = ge ? dum : aamax;
      }
    }
    if (j != imax) {
      // interchange rows j and imax: column j in registers, the others in LDS (lane = column)
      const double cj = lane_bcast(c, j), ci = lane_bcast(c, imax);
      if (row && lane != j) {
        const double t = a[imax + lane * m];
        a[imax + lane * m] = a[j + lane * m];
        a[j + lane * m] = t;
      }
      if (lane == j) c = ci;
      if (lane == imax) c = cj;
      const double vj = lane_bcast(vv, j);
      if (lane == imax) vv = vj;
    }
    if (lane == j) piv_lane = imax + 1;
    double piv = lane_bcast(c, j);
    if (piv == 0.0) {
      piv = 1.0e-20;
      if (lane == j) c = piv;
    }
    if (j != m - 1) {
      const double dinv = 1.0 / piv;
      if (row && lane > j) c = c * dinv;
    }
    if (row) a[lane + j * m] = c;
    wave_lds_sync();
  }
  if (row) indx[lane] = piv_lane;
  // forward substitution: the interchanges first (each touches entries >= i only,
  // so applying them up front gives every step the values lubksb reads), then
  // y_i -= a(i,j) y_j for j = ii, ii+1, ... with ii the first nonzero entry
  double y = row ? b[lane] : 0.0;
  for (int i = 0; i < m; ++i) {
    const int ll = __builtin_amdgcn_readlane(piv_lane, i) - 1;
    const double tl = lane_bcast(y, ll), ti = lane_bcast(y, i);
    if (lane == ll) y = ti;
    if (lane == i) y = tl;
  }
  int first = m;
  for (int i = 0; i < m; ++i)
    if (lane_bcast(y, i) != 0.0) {
      first = i;
      break;
    }
  for (int j0 = first; j0 < m - 1; j0 += kLuBatch) {
    double aj[kLuBatch];
#pragma unroll
    for (int u = 0; u < kLuBatch; ++u) aj[u] = a[rl + min(j0 + u, m - 2) * m];
#pragma unroll
    for (int u = 0; u < kLuBatch; ++u) {
      const int j = j0 + u;
      const double yj = lane_bcast(y, j);
      if (row && lane > j && j < m - 1) y = y - aj[u] * yj;
    }
  }
  // backward substitution (:329-335): the products a(i,j) x_j in parallel (lane j),
  // then the reference's increasing-j chain of subtractions
  const double diag = row ? a[lane + lane * m] : 1.0;
  double x = 0.0;
  double arow = row ? a[(m - 1) + lane * m] : 0.0;  // a(i, lane), one row ahead
  for (int i = m - 1; i >= 0; --i) {
    const double t = (lane > i && row) ? arow * x : 0.0;
    if (i > 0 && row) arow = a[(i - 1) + lane * m];
    double sacc = lane_bcast(y, i);
    for (int j0 = i + 1; j0 < m; j0 += kLuBatch)
#pragma unroll
      for (int u = 0; u < kLuBatch; ++u) sacc = sacc - lane_bcast(t, j0 + u);
    const double xi = sacc / lane_bcast(diag, i);
    if (lane == i) x = xi;
  }
  if (row) b[lane] = x;
  wave_lds_sync();
  return singular;
}

struct SolveArgs {
  int m;
  double lam_mix;
  int nblk_contour;
  const double* rpart;
  const double* spart;
  const double* gsum;  // np reduced Gram sums
  double* resid_hist;  // residual per iteration, or null
  int resid_cap;
  double dvol, volume, ds;
  int ns, ensemble;
  int force_anderson;  // compat: caller decides via iter
  int iter_override;   // compat: iteration number passed by the caller
  int measure;         // 0: accuracy = max|dfield|; 1: sum|dfield| * dvol; 2: sum (phi_old - phi)^2
  int mixing;          // 3 Anderson (default 0 reads as 3), 1 simple, 2 lambda
  const double* gsum_w;  // lambda mixing: Gram sums of the field ring
  ScfState* st;
  int resume;          // test st->accuracy of the last evaluation again (no new one), then mix
};

// Convergence test (main_multiblock.f90:676-685), grand canonical rho/energy
// (mod_multiblock.f90:179-180), and the Anderson solve (mod_iterate.f90:94-139).
__global__ void k_solve(SolveArgs a) {
  ScfState* st = a.st;
  if (st->done) return;
  __shared__ double sh[kThreads];
  __shared__ double U[kMaxHist * kMaxHist];
  __shared__ double V[kMaxHist];
  __shared__ int indx[kMaxHist];
  __shared__ int done_now;
  const int tid = threadIdx.x;
  const int m = a.m;
  int it;
  if (a.iter_override > 0) {
    it = a.iter_override;
  } else if (a.resume) {
    it = st->iter;
    if (tid == 0) {
      done_now = !(st->accuracy > st->tol);
      if (done_now) st->done = 1;
    }
    __syncthreads();
    if (done_now) return;
  } else {
    double r = 0.0;
    if (a.measure) {
      // block partials summed in a fixed order, then times dvol (main_blend.f90:567); or,
      // measure 2, the sum of squared density changes (main_diblock.f90:387)
      for (int b = tid; b < a.nblk_contour; b += blockDim.x) r += a.rpart[b];
      r = block_sum(r, sh);
      if (a.measure == 1) r = r * a.dvol;
    } else {
      for (int b = tid; b < a.nblk_contour; b += blockDim.x) r = nanmax(r, a.rpart[b]);
      r = block_max(r, sh);
    }
    double s = 0.0;
    for (int b = tid; b < a.nblk_contour; b += blockDim.x) s += a.spart[b];
    s = block_sum(s, sh);
    it = st->iter + 1;
    if (tid == 0) {
      st->iter = it;
      st->accuracy = r;
      if (a.resid_hist && it - 1 < a.resid_cap) a.resid_hist[it - 1] = r;
      if (a.ensemble == 2) {
        st->rho = s * a.dvol / a.volume / (a.ds * a.ns);
        st->energy = -st->rho * a.volume;
      }
      done_now = !(r > st->tol);
      if (done_now) st->done = 1;
    }
    __syncthreads();
    if (done_now) return;
  }
  if (a.mixing == 1 || a.mixing == 2) {
    // Mixing_Simple (mod_iterate.f90:27-38): a fixed step.  Mixing_Lambda (:46-84): the
    // step sqrt(sum (W0-W1)^2 / sum (D0-D1)^2) clamped to [1e-5, lammax], 1e-5 at
    // iteration 1 or for a vanishing denominator; pair (1,1) of both Gram sums is
    // <y_1, y_1> with y_1 = X_0 - X_1.
    if (tid == 0) {
      double lam = a.lam_mix;
      if (a.mixing == 2) {
        const double lammin = 0.00001;
        lam = lammin;
        if (it != 1) {
          const double num = a.gsum_w[m + 1], nom = a.gsum[m + 1];
          if (nom > 1.0e-50) {
            lam = sqrt(num / nom);
            if (lam < lammin) lam = lammin;
            else if (lam > a.lam_mix) lam = a.lam_mix;
          }
        }
      }
      st->mode = 0;
      st->lam = lam;
    }
    return;
  }
  const int status = it < 1000 ? it : 1000;
  if (status <= m) {
    if (tid == 0) {
      st->mode = 0;
      st->lam = 0.001;  // lam00 (mod_iterate.f90:94)
    }
    return;
  }
  const int np = (m + 1) * (m + 2) / 2;
  for (int p = tid; p < np; p += blockDim.x) {
    const double s = a.gsum[p];
    int i = 0, rem = p;
    while (rem >= m + 1 - i) { rem -= m + 1 - i; ++i; }
    const int j = i + rem;
    if (i == 0) {
      if (j > 0) V[j - 1] = s;  // v_j = <D0 - Dj, D0>
    } else {
      U[(i - 1) + (j - 1) * m] = s;
      U[(j - 1) + (i - 1) * m] = s;
    }
  }
  __syncthreads();
  if (tid < 64) {
    const int sing = lu_solve_wave(U, V, m, indx);
    if (tid < m) st->x[tid] = V[tid];
    if (tid == 0) {
      st->singular = sing;
      st->mode = 1;
      st->lam = a.lam_mix;
    }
  }
}

// The wave LU of k_solve on one n x n system (column-major a, right-hand side b), for
// checking it against the host ludcmp / lubksb: out = [x (n), factors (n*n)],
// ipiv = [indx (n, 1-based), singular].
__global__ void k_lu_check(const double* __restrict__ a_in, const double* __restrict__ b_in, int n,
                           double* __restrict__ out, int* __restrict__ ipiv) {
  __shared__ double A[kMaxHist * kMaxHist];
  __shared__ double B[kMaxHist];
  __shared__ int indx[kMaxHist];
  __shared__ int sing;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) A[i] = a_in[i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) B[i] = b_in[i];
  __syncthreads();
  if (threadIdx.x < 64) {
    const int s = lu_solve_wave(A, B, n, indx);
    if (threadIdx.x == 0) sing = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    out[i] = B[i];
    ipiv[i] = indx[i];
  }
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) out[n + i] = A[i];
  if (threadIdx.x == 0) ipiv[n] = sing;
}

constexpr int kMixBatch = 8;

// Mixing update (mod_iterate.f90:125-148) over the ring: the new field goes
// to the slot of the discarded oldest entry, which becomes logical index 0
// after the rotation in k_schedule.  E of the new field and its E^2 partials
// are produced in the same pass.  Grid row y covers the types y, y + gridDim.y, ...
template <typename T>
__global__ void __launch_bounds__(kThreads) k_mix(double* __restrict__ wring, const double* __restrict__ dring, long ng,
                                                  int nt, int m, const ScfState* st, double ds, T* __restrict__ E,
                                                  double* __restrict__ e2part) {
  if (st->done) return;
  __shared__ double sh[kThreads];
  const long nvar = ng * nt;
  const int head = st->head;
  const int mode = st->mode;
  const double lam = st->lam;
  const int dst = (head + m) % (m + 1);
  __shared__ double x[kMaxHist];
  __shared__ long off[kMaxHist + 1];  // history entry j = ring slot (head + j) mod (m + 1)
  if (mode == 1)
    for (int i = threadIdx.x; i < m; i += blockDim.x) x[i] = st->x[i];
  for (int j = threadIdx.x; j <= m; j += blockDim.x) off[j] = (long)((head + j) % (m + 1)) * nvar;
  __syncthreads();
  const double* w0 = wring + (long)head * nvar;
  const double* d0 = dring + (long)head * nvar;
  double* wn = wring + (long)dst * nvar;
  for (int t = blockIdx.y; t < nt; t += gridDim.y) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
      const long k = (long)t * ng + i;
      const double W0 = w0[k], D0 = d0[k];
      double f = W0, d = D0;
      if (mode == 1) {
        // kMixBatch history entries are loaded before their updates (entries past m
        // re-read entry m and are not used)
        for (int j0 = 1; j0 <= m; j0 += kMixBatch) {
          double wv[kMixBatch], dv[kMixBatch];
#pragma unroll
          for (int u = 0; u < kMixBatch; ++u) {
            const long o = off[min(j0 + u, m)] + k;
            wv[u] = wring[o];
            dv[u] = dring[o];
          }
#pragma unroll
          for (int u = 0; u < kMixBatch; ++u)
            if (j0 + u <= m) {
              f = f - x[j0 + u - 1] * (W0 - wv[u]);
              d = d - x[j0 + u - 1] * (D0 - dv[u]);
            }
        }
      }
      const double wnew = f + lam * d;
      wn[k] = wnew;
      const T e = (T)exp(-0.5 * ds * wnew);
      E[k] = e;
      acc += (double)e * (double)e;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) e2part[blockIdx.x * kMaxTypes + t] = acc;
  }
}

// Partial sums for Calculate_Energy (main_multiblock.f90:552-553) and
// Interaction_Energy (mod_FloryHuggins.f90:82-94): [sum phi.w, sum (tot-1)^2,
// sum vf_i vf_j for i<j].
__global__ void __launch_bounds__(kThreads) k_energy_parts(const double* __restrict__ phi, const double* __restrict__ w,
                                                           long ng, FHDesc fh, double* __restrict__ part, int npart) {
  __shared__ double sh[kThreads];
  const int nt = fh.nt;
  double acc[2 + kMaxTypes * (kMaxTypes - 1) / 2];
  const int nq = 2 + nt * (nt - 1) / 2;
  for (int q = 0; q < nq; ++q) acc[q] = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    double vf[kMaxTypes], tot = 0.0;
    for (int t = 0; t < nt; ++t) {
      const double p = phi[(long)t * ng + i];
      if (w) acc[0] += p * w[(long)t * ng + i];
      vf[t] = p * fh.vol[t];
      tot += vf[t];
    }
    acc[1] += (tot - 1.0) * (tot - 1.0);
    int q = 2;
    for (int t = 0; t < nt; ++t)
      for (int t2 = t + 1; t2 < nt; ++t2) acc[q++] += vf[t] * vf[t2];
  }
  for (int q = 0; q < nq; ++q) {
    const double s = block_sum(acc[q], sh);
    if (threadIdx.x == 0) part[(long)blockIdx.x * npart + q] = s;
  }
}

// Calculate_Mean_Fields for host-supplied densities (mod_FloryHuggins.f90:104-129)
__global__ void __launch_bounds__(kThreads) k_mean_fields(const double* __restrict__ den, double* __restrict__ field,
                                                          long ng, FHDesc fh) {
  const int nt = fh.nt;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    double vf[kMaxTypes], tot = 0.0;
    for (int t = 0; t < nt; ++t) vf[t] = den[(long)t * ng + i] * fh.vol[t];
    for (int t = 0; t < nt; ++t) tot += vf[t];
    for (int t = 0; t < nt; ++t) {
      double f = fh.kappa * (tot - 1.0);
      for (int t2 = 0; t2 < nt; ++t2) f = f + fh.chi[t + t2 * nt] * vf[t2];
      field[(long)t * ng + i] = f * fh.vol[t];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_convert(const double* __restrict__ a, T* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (T)a[i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_widen(const T* __restrict__ a, double* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (double)a[i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill(T* __restrict__ a, T v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = v;
}

// A new solve on the current field: the iteration count and the stop flag start
// again (Get_Conjugate_Fields restarts iter_status at 1, main_multiblock.f90:594).
// The ring keeps its entries; Anderson reads history only from iteration m+1 on,
// when every entry it reads belongs to the new solve.
// The chain scalars of a Propagate_* evaluation into out[4] = {Qsum, mu, rho, energy}
// (mod_multiblock.f90:155-186): canonical from the state k_qfinal wrote; grand canonical
// rho = sum(density) dvol / V / (ds Ns) from the contour kernel's per-block sums, added in
// block order as the host does, and energy = -rho V.
__global__ void k_chain_scalars(const ScfState* __restrict__ st, const double* __restrict__ spart, int nblk,
                                int ensemble, double rho_in, double dvol, double volume, double ds, int ns,
                                double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  out[0] = st->qsum;
  if (ensemble == 1) {
    out[1] = st->mu;
    out[2] = rho_in;
    out[3] = st->energy;
  } else {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += spart[b];
    const double rho = s * dvol / volume / (ds * ns);
    out[1] = st->mu;
    out[2] = rho;
    out[3] = -rho * volume;
  }
}

// the loop's tolerance (set_tol) and a cleared stop flag (clear_done), from kernel arguments
__global__ void k_state_update(ScfState* st, double tol, int set_tol, int clear_done) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (set_tol) st->tol = tol;
    if (clear_done) st->done = 0;
  }
}

__global__ void k_scf_restart(ScfState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->iter = 0;
    st->done = 0;
  }
}

// ---------------------------------------------------------------------------
// DDFT_Timestep_Euler (Main_Multiblock/main_multiblock.f90:794-880)
// ---------------------------------------------------------------------------

// mulocal = mean fields of the density - field (:824-825); the field is ring slot `head`
// of wring, or wring itself when st is null
__global__ void __launch_bounds__(kThreads) k_ddft_mulocal(const double* __restrict__ den,
                                                           const double* __restrict__ wring, const ScfState* st,
                                                           FHDesc fh, long ng, double* __restrict__ mu) {
  const int nt = fh.nt;
  const double* w = st ? wring + (long)st->head * nt * ng : wring;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    double vf[kMaxTypes], tot = 0.0;
    for (int t = 0; t < nt; ++t) vf[t] = den[(long)t * ng + i] * fh.vol[t];
    for (int t = 0; t < nt; ++t) tot += vf[t];
    for (int t = 0; t < nt; ++t) {
      double f = fh.kappa * (tot - 1.0);
      for (int t2 = 0; t2 < nt; ++t2) f = f + fh.chi[t + t2 * nt] * vf[t2];
      f = f * fh.vol[t];
      mu[(long)t * ng + i] = f - w[(long)t * ng + i];
    }
  }
}

// ddensity_q(t1) = 0 - sum over t2 of mobility_qq(:,t2,t1) * mu_q(t2), accumulated in
// type order (:843-850).  muq, ddq: nt half spectra of 2*nk doubles; mob: (nk, nt, nt).
__global__ void __launch_bounds__(kThreads) k_ddft_mobility(const double* __restrict__ muq,
                                                            const double* __restrict__ mob, long nk, int nt,
                                                            double* __restrict__ ddq) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x) {
    for (int t1 = 0; t1 < nt; ++t1) {
      double re = 0.0, im = 0.0;
      for (int t2 = 0; t2 < nt; ++t2) {
        const double m = mob[(long)(t2 + t1 * nt) * nk + i];
        re = re - m * muq[2 * ((long)t2 * nk + i)];
        im = im - m * muq[2 * ((long)t2 * nk + i) + 1];
      }
      ddq[2 * ((long)t1 * nk + i)] = re;
      ddq[2 * ((long)t1 * nk + i) + 1] = im;
    }
  }
}

// density = density + timestep*ddensity, then max(density, floor) (:866-871).  The
// reference clamps only when minval(density) <= floor; for finite densities the
// unconditional max gives the same values (nothing is below the floor otherwise).
__global__ void __launch_bounds__(kThreads) k_ddft_euler(double* __restrict__ den, const double* __restrict__ dd, long n,
                                                         double dt, double floor, int clamp) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double d = den[i] + dt * dd[i];
    if (clamp) d = d > floor ? d : floor;
    den[i] = d;
  }
}

// ---------------------------------------------------------------------------
// Chain_Dynamics_Timestep_Euler (Main_Blend/main_blend.f90:605-667)
// ---------------------------------------------------------------------------

// nabla_a * spectrum with nabla_a = i b (mod_fourier.f90:189-191), as Fortran's complex
// product (0 + i b)(x + i y) = (0 x - b y) + i (0 y + b x)
__global__ void __launch_bounds__(kThreads) k_nabla_mul(const double* __restrict__ in, const double* __restrict__ nab,
                                                        long nk, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x) {
    const double b = nab[i], x = in[2 * i], y = in[2 * i + 1];
    out[2 * i] = 0.0 * x - b * y;
    out[2 * i + 1] = 0.0 * y + b * x;
  }
}

// newfield = field - eps*dfield (:647); the field is ring slot `head`
__global__ void __launch_bounds__(kThreads) k_field_step(const double* __restrict__ wring, const ScfState* st,
                                                         const double* __restrict__ d, long nvar, double eps,
                                                         double* __restrict__ out) {
  const double* w = wring + (long)st->head * nvar;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nvar; i += (long)gridDim.x * blockDim.x)
    out[i] = w[i] - eps * d[i];
}

// out = a - b
__global__ void __launch_bounds__(kThreads) k_sub(const double* __restrict__ a, const double* __restrict__ b, long n,
                                                  double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}

// density = density + sum(ddensity, dim=5) * timestep/eps (:662): the three axes summed in
// order, then times timestep, then divided by eps
__global__ void __launch_bounds__(kThreads) k_epd_update(double* __restrict__ den, const double* __restrict__ dd,
                                                         long n, double dt, double eps) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double sum = dd[i];
    sum = sum + dd[n + i];
    sum = sum + dd[2 * n + i];
    den[i] = den[i] + sum * dt / eps;
  }
}

// ---------------------------------------------------------------------------
// EPD_Timestep_Euler / EPD_Timestep_RungeKutta (Main_Blend/main_blend.f90:580-740)
// ---------------------------------------------------------------------------

// dummyfourier = - laplace*dummyfourier (:735): Fortran's complex product with
// laplace = (lr, 0), then the negation of both parts
__global__ void __launch_bounds__(kThreads) k_neg_laplace_mul(const double* __restrict__ in,
                                                              const double* __restrict__ lr, long nk,
                                                              double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nk; i += (long)gridDim.x * blockDim.x) {
    const double a = lr[i], x = in[2 * i], y = in[2 * i + 1];
    out[2 * i] = -(a * x - 0.0 * y);
    out[2 * i + 1] = -(a * y + 0.0 * x);
  }
}

// out = x + a*d (the stage fields of :691, :693, :695 and the Euler update of :592)
__global__ void __launch_bounds__(kThreads) k_axpy(const double* __restrict__ x, double a,
                                                   const double* __restrict__ d, long n, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = x[i] + a * d[i];
}

// field = field + timestep*(df1/6 + df2/3 + df3/3 + df4/6) (:698), summed left to right
__global__ void __launch_bounds__(kThreads) k_rk4_update(double* __restrict__ f, const double* __restrict__ df,
                                                         long n, double dt) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = df[i] / 6.0;
    s = s + df[n + i] / 3.0;
    s = s + df[2 * n + i] / 3.0;
    s = s + df[3 * n + i] / 6.0;
    f[i] = f[i] + dt * s;
  }
}

}  // namespace adscf
