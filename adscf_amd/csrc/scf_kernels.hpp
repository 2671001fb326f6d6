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
  double* dist;       // CONTOUR_BLOCKS: (nb, ng) or null
  double* dens;       // CONTOUR_BLOCKS: (nb, ng) or null
  const double* w0;   // CONTOUR_SCF: field ring base; slot `head` is the current field (nt, ng)
  double* phi;        // CONTOUR_SCF: species densities (nt, ng)
  double* dring;      // CONTOUR_SCF: residual history ring (m+1, nt*ng) or null
  int m;
  double* rpart;      // per-block max|dfield|
  double* spart;      // per-block sum of densities (grand canonical rho)
  const int* done;
};

template <typename T, int VP>
struct alignas(sizeof(T) * VP) PackT {
  T v[VP];
};

// Trapezoid contour integral (mod_multiblock.f90:157-168) in the reference's
// summation order, then the ensemble normalisation (173-180).  In SCF mode the
// species densities (main_multiblock.f90:530-534), the mean fields
// (mod_FloryHuggins.f90:104-129), dfield = new - field and max|dfield|
// (main_multiblock.f90:674-676, 780) follow in the same pass.
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
  auto ld = [&](const T* base, int s, long i0) { return *reinterpret_cast<const P*>(base + (long)s * ng + i0); };
  auto scl = [&](int s) { return ldexp(1.0, -(a.e1[s] + a.e2[ns - s])); };
  const long npk = ng / VP;
  for (long ip = blockIdx.x * (long)blockDim.x + threadIdx.x; ip < npk; ip += (long)gridDim.x * blockDim.x) {
    const long i0 = ip * VP;
    double phi[kMaxTypes][VP];
    for (int t = 0; t < a.fh.nt; ++t)
#pragma unroll
      for (int v = 0; v < VP; ++v) phi[t][v] = 0.0;
    for (int b = 0; b < a.cd.nblocks; ++b) {
      const int s0 = a.cd.start[b], nb = a.cd.seg[b];
      double acc[VP];
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
    if (a.mode == CONTOUR_SCF) {
      const int nt = a.fh.nt;
      double* d0 = a.dring ? a.dring + (long)a.st->head * nt * ng : nullptr;
      const double* w0 = a.w0 + (long)a.st->head * nt * ng;  // current field: ring slot `head`
#pragma unroll
      for (int v = 0; v < VP; ++v) {
        const long i = i0 + v;
        double vf[kMaxTypes], tot = 0.0;
        for (int t = 0; t < nt; ++t) {
          a.phi[(long)t * ng + i] = phi[t][v];
          vf[t] = phi[t][v] * a.fh.vol[t];
        }
        for (int t = 0; t < nt; ++t) tot += vf[t];
        for (int t = 0; t < nt; ++t) {
          double f = a.fh.kappa * (tot - 1.0);
          for (int t2 = 0; t2 < nt; ++t2) f = f + a.fh.chi[t + t2 * nt] * vf[t2];
          f = f * a.fh.vol[t];
          const double df = f - w0[(long)t * ng + i];
          if (d0) d0[(long)t * ng + i] = df;
          rmax = nanmax(rmax, fabs(df));
        }
      }
    }
  }
  rmax = block_max(rmax, sh);
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
// them contiguously).
__global__ void __launch_bounds__(kThreads) k_gram(const double* __restrict__ dring, long nvar, int m,
                                                   const ScfState* st, int force, double* __restrict__ part,
                                                   const int* done) {
  if (done && *done) return;
  if (!force) {
    const int it = st->iter + 1;
    if ((it < 1000 ? it : 1000) <= m) return;  // this iteration mixes without the history
  }
  extern __shared__ double ys[];  // (m+1) x (kGramTile+1)
  const int ld = kGramTile + 1;
  const int head = st ? st->head : 0;
  const int np = (m + 1) * (m + 2) / 2;
  const int tid = threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int pi[4], pj[4];
  for (int r = 0; r < 4; ++r) {
    const int p = tid + r * blockDim.x;
    pi[r] = pj[r] = -1;
    if (p < np) {
      int i = 0, rem = p;
      while (rem >= m + 1 - i) { rem -= m + 1 - i; ++i; }
      pi[r] = i;
      pj[r] = i + rem;
    }
  }
  const long ntile = (nvar + kGramTile - 1) / kGramTile;
  const double* d0 = dring + (long)head * nvar;
  for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long k0 = tile * kGramTile;
    for (int e = tid; e < (m + 1) * kGramTile; e += blockDim.x) {
      const int i = e / kGramTile, c = e % kGramTile;
      const long k = k0 + c;
      double y = 0.0;
      if (k < nvar) {
        const double x0 = d0[k];
        y = i == 0 ? x0 : x0 - dring[(long)((head + i) % (m + 1)) * nvar + k];
      }
      ys[i * ld + c] = y;
    }
    __syncthreads();
    for (int r = 0; r < 4; ++r) {
      if (pi[r] < 0) continue;
      const double* a = ys + pi[r] * ld;
      const double* b = ys + pj[r] * ld;
      double s = acc[r];
      for (int c = 0; c < kGramTile; ++c) s += a[c] * b[c];
      acc[r] = s;
    }
    __syncthreads();
  }
  for (int r = 0; r < 4; ++r) {
    const int p = tid + r * blockDim.x;
    if (p < np) part[(long)p * gridDim.x + blockIdx.x] = acc[r];
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
  ScfState* st;
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
  __shared__ double vv[kMaxHist + 1];
  __shared__ int done_now;
  const int tid = threadIdx.x;
  const int m = a.m;
  int it;
  if (a.iter_override > 0) {
    it = a.iter_override;
  } else {
    double r = 0.0;
    for (int b = tid; b < a.nblk_contour; b += blockDim.x) r = nanmax(r, a.rpart[b]);
    r = block_max(r, sh);
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
  if (tid == 0) {
    double d;
    st->singular = lu_decompose(U, m, indx, &d, vv);
    lu_backsubstitute(U, m, indx, V);
    for (int i = 0; i < m; ++i) st->x[i] = V[i];
    st->mode = 1;
    st->lam = a.lam_mix;
  }
}

// Mixing update (mod_iterate.f90:125-148) over the ring: the new field goes
// to the slot of the discarded oldest entry, which becomes logical index 0
// after the rotation in k_schedule.  E of the new field and its E^2 partials
// are produced in the same pass.
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
  if (mode == 1)
    for (int i = threadIdx.x; i < m; i += blockDim.x) x[i] = st->x[i];
  __syncthreads();
  const double* w0 = wring + (long)head * nvar;
  const double* d0 = dring + (long)head * nvar;
  double* wn = wring + (long)dst * nvar;
  for (int t = 0; t < nt; ++t) {
    double acc = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
      const long k = (long)t * ng + i;
      const double W0 = w0[k], D0 = d0[k];
      double f = W0, d = D0;
      if (mode == 1) {
        for (int j = 1; j <= m; ++j) {
          const long o = (long)((head + j) % (m + 1)) * nvar + k;
          f = f - x[j - 1] * (W0 - wring[o]);
          d = d - x[j - 1] * (D0 - dring[o]);
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

}  // namespace adscf
