// Pointwise kernels of the adjoint (reverse-mode) pass through one density
// evaluation.  The reference has no adjoint; this carries the loss
//   L = sum_r (phi_T(r) - phi_target(r))^2 + lam * F
// (canonical ensemble, F of main_multiblock.f90:548-557) back to gradients in
// the fields w and the box lengths.
//
// Forward model per chain: q(s) = T_s q(s-1), T_s = E L E with E = exp(-w ds/2)
// and L the diffusion operator; T_s is symmetric, so the adjoint recurrence
//   mu(s-1) = dL/dq(s-1) + T_s mu(s)
// reuses the forward line passes.  Per step the E gradient collects
//   mu(s) * q(s)/E + q(s-1) * L(E mu(s)),
// and the box gradient collects (2 ds / L_a) <E mu(s), D_a q(s)/E> with D_a the
// k_a^2 line operator (it commutes with L), plus lam * F / L_a from F being
// proportional to the volume at fixed propagators.
#pragma once

#include "kernels.hpp"
#include "scf_kernels.hpp"

namespace adscf {

// dist_b += w1 q p (and dist_b2 += w2 q p at a block junction)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_adj_contour(const T* __restrict__ q, const T* __restrict__ p, long ng,
                                                          double w1, int b1, double w2, int b2,
                                                          double* __restrict__ dist) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    const double prod = (double)q[i] * (double)p[i];
    dist[(long)b1 * ng + i] += w1 * prod;
    if (b2 >= 0) dist[(long)b2 * ng + i] += w2 * prod;
  }
}

// phi_t = A * sum_{b of type t} dist_b (A = rho Ng / S); partial sums of the target misfit.
__global__ void __launch_bounds__(kThreads) k_adj_phi(const double* __restrict__ dist, long ng, ChainDesc cd, int nt,
                                                      double A, int target_type, const double* __restrict__ phit,
                                                      double* __restrict__ phi, double* __restrict__ part) {
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    double ph[kMaxTypes];
    for (int t = 0; t < nt; ++t) ph[t] = 0.0;
    for (int b = 0; b < cd.nblocks; ++b) ph[cd.type[b]] += dist[(long)b * ng + i];
    for (int t = 0; t < nt; ++t) phi[(long)t * ng + i] = A * ph[t];
    const double d = A * ph[target_type] - phit[i];
    acc += d * d;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// g_t = dL/dphi_t: target misfit + lam * (-dvol w_t + dE_int/dphi_t); partial sums of sum_t g_t phi_t
__global__ void __launch_bounds__(kThreads) k_adj_g(const double* __restrict__ phi, const double* __restrict__ w,
                                                    long ng, FHDesc fh, double lam, double dvol, int target_type,
                                                    const double* __restrict__ phit, double* __restrict__ g,
                                                    double* __restrict__ part) {
  __shared__ double sh[kThreads];
  const int nt = fh.nt;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    double vf[kMaxTypes], tot = 0.0;
    for (int t = 0; t < nt; ++t) {
      vf[t] = phi[(long)t * ng + i] * fh.vol[t];
      tot += vf[t];
    }
    for (int t = 0; t < nt; ++t) {
      double de = fh.kappa * (tot - 1.0) * fh.vol[t];
      for (int j = 0; j < nt; ++j)
        if (j != t) de += fh.chi[t + j * nt] * fh.vol[t] * vf[j];
      double gt = lam * dvol * (de - w[(long)t * ng + i]);
      if (t == target_type) gt += 2.0 * (phi[(long)t * ng + i] - phit[i]);
      g[(long)t * ng + i] = gt;
      acc += gt * phi[(long)t * ng + i];
    }
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Seed a(r, s) = A * sum over the (at most two) blocks containing contour point s
// of weight * g_type(r).
struct Seed {
  double w1;
  int t1;
  double w2;
  int t2;  // -1 when s lies inside one block
};

__device__ __forceinline__ double seed_at(const Seed& sd, const double* g, long ng, long i) {
  double a = sd.w1 * g[(long)sd.t1 * ng + i];
  if (sd.t2 >= 0) a += sd.w2 * g[(long)sd.t2 * ng + i];
  return a;
}

// mu(Ns) = a(Ns) p(0) + dL/dS, nu(Ns) = a(0) q(0); q(0) = p(0) = 1
template <typename T>
__global__ void __launch_bounds__(kThreads) k_adj_init(const double* __restrict__ g, long ng, double A, Seed sN,
                                                       Seed s0, double dLdS, T* __restrict__ mu, T* __restrict__ nu) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ng; i += (long)gridDim.x * blockDim.x) {
    mu[i] = (T)(A * seed_at(sN, g, ng, i) + dLdS);
    nu[i] = (T)(A * seed_at(s0, g, ng, i));
  }
}

template <typename T>
struct AdjStepArgs {
  long ng;
  double A;
  Seed seed_q;     // seed at contour point s-1 (forward chain source)
  Seed seed_p;     // seed at contour point Ns-s+1 (complementary chain source)
  const double* g;
  const T* Eq;     // E of forward step s
  const T* Ep;     // E of complementary step s
  const T* y;      // L(E mu(s))
  const T* yp;     // L(E nu(s))
  const T* mu;     // mu(s)
  const T* nu;     // nu(s)
  const T* q_s;    // q(s)
  const T* q_m;    // q(s-1)
  const T* p_s;    // p(s)
  const T* p_m;    // p(s-1)
  const T* p_f;    // p(Ns-s+1)
  const T* q_f;    // q(Ns-s+1)
  T* mu_out;       // mu(s-1)
  T* nu_out;       // nu(s-1)
  double* gEq;     // dL/dE of the forward step's type
  double* gEp;     // dL/dE of the complementary step's type
};

template <typename T>
__global__ void __launch_bounds__(kThreads) k_adj_step(AdjStepArgs<T> a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.ng; i += (long)gridDim.x * blockDim.x) {
    const double eq = (double)a.Eq[i], ep = (double)a.Ep[i];
    const double y = (double)a.y[i], yp = (double)a.yp[i];
    const double mu = (double)a.mu[i], nu = (double)a.nu[i];
    a.gEq[i] += mu * ((double)a.q_s[i] / eq) + (double)a.q_m[i] * y;
    a.gEp[i] += nu * ((double)a.p_s[i] / ep) + (double)a.p_m[i] * yp;
    a.mu_out[i] = (T)(a.A * seed_at(a.seed_q, a.g, a.ng, i) * (double)a.p_f[i] + eq * y);
    a.nu_out[i] = (T)(a.A * seed_at(a.seed_p, a.g, a.ng, i) * (double)a.q_f[i] + ep * yp);
  }
}

// dL/dw_t = -(ds/2) E_t dL/dE_t - lam dvol phi_t
template <typename T>
__global__ void __launch_bounds__(kThreads) k_adj_gradw(const double* __restrict__ gE, const T* __restrict__ E,
                                                        const double* __restrict__ phi, long n, double ds, double lam,
                                                        double dvol, double* __restrict__ gw) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    gw[i] = -0.5 * ds * (double)E[i] * gE[i] - lam * dvol * phi[i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_copy(const T* __restrict__ a, T* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace adscf
