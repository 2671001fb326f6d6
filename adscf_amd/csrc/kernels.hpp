// Device kernels of the SCFT propagator path (HIP, gfx950).
//
// The diffusion operator of one contour step, F^-1 exp(-ds k^2) F, is the
// tensor product Lx Ly Lz of three 1-D circulant operators because
// exp(-ds k^2) factorises over the axes.  A step is therefore applied as
// E * L_a L_b L_c * E with one "line pass" per axis; each pass loads a tile of
// lines, runs a forward FFT, multiplies by K_axis/n, runs the inverse FFT and
// writes the lines back.  Lengths with a four-step shape (fft_reg.hpp) keep
// each line in the registers of a few threads of one wave and exchange through
// LDS twice per operator; other lengths use LDS-staged mixed-radix Stockham
// passes (fft_line.hpp).  Steps alternate the axis order
// (x,y,z then z,y,x), so the last pass of a step and the first pass of the
// next act on the same axis and merge into one pass that also multiplies by
// E, stores q(s) and multiplies by the next step's E.  A contour step thus
// costs two passes over the grid.  The forward propagator q and the
// complementary propagator q+ travel together as the real and imaginary parts
// of one complex line (both operators are real and even in k).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "fft_cube.hpp"
#include "fft_line.hpp"
#include "fft_reg.hpp"

namespace adscf {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 64;  // chain blocks
constexpr int kMaxTypes = 4;    // monomer types
constexpr int kMaxHist = 40;    // Anderson history depth

// ---------------------------------------------------------------------------
// line pass
// ---------------------------------------------------------------------------
enum InKind : int { IN_TMP = 0, IN_ONES = 1, IN_PLANAR = 2 };
enum OutKind : int { OUT_TMP = 0, OUT_PLANAR = 1, OUT_DOT = 2 };

template <typename T>
struct PassArgs {
  int nx, ny, nz;
  int axis;     // 0 x, 1 y, 2 z
  int L;        // lines per tile
  int ldn;      // LDS leading dimension (n + 1)
  int gx;       // tiles along x for the y/z axes: ceil(nx / L)
  long stride;  // point stride along the line
  LinePlan plan;
  const C2<T>* tw;  // n forward twiddles
  const T* kf;      // n multipliers exp(-ds q_k^2) / n
  const C2<T>* tw2; // plane passes (axis 3): twiddles and multipliers of the y operator
  const T* kf2;
  int lds_line;     // LDS elements between the exchange areas of consecutive lines (register
  int lds_line2;    // engines; lds_line2: the column layout of a plane pass), chosen by the host
  int nops;         // 1, or 2 for a merged pass (op, mid store, op)
  int in_kind;
  int pre_E;        // +1: multiply the loaded lines by (Ea1, Ea2); -1: divide by them
  int post_E;       // OUT_PLANAR: multiply by (Ea1, Ea2) and the rescale before storing
  int out_kind;
  const C2<T>* tmp_in;
  const T* q1_in;
  const T* q2_in;   // may be null (treated as 0)
  const T* Ea1;     // E of the step that ends in this pass (and starts, with pre_E)
  const T* Ea2;
  const T* Eb1;     // E of the step that starts after the mid store
  const T* Eb2;
  T* h1;            // planar output q(s) / q+(s)
  T* h2;            // may be null
  const int* dexp1; // power-of-two rescale exponents, indexed by sidx
  const int* dexp2;
  int sidx;
  C2<T>* tmp_out;
  const T* d1;      // OUT_DOT: sum over points of Re*Ea1*d1 + Im*Ea2*d2 -> dot_out[block]
  const T* d2;
  double* dot_out;
  const int* done;  // skip all work once the SCF loop has converged
};

template <typename T>
__device__ __forceinline__ void apply_pre_E(const PassArgs<T>& a, long g, C2<T>& v) {
  if (a.pre_E > 0) {
    v.x *= a.Ea1[g];
    v.y *= a.Ea2[g];
  } else if (a.pre_E < 0) {
    v.x /= a.Ea1[g];
    v.y /= a.Ea2[g];
  }
}

template <typename T>
__device__ __forceinline__ double dot_term(const PassArgs<T>& a, long g, C2<T> v) {
  return (double)v.x * (double)a.Ea1[g] * (double)a.d1[g] + (double)v.y * (double)a.Ea2[g] * (double)a.d2[g];
}

__device__ __forceinline__ void block_dot_out(double acc, double* out) {
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  red[tid] = acc;
  __syncthreads();
  const int nthr = (int)blockDim.x;
  for (int s = 1; s < nthr; s <<= 1) {
    if ((tid % (2 * s)) == 0 && tid + s < nthr) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) out[blockIdx.x] = red[0];
}

// read in place of PassArgs::done when a pass has no flag
__device__ const int g_no_stop = 0;

template <typename T>
__device__ __forceinline__ long pass_line_base(const PassArgs<T>& a, int item, int l, bool& valid) {
  if (a.axis == 0) {
    const long r = (long)item * a.L + l;
    valid = r < (long)a.ny * a.nz;
    return r * a.nx;
  }
  const int x0 = (item % a.gx) * a.L;
  const int o = item / a.gx;
  valid = (x0 + l) < a.nx;
  if (a.axis == 1) return (long)o * a.nx * a.ny + x0 + l;  // o = z
  return (long)o * a.nx + x0 + l;                            // o = y
}

template <typename T>
__device__ __forceinline__ void pass_elem(const PassArgs<T>& a, int e, int n, int& l, int& j) {
  if (a.axis == 0) {
    l = e / n;
    j = e - l * n;
  } else {
    l = e % a.L;
    j = e / a.L;
  }
}

template <typename T>
__device__ __forceinline__ T exp2i(int e) {
  return (T)ldexp(1.0, e);
}

template <typename T>
__device__ C2<T>* pass_line_op(const PassArgs<T>& a, C2<T>* b0, C2<T>* b1, const C2<T>* tw, const T* kf) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int n = a.plan.n;
  auto sync = [] { __syncthreads(); };
  C2<T>* s = line_fft<T>(b0, b1, a.plan, a.ldn, a.L, tw, -1, tid, nthr, sync);
  const int tot = a.L * n;
  for (int e = tid; e < tot; e += nthr) {
    const int l = e / n, k = e - l * n;
    const T f = kf[k];
    C2<T> v = s[l * a.ldn + k];
    s[l * a.ldn + k] = {v.x * f, v.y * f};
  }
  __syncthreads();
  C2<T>* o = (s == b0) ? b1 : b0;
  return line_fft<T>(s, o, a.plan, a.ldn, a.L, tw, +1, tid, nthr, sync);
}

// LDS size of the Stockham kernels: two tiles of L lines (leading dimension ldn) and
// the n twiddles of the axis
template <typename T>
__host__ __device__ constexpr size_t stockham_lds_bytes(int L, int ldn, int n) {
  return (2 * (size_t)L * ldn + (size_t)n) * sizeof(C2<T>);
}
// k_line_pass also keeps the axis's n K multipliers in LDS
template <typename T>
__host__ __device__ constexpr size_t line_pass_lds_bytes(int L, int ldn, int n) {
  return stockham_lds_bytes<T>(L, ldn, n) + (size_t)n * sizeof(T);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_line_pass(PassArgs<T> a) {
  if (a.done && *a.done) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* b0 = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* b1 = b0 + a.L * a.ldn;
  C2<T>* twl = b1 + a.L * a.ldn;  // the axis twiddles: every butterfly reads them
  const int n = a.plan.n;
  T* kfl = reinterpret_cast<T*>(twl + n);  // the K multipliers, read between the transforms
  const int tot = a.L * n;
  const int item = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int e = tid; e < n; e += nthr) {
    twl[e] = a.tw[e];
    kfl[e] = a.kf[e];
  }
  // ---- load
  for (int e = tid; e < tot; e += nthr) {
    int l, j;
    pass_elem(a, e, n, l, j);
    bool valid;
    const long base = pass_line_base(a, item, l, valid);
    C2<T> v = {T(0), T(0)};
    if (valid) {
      const long g = base + (long)j * a.stride;
      if (a.in_kind == IN_TMP) {
        v = a.tmp_in[g];
      } else if (a.in_kind == IN_ONES) {
        v = {T(1), a.h2 ? T(1) : T(0)};
      } else {
        v.x = a.q1_in[g];
        v.y = a.q2_in ? a.q2_in[g] : T(0);
      }
      apply_pre_E(a, g, v);
    }
    b0[l * a.ldn + j] = v;
  }
  __syncthreads();
  C2<T>* r = pass_line_op(a, b0, b1, twl, kfl);
  if (a.nops == 2) {
    const T sc1 = exp2i<T>(a.dexp1[a.sidx]);
    const T sc2 = exp2i<T>(a.dexp2[a.sidx]);
    for (int e = tid; e < tot; e += nthr) {
      int l, j;
      pass_elem(a, e, n, l, j);
      bool valid;
      const long base = pass_line_base(a, item, l, valid);
      if (!valid) continue;
      const long g = base + (long)j * a.stride;
      C2<T> v = r[l * a.ldn + j];
      v.x = v.x * a.Ea1[g] * sc1;
      v.y = v.y * a.Ea2[g] * sc2;
      a.h1[g] = v.x;
      a.h2[g] = v.y;
      v.x *= a.Eb1[g];
      v.y *= a.Eb2[g];
      r[l * a.ldn + j] = v;
    }
    __syncthreads();
    C2<T>* other = (r == b0) ? b1 : b0;
    r = pass_line_op(a, r, other, twl, kfl);
  }
  // ---- store
  if (a.out_kind == OUT_TMP) {
    for (int e = tid; e < tot; e += nthr) {
      int l, j;
      pass_elem(a, e, n, l, j);
      bool valid;
      const long base = pass_line_base(a, item, l, valid);
      if (valid) a.tmp_out[base + (long)j * a.stride] = r[l * a.ldn + j];
    }
  } else if (a.out_kind == OUT_DOT) {
    double acc = 0.0;
    for (int e = tid; e < tot; e += nthr) {
      int l, j;
      pass_elem(a, e, n, l, j);
      bool valid;
      const long base = pass_line_base(a, item, l, valid);
      if (valid) acc += dot_term(a, base + (long)j * a.stride, r[l * a.ldn + j]);
    }
    block_dot_out(acc, a.dot_out);
  } else {
    const T sc1 = a.dexp1 ? exp2i<T>(a.dexp1[a.sidx]) : T(1);
    const T sc2 = a.dexp2 ? exp2i<T>(a.dexp2[a.sidx]) : T(1);
    for (int e = tid; e < tot; e += nthr) {
      int l, j;
      pass_elem(a, e, n, l, j);
      bool valid;
      const long base = pass_line_base(a, item, l, valid);
      if (!valid) continue;
      const long g = base + (long)j * a.stride;
      C2<T> v = r[l * a.ldn + j];
      if (a.post_E) {
        v.x = v.x * a.Ea1[g] * sc1;
        v.y = v.y * a.Ea2[g] * sc2;
      }
      a.h1[g] = v.x;
      if (a.h2) a.h2[g] = v.y;
    }
  }
}

// ---------------------------------------------------------------------------
// register-resident line pass for n = N1 * N2 (fft_reg.hpp)
// ---------------------------------------------------------------------------

// The exchanges inside one line's operator.  A wavefront's LDS operations are performed
// in issue order, so when every line of the workgroup lies inside one wavefront the
// phases need only a compiler fence at wavefront scope: no workgroup barrier and no wait
// for the LDS stores to complete before the loads that read them are issued.  Otherwise
// (lines spread over the waves of a multi-wave workgroup) a workgroup barrier.
__device__ __forceinline__ void line_sync(bool wave_local) {
  if (wave_local) {  // uniform over the block
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Whether every line of the block lies inside one wavefront: a one-wave block, or an x
// pass (thread t of line l at lane l TPL + t) whose threads per line divide 64.  The y and
// z passes put thread t of line l at l + t L, which spreads a line over the block.
template <int TPL>
__device__ __forceinline__ bool lines_in_wave(int axis) {
  return blockDim.x <= 64 || (axis == 0 && 64 % TPL == 0);
}

template <typename T, int N1, int N2>
__device__ __forceinline__ void reg_line_op(C2<T> (&v)[N2], C2<T>* lds, int t, const RegTables<T, N1, N2>& R,
                                            bool wl) {
  if constexpr (N1 == 1) {
    line_op_single<T, N2>(v, R);
  } else {
    // every caller runs exactly max(N1, N2) threads per line (t < TPL), so the guard of the
    // phase whose thread count is TPL is always true and is resolved at compile time
    constexpr bool allA = N1 >= N2, allB = N2 >= N1;
    if (allA || t < N1) fs_phase_a<T, N1, N2>(v, lds, t, R);
    line_sync(wl);
    if (allB || t < N2) fs_phase_b<T, N1, N2>(lds, t, R);
    line_sync(wl);
    if (allA || t < N1) fs_phase_c<T, N1, N2>(v, lds, t);
  }
}

template <int N1, int N2>
constexpr int reg_tpl() {
  return N1 == 1 ? 1 : FourStep<N1, N2>::tpl;
}

// Register line engines of a pass: TPL threads per line, NP points per thread,
// thread t (t < NACT) holds points t + STRIDE * j (j < NP); LDS elements per line.
template <typename T, int N1, int N2>
struct EngFour {  // n = N1 * N2 four-step (fft_reg.hpp)
  static constexpr int NP = N2, TPL = reg_tpl<N1, N2>(), NACT = N1, STRIDE = N1;
  static constexpr int LDS = N1 == 1 ? 0 : FourStep<N1, N2>::lds_elems;
  using Tables = RegTables<T, N1, N2>;
  static __device__ __forceinline__ void op(C2<T> (&v)[NP], C2<T>* lds, int t, const Tables& R, bool wl) {
    reg_line_op<T, N1, N2>(v, lds, t, R, wl);
  }
};

template <typename T, int R>
struct EngCube {  // n = R^3 on R^2 threads (fft_cube.hpp)
  static constexpr int NP = R, TPL = R * R, NACT = R * R, STRIDE = R * R;
  static constexpr int LDS = Cube<R>::lds_elems;
  using Tables = CubeTables<T, R>;
  static __device__ __forceinline__ void op(C2<T> (&v)[NP], C2<T>* lds, int t, const Tables& W, bool wl) {
    cube_a<T, R>(v, lds, t, W);
    line_sync(wl);
    cube_b<T, R>(lds, t, W);
    line_sync(wl);
    cube_c<T, R>(lds, t, W);
    line_sync(wl);
    cube_bi<T, R>(lds, t, W);
    line_sync(wl);
    cube_ai<T, R>(v, lds, t, W);
  }
};

template <typename T, int A, int R>
struct EngTri {  // n = A R^2 on R^2 threads (fft_cube.hpp, tri_*)
  static constexpr int NP = A, TPL = R * R, NACT = R * R, STRIDE = R * R;
  static constexpr int LDS = Tri<A, R>::lds_elems;
  using Tables = TriTables<T, A, R>;
  static __device__ __forceinline__ void op(C2<T> (&v)[NP], C2<T>* lds, int t, const Tables& W, bool wl) {
    tri_a<T, A, R>(v, lds, t, W);
    line_sync(wl);
    tri_b<T, A, R>(lds, t, W);
    line_sync(wl);
    tri_c<T, A, R>(lds, t, W);
    line_sync(wl);
    tri_bi<T, A, R>(lds, t, W);
    line_sync(wl);
    tri_ai<T, A, R>(v, lds, t, W);
  }
};

// shape (N1, N2) of the schedule tables: N2 == 0 selects the cube engine with R = N1,
// N2 < 0 the A R^2 engine with A = -N2 and R = N1, N2 > 0 the N1 x N2 four-step
template <typename T, int N1, int N2>
using PassEngine = typename std::conditional<
    N2 == 0, EngCube<T, N1>,
    typename std::conditional<(N2 < 0), EngTri<T, (N2 < 0 ? -N2 : N1), N1>, EngFour<T, N1, N2 <= 0 ? 1 : N2>>::type>::type;

// Pass variants fixed at compile time.  The four shapes of a propagator sweep
// (first pass from q = 1, plain single-operator pass, merged pass with the mid
// store, last pass with planar output) get kernels without runtime branches on
// the pass arguments, so that every argument load is issued at entry;
// PM_GENERIC reads the kind fields at run time (adjoint, compat, batched).
enum PassMode : int { PM_GENERIC = 0, PM_FIRST = 1, PM_PLAIN = 2, PM_MERGED = 3, PM_LAST = 4 };

// which fixed variant a pass matches (host side); PM_GENERIC if none
template <typename T>
inline int pass_mode(const PassArgs<T>& p) {
  const bool both = p.h2 != nullptr;
  if (p.in_kind == IN_ONES && both && p.pre_E == 1 && p.nops == 1 && p.out_kind == OUT_TMP) return PM_FIRST;
  if (p.in_kind == IN_TMP && p.pre_E == 0 && p.nops == 1 && p.out_kind == OUT_TMP) return PM_PLAIN;
  if (p.in_kind == IN_TMP && p.pre_E == 0 && p.nops == 2 && p.out_kind == OUT_TMP && both && p.dexp1 && p.dexp2)
    return PM_MERGED;
  if (p.in_kind == IN_TMP && p.pre_E == 0 && p.nops == 1 && p.out_kind == OUT_PLANAR && p.post_E && both &&
      p.dexp1 && p.dexp2)
    return PM_LAST;
  return PM_GENERIC;
}

// Same contract as k_line_pass; a line lives in registers of ENG::TPL threads,
// thread t < ENG::NACT holding the points t + STRIDE * j.  blockDim = TPL * L.
// At the small grids of the SCF loop a pass has few waves per SIMD, so its
// time is mostly memory latency: every global load of the pass (the stop flag,
// the lines, the E factors of both steps, twiddles, multipliers and rescale
// exponents) is issued before the first use, and the pass pays one latency.
// Loads and stores at a 32-bit byte offset from a uniform base pointer: the
// compiler then addresses with the scalar base plus one vector offset register
// instead of 64-bit vector address arithmetic per access.  Every array a pass
// touches is below 2 GiB (the context limits the grid size).
template <typename V>
__device__ __forceinline__ V ldo(const V* base, unsigned off) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + off);
}
template <typename V>
__device__ __forceinline__ void sto(V* base, unsigned off, V x) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + off) = x;
}
// Nontemporal store, for the q(s) slices: no later pass of the sweep reads them (the contour
// integral does, after the sweep), so they are written without being kept in the caches.
template <typename V>
__device__ __forceinline__ void sto_nt(V* base, unsigned off, V x) {
  static_assert(std::is_arithmetic<V>::value, "slices are stored one real value at a time");
  __builtin_nontemporal_store(x, reinterpret_cast<V*>(reinterpret_cast<char*>(base) + off));
}
// Line passes over lines of up to this many points store the slices nontemporally: measured
// faster on 64^3 grids (4% more iterations/s), while on 128^3 the line passes' slice stores,
// a few consecutive lines wide, went out as partial cache lines and cost 11% of the sweep.
// Plane passes, whose stores cover whole cache lines, always store the slices so.
constexpr int kNtSliceMaxLine = 64;
template <bool NT, typename V>
__device__ __forceinline__ void sto_slice(V* base, unsigned off, V x) {
  if constexpr (NT) sto_nt(base, off, x);
  else sto(base, off, x);
}

// Coherent access to the packed line buffer for the persistent sweep, where the pass that
// wrote a line and the pass that reads it are workgroups of one kernel on different XCDs:
// relaxed atomic loads and stores at agent scope, 64 bits at a time, which bypass the
// non-coherent caches (an XCD's L2 may hold a stale copy of the line from an earlier pass)
// and write through to the device-wide coherence point.
template <typename T>
__device__ __forceinline__ C2<T> ld_co(const C2<T>* base, unsigned off) {
  const char* p = reinterpret_cast<const char*>(base) + off;
  if constexpr (sizeof(T) == 4) {
    const unsigned long long u =
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    C2<T> v;
    __builtin_memcpy(&v, &u, sizeof v);
    return v;
  } else {
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
    const unsigned long long ux = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long uy = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    C2<T> v;
    __builtin_memcpy(&v.x, &ux, sizeof ux);
    __builtin_memcpy(&v.y, &uy, sizeof uy);
    return v;
  }
}
template <typename T>
__device__ __forceinline__ void st_co(C2<T>* base, unsigned off, C2<T> v) {
  char* p = reinterpret_cast<char*>(base) + off;
  if constexpr (sizeof(T) == 4) {
    unsigned long long u;
    __builtin_memcpy(&u, &v, sizeof u);
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    unsigned long long ux, uy;
    __builtin_memcpy(&ux, &v.x, sizeof ux);
    __builtin_memcpy(&uy, &v.y, sizeof uy);
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    __hip_atomic_store(q, ux, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, uy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T, class ENG, int MODE = PM_GENERIC, bool CO = false>
__device__ __forceinline__ void reg_pass_body(const PassArgs<T>& a, int item, unsigned char* smem_raw) {
  constexpr int NP = ENG::NP, STRIDE = ENG::STRIDE;
  constexpr bool kNtH = NP * STRIDE <= kNtSliceMaxLine;
  constexpr bool G = MODE == PM_GENERIC;
  const int in_kind = G ? a.in_kind : (MODE == PM_FIRST ? (int)IN_ONES : (int)IN_TMP);
  const int nops = G ? a.nops : (MODE == PM_MERGED ? 2 : 1);
  const int out_kind = G ? a.out_kind : (MODE == PM_LAST ? (int)OUT_PLANAR : (int)OUT_TMP);
  const int pre_E = G ? a.pre_E : (MODE == PM_FIRST ? 1 : 0);
  const int post_E = G ? a.post_E : (MODE == PM_LAST ? 1 : 0);
  const bool two = G ? a.h2 != nullptr : true;  // second chain present
  constexpr int TPL = ENG::TPL;
  const int tid = threadIdx.x;
  int l, t;
  if (a.axis == 0) {
    t = tid % TPL;
    l = tid / TPL;
  } else {
    l = tid % a.L;
    t = tid / a.L;
  }
  bool valid;
  const long base = pass_line_base(a, item, l, valid);
  // byte offsets of this thread's points j2 in a real (fo) and a complex (fc) array
  const unsigned f0 = (unsigned)(base + (long)t * a.stride) * (unsigned)sizeof(T);
  const unsigned fstep = (unsigned)(STRIDE * a.stride) * (unsigned)sizeof(T);
  auto fo = [&](int j2) { return f0 + (unsigned)j2 * fstep; };
  auto fc = [&](int j2) { return 2u * (f0 + (unsigned)j2 * fstep); };
  C2<T>* lds = reinterpret_cast<C2<T>*>(smem_raw) + l * a.lds_line;
  const bool wl = lines_in_wave<TPL>(a.axis);
  const bool act = t < ENG::NACT && valid;
  const int stop = *(a.done ? a.done : &g_no_stop);
  const bool useEa = pre_E != 0 || nops == 2 || out_kind == OUT_DOT || (out_kind == OUT_PLANAR && post_E);
  const bool useEb = nops == 2;
  // FP64 engines with 8 or more points per thread would hold 4 NP E factors (8 NP VGPRs)
  // through the first operator and spill there (16x8 merged: 56 VGPRs spilled); they load
  // the E factors that operator does not need after it instead (one more memory latency per
  // pass, hidden by the other wavefronts at the grid sizes these shapes serve: 128^3 FP64
  // 27.4 against 31.5 us per pass).  The FP32 128-point engines gain nothing from it, and
  // the engines of the 64^3 sweep keep every load up front.
  constexpr bool LEAN = sizeof(T) == 8 && NP >= 8;
  const bool earlyEa = useEa && (!LEAN || pre_E != 0);
  const bool earlyEb = useEb && !LEAN;
  C2<T> v[NP];
  T ea1[NP], ea2[NP], eb1[NP], eb2[NP];
#pragma unroll
  for (int j2 = 0; j2 < NP; ++j2) {
    C2<T> x = {T(0), T(0)};
    ea1[j2] = ea2[j2] = eb1[j2] = eb2[j2] = T(1);
    if (act) {
      if (in_kind == IN_TMP) {
        x = CO ? ld_co(a.tmp_in, fc(j2)) : ldo(a.tmp_in, fc(j2));
      } else if (in_kind == IN_ONES) {
        x = {T(1), two ? T(1) : T(0)};
      } else {
        x.x = ldo(a.q1_in, fo(j2));
        x.y = a.q2_in ? ldo(a.q2_in, fo(j2)) : T(0);
      }
      if (earlyEa) {
        ea1[j2] = ldo(a.Ea1, fo(j2));
        ea2[j2] = ldo(a.Ea2, fo(j2));
      }
      if (earlyEb) {
        eb1[j2] = ldo(a.Eb1, fo(j2));
        eb2[j2] = ldo(a.Eb2, fo(j2));
      }
    }
    v[j2] = x;
  }
  typename ENG::Tables R;
  R.load(t, a.tw, a.kf);
  const bool scaled = G ? a.dexp1 != nullptr : (MODE == PM_MERGED || MODE == PM_LAST);
  const T sc1 = scaled ? exp2i<T>(a.dexp1[a.sidx]) : T(1);
  const T sc2 = (G ? a.dexp2 != nullptr : scaled) ? exp2i<T>(a.dexp2[a.sidx]) : T(1);
  if (pre_E > 0) {
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      v[j2].x *= ea1[j2];
      v[j2].y *= ea2[j2];
    }
  } else if (pre_E < 0) {
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      v[j2].x /= ea1[j2];
      v[j2].y /= ea2[j2];
    }
  }
  ENG::op(v, lds, t, R, wl);
  // The stop flag is tested only here, after the first operator, so its load flies with
  // every load above, and nothing has been stored yet.
  if (stop) return;  // uniform over the block
  if (LEAN && act) {
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      if (useEa && !earlyEa) {
        ea1[j2] = ldo(a.Ea1, fo(j2));
        ea2[j2] = ldo(a.Ea2, fo(j2));
      }
      if (useEb) {
        eb1[j2] = ldo(a.Eb1, fo(j2));
        eb2[j2] = ldo(a.Eb2, fo(j2));
      }
    }
  }
  if (nops == 2) {
    if (act) {
#pragma unroll
      for (int j2 = 0; j2 < NP; ++j2) {
        C2<T> x = v[j2];
        x.x = x.x * ea1[j2] * sc1;
        x.y = x.y * ea2[j2] * sc2;
        sto_slice<kNtH>(a.h1, fo(j2), x.x);
        sto_slice<kNtH>(a.h2, fo(j2), x.y);
        x.x *= eb1[j2];
        x.y *= eb2[j2];
        v[j2] = x;
      }
    }
    ENG::op(v, lds, t, R, wl);
  }
  if (out_kind == OUT_DOT) {
    double acc = 0.0;
    if (act) {
#pragma unroll
      for (int j2 = 0; j2 < NP; ++j2) acc += dot_term(a, base + (long)(t + STRIDE * j2) * a.stride, v[j2]);
    }
    block_dot_out(acc, a.dot_out);
    return;
  }
  if (!act) return;
  if (out_kind == OUT_TMP) {
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      if constexpr (CO) st_co(a.tmp_out, fc(j2), v[j2]);
      else sto(a.tmp_out, fc(j2), v[j2]);
    }
  } else {
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      C2<T> x = v[j2];
      if (post_E) {
        x.x = x.x * ea1[j2] * sc1;
        x.y = x.y * ea2[j2] * sc2;
      }
      sto_slice<kNtH>(a.h1, fo(j2), x.x);
      if (two) sto_slice<kNtH>(a.h2, fo(j2), x.y);
    }
  }
}

// Tile of the K-tile pass: block g (on XCD g % 8) takes tiles j = 0..K-1 of its XCD's
// contiguous range, so that the tiles in flight together on one XCD stay neighbours.
// Returns -1 past the range.  gridDim.x is a multiple of kXcds (launch_reg sizes it).
__device__ __forceinline__ int xcd_tile_k(int g, int j, int nb, int G) {
  const int x = g % 8, i = g / 8, gx = G / 8;
  const int q = nb / 8, r = nb % 8;
  const int start = x * q + (x < r ? x : r), len = q + (x < r ? 1 : 0);
  const int k = i + j * gx;
  return k < len ? start + k : -1;
}

// A pass whose workgroups each take K tiles in turn (fixed sweep modes only): the loads of
// tile j+1 are issued before tile j is transformed and stored, so that a wave's memory
// traffic overlaps its own arithmetic and the reads of some waves overlap the writes of
// others (with one tile per wave every wave of a large grid reads, computes and writes in
// step).  Every point takes the operations of reg_pass_body in the same order, so results
// are bit-identical to the one-tile pass.
template <typename T, class ENG, int MODE, int K>
__device__ __forceinline__ void reg_pass_tiles(const PassArgs<T>& a, int nitems, unsigned char* smem_raw) {
  static_assert(MODE != PM_GENERIC, "fixed modes only");
  constexpr int NP = ENG::NP, STRIDE = ENG::STRIDE, TPL = ENG::TPL;
  constexpr bool kNtH = NP * STRIDE <= kNtSliceMaxLine;
  constexpr bool FIRST = MODE == PM_FIRST, MERGED = MODE == PM_MERGED, LAST = MODE == PM_LAST;
  const int tid = threadIdx.x;
  int l, t;
  if (a.axis == 0) {
    t = tid % TPL;
    l = tid / TPL;
  } else {
    l = tid % a.L;
    t = tid / a.L;
  }
  C2<T>* lds = reinterpret_cast<C2<T>*>(smem_raw) + l * a.lds_line;
  const bool wl = lines_in_wave<TPL>(a.axis);
  const int stop = *(a.done ? a.done : &g_no_stop);
  const unsigned fstep = (unsigned)(STRIDE * a.stride) * (unsigned)sizeof(T);
  struct Tile {
    C2<T> v[NP];
    T ea1[NP], ea2[NP], eb1[NP], eb2[NP];
    unsigned f0;
    bool act;
  };
  auto load = [&](int item, Tile& w) {
    bool valid = false;
    long base = 0;
    if (item >= 0) base = pass_line_base(a, item, l, valid);
    w.act = item >= 0 && t < ENG::NACT && valid;
    w.f0 = (unsigned)(base + (long)t * a.stride) * (unsigned)sizeof(T);
#pragma unroll
    for (int j2 = 0; j2 < NP; ++j2) {
      const unsigned fo = w.f0 + (unsigned)j2 * fstep;
      C2<T> x = {T(0), T(0)};
      w.ea1[j2] = w.ea2[j2] = w.eb1[j2] = w.eb2[j2] = T(1);
      if (w.act) {
        if (FIRST) x = {T(1), T(1)};
        else x = ldo(a.tmp_in, 2u * fo);
        if (FIRST || MERGED || LAST) {
          w.ea1[j2] = ldo(a.Ea1, fo);
          w.ea2[j2] = ldo(a.Ea2, fo);
        }
        if (MERGED) {
          w.eb1[j2] = ldo(a.Eb1, fo);
          w.eb2[j2] = ldo(a.Eb2, fo);
        }
      }
      w.v[j2] = x;
    }
  };
  typename ENG::Tables R;
  R.load(t, a.tw, a.kf);
  const T sc1 = (MERGED || LAST) ? exp2i<T>(a.dexp1[a.sidx]) : T(1);
  const T sc2 = (MERGED || LAST) ? exp2i<T>(a.dexp2[a.sidx]) : T(1);
  const int G = gridDim.x;
  Tile cur, nxt;
  int it_cur = xcd_tile_k(blockIdx.x, 0, nitems, G);
  load(it_cur, cur);
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int it_nxt = j + 1 < K ? xcd_tile_k(blockIdx.x, j + 1, nitems, G) : -1;
    if (j + 1 < K) load(it_nxt, nxt);
    if (it_cur >= 0) {  // uniform over the block
      if (FIRST) {
#pragma unroll
        for (int j2 = 0; j2 < NP; ++j2) {
          cur.v[j2].x *= cur.ea1[j2];
          cur.v[j2].y *= cur.ea2[j2];
        }
      }
      ENG::op(cur.v, lds, t, R, wl);
      if (stop) return;  // uniform over the block
      if (MERGED) {
        if (cur.act) {
#pragma unroll
          for (int j2 = 0; j2 < NP; ++j2) {
            const unsigned fo = cur.f0 + (unsigned)j2 * fstep;
            C2<T> x = cur.v[j2];
            x.x = x.x * cur.ea1[j2] * sc1;
            x.y = x.y * cur.ea2[j2] * sc2;
            sto_slice<kNtH>(a.h1, fo, x.x);
            sto_slice<kNtH>(a.h2, fo, x.y);
            x.x *= cur.eb1[j2];
            x.y *= cur.eb2[j2];
            cur.v[j2] = x;
          }
        }
        ENG::op(cur.v, lds, t, R, wl);
      }
      if (cur.act) {
#pragma unroll
        for (int j2 = 0; j2 < NP; ++j2) {
          const unsigned fo = cur.f0 + (unsigned)j2 * fstep;
          if (LAST) {
            C2<T> x = cur.v[j2];
            x.x = x.x * cur.ea1[j2] * sc1;
            x.y = x.y * cur.ea2[j2] * sc2;
            sto_slice<kNtH>(a.h1, fo, x.x);
            sto_slice<kNtH>(a.h2, fo, x.y);
          } else {
            sto(a.tmp_out, 2u * fo, cur.v[j2]);
          }
        }
      }
    }
    if (j + 1 < K) {
      cur = nxt;
      it_cur = it_nxt;
    }
  }
}

// Makes every kernel-argument field the pass reads live in scalar registers at
// entry, so the compiler issues their constant-cache loads as one batch rather
// than one batch per branch of the pass body.  (The empty asm emits nothing.)
template <typename T>
__device__ __forceinline__ void pin_pass_args(const PassArgs<T>& a) {
  asm volatile("" ::"s"(a.nx), "s"(a.ny), "s"(a.nz), "s"(a.axis), "s"(a.L), "s"(a.gx), "s"(a.stride), "s"(a.tw),
               "s"(a.kf), "s"(a.tmp_in), "s"(a.Ea1), "s"(a.Ea2), "s"(a.Eb1), "s"(a.Eb2), "s"(a.h1), "s"(a.h2),
               "s"(a.dexp1), "s"(a.dexp2), "s"(a.sidx), "s"(a.tmp_out), "s"(a.done));
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Giving XCD k a contiguous range of tiles keeps tiles that share 128-byte lines
// (neighbouring x tiles of the y and z passes) in one L2, and gives the x and y
// passes of the same z-planes the same XCD, so a line fetched or written by one
// pass is found in that XCD's L2 by the next.
constexpr int kXcds = 8;

__host__ __device__ __forceinline__ int xcd_tile(int b, int nb) {
  if (nb < kXcds) return b;
  const int q = nb / kXcds, r = nb % kXcds;
  const int k = b % kXcds, i = b / kXcds;
  return k * q + (k < r ? k : r) + i;
}

template <typename T, int N1, int N2, int MODE>
__global__ void __launch_bounds__(kThreads) k_line_pass_reg(PassArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  if constexpr (MODE != PM_GENERIC) pin_pass_args(a);
  reg_pass_body<T, PassEngine<T, N1, N2>, MODE>(a, xcd_tile(blockIdx.x, gridDim.x), smem_raw);
}

// K tiles per workgroup (reg_pass_tiles); nitems is the tile count of the pass
template <typename T, int N1, int N2, int MODE, int K>
__global__ void __launch_bounds__(kThreads) k_line_pass_tiles(PassArgs<T> a, int nitems) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  pin_pass_args(a);
  reg_pass_tiles<T, PassEngine<T, N1, N2>, MODE, K>(a, nitems, smem_raw);
}

// One pass of many independent instances that share a register shape, in one
// launch: blk[b] packs the instance slot of block b (bits 20 and up) and the
// block's item within that instance (low 20 bits); every instance carries its own
// arguments (grid, buffers, step types).  Line-pass blocks are kBatchLineWaves
// wavefronts of lines, taken in xcd_tile order by the host's table.  A pass of a
// batch is one such launch per shape present (the shape is a template argument,
// so each launch has the register budget of its own engine).
template <typename T>
struct BatchGroup {
  const PassArgs<T>* args;
  const int* blk;  // one entry per block
};

constexpr int kBatchItemBits = 20;
// wavefronts per workgroup of the batched line passes (lines per tile: 64 kBatchLineWaves / TPL)
constexpr int kBatchLineWaves = 4;

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(64 * kBatchLineWaves) k_line_pass_batch(BatchGroup<T> g) {
  const int code = g.blk[blockIdx.x];
  const int i = code >> kBatchItemBits;
  const int item = code & ((1 << kBatchItemBits) - 1);
  const PassArgs<T> a = g.args[i];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  reg_pass_body<T, PassEngine<T, N1, N2>>(a, item, smem_raw);
}

// ---------------------------------------------------------------------------
// persistent batched sweep: every pass of every instance in one launch
// ---------------------------------------------------------------------------
// A work item is one 256-thread tile of one pass of one instance.  Items are ordered pass
// by pass (then instance, then tile) and dealt to the resident workgroups round-robin; each
// workgroup runs its items in that order.  An item waits only until every tile of the
// previous pass of its own instance is done (a per-pass completion counter), so instances
// advance independently and a workgroup never idles on a global barrier.  An item depends
// only on items earlier in the order, and every workgroup is resident (the grid is sized by
// the occupancy), so the earliest unfinished item can always run.  Waits are bounded: a
// workgroup that waits past the spin limit records it in *err and leaves the kernel, and
// the host reports the sweep as failed.
constexpr int kPersistThreads = 256;
constexpr int kPersistSpinLimit = 2000000;  // polls (~1 us each) before a wait gives up

struct PersistItem {
  int pass;  // flat pass index: instance base + pass number
  int tile;  // tile of that pass
};

template <typename T>
struct PersistArgs {
  const PassArgs<T>* args;  // one per flat pass (256-thread tile geometry)
  const int* shape;         // shape id of each flat pass (index into the kernel's shape list)
  const int* need;          // tiles of the previous pass of the same instance (0: first pass)
  const PersistItem* items;
  int nitems;
  int* done;                // tiles completed per flat pass; zeroed before each sweep
  int* err;
  int spin_limit;
};

template <typename T, int... S>
struct PersistShapes;
template <typename T>
struct PersistShapes<T> {
  static __device__ __forceinline__ void run(int, const PassArgs<T>&, int, unsigned char*) {}
};
template <typename T, int A, int B, int... R>
struct PersistShapes<T, A, B, R...> {
  static __device__ __forceinline__ void run(int id, const PassArgs<T>& a, int tile, unsigned char* sm) {
    if (id == 0) reg_pass_body<T, PassEngine<T, A, B>, PM_GENERIC, true>(a, tile, sm);
    else PersistShapes<T, R...>::run(id - 1, a, tile, sm);
  }
};

template <typename T, int... S>
__global__ void __launch_bounds__(kPersistThreads) k_batch_persist(PersistArgs<T> pa) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int s_fail;
  for (int k = blockIdx.x; k < pa.nitems; k += gridDim.x) {
    const PersistItem it = pa.items[k];
    if (threadIdx.x == 0) {
      int fail = 0;
      const int need = pa.need[it.pass];
      if (need > 0) {
        int spins = 0;
        while (__hip_atomic_load(pa.done + it.pass - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++spins > pa.spin_limit) {
            fail = 1;
            __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      s_fail = fail;
    }
    __syncthreads();
    if (s_fail) return;  // uniform: the previous pass never completed
    // the line buffer is read and written with agent-scope atomics (ld_co / st_co), so the
    // handshake needs no cache-wide fence: each wave waits until its stores are complete
    // (written through to the coherence point) before the workgroup reports the tile done,
    // and the next pass issues its loads only after it has seen the count
    const PassArgs<T> a = pa.args[it.pass];
    PersistShapes<T, S...>::run(pa.shape[it.pass], a, it.tile, smem_raw);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(pa.done + it.pass, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// the shapes one persistent kernel carries: (N1, N2) pairs, N2 == 0 the cube engine
constexpr int kPersistShapes[][2] = {{4, 8}, {8, 4}, {8, 6}, {6, 8}, {8, 8}, {4, 0}};
constexpr int kNumPersistShapes = 6;

constexpr int kBatchShapes[][2] = {{4, 8}, {8, 4}, {2, 16}, {8, 6}, {6, 8}, {16, 3}, {8, 8},
                                   {4, 16}, {16, 4}, {8, 16}, {16, 8}, {1, 16}, {1, 8}, {4, 0}};

// ---------------------------------------------------------------------------
// generic line FFT along one axis of a complex array (compat r2c / c2r)
// ---------------------------------------------------------------------------
enum FftLoad : int { LOAD_CPLX = 0, LOAD_REAL = 1, LOAD_HERM = 2 };
enum FftStore : int { STORE_CPLX = 0, STORE_HALF = 1, STORE_REAL = 2 };

template <typename T>
struct FftArgs {
  int d0, d1, d2;  // array dims of the complex input, x fastest
  int axis;
  int n;           // transform length
  int L, ldn, gx;
  long stride;
  LinePlan plan;
  const C2<T>* tw;
  int dir;
  T scale;
  int load, store;
  const T* rin;       // LOAD_REAL: real rows of length n (x axis)
  const C2<T>* cin;   // LOAD_CPLX / LOAD_HERM (rows of length d0 = n/2+1)
  C2<T>* cout;        // STORE_CPLX / STORE_HALF (rows of length n/2+1)
  T* rout;            // STORE_REAL
};

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fft_axis(FftArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* b0 = reinterpret_cast<C2<T>*>(smem_raw);
  C2<T>* b1 = b0 + a.L * a.ldn;
  C2<T>* twl = b1 + a.L * a.ldn;
  const int n = a.n, tot = a.L * n, item = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nh = n / 2 + 1;
  for (int e = tid; e < n; e += nthr) twl[e] = a.tw[e];
  for (int e = tid; e < tot; e += nthr) {
    int l, j;
    if (a.axis == 0) { l = e / n; j = e - l * n; } else { l = e % a.L; j = e / a.L; }
    C2<T> v = {T(0), T(0)};
    if (a.axis == 0) {
      const long r = (long)item * a.L + l;
      if (r < (long)a.d1 * a.d2) {
        if (a.load == LOAD_REAL) v = {a.rin[r * n + j], T(0)};
        else if (a.load == LOAD_HERM) {
          if (j < nh) v = a.cin[r * nh + j];
          else { C2<T> c = a.cin[r * nh + (n - j)]; v = {c.x, -c.y}; }
        } else v = a.cin[r * n + j];
      }
    } else {
      const int x0 = (item % a.gx) * a.L, o = item / a.gx;
      if (x0 + l < a.d0) {
        const long base = (a.axis == 1) ? (long)o * a.d0 * a.d1 + x0 + l : (long)o * a.d0 + x0 + l;
        v = a.cin[base + (long)j * a.stride];
      }
    }
    b0[l * a.ldn + j] = v;
  }
  __syncthreads();
  auto sync = [] { __syncthreads(); };
  C2<T>* r = line_fft<T>(b0, b1, a.plan, a.ldn, a.L, twl, a.dir, tid, nthr, sync);
  for (int e = tid; e < tot; e += nthr) {
    int l, j;
    if (a.axis == 0) { l = e / n; j = e - l * n; } else { l = e % a.L; j = e / a.L; }
    C2<T> v = r[l * a.ldn + j];
    v = {v.x * a.scale, v.y * a.scale};
    if (a.axis == 0) {
      const long rr = (long)item * a.L + l;
      if (rr < (long)a.d1 * a.d2) {
        if (a.store == STORE_HALF) { if (j < nh) a.cout[rr * nh + j] = v; }
        else if (a.store == STORE_REAL) a.rout[rr * n + j] = v.x;
        else a.cout[rr * n + j] = v;
      }
    } else {
      const int x0 = (item % a.gx) * a.L, o = item / a.gx;
      if (x0 + l < a.d0) {
        const long base = (a.axis == 1) ? (long)o * a.d0 * a.d1 + x0 + l : (long)o * a.d0 + x0 + l;
        a.cout[base + (long)j * a.stride] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// device state of the SCF loop and chain description
// ---------------------------------------------------------------------------
struct ChainDesc {
  int nblocks;
  int ns;                      // total segments
  int type[kMaxBlocks];        // 0-based monomer type per block
  int seg[kMaxBlocks];         // block_segments
  int start[kMaxBlocks];       // first segment index of each block
  int ensemble;                // 1 canonical, 2 grand canonical
};

struct FHDesc {
  int nt;
  double kappa;
  double chi[kMaxTypes * kMaxTypes];  // column-major
  double vol[kMaxTypes];
};

// Scalars shared by the kernels of one SCF iteration; lives in device memory.
struct ScfState {
  int head;        // physical ring slot of logical history index 0
  int iter;        // iterations evaluated (reference loop index of the last one)
  int done;        // converged: every later kernel returns at once
  int mode;        // 0 simple (lam00), 1 Anderson
  int singular;    // ludcmp saw an all-zero row
  int pad_;
  double tol;
  double accuracy; // max|dfield| of the last evaluation
  double lam;      // step used by the last mix
  double qsum;     // Qsum of the last evaluation
  double norm;     // density = norm * distribution
  double mu, rho, energy;  // chain scalars of the last evaluation
  double x[kMaxHist];      // Anderson coefficients
};

}  // namespace adscf
