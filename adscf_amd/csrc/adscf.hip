// libadscf_hip.so: host side of the MI355X SCFT propagator path.
//
// The context owns every device buffer and one HIP stream.  Entry points that
// take host arrays (the reference-compatible ones) copy in, run on the GPU and
// copy out; the device-resident SCF loop keeps fields, propagators and the
// Anderson history in HBM and replays one captured HIP graph per iteration.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/adscf.h"
#include "kernels.hpp"
#include "plane.hpp"
#include "lu.hpp"
#include "scf_kernels.hpp"
#include "adjoint.hpp"

using namespace adscf;

// Device scratch of adscf_propagate_chain_noalloc: the propagator history, the packed line
// buffer, rescale exponents, reduction partials and the chain state of one evaluation live
// here instead of in allocations.  Each GPU the library runs on holds its own copy for the
// life of the process.  Host code never passes this name to a kernel (on the host it is a
// stand-in, not a device address): it uses the per-device address from hipGetSymbolAddress.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kRedBlocks = 512;  // blocks of the grid-stride reduction kernels
constexpr long kMaxGridPoints = 1L << 27;  // 512^3; a complex FP64 grid array stays below 2 GiB

struct HipError {
  hipError_t e;
  const char* what;
  int line;
};

#define HIPCK(x)                                      \
  do {                                                \
    hipError_t e__ = (x);                             \
    if (e__ != hipSuccess) throw HipError{e__, #x, __LINE__}; \
  } while (0)

struct ArgError {
  std::string msg;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t want) {
    if (want <= n) return;
    release();
    HIPCK(hipMalloc((void**)&p, std::max<size_t>(want, 1) * sizeof(T)));
    n = want;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

struct AxisTables {
  int n = 1;
  LinePlan plan{};
  int L = 1;      // Stockham tile width
  int N1 = 0;     // register-resident shape (0: Stockham kernel)
  int N2 = 0;
  int waves = 1;  // wavefronts per workgroup of the register kernel
  int tiles = 1;  // tiles per workgroup of the sweep's fixed-mode passes (k_line_pass_tiles)
};

// Captured graphs live until the context (or batch) is torn down: an executable graph that
// is replaced (new buffers, another loop, a tuned schedule) is retired rather than destroyed,
// and the sweep graphs of the timing entry points are kept per schedule and reused.  No
// graph object is destroyed while the process may still hold work or tool records of it,
// and a new executable graph never takes the address of one destroyed earlier in the
// context's life.
struct GraphPair {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
};
struct GraphKeep {
  std::vector<GraphPair> retired;
  std::vector<std::pair<std::string, GraphPair>> sweeps;  // timing sweeps by schedule key
  void retire(hipGraph_t& g, hipGraphExec_t& e) {
    if (g || e) retired.push_back({g, e});
    g = nullptr;
    e = nullptr;
  }
  GraphPair* find(const std::string& key) {
    for (auto& kv : sweeps)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  // after the owning stream has been synchronised
  void release() {
    for (GraphPair& p : retired) {
      if (p.e) (void)hipGraphExecDestroy(p.e);
      if (p.g) (void)hipGraphDestroy(p.g);
    }
    for (auto& kv : sweeps) {
      if (kv.second.e) (void)hipGraphExecDestroy(kv.second.e);
      if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
    }
    retired.clear();
    sweeps.clear();
  }
};

// Per-precision engine: the propagator arrays and FFT tables in T.
template <typename T>
struct Engine {
  DevBuf<C2<T>> tw[3];
  DevBuf<T> kf[3];
  DevBuf<T> E;       // nt x ng
  DevBuf<C2<T>> tmp; // ng complex, in-place line passes
  DevBuf<T> h1, h2;  // (ns+1) x ng each
  size_t hist_slices = 0;
  DevBuf<T> cq;      // compat single-field propagate
  DevBuf<C2<T>> ch;  // compat half spectrum
  DevBuf<T> cmixE;   // E output of the compat mixing kernel (unused values)
  // adjoint: contour checkpoints, recompute segments, adjoint states
  DevBuf<T> ck1, ck2, sa1, sa2, sb1, sb2, mu, nu, y, ones, scr;
  DevBuf<T> kd[3];   // k_a^2 / n_a multipliers of the D_a line operators
};

}  // namespace

struct adscf_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  int nx = 1, ny = 1, nz = 1, nt = 2;
  long ng = 1;
  double lx = 1, ly = 1, lz = 1, ds = 0.01, volume = 1, dvol = 1;
  int prec = ADSCF_FP64;
  std::string err;
  AxisTables ax[3];
  Engine<double> e64;
  Engine<float> e32;
  bool tables64 = false, tables32 = false;
  // shared FP64 / scalar state
  FHDesc fh{};
  bool fh_set = false;
  DevBuf<double> wring, dring, phi, dens_b, dist_b, io, io2;
  DevBuf<double> qpart, rpart, spart, e2part, gpart, gsum, epart, resid;
  DevBuf<int> e1, e2, d1, d2;
  DevBuf<ScfState> st, cst;
  // SCF configuration
  bool scf_ready = false, field_set = false, phi_fresh = false;
  ChainDesc cd{};
  int m = 10;
  double lam_mix = 0.5;
  double rho = 1.0, mu = 0.0;
  hipGraphExec_t graph = nullptr;
  GraphKeep keep;         // retired and timing graphs, destroyed at adscf_finalize
  unsigned char* scratch_dev = nullptr;  // this device's copy of `scratch` (adscf_init)
  bool use_graph = true;  // ADSCF_GRAPH=0 in the environment: stream launches (profiling aid)
  hipGraph_t graph_src = nullptr;
  int graph_prec = -1;
  std::vector<const void*> graph_key;  // buffers baked into the captured iteration
  double last_ms_iter = 0, last_ms_sweep = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // residual of the SCF loop: 0 mean field - field, 1 density - target (conjugate fields);
  // measure 0 max|dfield|, 1 sum|dfield| dvol
  int resid_kind = 0, measure = 0, graph_mode = -1;
  int user_kind = 0, user_measure = 0;  // the loop adscf_scf_iterate runs (adscf_scf_set_residual)
  int user_mixing = 3, mixing = 3;       // 1 simple, 2 lambda, 3 Anderson (adscf_scf_set_mixing)
  int mid_axis = -1;  // axis of the single-operator passes of a 3-D sweep (-1: y); the others merge
  // xy-plane passes (plane.hpp): -1 where the grid allows them, 0 never (line passes only), 1 asked
  // for (an error where the grid does not allow them); the engine shape of the plane operators
  int plane = -1;
  bool plane_shape_set = false;  // plane_N1/N2 chosen by the caller or the tuner (else per precision)
  int plane_N1 = 0, plane_N2 = 0;
  int done_loop = -1;  // loop kind ((mixing*2 + resid_kind)*4 + measure) of the call that converged
  int last_loop = -1;  // loop kind of the last call since adscf_scf_set_field (-1: none yet)
  int contour_rule = RULE_TRAPEZOID;  // adscf_set_contour_rule
  DevBuf<double> gsum_w;                 // lambda mixing: Gram sums of the field ring
  DevBuf<double> target;             // target densities (nt, ng): the DDFT density
  DevBuf<double> mob;                // mobility_qq (nk, nt, nt)
  DevBuf<double> dmu, dmuq, ddq, dd; // DDFT step work arrays
  bool target_set = false, mob_set = false;
  // homopolymer blend (adscf_blend_setup): one single-block chain per polymer, canonical;
  // empty for a single copolymer chain (adscf_scf_setup)
  std::vector<ChainDesc> bl_cd;
  std::vector<double> bl_rho;
  DevBuf<ScfState> pst;              // chain scalars (Q, norm, mu, energy) of each polymer
  DevBuf<double> nab, dold, dfa;     // EPD: nabla (imaginary parts), old densities, gradients
  DevBuf<double> lapr, epdw;         // EPD field steps: laplace's real parts; field, stage, 4 derivatives
};

namespace {

void set_err(adscf_ctx* c, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
}

template <typename F>
int guarded(adscf_ctx* c, F&& f) {
  if (!c) return ADSCF_ERR_ARG;
  try {
    return f();
  } catch (const HipError& e) {
    set_err(c, "HIP error %d (%s) in %s [adscf.hip:%d]", (int)e.e, hipGetErrorString(e.e), e.what, e.line);
    return ADSCF_ERR_HIP;
  } catch (const ArgError& e) {
    c->err = e.msg;
    return ADSCF_ERR_ARG;
  } catch (const std::bad_alloc&) {
    set_err(c, "host allocation failed");
    return ADSCF_ERR_HIP;
  }
}

int grid_for(long n) {
  long b = (n + kThreads - 1) / kThreads;
  return (int)std::max<long>(1, std::min<long>(b, kRedBlocks));
}

// contour kernel: 16-byte point packs when ng allows and gives at least 128 blocks of them,
// else one point per thread (more threads in flight on small grids); returns the blocks
// launched (one rpart/spart slot each)
template <typename T>
int launch_contour(hipStream_t stream, const ContourArgs<T>& a) {
  constexpr int VP = 16 / sizeof(T);
  if (a.ng % VP == 0 && a.ng / VP >= (long)kThreads * 128) {
    const int nb = grid_for(a.ng / VP);
    hipLaunchKernelGGL((k_contour<T, VP>), dim3(nb), dim3(kThreads), 0, stream, a);
    return nb;
  }
  const int nb = grid_for(a.ng);
  hipLaunchKernelGGL((k_contour<T, 1>), dim3(nb), dim3(kThreads), 0, stream, a);
  return nb;
}

// Anderson Gram sums of a history ring (m+1 slots of nvar doubles): partials over up to
// kGramBlocks blocks, then one block per pair; gsum receives the (m+1)(m+2)/2 sums.
void launch_gram(hipStream_t stream, const double* ring, long nvar, int m, const ScfState* st, int force,
                 double* part, double* gsum, const int* done) {
  const long ntile = (nvar + kGramTile - 1) / kGramTile;
  const int nblk = (int)std::max<long>(1, std::min<long>(ntile, kGramBlocks));
  const int np = (m + 1) * (m + 2) / 2;
  const size_t gsm = (size_t)gram_lds_doubles(m) * sizeof(double);
  hipLaunchKernelGGL(k_gram, dim3(nblk), dim3(kThreads), gsm, stream, ring, nvar, m, st, force, part, done);
  hipLaunchKernelGGL(k_gram_reduce, dim3(np), dim3(kThreads), 0, stream, (const double*)part, nblk, m, st, force,
                     gsum, done);
}

// Wave number of FFT bin k on an axis of n points and length len, folded as
// Calculate_Laplace folds y and z (mod_fourier.f90:181-186); for x the bins
// above n/2 mirror the Hermitian half, so the same q^2 applies.
double wavenumber(int k, int n, double len) {
  const double pi = 3.141592653589793;  // Global: pi (Main_Diblock_Melt/mod_global.f90:10)
  const int kk = (k + 1 > n / 2 + 1) ? k - n : k;
  return 2.0 * pi * (double)kk / len;
}

// Lines per workgroup of the Stockham kernels: up to 1024 points, on y and z up to 16
// lines adjacent in x (coalesced rows), then halved until the pass has at least 512
// workgroups or one line each (small grids such as Main_Multiblock's 86 x 1 x 75 would
// otherwise run on a handful of CUs).
int choose_tile(int axis, int n, int nx, int ny, int nz) {
  n = std::max(n, 1);
  int L = std::max(1, 1024 / n);
  if (axis == 0) {
    L = std::min(L, ny * nz);
  } else {
    L = std::max(L, 4);
    L = std::min(L, 16);
    L = std::min(L, nx);
  }
  L = std::max(L, 1);
  const long lines = (long)nx * ny * nz / n;
  while (L > 1 && (lines + L - 1) / L < 512) L /= 2;
  return L;
}

template <typename T>
Engine<T>& eng(adscf_ctx* c);
template <>
Engine<double>& eng<double>(adscf_ctx* c) { return c->e64; }
template <>
Engine<float>& eng<float>(adscf_ctx* c) { return c->e32; }

template <typename T>
void upload_tables(adscf_ctx* c) {
  Engine<T>& E = eng<T>(c);
  const int dims[3] = {c->nx, c->ny, c->nz};
  const double lens[3] = {c->lx, c->ly, c->lz};
  for (int a = 0; a < 3; ++a) {
    const int n = dims[a];
    std::vector<C2<T>> tw(n);
    std::vector<T> kf(n);
    for (int t = 0; t < n; ++t) {
      const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
      tw[t] = {(T)cosl(ang), (T)sinl(ang)};
      const double q = wavenumber(t, n, lens[a]);
      kf[t] = (T)(exp(-c->ds * q * q) / (double)n);
    }
    E.tw[a].ensure(n);
    E.kf[a].ensure(n);
    HIPCK(hipMemcpyAsync(E.tw[a].p, tw.data(), sizeof(C2<T>) * n, hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(E.kf[a].p, kf.data(), sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
  }
  HIPCK(hipStreamSynchronize(c->stream));
}

void invalidate_tables(adscf_ctx* c) {
  c->tables64 = c->tables32 = false;
  c->keep.retire(c->graph_src, c->graph);
}

template <typename T>
void ensure_tables(adscf_ctx* c) {
  bool& ok = std::is_same<T, double>::value ? c->tables64 : c->tables32;
  if (!ok) {
    upload_tables<T>(c);
    ok = true;
  }
}

// the transform tables of the context's precision, uploaded at set-up so that the
// no-allocation entry point finds them
void ensure_current_tables(adscf_ctx* c) {
  if (c->prec == ADSCF_FP64) ensure_tables<double>(c);
  else ensure_tables<float>(c);
}

template <typename T>
PassArgs<T> base_pass(adscf_ctx* c, int axis) {
  Engine<T>& E = eng<T>(c);
  PassArgs<T> p{};
  p.nx = c->nx;
  p.ny = c->ny;
  p.nz = c->nz;
  p.axis = axis;
  if (axis == kAxisPlane) {  // the x and y operators of a step over one z-plane (plane.hpp)
    p.plan = c->ax[0].plan;
    p.L = 1;
    p.ldn = c->nx + 1;
    p.gx = 1;
    p.stride = 1;
    p.tw = E.tw[0].p;
    p.kf = E.kf[0].p;
    p.tw2 = E.tw[1].p;
    p.kf2 = E.kf[1].p;
    p.nops = 1;
    p.post_E = 1;
    return p;
  }
  const int n = c->ax[axis].n;
  p.plan = c->ax[axis].plan;
  p.L = c->ax[axis].L;
  p.ldn = n + 1;
  p.gx = (c->nx + p.L - 1) / p.L;
  p.stride = axis == 0 ? 1 : (axis == 1 ? (long)c->nx : (long)c->nx * c->ny);
  p.tw = E.tw[axis].p;
  p.kf = E.kf[axis].p;
  p.nops = 1;
  p.post_E = 1;
  return p;
}

// Register-resident schedules: (n, N1, N2).  Several shapes per length are
// compiled; the autotuner picks one per axis (and the waves per workgroup).
template <typename T>
using RegLauncher = int (*)(adscf_ctx*, PassArgs<T>, int);

struct RegShape {
  int n, N1, N2;
};

// N2 == 0: the R^3 cube engine with R = N1; N2 < 0: the A R^2 engine with A = -N2, R = N1
// (fft_cube.hpp)
const RegShape kRegShapes[] = {
    {2, 1, 2},    {3, 1, 3},   {4, 1, 4},   {5, 1, 5},    {6, 1, 6},    {8, 1, 8},   {16, 1, 16},  {10, 2, 5},
    {12, 4, 3},   {20, 4, 5},  {24, 8, 3},  {24, 4, 6},   {32, 4, 8},   {32, 8, 4},  {32, 2, 16},  {36, 6, 6},
    {40, 8, 5},   {48, 8, 6},  {48, 6, 8},  {48, 16, 3},  {64, 8, 8},   {64, 4, 0},  {64, 4, 16},  {64, 16, 4},
    {80, 16, 5},  {96, 16, 6}, {128, 8, 16}, {128, 16, 8}, {128, 4, -8}, {256, 16, 16}, {256, 4, -16}};

std::string shape_name(int N1, int N2) {
  if (N1 == 0) return "stockham";
  if (N2 == 0) return "cube" + std::to_string(N1);
  if (N2 < 0) return "tri" + std::to_string(-N2) + "x" + std::to_string(N1) + "x" + std::to_string(N1);
  return std::to_string(N1) + "x" + std::to_string(N2);
}

int reg_tpl_rt(int N1, int N2) { return N2 <= 0 ? N1 * N1 : (N1 == 1 ? 1 : std::max(N1, N2)); }

// ---------------------------------------------------------------------------
// LDS line strides.  A workgroup keeps one exchange area per line; where the areas start
// decides which banks the lines of a wave hit together.  The best offset depends on the
// engine's access pattern and on how lines map to lanes (x passes and the row layout of a
// plane pass: a line's threads on consecutive lanes; y/z passes and the column layout: lines
// on consecutive lanes), so the host picks it per pass from a model of the LDS: per
// instruction, the lanes of a wave go through in groups that fill the banks once, and a group
// costs the largest number of its accesses that fall in one bank.  The model is evaluated for
// 64 banks (the tie broken by 32) over strides from the engine's element count up.
// ---------------------------------------------------------------------------

// pat[i][t]: element offset in a line's exchange area of thread t's access in LDS instruction
// i of one line operator (-1: idle), as the engines of fft_reg.hpp and fft_cube.hpp issue them;
// each listed pattern is issued twice per operator (a store and a load of the same slots)
std::vector<std::vector<int>> engine_lds_pattern(int N1, int N2, int* tpl) {
  std::vector<std::vector<int>> pat;
  const int TPL = reg_tpl_rt(N1, N2);
  *tpl = TPL;
  auto add = [&](auto f) {
    std::vector<int> r(TPL);
    for (int t = 0; t < TPL; ++t) r[t] = f(t);
    pat.push_back(r);
  };
  if (N2 > 0) {  // four-step (fft_reg.hpp): rows of N1 + 1
    if (N1 == 1) return pat;
    const int n1p = N1 + 1;
    for (int k2 = 0; k2 < N2; ++k2) add([&](int t) { return t < N1 ? t + n1p * k2 : -1; });  // A, C
    for (int j1 = 0; j1 < N1; ++j1) add([&](int t) { return t < N2 ? j1 + n1p * t : -1; });  // B
  } else if (N2 == 0) {  // cube (fft_cube.hpp): S then S2, rows of R + 1
    const int R = N1, sp = R + 1, S2 = R * R * sp;
    for (int q = 0; q < R; ++q) add([&](int t) { return t * sp + q; });                            // a, ai
    for (int b = 0; b < R; ++b) add([&](int t) { return (t % R + R * b) * sp + t / R; });           // b, bi
    for (int c = 0; c < R; ++c) add([&](int t) { return S2 + ((t / R) * R + c) * sp + t % R; });   // b, bi
    for (int q = 0; q < R; ++q) add([&](int t) { return S2 + t * sp + q; });                        // c
  } else {  // A R^2 (fft_cube.hpp, tri_*): S rows of A + 1, S2 rows of R + 1
    const int A = -N2, R = N1, sa = A + 1, sb = R + 1, TT = R * R, S2 = TT * sa, P = A / R;
    for (int q = 0; q < A; ++q) add([&](int t) { return t * sa + q; });
    for (int i = 0; i < P; ++i) {
      for (int b = 0; b < R; ++b)
        add([&](int t) {
          const int p = t + TT * i;
          return (p % R + R * b) * sa + p / R;
        });
      for (int c = 0; c < R; ++c)
        add([&](int t) {
          const int p = t + TT * i;
          return S2 + ((p / R) * R + c) * sb + p % R;
        });
      for (int q = 0; q < R; ++q) add([&](int t) { return S2 + (t + TT * i) * sb + q; });
    }
  }
  return pat;
}

// Model cost of the exchanges of `lines` lines with their areas `stride` elements apart;
// thread t of line l is block thread l TPL + t (across = false) or l + lines t (across = true).
long lds_model_cost(const std::vector<std::vector<int>>& pat, int tpl, int lines, bool across, int stride,
                    int words, int banks) {
  const int group = std::max(1, banks / words);  // lanes per bank cycle
  const int nthreads = tpl * lines;
  long cost = 0;
  std::vector<int> cnt(banks);
  for (const auto& ins : pat)
    for (int w0 = 0; w0 < nthreads; w0 += 64)
      for (int g0 = w0; g0 < std::min(w0 + 64, nthreads); g0 += group) {
        std::fill(cnt.begin(), cnt.end(), 0);
        int worst = 0;
        for (int i = g0; i < std::min(g0 + group, nthreads); ++i) {
          const int l = across ? i % lines : i / tpl, t = across ? i / lines : i % tpl;
          const int e = ins[t];
          if (e < 0) continue;
          const long addr = ((long)l * stride + e) * words;
          for (int k = 0; k < words; ++k) worst = std::max(worst, ++cnt[(addr + k) % banks]);
        }
        cost += worst;
      }
  return cost;
}

// stride (elements, at least the engine's count) of the exchange areas of a pass, cached
int line_stride(int N1, int N2, int elem_bytes, int lines, bool across, int lds) {
  if (lds <= 0) return 0;
  static std::mutex mu;
  static std::map<std::vector<int>, int> cache;
  const std::vector<int> key = {N1, N2, elem_bytes, lines, across ? 1 : 0, lds};
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int tpl = 1;
  const auto pat = engine_lds_pattern(N1, N2, &tpl);
  const int words = elem_bytes / 4;
  int best = lds;
  long b64 = -1, b32 = -1;
  for (int s = lds; s < lds + 64; ++s) {
    const long c64 = lds_model_cost(pat, tpl, lines, across, s, words, 64);
    const long c32 = lds_model_cost(pat, tpl, lines, across, s, words, 32);
    if (b64 < 0 || c64 < b64 || (c64 == b64 && c32 < b32)) {
      b64 = c64;
      b32 = c32;
      best = s;
    }
  }
  cache[key] = best;
  return best;
}

// LDS elements of one line's exchange area in the register engine (N1, N2)
int reg_lds_elems(int N1, int N2) {
  return N2 == 0  ? 2 * N1 * N1 * (N1 + 1)
         : N2 < 0 ? N1 * N1 * (1 - N2) + (-N2) * N1 * (N1 + 1)
                  : (N1 == 1 ? 0 : N2 * (N1 + 1));
}

// Tile geometry of a register-resident pass: `waves` wavefronts per block, L lines
// per block; fills p.L and p.gx and returns the number of blocks.
template <typename T>
int prep_reg(const adscf_ctx* c, PassArgs<T>& p, int N1, int N2, int waves, int* threads, size_t* smem) {
  const int TPL = reg_tpl_rt(N1, N2);
  int L = std::max(1, 64 * waves / TPL);
  if (p.axis != 0) L = std::min(L, c->nx);
  p.L = L;
  p.gx = (c->nx + L - 1) / L;
  int items;
  if (p.axis == 0) items = (int)(((long)c->ny * c->nz + L - 1) / L);
  else if (p.axis == 1) items = c->nz * p.gx;
  else items = c->ny * p.gx;
  if (threads) *threads = TPL * L;
  p.lds_line = line_stride(N1, N2, (int)sizeof(C2<T>), L, p.axis != 0, reg_lds_elems(N1, N2));
  if (smem) *smem = (size_t)L * p.lds_line * sizeof(C2<T>);
  return items;
}

// points of a line of the register shape (N1, N2)
constexpr int reg_shape_n(int N1, int N2) { return N2 == 0 ? N1 * N1 * N1 : (N2 < 0 ? -N2 * N1 * N1 : N1 * N2); }

// the K-tile variants (k_line_pass_tiles) of the fixed sweep modes; lengths of 32 and more
template <typename T, int N1, int N2, int K>
bool launch_tiles(adscf_ctx* c, const PassArgs<T>& p, int items, int threads, size_t smem) {
  // FP64 with four tiles spills registers on the long-line shapes: not compiled
  if constexpr (reg_shape_n(N1, N2) < 32 || (sizeof(T) == 8 && K == 4)) {
    return false;
  } else {
    if (items < kXcds) return false;
    const int per_xcd = (items + kXcds - 1) / kXcds;
    const dim3 grid(kXcds * ((per_xcd + K - 1) / K)), block(threads);
    switch (pass_mode(p)) {
      case PM_FIRST:
        hipLaunchKernelGGL((k_line_pass_tiles<T, N1, N2, PM_FIRST, K>), grid, block, smem, c->stream, p, items);
        break;
      case PM_PLAIN:
        hipLaunchKernelGGL((k_line_pass_tiles<T, N1, N2, PM_PLAIN, K>), grid, block, smem, c->stream, p, items);
        break;
      case PM_MERGED:
        hipLaunchKernelGGL((k_line_pass_tiles<T, N1, N2, PM_MERGED, K>), grid, block, smem, c->stream, p, items);
        break;
      case PM_LAST:
        hipLaunchKernelGGL((k_line_pass_tiles<T, N1, N2, PM_LAST, K>), grid, block, smem, c->stream, p, items);
        break;
      default:
        return false;
    }
    HIPCK(hipGetLastError());
    return true;
  }
}

template <typename T, int N1, int N2>
int launch_reg(adscf_ctx* c, PassArgs<T> p, int waves) {
  int threads;
  size_t smem;
  const int items = prep_reg<T>(c, p, N1, N2, waves, &threads, &smem);
  const int tiles = c->ax[p.axis].tiles;
  if (tiles == 2 && launch_tiles<T, N1, N2, 2>(c, p, items, threads, smem)) return items;
  if (tiles == 4 && launch_tiles<T, N1, N2, 4>(c, p, items, threads, smem)) return items;
  const dim3 grid(items), block(threads);
  switch (pass_mode(p)) {
    case PM_FIRST: hipLaunchKernelGGL((k_line_pass_reg<T, N1, N2, PM_FIRST>), grid, block, smem, c->stream, p); break;
    case PM_PLAIN: hipLaunchKernelGGL((k_line_pass_reg<T, N1, N2, PM_PLAIN>), grid, block, smem, c->stream, p); break;
    case PM_MERGED:
      hipLaunchKernelGGL((k_line_pass_reg<T, N1, N2, PM_MERGED>), grid, block, smem, c->stream, p);
      break;
    case PM_LAST: hipLaunchKernelGGL((k_line_pass_reg<T, N1, N2, PM_LAST>), grid, block, smem, c->stream, p); break;
    default: hipLaunchKernelGGL((k_line_pass_reg<T, N1, N2, PM_GENERIC>), grid, block, smem, c->stream, p);
  }
  HIPCK(hipGetLastError());
  return items;
}

template <typename T>
RegLauncher<T> reg_launcher(int n, int N1, int N2) {
#define ADSCF_SHAPE(a, b, cc) \
  if (n == a && N1 == b && N2 == cc) return &launch_reg<T, b, cc>;
  ADSCF_SHAPE(2, 1, 2) ADSCF_SHAPE(3, 1, 3) ADSCF_SHAPE(4, 1, 4) ADSCF_SHAPE(5, 1, 5) ADSCF_SHAPE(6, 1, 6)
  ADSCF_SHAPE(8, 1, 8) ADSCF_SHAPE(16, 1, 16) ADSCF_SHAPE(10, 2, 5) ADSCF_SHAPE(12, 4, 3) ADSCF_SHAPE(20, 4, 5)
  ADSCF_SHAPE(24, 8, 3) ADSCF_SHAPE(24, 4, 6) ADSCF_SHAPE(32, 4, 8) ADSCF_SHAPE(32, 8, 4) ADSCF_SHAPE(32, 2, 16)
  ADSCF_SHAPE(36, 6, 6) ADSCF_SHAPE(40, 8, 5) ADSCF_SHAPE(48, 8, 6) ADSCF_SHAPE(48, 6, 8) ADSCF_SHAPE(48, 16, 3)
  ADSCF_SHAPE(64, 8, 8) ADSCF_SHAPE(64, 4, 0) ADSCF_SHAPE(64, 4, 16) ADSCF_SHAPE(64, 16, 4) ADSCF_SHAPE(80, 16, 5)
  ADSCF_SHAPE(96, 16, 6) ADSCF_SHAPE(128, 8, 16) ADSCF_SHAPE(128, 16, 8) ADSCF_SHAPE(128, 4, -8)
  ADSCF_SHAPE(256, 16, 16) ADSCF_SHAPE(256, 4, -16)
#undef ADSCF_SHAPE
  return nullptr;
}

// xy-plane passes (plane.hpp): the engine shapes compiled for the plane operators, n = nx = ny
struct PlaneShape {
  int n, N1, N2;
};
const PlaneShape kPlaneShapes[] = {{32, 4, 8}, {32, 8, 4}, {48, 8, 6}, {48, 6, 8}, {64, 4, 0}, {64, 8, 8}, {128, 4, -8}};

// exchange-area strides of a plane pass (row and column layouts) and its LDS bytes
template <typename T, int N1, int N2>
size_t plane_strides(int* sx, int* sy) {
  using G = PlaneGeom<T, PassEngine<T, N1, N2>>;
  const int lds = reg_lds_elems(N1, N2);
  *sx = line_stride(N1, N2, (int)sizeof(C2<T>), G::NLT, false, lds);
  *sy = line_stride(N1, N2, (int)sizeof(C2<T>), G::NLT, true, lds);
  return G::lds_bytes(std::max(*sx, *sy));
}

template <typename T, int N1, int N2>
int launch_plane_shape(adscf_ctx* c, const PassArgs<T>& p0) {
  using G = PlaneGeom<T, PassEngine<T, N1, N2>>;
  PassArgs<T> p = p0;
  const size_t smem = plane_strides<T, N1, N2>(&p.lds_line, &p.lds_line2);
  const dim3 grid(16 * ((c->nz + 7) / 8)), block(G::threads);
  switch (pass_mode(p)) {
    case PM_FIRST:
      hipLaunchKernelGGL((k_plane_pass<T, N1, N2, PM_FIRST>), grid, block, smem, c->stream, p);
      break;
    case PM_MERGED:
      hipLaunchKernelGGL((k_plane_pass<T, N1, N2, PM_MERGED>), grid, block, smem, c->stream, p);
      break;
    case PM_LAST:
      hipLaunchKernelGGL((k_plane_pass<T, N1, N2, PM_LAST>), grid, block, smem, c->stream, p);
      break;
    default:
      throw ArgError{"plane pass of a kind the sweep does not build"};
  }
  HIPCK(hipGetLastError());
  return (int)grid.x;
}

template <typename T>
using PlaneLauncher = int (*)(adscf_ctx*, const PassArgs<T>&);

// The compiled plane shapes, (n, N1, N2); 128-point planes in FP32 only (the 8 x 4 x 4
// engine, two lines per thread): an FP64 plane of 128 x 128 and its exchange areas exceed the
// workgroup's LDS.
#define ADSCF_PLANE_SHAPES(X)                                                          \
  X(32, 4, 8) X(32, 8, 4) X(48, 8, 6) X(48, 6, 8) X(64, 4, 0) X(64, 8, 8)              \
  if constexpr (sizeof(T) == 4) {                                                      \
    X(128, 4, -8)                                                                      \
  }

// launcher and LDS bytes of a plane shape, or null / 0 if it is not compiled
template <typename T>
PlaneLauncher<T> plane_launcher(int n, int N1, int N2, size_t* lds) {
#define ADSCF_PLANE(a, b, cc)                          \
  if (n == a && N1 == b && N2 == cc) {                 \
    int sx, sy;                                        \
    if (lds) *lds = plane_strides<T, b, cc>(&sx, &sy); \
    return &launch_plane_shape<T, b, cc>;              \
  }
  ADSCF_PLANE_SHAPES(ADSCF_PLANE)
#undef ADSCF_PLANE
  if (lds) *lds = 0;
  return nullptr;
}

// Batched plane passes: exchange strides, LDS bytes and threads of a plane shape
template <typename T>
bool plane_batch_geom(int n, int N1, int N2, int* sx, int* sy, size_t* smem, int* threads) {
#define ADSCF_PLANE(a, b, cc)                                   \
  if (n == a && N1 == b && N2 == cc) {                          \
    *smem = plane_strides<T, b, cc>(sx, sy);                    \
    *threads = PlaneGeom<T, PassEngine<T, b, cc>>::threads;     \
    return true;                                                \
  }
  ADSCF_PLANE_SHAPES(ADSCF_PLANE)
#undef ADSCF_PLANE
  return false;
}

size_t device_lds_max() {
  static size_t v = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0, b = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
      v = (size_t)b;
  });
  return v;
}

// Whether the grid and shape allow plane passes in the precision of T: square xy planes, a
// third axis to alternate with, a compiled plane shape and its LDS within the device's limit.
template <typename T>
bool plane_allowed(const adscf_ctx* c, int N1, int N2) {
  if (c->nx != c->ny || c->nz < 2 || c->nx < 2) return false;
  size_t lds = 0;
  if (!plane_launcher<T>(c->nx, N1, N2, &lds)) return false;
  return lds <= device_lds_max();
}

// Engine shape of the plane operators in the precision of T: the one the caller or the tuner
// set, else the default, which is the tuner's choice at 64^3 x 200 on the MI355X (FP32 the cube
// engine, FP64 8 x 8) and the first compiled shape at other lengths.
template <typename T>
void plane_shape(const adscf_ctx* c, int& N1, int& N2) {
  N1 = N2 = 0;
  if (c->plane_shape_set) {
    N1 = c->plane_N1;
    N2 = c->plane_N2;
    return;
  }
  for (const PlaneShape& s : kPlaneShapes)
    if (s.n == c->nx) {
      N1 = s.N1;
      N2 = s.N2;
      break;
    }
  if (c->nx == 64) {
    N1 = sizeof(T) == 8 ? 8 : 4;
    N2 = sizeof(T) == 8 ? 8 : 0;
  }
}

// whether build_sweep alternates plane passes with z passes
template <typename T>
bool plane_sweep(const adscf_ctx* c) {
  if (c->plane == 0) return false;
  int N1, N2;
  plane_shape<T>(c, N1, N2);
  return plane_allowed<T>(c, N1, N2);
}

// Default schedule for an axis of n points: the first registered shape, one wave
// per block; at n = 64 and 128 the choice the autotuner makes on the MI355X for the
// SCF sweep (n = 64: the cube engine on every axis, FP32 y and z with two waves per
// workgroup; n = 128: 16 x 8 with the exceptions below).
void default_schedule(AxisTables& ax, int axis, int prec) {
  ax.N1 = ax.N2 = 0;
  ax.waves = 1;
  ax.tiles = 1;
  for (const RegShape& s : kRegShapes)
    if (s.n == ax.n) {
      ax.N1 = s.N1;
      ax.N2 = s.N2;
      break;
    }
  if (ax.n == 64) {  // the tuner's choice at 64^3 x 200: the cube engine, FP32 y/z with 2 waves
    ax.N1 = 4;
    ax.N2 = 0;
    if (prec == ADSCF_FP32 && axis != 0) ax.waves = 2;
  }
  if (ax.n == 128) {  // the tuner's choice at 128^3 x 400 on the MI355X
    ax.N1 = 16;
    ax.N2 = 8;
    if (prec == ADSCF_FP32) {
      ax.waves = axis == 0 ? 1 : 2;
      ax.tiles = axis == 2 ? 2 : 1;
    } else if (axis == 1) {
      ax.waves = 2;  // 16 x 8, two tiles per workgroup
      ax.tiles = 2;
    } else {
      ax.N1 = 4;  // the 8 x 4 x 4 engine; four waves on z
      ax.N2 = -8;
      if (axis == 2) ax.waves = 4;
    }
  }
}

template <typename T>
int launch_pass_reg(adscf_ctx* c, const PassArgs<T>& p) {
  const AxisTables& ax = c->ax[p.axis];
  if (ax.N1 == 0) return 0;
  RegLauncher<T> f = reg_launcher<T>(ax.n, ax.N1, ax.N2);
  return f ? f(c, p, ax.waves) : 0;
}

// launches one line pass; returns the number of workgroups (OUT_DOT writes one partial each)
template <typename T>
int launch_pass(adscf_ctx* c, const PassArgs<T>& p) {
  if (p.axis == kAxisPlane) {
    int N1, N2;
    plane_shape<T>(c, N1, N2);
    PlaneLauncher<T> f = plane_launcher<T>(c->nx, N1, N2, nullptr);
    if (!f) throw ArgError{"no plane kernel for this shape"};
    return f(c, p);
  }
  {
    const int items = launch_pass_reg<T>(c, p);
    if (items > 0) return items;
  }
  int items;
  if (p.axis == 0) items = (int)(((long)c->ny * c->nz + p.L - 1) / p.L);
  else if (p.axis == 1) items = c->nz * p.gx;
  else items = c->ny * p.gx;
  const size_t smem = line_pass_lds_bytes<T>(p.L, p.ldn, p.plan.n);
  hipLaunchKernelGGL(k_line_pass<T>, dim3(items), dim3(kThreads), smem, c->stream, p);
  HIPCK(hipGetLastError());
  return items;
}

std::vector<int> active_axes(const adscf_ctx* c) {
  std::vector<int> A;
  const int d[3] = {c->nx, c->ny, c->nz};
  for (int a = 0; a < 3; ++a)
    if (d[a] > 1) A.push_back(a);
  if (A.empty()) A.push_back(0);
  return A;
}

// Block type of contour step s (1..ns) for the forward chain (c == 0) and
// the complementary chain (blocks reversed, mod_multiblock.f90:145).
int step_type(const ChainDesc& cd, int chain, int s) {
  int acc = 0;
  for (int bb = 0; bb < cd.nblocks; ++bb) {
    const int b = chain == 0 ? bb : cd.nblocks - 1 - bb;
    acc += cd.seg[b];
    if (s <= acc) return cd.type[b];
  }
  return cd.type[chain == 0 ? cd.nblocks - 1 : 0];
}

// Build the line passes of a forward + complementary sweep.
template <typename T>
std::vector<PassArgs<T>> build_sweep(adscf_ctx* c, const ChainDesc& cd, const T* E, T* h1, T* h2, const int* dexp1,
                                     const int* dexp2, const int* done, C2<T>* tmp = nullptr) {
  C2<T>* const tmpp = tmp ? tmp : eng<T>(c).tmp.p;  // packed line buffer between passes
  const long ng = c->ng;
  // Steps alternate the axis order (a, b, c then c, b, a), so the operators of the outer axes
  // of consecutive steps meet and share a pass; the middle axis gets one pass per step.  The
  // line operators of different axes commute, so any order gives the same propagator.
  std::vector<int> A = active_axes(c);
  if (A.size() == 3 && plane_sweep<T>(c)) A = {kAxisPlane, 2};  // plane passes alternate with z passes
  else if (A.size() == 3 && c->mid_axis >= 0 && c->mid_axis != A[1]) {
    std::vector<int> o;
    for (int a : A)
      if (a != c->mid_axis) o.push_back(a);
    A = {o[0], c->mid_axis, o[1]};
  }
  const int ns = cd.ns;
  auto Ep = [&](int chain, int s) { return E + (long)step_type(cd, chain, s) * ng; };
  std::vector<PassArgs<T>> passes;
  if (A.size() == 1) {
    for (int s = 1; s <= ns; ++s) {
      PassArgs<T> p = base_pass<T>(c, A[0]);
      p.in_kind = s == 1 ? IN_ONES : IN_PLANAR;
      p.q1_in = h1 + (long)(s - 1) * ng;
      p.q2_in = h2 + (long)(s - 1) * ng;
      p.pre_E = 1;
      p.Ea1 = Ep(0, s);
      p.Ea2 = Ep(1, s);
      p.out_kind = OUT_PLANAR;
      p.h1 = h1 + (long)s * ng;
      p.h2 = h2 + (long)s * ng;
      p.dexp1 = dexp1;
      p.dexp2 = dexp2;
      p.sidx = s;
      p.done = done;
      passes.push_back(p);
    }
    return passes;
  }
  for (int s = 1; s <= ns; ++s) {
    std::vector<int> order = A;
    if ((s - 1) % 2 == 1) std::reverse(order.begin(), order.end());
    for (size_t idx = 0; idx < order.size(); ++idx) {
      const int a = order[idx];
      if (idx == 0 && s > 1) {
        PassArgs<T>& p = passes.back();  // last op of step s-1, same axis
        p.nops = 2;
        p.out_kind = OUT_TMP;
        p.tmp_out = tmpp;
        p.h1 = h1 + (long)(s - 1) * ng;
        p.h2 = h2 + (long)(s - 1) * ng;
        p.Ea1 = Ep(0, s - 1);
        p.Ea2 = Ep(1, s - 1);
        p.Eb1 = Ep(0, s);
        p.Eb2 = Ep(1, s);
        p.dexp1 = dexp1;
        p.dexp2 = dexp2;
        p.sidx = s - 1;
        continue;
      }
      PassArgs<T> p = base_pass<T>(c, a);
      p.done = done;
      if (idx == 0) {  // s == 1
        p.in_kind = IN_ONES;
        p.pre_E = 1;
        p.Ea1 = Ep(0, 1);
        p.Ea2 = Ep(1, 1);
        p.h2 = h2;  // marks the second chain present for IN_ONES
      } else {
        p.in_kind = IN_TMP;
        p.tmp_in = tmpp;
      }
      if (idx + 1 == order.size() && s == ns) {
        p.out_kind = OUT_PLANAR;
        p.h1 = h1 + (long)ns * ng;
        p.h2 = h2 + (long)ns * ng;
        p.Ea1 = Ep(0, ns);
        p.Ea2 = Ep(1, ns);
        p.dexp1 = dexp1;
        p.dexp2 = dexp2;
        p.sidx = ns;
      } else {
        p.out_kind = OUT_TMP;
        p.tmp_out = tmpp;
      }
      passes.push_back(p);
    }
  }
  return passes;
}

ChainDesc make_chain(adscf_ctx* c, int nblocks, const int* monomer_type, const int* block_segments, int ensemble) {
  if (nblocks < 1 || nblocks > kMaxBlocks) throw ArgError{"nblocks must be in [1, 64]"};
  ChainDesc cd{};
  cd.nblocks = nblocks;
  cd.ensemble = ensemble;
  int s = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (monomer_type[b] < 1 || monomer_type[b] > c->nt) throw ArgError{"monomer_type out of range 1..ntypes"};
    if (block_segments[b] < 0) throw ArgError{"negative block_segments"};
    cd.type[b] = monomer_type[b] - 1;
    cd.seg[b] = block_segments[b];
    cd.start[b] = s;
    s += block_segments[b];
  }
  if (s < 1) throw ArgError{"chain has no segments"};
  cd.ns = s;
  return cd;
}

template <typename T>
void ensure_hist(adscf_ctx* c, int ns) {
  Engine<T>& E = eng<T>(c);
  const size_t want = (size_t)(ns + 1) * c->ng;
  if (E.h1.n < want || E.h2.n < want) {
    E.h1.ensure(want);
    E.h2.ensure(want);
  }
  E.tmp.ensure(c->ng);
  E.E.ensure((size_t)c->nt * c->ng);
  c->e1.ensure(ns + 1);
  c->e2.ensure(ns + 1);
  c->d1.ensure(ns + 1);
  c->d2.ensure(ns + 1);
}

void ensure_partials(adscf_ctx* c) {
  c->qpart.ensure(kRedBlocks);
  c->rpart.ensure(kRedBlocks);
  c->spart.ensure(kRedBlocks);
  c->e2part.ensure((size_t)kRedBlocks * kMaxTypes);
  c->epart.ensure((size_t)kRedBlocks * 16);
  c->st.ensure(1);
  c->cst.ensure(1);
}

// sweep + Q + normalisation for the chain described by cd, E already in place
template <typename T>
void run_chain(adscf_ctx* c, const ChainDesc& cd, ScfState* st, const int* done, double rho, double mu,
               bool passes_too = true) {
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  if (passes_too) {
    auto passes = build_sweep<T>(c, cd, En.E.p, En.h1.p, En.h2.p, c->d1.p, c->d2.p, done);
    for (auto& p : passes) launch_pass<T>(c, p);
  }
  const int nb = grid_for(ng);
  hipLaunchKernelGGL(k_qsum<T>, dim3(nb), dim3(kThreads), 0, c->stream, En.h1.p + (long)cd.ns * ng, ng, c->qpart.p,
                     done);
  hipLaunchKernelGGL(k_qfinal, dim3(1), dim3(kThreads), 0, c->stream, c->qpart.p, nb, c->e1.p, cd.ns, c->dvol,
                     c->volume, cd.ensemble, rho, mu, st, done);
  HIPCK(hipGetLastError());
}

template <typename T>
ContourArgs<T> contour_args(adscf_ctx* c, const ChainDesc& cd, const ScfState* st) {
  Engine<T>& En = eng<T>(c);
  ContourArgs<T> a{};
  a.ng = c->ng;
  a.cd = cd;
  a.fh = c->fh;
  a.fh.nt = c->nt;
  a.ds = c->ds;
  a.rule = c->contour_rule;
  a.h1 = En.h1.p;
  a.h2 = En.h2.p;
  a.e1 = c->e1.p;
  a.e2 = c->e2.p;
  a.st = st;
  a.m = c->m;
  a.rpart = c->rpart.p;
  a.spart = c->spart.p;
  return a;
}

template <typename T>
void fill_ones_slot0(adscf_ctx* c) {
  Engine<T>& En = eng<T>(c);
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(c->ng)), dim3(kThreads), 0, c->stream, En.h1.p, T(1), c->ng);
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(c->ng)), dim3(kThreads), 0, c->stream, En.h2.p, T(1), c->ng);
  HIPCK(hipGetLastError());
}

// contour slices the history must hold: the chain, or the longest polymer of a blend
int hist_ns(const adscf_ctx* c) {
  int ns = c->cd.ns;
  for (const ChainDesc& cd : c->bl_cd) ns = std::max(ns, cd.ns);
  return ns;
}

// Calculate_Densities of a homopolymer blend (Main_Blend/main_blend.f90:321-345): per
// polymer the rescale schedule, the propagator sweep, Q and the normalisation
// (Propagate_Homopolymer, mod_homopolymer.f90:90-141, as a one-block chain whose packed
// complementary sweep is the same propagator), then its densities added to phi in polymer
// order.  With `resid` the last polymer's contour also forms the residual of the loop.
// Returns the contour's block count.
template <typename T>
int enqueue_blend_densities(adscf_ctx* c, const int* done, bool resid) {
  const long ng = c->ng;
  const int nb = grid_for(ng);
  int nbc = 0;
  for (size_t i = 0; i < c->bl_cd.size(); ++i) {
    const ChainDesc& cd = c->bl_cd[i];
    hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, cd, (const double*)c->e2part.p, nb, ng,
                       c->nt, c->e1.p, c->e2.p, c->d1.p, c->d2.p, (ScfState*)nullptr, 0, c->m, done);
    run_chain<T>(c, cd, c->pst.p + i, done, c->bl_rho[i], 0.0);
    ContourArgs<T> a = contour_args<T>(c, cd, c->pst.p + i);
    a.mode = CONTOUR_SCF;
    a.rst = c->st.p;
    a.phi = c->phi.p;
    a.w0 = c->wring.p;
    a.done = done;
    a.accumulate = i > 0;
    const bool last = i + 1 == c->bl_cd.size();
    a.no_resid = !(resid && last);
    if (resid && last) {
      a.dring = c->dring.p;
      a.target = c->resid_kind ? c->target.p : nullptr;
      a.measure = c->measure;
    }
    nbc = launch_contour<T>(c->stream, a);
  }
  HIPCK(hipGetLastError());
  return nbc;
}

// ---------------------------------------------------------------------------
// compat: Propagate_Multiblock
// ---------------------------------------------------------------------------
template <typename T>
int propagate_chain_impl(adscf_ctx* c, const ChainDesc& cd, double* rho, double* mu, const double* expfield,
                         double* distribution, double* density, double* qsum, double* energy) {
  ensure_tables<T>(c);
  ensure_hist<T>(c, cd.ns);
  ensure_partials(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const size_t nE = (size_t)c->nt * ng;
  c->io.ensure(nE);
  HIPCK(hipMemcpyAsync(c->io.p, expfield, nE * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_convert<T>, dim3(grid_for((long)nE)), dim3(kThreads), 0, c->stream, c->io.p, En.E.p, (long)nE);
  const int nb = grid_for(ng);
  hipLaunchKernelGGL(k_e2<T>, dim3(nb), dim3(kThreads), 0, c->stream, En.E.p, ng, c->nt, c->e2part.p);
  hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, cd, c->e2part.p, nb, ng, c->nt, c->e1.p, c->e2.p,
                     c->d1.p, c->d2.p, (ScfState*)nullptr, 0, 0, (const int*)nullptr);
  fill_ones_slot0<T>(c);
  ScfState init{};
  HIPCK(hipMemcpyAsync(c->cst.p, &init, sizeof init, hipMemcpyHostToDevice, c->stream));
  run_chain<T>(c, cd, c->cst.p, nullptr, *rho, *mu);
  const size_t nbl = (size_t)cd.nblocks * ng;
  c->dist_b.ensure(nbl);
  c->dens_b.ensure(nbl);
  ContourArgs<T> a = contour_args<T>(c, cd, c->cst.p);
  a.mode = CONTOUR_BLOCKS;
  a.dist = c->dist_b.p;
  a.dens = c->dens_b.p;
  a.rpart = nullptr;
  const int nbc = launch_contour<T>(c->stream, a);
  HIPCK(hipGetLastError());
  ScfState st{};
  std::vector<double> sp(nbc);
  HIPCK(hipMemcpyAsync(&st, c->cst.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipMemcpyAsync(sp.data(), c->spart.p, nbc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (distribution)
    HIPCK(hipMemcpyAsync(distribution, c->dist_b.p, nbl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (density) HIPCK(hipMemcpyAsync(density, c->dens_b.p, nbl * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  *qsum = st.qsum;
  if (cd.ensemble == 1) {
    *mu = st.mu;
    *energy = st.energy;
  } else {
    double s = 0.0;
    for (int b = 0; b < nbc; ++b) s += sp[b];
    *rho = s * c->dvol / c->volume / (c->ds * cd.ns);
    *energy = -*rho * c->volume;
  }
  return ADSCF_OK;
}

// ---------------------------------------------------------------------------
// compat: Propagate_Step on one field
// ---------------------------------------------------------------------------
template <typename T>
int propagate_step_impl(adscf_ctx* c, double* prop, const double* Et) {
  ensure_tables<T>(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  En.cq.ensure(2 * ng);
  En.tmp.ensure(ng);
  c->io.ensure(2 * ng);
  HIPCK(hipMemcpyAsync(c->io.p, prop, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCK(hipMemcpyAsync(c->io.p + ng, Et, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_convert<T>, dim3(grid_for(2 * ng)), dim3(kThreads), 0, c->stream, c->io.p, En.cq.p, 2 * ng);
  T* q = En.cq.p;
  const T* E = En.cq.p + ng;
  const std::vector<int> A = active_axes(c);
  for (size_t i = 0; i < A.size(); ++i) {
    PassArgs<T> p = base_pass<T>(c, A[i]);
    p.Ea1 = p.Ea2 = E;
    if (i == 0) {
      p.in_kind = IN_PLANAR;
      p.q1_in = q;
      p.pre_E = 1;
    } else {
      p.in_kind = IN_TMP;
      p.tmp_in = En.tmp.p;
    }
    if (i + 1 == A.size()) {
      p.out_kind = OUT_PLANAR;
      p.h1 = q;
    } else {
      p.out_kind = OUT_TMP;
      p.tmp_out = En.tmp.p;
    }
    launch_pass<T>(c, p);
  }
  hipLaunchKernelGGL(k_widen<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, q, c->io.p, ng);
  HIPCK(hipMemcpyAsync(prop, c->io.p, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return ADSCF_OK;
}

// ---------------------------------------------------------------------------
// compat: Real2Fourier / Fourier2Real
// ---------------------------------------------------------------------------
template <typename T>
void launch_fft(adscf_ctx* c, FftArgs<T> a) {
  int items;
  if (a.axis == 0) items = (int)(((long)a.d1 * a.d2 + a.L - 1) / a.L);
  else if (a.axis == 1) items = a.d2 * a.gx;
  else items = a.d1 * a.gx;
  const size_t smem = stockham_lds_bytes<T>(a.L, a.ldn, a.n);
  hipLaunchKernelGGL(k_fft_axis<T>, dim3(items), dim3(kThreads), smem, c->stream, a);
  HIPCK(hipGetLastError());
}

template <typename T>
FftArgs<T> fft_args(adscf_ctx* c, int axis, int d0, int dir) {
  Engine<T>& En = eng<T>(c);
  FftArgs<T> a{};
  a.d0 = d0;
  a.d1 = c->ny;
  a.d2 = c->nz;
  a.axis = axis;
  a.n = c->ax[axis].n;
  a.plan = c->ax[axis].plan;
  a.L = axis == 0 ? std::max(1, std::min(1024 / a.n, c->ny * c->nz)) : std::max(1, std::min(16, d0));
  {  // at least 512 workgroups where the lines allow it (choose_tile)
    const long lines = axis == 0 ? (long)c->ny * c->nz : (long)d0 * (axis == 1 ? c->nz : c->ny);
    while (a.L > 1 && (lines + a.L - 1) / a.L < 512) a.L /= 2;
  }
  a.ldn = a.n + 1;
  a.gx = (d0 + a.L - 1) / a.L;
  a.stride = axis == 1 ? (long)d0 : (long)d0 * c->ny;
  a.tw = En.tw[axis].p;
  a.dir = dir;
  a.scale = T(1);
  return a;
}

// Real2Fourier on device arrays: din (nx,ny,nz) doubles -> dout (nx/2+1,ny,nz) complex
// as interleaved doubles, divided by nx*ny*nz (mod_fourier.f90:94-105).  din and dout may
// be the same buffer.
template <typename T>
void enqueue_r2c(adscf_ctx* c, const double* din, double* dout) {
  ensure_tables<T>(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const int nh = c->nx / 2 + 1;
  const long nk = (long)nh * c->ny * c->nz;
  En.cq.ensure(2 * ng);
  En.ch.ensure(nk);
  hipLaunchKernelGGL(k_convert<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, din, En.cq.p, ng);
  FftArgs<T> a = fft_args<T>(c, 0, nh, -1);
  a.load = LOAD_REAL;
  a.rin = En.cq.p;
  a.store = STORE_HALF;
  a.cout = En.ch.p;
  launch_fft<T>(c, a);
  const T inv = (T)(1.0 / (double)(c->nx * c->ny * c->nz));  // mod_fourier.f90:102
  const bool doy = c->ny > 1, doz = c->nz > 1;
  if (doy) {
    FftArgs<T> b = fft_args<T>(c, 1, nh, -1);
    b.load = LOAD_CPLX;
    b.cin = En.ch.p;
    b.store = STORE_CPLX;
    b.cout = En.ch.p;
    if (!doz) b.scale = inv;
    launch_fft<T>(c, b);
  }
  if (doz || !doy) {
    // the z pass applies the scale; with ny = nz = 1 it is a trivial pass of length 1
    FftArgs<T> b = fft_args<T>(c, 2, nh, -1);
    b.load = LOAD_CPLX;
    b.cin = En.ch.p;
    b.store = STORE_CPLX;
    b.cout = En.ch.p;
    b.scale = inv;
    launch_fft<T>(c, b);
  }
  hipLaunchKernelGGL(k_widen<T>, dim3(grid_for(2 * nk)), dim3(kThreads), 0, c->stream, (const T*)En.ch.p, dout,
                     2 * nk);
  HIPCK(hipGetLastError());
}

// Fourier2Real on device arrays (unnormalised c2r, mod_fourier.f90:75-86); din and dout
// may be the same buffer.
template <typename T>
void enqueue_c2r(adscf_ctx* c, const double* din, double* dout) {
  ensure_tables<T>(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const int nh = c->nx / 2 + 1;
  const long nk = (long)nh * c->ny * c->nz;
  En.cq.ensure(2 * ng);
  En.ch.ensure(nk);
  hipLaunchKernelGGL(k_convert<T>, dim3(grid_for(2 * nk)), dim3(kThreads), 0, c->stream, din, (T*)En.ch.p, 2 * nk);
  if (c->nz > 1) {
    FftArgs<T> b = fft_args<T>(c, 2, nh, +1);
    b.load = LOAD_CPLX;
    b.cin = En.ch.p;
    b.store = STORE_CPLX;
    b.cout = En.ch.p;
    launch_fft<T>(c, b);
  }
  if (c->ny > 1) {
    FftArgs<T> b = fft_args<T>(c, 1, nh, +1);
    b.load = LOAD_CPLX;
    b.cin = En.ch.p;
    b.store = STORE_CPLX;
    b.cout = En.ch.p;
    launch_fft<T>(c, b);
  }
  FftArgs<T> a = fft_args<T>(c, 0, nh, +1);
  a.load = LOAD_HERM;
  a.cin = En.ch.p;
  a.store = STORE_REAL;
  a.rout = En.cq.p;
  launch_fft<T>(c, a);
  hipLaunchKernelGGL(k_widen<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, (const T*)En.cq.p, dout, ng);
  HIPCK(hipGetLastError());
}

template <typename T>
int r2c_impl(adscf_ctx* c, const double* prop, double* out) {
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  c->io.ensure(std::max(c->ng, 2 * nk));
  HIPCK(hipMemcpyAsync(c->io.p, prop, c->ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
  enqueue_r2c<T>(c, c->io.p, c->io.p);
  HIPCK(hipMemcpyAsync(out, c->io.p, 2 * nk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return ADSCF_OK;
}

template <typename T>
int c2r_impl(adscf_ctx* c, const double* in, double* prop) {
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  c->io.ensure(std::max(c->ng, 2 * nk));
  HIPCK(hipMemcpyAsync(c->io.p, in, 2 * nk * sizeof(double), hipMemcpyHostToDevice, c->stream));
  enqueue_c2r<T>(c, c->io.p, c->io.p);
  HIPCK(hipMemcpyAsync(prop, c->io.p, c->ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return ADSCF_OK;
}

// ---------------------------------------------------------------------------
// device-resident SCF iteration
// ---------------------------------------------------------------------------
// The mixing half of an iteration: Gram sums, convergence test and solve, the mix with E
// of the new field, the rescale schedule and the ring rotation.  resume: the loop of an
// earlier call stopped at convergence; the test uses that evaluation's accuracy against the
// current tolerance (no new evaluation), so a continued loop is the one that never stopped.
template <typename T>
void enqueue_mix_tail(adscf_ctx* c, int nbc, bool resume) {
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const long nvar = ng * c->nt;
  const int* done = &c->st.p->done;
  const int nb = grid_for(ng);
  const int m = c->m;
  if (c->mixing == 2) {
    // Mixing_Lambda needs <D0-D1, D0-D1> and <W0-W1, W0-W1> from iteration 2 on
    launch_gram(c->stream, c->dring.p, nvar, m, (const ScfState*)c->st.p, 1, c->gpart.p, c->gsum.p, done);
    launch_gram(c->stream, c->wring.p, nvar, m, (const ScfState*)c->st.p, 1, c->gpart.p, c->gsum_w.p, done);
  } else if (c->mixing == 3) {
    launch_gram(c->stream, c->dring.p, nvar, m, (const ScfState*)c->st.p, 0, c->gpart.p, c->gsum.p, done);
  }
  SolveArgs s{};
  s.m = m;
  s.lam_mix = c->lam_mix;
  s.nblk_contour = nbc;
  s.rpart = c->rpart.p;
  s.spart = c->spart.p;
  s.gsum = c->gsum.p;
  s.resid_hist = c->resid.p;
  s.resid_cap = (int)c->resid.n;
  s.dvol = c->dvol;
  s.volume = c->volume;
  s.ds = c->ds;
  s.ns = c->cd.ns;
  s.ensemble = c->cd.ensemble;
  s.measure = c->measure;
  s.mixing = c->mixing;
  s.gsum_w = c->gsum_w.p;
  s.st = c->st.p;
  s.resume = resume ? 1 : 0;
  hipLaunchKernelGGL(k_solve, dim3(1), dim3(kThreads), 0, c->stream, s);
  hipLaunchKernelGGL(k_mix<T>, dim3(nb, c->nt), dim3(kThreads), 0, c->stream, c->wring.p, (const double*)c->dring.p, ng,
                     c->nt, m, (const ScfState*)c->st.p, c->ds, En.E.p, c->e2part.p);
  hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, c->cd, (const double*)c->e2part.p, nb, ng, c->nt,
                     c->e1.p, c->e2.p, c->d1.p, c->d2.p, c->st.p, 1, m, done);
  HIPCK(hipGetLastError());
}


template <typename T>
void enqueue_iteration(adscf_ctx* c, bool passes_too = true) {
  const int* done = &c->st.p->done;
  int nbc;
  if (!c->bl_cd.empty()) {
    nbc = enqueue_blend_densities<T>(c, done, true);
  } else {
    run_chain<T>(c, c->cd, c->st.p, done, c->rho, c->mu, passes_too);
    ContourArgs<T> a = contour_args<T>(c, c->cd, c->st.p);
    a.mode = CONTOUR_SCF;
    a.phi = c->phi.p;
    a.dring = c->dring.p;
    a.done = done;
    a.w0 = c->wring.p;  // ring base; the kernel reads slot `head` from the device state
    a.target = c->resid_kind ? c->target.p : nullptr;
    a.measure = c->measure;
    nbc = launch_contour<T>(c->stream, a);
    HIPCK(hipGetLastError());
  }
  enqueue_mix_tail<T>(c, nbc, false);
}

template <typename T>
void enqueue_evaluate(adscf_ctx* c) {
  if (!c->bl_cd.empty()) {
    enqueue_blend_densities<T>(c, nullptr, false);
    return;
  }
  run_chain<T>(c, c->cd, c->st.p, nullptr, c->rho, c->mu);
  ContourArgs<T> a = contour_args<T>(c, c->cd, c->st.p);
  a.mode = CONTOUR_SCF;
  a.w0 = c->wring.p;
  a.phi = c->phi.p;
  a.dring = nullptr;
  a.done = nullptr;
  launch_contour<T>(c->stream, a);
  HIPCK(hipGetLastError());
}

template <typename T>
void scf_set_field_impl(adscf_ctx* c, const double* field, hipMemcpyKind kind = hipMemcpyHostToDevice) {
  ensure_tables<T>(c);
  ensure_hist<T>(c, hist_ns(c));
  ensure_partials(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng, nvar = ng * c->nt;
  const int m = c->m;
  c->wring.ensure((size_t)(m + 1) * nvar);
  c->dring.ensure((size_t)(m + 1) * nvar);
  c->phi.ensure(nvar);
  c->gpart.ensure((size_t)kGramBlocks * (m + 1) * (m + 2) / 2);
  c->gsum.ensure((size_t)(m + 1) * (m + 2) / 2);
  c->gsum_w.ensure((size_t)(m + 1) * (m + 2) / 2);
  c->resid.ensure(200000);
  HIPCK(hipMemsetAsync(c->wring.p, 0, sizeof(double) * (m + 1) * nvar, c->stream));
  HIPCK(hipMemsetAsync(c->dring.p, 0, sizeof(double) * (m + 1) * nvar, c->stream));
  HIPCK(hipMemcpyAsync(c->wring.p, field, nvar * sizeof(double), kind, c->stream));
  ScfState init{};
  init.head = 0;
  HIPCK(hipMemcpyAsync(c->st.p, &init, sizeof init, hipMemcpyHostToDevice, c->stream));
  fill_ones_slot0<T>(c);
  const int nb = grid_for(ng);
  hipLaunchKernelGGL(k_expfield<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)c->wring.p, En.E.p, ng,
                     c->nt, c->ds, c->e2part.p, (const int*)nullptr);
  hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, c->cd, (const double*)c->e2part.p, nb, ng, c->nt,
                     c->e1.p, c->e2.p, c->d1.p, c->d2.p, (ScfState*)nullptr, 0, m, (const int*)nullptr);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
}

// The loop an iteration runs: mixing type (1..3), residual kind (0..1) and accuracy
// measure (0..2), one distinct code each.  It keys the captured iteration graph (the
// three are baked into its kernel arguments) and the continuation of a converged loop.
int loop_kind(const adscf_ctx* c) { return (c->mixing * 2 + c->resid_kind) * 4 + c->measure; }

template <typename T>
std::vector<const void*> graph_key(adscf_ctx* c) {
  Engine<T>& En = eng<T>(c);
  return {En.E.p, En.tmp.p, En.h1.p, En.h2.p, En.tw[0].p, En.tw[1].p, En.tw[2].p, En.kf[0].p, En.kf[1].p,
          En.kf[2].p, c->wring.p, c->dring.p, c->phi.p, c->qpart.p, c->rpart.p, c->spart.p, c->e2part.p,
          c->gpart.p, c->gsum.p, c->resid.p, c->e1.p, c->e2.p, c->d1.p, c->d2.p, c->st.p, c->target.p,
          c->pst.p, c->gsum_w.p};
}

// E and the rescale schedule of the current field in precision T (the history is kept)
template <typename T>
void refresh_E(adscf_ctx* c) {
  ensure_tables<T>(c);
  ensure_hist<T>(c, hist_ns(c));
  ensure_partials(c);
  fill_ones_slot0<T>(c);
  ScfState st{};
  HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const int nb = grid_for(ng);
  hipLaunchKernelGGL(k_expfield<T>, dim3(nb), dim3(kThreads), 0, c->stream,
                     (const double*)(c->wring.p + (size_t)st.head * ng * c->nt), En.E.p, ng, c->nt, c->ds,
                     c->e2part.p, (const int*)nullptr);
  hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, c->cd, (const double*)c->e2part.p, nb, ng, c->nt,
                     c->e1.p, c->e2.p, c->d1.p, c->d2.p, (ScfState*)nullptr, 0, c->m, (const int*)nullptr);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
}

template <typename T>
void build_graph(adscf_ctx* c) {
  ensure_tables<T>(c);
  std::vector<const void*> key = graph_key<T>(c);
  const int mode = loop_kind(c);
  if (c->graph && c->graph_prec == c->prec && c->graph_mode == mode && key == c->graph_key) return;
  c->keep.retire(c->graph_src, c->graph);
  HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_iteration<T>(c);
  } catch (...) {
    hipGraph_t g;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCK(hipStreamEndCapture(c->stream, &c->graph_src));
  HIPCK(hipGraphInstantiate(&c->graph, c->graph_src, nullptr, nullptr, 0));
  c->graph_prec = c->prec;
  c->graph_mode = mode;
  c->graph_key = key;
}

template <typename T>
int scf_iterate_impl(adscf_ctx* c, int niter, double tol, int* iters, double* acc, int* conv) {
  if (c->use_graph) build_graph<T>(c);
  ScfState st{};
  hipLaunchKernelGGL(k_state_update, dim3(1), dim3(64), 0, c->stream, c->st.p, tol, 1, 0);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  const int loop = loop_kind(c);
  if (c->last_loop >= 0 && loop != c->last_loop) {
    // another loop than the last call's: it starts at iteration 1 on the current field, as
    // each saddle-point or conjugate-field loop of the reference starts its count
    hipLaunchKernelGGL(k_scf_restart, dim3(1), dim3(64), 0, c->stream, c->st.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  c->last_loop = loop;
  if (st.done && st.iter > 0 && niter > 0) {
    // an earlier call stopped at convergence: the same loop continues under the current
    // tolerance from the mixing step of its last iteration; another loop starts with a
    // fresh evaluation of the current field
    hipLaunchKernelGGL(k_state_update, dim3(1), dim3(64), 0, c->stream, c->st.p, 0.0, 0, 1);
    HIPCK(hipGetLastError());
    if (loop == c->done_loop) enqueue_mix_tail<T>(c, 0, true);
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (st.done) {
      c->phi_fresh = true;
      if (iters) *iters = st.iter;
      if (acc) *acc = st.accuracy;
      if (conv) *conv = 1;
      return ADSCF_OK;
    }
  }
  int done_iters = 0;
  float ms_total = 0.f;
  while (done_iters < niter) {
    // with a tolerance, the stop flag is read after 1, 1, 2, 4, ... 32 iterations, so a
    // short solve (the conjugate-field solves of a DDFT step take ~10-25) replays few
    // graphs past its exit; without one the whole count is queued at once
    const int chunk = std::min(niter - done_iters, tol > 0 ? std::max(1, std::min(32, done_iters)) : niter);
    HIPCK(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < chunk; ++i) {
      if (c->use_graph) HIPCK(hipGraphLaunch(c->graph, c->stream));
      else enqueue_iteration<T>(c);
    }
    HIPCK(hipEventRecord(c->ev1, c->stream));
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ms_total += ms;
    done_iters += chunk;
    if (st.done) break;
  }
  c->last_ms_iter = done_iters > 0 ? ms_total / done_iters : 0.0;
  c->done_loop = st.done ? loop : -1;
  c->phi_fresh = st.done != 0;
  if (iters) *iters = st.iter;
  if (acc) *acc = st.accuracy;
  if (conv) *conv = st.done;
  return ADSCF_OK;
}

// Get_Conjugate_Fields (main_multiblock.f90:566-640; main_blend.f90:376-428) on the
// device-resident loop: the current field is iterated until the densities it produces
// match the target (resid_kind 1), from iteration count 1.
template <typename T>
int conjugate_impl(adscf_ctx* c, double thresh, int maxit, int measure, int* iters, double* acc) {
  c->resid_kind = 1;
  c->measure = measure;
  c->mixing = 3;  // both drivers solve with mixing_type = 3 (main_multiblock.f90:823, main_blend.f90:390)
  hipLaunchKernelGGL(k_scf_restart, dim3(1), dim3(64), 0, c->stream, c->st.p);
  HIPCK(hipGetLastError());
  int conv = 0;
  scf_iterate_impl<T>(c, maxit, thresh, iters, acc, &conv);
  // the reference's loop index at exit: the converging iteration, or maxit + 1
  if (iters && !conv) *iters = maxit + 1;
  return ADSCF_OK;
}

// main_multiblock.f90:619-634
int gcfstat_of(double acc, double accuracy_scf) {
  if (!(acc < accuracy_scf * 0.5)) {
    if (!(acc < accuracy_scf)) return !(acc < 0.1) ? 3 : 2;
    return 1;
  }
  return 0;
}

// DDFT_Timestep_Euler (main_multiblock.f90:794-880) on device-resident state: the density
// is the target buffer, the field the SCF ring.
template <typename T>
int ddft_step_impl(adscf_ctx* c, double dt, int qdmin, double accuracy_scf, int itermax, int* gcf, int* its) {
  const long ng = c->ng, nvar = ng * c->nt;
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  double acc = 0.0;
  int it = 0;
  conjugate_impl<T>(c, accuracy_scf * 0.5, itermax, 0, &it, &acc);  // :823
  int st = gcfstat_of(acc, accuracy_scf);
  if (its) its[0] = it;
  if (gcf) gcf[0] = st;
  if (st == 3) {
    set_err(c, "Get_Conjugate_Fields: accuracy %.6e after %d iterations (the reference stops)", acc, it);
    return ADSCF_ERR_STATE;
  }
  c->dmu.ensure(nvar);
  c->dmuq.ensure(2 * nk * c->nt);
  c->ddq.ensure(2 * nk * c->nt);
  c->dd.ensure(nvar);
  FHDesc fh = c->fh;
  fh.nt = c->nt;
  hipLaunchKernelGGL(k_ddft_mulocal, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, (const double*)c->target.p,
                     (const double*)c->wring.p, (const ScfState*)c->st.p, fh, ng, c->dmu.p);  // :824-825
  for (int t = 0; t < c->nt; ++t) enqueue_r2c<T>(c, c->dmu.p + t * ng, c->dmuq.p + 2 * t * nk);  // :834-839
  hipLaunchKernelGGL(k_ddft_mobility, dim3(grid_for(nk)), dim3(kThreads), 0, c->stream, (const double*)c->dmuq.p,
                     (const double*)c->mob.p, nk, c->nt, c->ddq.p);  // :843-850
  for (int t = 0; t < c->nt; ++t) enqueue_c2r<T>(c, c->ddq.p + 2 * t * nk, c->dd.p + t * ng);  // :854-858
  hipLaunchKernelGGL(k_ddft_euler, dim3(grid_for(nvar)), dim3(kThreads), 0, c->stream, c->target.p,
                     (const double*)c->dd.p, nvar, dt, 1.0e-4, qdmin ? 1 : 0);  // :865-871
  HIPCK(hipGetLastError());
  conjugate_impl<T>(c, accuracy_scf * 0.5, itermax, 0, &it, &acc);  // :873
  st = gcfstat_of(acc, accuracy_scf);
  if (its) its[1] = it;
  if (gcf) gcf[1] = st;
  if (st == 3) {
    set_err(c, "Get_Conjugate_Fields: accuracy %.6e after %d iterations (the reference stops)", acc, it);
    return ADSCF_ERR_STATE;
  }
  if (qdmin && st >= 1) {  // :874-876: the densities of the field replace the density
    enqueue_evaluate<T>(c);
    HIPCK(hipMemcpyAsync(c->target.p, c->phi.p, nvar * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    c->phi_fresh = true;
  }
  HIPCK(hipStreamSynchronize(c->stream));
  return ADSCF_OK;
}

// nabla's imaginary parts (nk, 3) on the device, from the host table of Calculate_Laplace
void ensure_nabla(adscf_ctx* c) {
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  if (c->nab.n >= (size_t)(3 * nk)) return;
  std::vector<double> nab(6 * nk), im(3 * nk);
  const int rc = adscf_laplace(c, nullptr, nullptr, nab.data());
  if (rc != ADSCF_OK) throw ArgError{"nabla table"};
  for (long i = 0; i < 3 * nk; ++i) im[i] = nab[2 * i + 1];
  c->nab.ensure(3 * nk);
  HIPCK(hipMemcpy(c->nab.p, im.data(), 3 * nk * sizeof(double), hipMemcpyHostToDevice));
}

// Chain_Dynamics_Timestep_Euler (Main_Blend/main_blend.f90:605-667) on the device-resident
// density (target buffer) and field (ring): densities and mean fields of the field, the
// gradient of mulocal along each axis, the density response to field - eps*gradient per
// axis and its divergence, the Euler update, then the conjugate fields of the new density
// (sum|dfield| dvol accuracy, main_blend.f90:567).
template <typename T>
int epd_step_impl(adscf_ctx* c, double dt, double thresh, int itermax, int* iters) {
  const double eps = (double)0.0001f;  // eps = 0.0001 (:630): a single-precision literal
  const long ng = c->ng, nvar = ng * c->nt;
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  const int nbv = grid_for(nvar), nbk = grid_for(nk);
  ensure_nabla(c);
  c->dold.ensure(nvar);
  c->dmu.ensure(nvar);
  c->dfa.ensure(3 * nvar);
  c->dd.ensure(3 * nvar);
  c->dmuq.ensure(2 * nk);
  c->ddq.ensure(2 * nk);
  Engine<T>& En = eng<T>(c);
  FHDesc fh = c->fh;
  fh.nt = c->nt;
  enqueue_evaluate<T>(c);  // olddensity (:632); E is that of the ring's field
  HIPCK(hipMemcpyAsync(c->dold.p, c->phi.p, nvar * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_ddft_mulocal, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, (const double*)c->dold.p,
                     (const double*)c->wring.p, (const ScfState*)c->st.p, fh, ng, c->dmu.p);  // :633-634
  for (int t = 0; t < c->nt; ++t) {  // :636-644
    enqueue_r2c<T>(c, c->dmu.p + t * ng, c->dmuq.p);
    for (int ax = 0; ax < 3; ++ax) {
      hipLaunchKernelGGL(k_nabla_mul, dim3(nbk), dim3(kThreads), 0, c->stream, (const double*)c->dmuq.p,
                         (const double*)(c->nab.p + ax * nk), nk, c->ddq.p);
      enqueue_c2r<T>(c, c->ddq.p, c->dfa.p + ax * nvar + t * ng);
    }
  }
  for (int ax = 0; ax < 3; ++ax) {  // :646-657
    hipLaunchKernelGGL(k_field_step, dim3(nbv), dim3(kThreads), 0, c->stream, (const double*)c->wring.p,
                       (const ScfState*)c->st.p, (const double*)(c->dfa.p + ax * nvar), nvar, eps, c->dmu.p);
    hipLaunchKernelGGL(k_expfield<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, (const double*)c->dmu.p,
                       En.E.p, ng, c->nt, c->ds, c->e2part.p, (const int*)nullptr);
    enqueue_evaluate<T>(c);
    double* dda = c->dd.p + ax * nvar;
    hipLaunchKernelGGL(k_sub, dim3(nbv), dim3(kThreads), 0, c->stream, (const double*)c->phi.p,
                       (const double*)c->dold.p, nvar, dda);
    for (int t = 0; t < c->nt; ++t) {
      enqueue_r2c<T>(c, dda + t * ng, c->dmuq.p);
      hipLaunchKernelGGL(k_nabla_mul, dim3(nbk), dim3(kThreads), 0, c->stream, (const double*)c->dmuq.p,
                         (const double*)(c->nab.p + ax * nk), nk, c->ddq.p);
      enqueue_c2r<T>(c, c->ddq.p, dda + t * ng);
    }
  }
  hipLaunchKernelGGL(k_epd_update, dim3(nbv), dim3(kThreads), 0, c->stream, c->target.p, (const double*)c->dd.p, nvar,
                     dt, eps);  // :662
  HIPCK(hipGetLastError());
  refresh_E<T>(c);  // E and the rescale schedule of the ring's field again
  c->phi_fresh = false;
  double acc = 0.0;
  return conjugate_impl<T>(c, thresh, itermax, 1, iters, &acc);  // :663
}

// laplace's real parts (-q^2 per half-spectrum point) on the device, from the host table
// of Calculate_Laplace
void ensure_lap(adscf_ctx* c) {
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  if (c->lapr.n >= (size_t)nk) return;
  std::vector<double> lap(2 * nk), re(nk);
  const int rc = adscf_laplace(c, lap.data(), nullptr, nullptr);
  if (rc != ADSCF_OK) throw ArgError{"laplace table"};
  for (long i = 0; i < nk; ++i) re[i] = lap[2 * i];
  c->lapr.ensure(nk);
  HIPCK(hipMemcpy(c->lapr.p, re.data(), nk * sizeof(double), hipMemcpyHostToDevice));
}

// Calculate_Field_Derivative (Main_Blend/main_blend.f90:709-740) of the device field
// `field` (nt, ng) into `dfield`: E and the rescale schedule of the field, the densities
// (Calculate_Densities, :727), mean fields minus field (:728-729), then per type the
// product with -laplace in Fourier space (:731-737).
template <typename T>
void enqueue_field_derivative(adscf_ctx* c, const double* field, double* dfield) {
  const long ng = c->ng;
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  const int nb = grid_for(ng);
  Engine<T>& En = eng<T>(c);
  hipLaunchKernelGGL(k_expfield<T>, dim3(nb), dim3(kThreads), 0, c->stream, field, En.E.p, ng, c->nt, c->ds,
                     c->e2part.p, (const int*)nullptr);
  if (c->bl_cd.empty())  // a blend's densities compute each polymer's schedule themselves
    hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, c->cd, (const double*)c->e2part.p, nb, ng,
                       c->nt, c->e1.p, c->e2.p, c->d1.p, c->d2.p, (ScfState*)nullptr, 0, c->m, (const int*)nullptr);
  enqueue_evaluate<T>(c);
  FHDesc fh = c->fh;
  fh.nt = c->nt;
  hipLaunchKernelGGL(k_ddft_mulocal, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)c->phi.p, field,
                     (const ScfState*)nullptr, fh, ng, dfield);
  for (int t = 0; t < c->nt; ++t) {
    enqueue_r2c<T>(c, dfield + t * ng, c->dmuq.p);
    hipLaunchKernelGGL(k_neg_laplace_mul, dim3(grid_for(nk)), dim3(kThreads), 0, c->stream, (const double*)c->dmuq.p,
                       (const double*)c->lapr.p, nk, c->ddq.p);
    enqueue_c2r<T>(c, c->ddq.p, dfield + t * ng);
  }
  HIPCK(hipGetLastError());
}

// the device-resident field (ring slot head) copied into `out`
void copy_ring_field(adscf_ctx* c, double* out) {
  ScfState st{};
  HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  const long nvar = c->ng * c->nt;
  HIPCK(hipMemcpyAsync(out, c->wring.p + (size_t)st.head * nvar, nvar * sizeof(double), hipMemcpyDeviceToDevice,
                       c->stream));
}

void ensure_epd_work(adscf_ctx* c) {
  const long nvar = c->ng * c->nt;
  const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
  ensure_lap(c);
  c->dmuq.ensure(2 * nk);
  c->ddq.ensure(2 * nk);
  c->epdw.ensure(6 * nvar);
}

// Calculate_Field_Derivative at the device-resident field, copied to the host
template <typename T>
void field_derivative_impl(adscf_ctx* c, double* dfield) {
  ensure_epd_work(c);
  const long nvar = c->ng * c->nt;
  double* f0 = c->epdw.p;
  double* d = c->epdw.p + 2 * nvar;
  copy_ring_field(c, f0);
  enqueue_field_derivative<T>(c, f0, d);
  HIPCK(hipMemcpyAsync(dfield, d, nvar * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  refresh_E<T>(c);  // E and the rescale schedule of the device-resident field again
  c->phi_fresh = false;
}

// EPD_Timestep_Euler (main_blend.f90:580-596, method 0) or EPD_Timestep_RungeKutta
// (:678-701, method 1) on the device-resident field; the new field then starts the SCF
// state afresh, as adscf_scf_set_field does.
template <typename T>
void epd_field_step_impl(adscf_ctx* c, double dt, int method) {
  ensure_epd_work(c);
  const long nvar = c->ng * c->nt;
  const int nbv = grid_for(nvar);
  double* f0 = c->epdw.p;
  double* stage = c->epdw.p + nvar;
  double* df = c->epdw.p + 2 * nvar;  // df1 .. df4
  copy_ring_field(c, f0);
  enqueue_field_derivative<T>(c, f0, df);  // :590 / :690
  if (method == 0) {
    hipLaunchKernelGGL(k_axpy, dim3(nbv), dim3(kThreads), 0, c->stream, (const double*)f0, dt, (const double*)df, nvar,
                       f0);  // :592
  } else {
    const double h[3] = {0.5 * dt, 0.5 * dt, dt};  // :691, :693, :695
    for (int k = 0; k < 3; ++k) {
      hipLaunchKernelGGL(k_axpy, dim3(nbv), dim3(kThreads), 0, c->stream, (const double*)f0, h[k],
                         (const double*)(df + k * nvar), nvar, stage);
      enqueue_field_derivative<T>(c, stage, df + (k + 1) * nvar);
    }
    hipLaunchKernelGGL(k_rk4_update, dim3(nbv), dim3(kThreads), 0, c->stream, f0, (const double*)df, nvar, dt);  // :698
  }
  HIPCK(hipGetLastError());
  scf_set_field_impl<T>(c, f0, hipMemcpyDeviceToDevice);
}

// the schedule and buffers a captured sweep bakes in
template <typename T>
std::string sweep_key(adscf_ctx* c) {
  int pN1, pN2;
  plane_shape<T>(c, pN1, pN2);
  std::string k = std::to_string(sizeof(T)) + "/" + std::to_string(c->mid_axis) + "/" + std::to_string(c->cd.ns) +
                  "/p" + std::to_string(plane_sweep<T>(c) ? 1 : 0) + "," + std::to_string(pN1) + "," +
                  std::to_string(pN2);
  for (int b = 0; b < c->cd.nblocks; ++b) k += "/" + std::to_string(c->cd.type[b]) + ":" + std::to_string(c->cd.seg[b]);
  for (const AxisTables& a : c->ax)
    k += "/" + std::to_string(a.N1) + "," + std::to_string(a.N2) + "," + std::to_string(a.waves) + "," +
         std::to_string(a.L) + "," + std::to_string(a.tiles);
  char buf[32];
  for (const void* p : graph_key<T>(c)) {
    snprintf(buf, sizeof buf, "/%p", p);
    k += buf;
  }
  return k;
}

template <typename T>
double sweep_timing(adscf_ctx* c, int reps) {
  // time the propagator sweep alone (the line passes of one SCF iteration); the captured
  // sweep of each schedule is kept and replayed by later calls
  const std::string key = sweep_key<T>(c);
  GraphPair* gp = c->keep.find(key);
  if (!gp) {
    Engine<T>& En = eng<T>(c);
    auto passes = build_sweep<T>(c, c->cd, En.E.p, En.h1.p, En.h2.p, c->d1.p, c->d2.p, nullptr);
    GraphPair np;
    HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      for (auto& p : passes) launch_pass<T>(c, p);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &np.g);
      if (np.g) (void)hipGraphDestroy(np.g);
      throw;
    }
    HIPCK(hipStreamEndCapture(c->stream, &np.g));
    const hipError_t e = hipGraphInstantiate(&np.e, np.g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(np.g);
      throw HipError{e, "hipGraphInstantiate(sweep)", __LINE__};
    }
    c->keep.sweeps.push_back({key, np});
    gp = &c->keep.sweeps.back().second;
  }
  HIPCK(hipGraphLaunch(gp->e, c->stream));
  HIPCK(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; ++r) HIPCK(hipGraphLaunch(gp->e, c->stream));
  HIPCK(hipEventRecord(c->ev1, c->stream));
  HIPCK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  return ms / reps;
}


// Autotuner: per active axis, time the captured sweep for every registered
// register-resident shape of that length and 1, 2 or 4 waves per workgroup, and
// the LDS Stockham kernel; keep the fastest.  Axes are tuned one after another.
template <typename T>
std::string autotune_impl(adscf_ctx* c, int reps) {
  std::string rep;
  char line[256];
  const int plane_req = c->plane;
  c->plane = 0;  // the line schedules first; plane passes are tried last
  double best_lines = 1e300;
  for (int a : active_axes(c)) {
    AxisTables& ax = c->ax[a];
    if (ax.n < 2) continue;
    AxisTables best = ax;
    double best_ms = 1e300;
    std::vector<std::pair<int, int>> shapes = {{0, 0}};
    for (const RegShape& s : kRegShapes)
      if (s.n == ax.n) shapes.push_back({s.N1, s.N2});
    // (waves per workgroup, tiles per workgroup) of a register shape; the multi-tile passes
    // exist for lines of 32 points or more, in FP64 with two tiles only
    const std::pair<int, int> geoms[] = {{1, 1}, {2, 1}, {4, 1}, {1, 2}, {1, 4}, {2, 2}};
    for (auto sh : shapes) {
      for (auto wt : geoms) {
        const int waves = wt.first, tiles = wt.second;
        if (sh.first == 0 && (waves > 1 || tiles > 1)) break;
        if (tiles > 1 && (ax.n < 32 || (sizeof(T) == 8 && tiles == 4))) continue;
        ax.N1 = sh.first;
        ax.N2 = sh.second;
        ax.waves = waves;
        ax.tiles = tiles;
        const double ms = sweep_timing<T>(c, reps);
        snprintf(line, sizeof line, "axis %d n=%d shape=%s waves=%d tiles=%d sweep=%.4f ms\n", a, ax.n,
                 shape_name(sh.first, sh.second).c_str(), waves, tiles, ms);
        rep += line;
        if (ms < best_ms) {
          best_ms = ms;
          best = ax;
        }
      }
    }
    ax = best;
    best_lines = best_ms;
    snprintf(line, sizeof line, "axis %d chosen: %s waves=%d tiles=%d (%.4f ms)\n", a,
             shape_name(best.N1, best.N2).c_str(), best.waves, best.tiles, best_ms);
    rep += line;
  }
  if (active_axes(c).size() == 3) {  // which axis takes the single-operator passes
    int best_mid = c->mid_axis;
    double best_ms = 1e300;
    for (int mid = 0; mid < 3; ++mid) {
      c->mid_axis = mid;
      const double ms = sweep_timing<T>(c, reps);
      snprintf(line, sizeof line, "middle axis %d sweep=%.4f ms\n", mid, ms);
      rep += line;
      if (ms < best_ms) {
        best_ms = ms;
        best_mid = mid;
      }
    }
    c->mid_axis = best_mid;
    best_lines = best_ms;
    snprintf(line, sizeof line, "middle axis chosen: %d (%.4f ms)\n", best_mid, best_ms);
    rep += line;
  }
  // plane passes (x and y of a step in one pass over each z-plane) against the line passes
  c->plane = plane_req;
  if (plane_req != 0 && active_axes(c).size() == 3) {
    int bN1, bN2;
    plane_shape<T>(c, bN1, bN2);
    const bool had_shape = c->plane_shape_set;
    double best_plane = 1e300;
    for (const PlaneShape& sh : kPlaneShapes) {
      if (sh.n != c->nx || !plane_allowed<T>(c, sh.N1, sh.N2)) continue;
      c->plane = 1;
      c->plane_shape_set = true;
      c->plane_N1 = sh.N1;
      c->plane_N2 = sh.N2;
      const double ms = sweep_timing<T>(c, reps);
      snprintf(line, sizeof line, "plane shape=%s sweep=%.4f ms\n", shape_name(sh.N1, sh.N2).c_str(), ms);
      rep += line;
      if (ms < best_plane) {
        best_plane = ms;
        bN1 = sh.N1;
        bN2 = sh.N2;
      }
    }
    c->plane_N1 = bN1;
    c->plane_N2 = bN2;
    c->plane_shape_set = had_shape || best_plane < 1e300;
    c->plane = best_plane < best_lines ? 1 : (plane_req == 1 ? 1 : 0);
    if (best_plane < 1e300)
      snprintf(line, sizeof line, "plane passes chosen: %s (%s, %.4f ms; line passes %.4f ms)\n",
               c->plane == 1 ? "yes" : "no", shape_name(bN1, bN2).c_str(), best_plane, best_lines);
    else
      snprintf(line, sizeof line, "plane passes chosen: no (no plane kernel fits this grid and precision)\n");
    rep += line;
  }
  c->keep.retire(c->graph_src, c->graph);
  return rep;
}


// ---------------------------------------------------------------------------
// hipFFT comparator: the reference's step structure (mod_propagate.f90:56-64)
// with vendor r2c/c2r transforms; used only to time against the line passes.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_cmp_mulE(T* __restrict__ q, const T* __restrict__ E, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) q[i] *= E[i];
}

template <typename T>
__global__ void k_cmp_mulK(C2<T>* __restrict__ s, const T* __restrict__ kx, const T* __restrict__ ky,
                           const T* __restrict__ kz, int nh, int ny, int nz) {
  const long n = (long)nh * ny * nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % nh), y = (int)((i / nh) % ny), z = (int)(i / ((long)nh * ny));
    const T f = kx[x] * ky[y] * kz[z];
    s[i].x *= f;
    s[i].y *= f;
  }
}

template <typename T>
double hipfft_step_us(adscf_ctx* c, int reps) {
  ensure_tables<T>(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const int nh = c->nx / 2 + 1;
  const long nk = (long)nh * c->ny * c->nz;
  DevBuf<T> q, E;
  DevBuf<C2<T>> s;
  q.ensure(ng);
  E.ensure(ng);
  s.ensure(nk);
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, q.p, T(1), ng);
  hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, E.p, T(0.999), ng);
  hipfftHandle fwd, inv;
  const bool dbl = std::is_same<T, double>::value;
  if (hipfftPlan3d(&fwd, c->nz, c->ny, c->nx, dbl ? HIPFFT_D2Z : HIPFFT_R2C) != HIPFFT_SUCCESS ||
      hipfftPlan3d(&inv, c->nz, c->ny, c->nx, dbl ? HIPFFT_Z2D : HIPFFT_C2R) != HIPFFT_SUCCESS)
    throw ArgError{"hipfftPlan3d failed"};
  hipfftSetStream(fwd, c->stream);
  hipfftSetStream(inv, c->stream);
  // K/n per axis: the product over axes is exp(-ds k^2)/Ng
  auto step = [&]() {
    hipLaunchKernelGGL(k_cmp_mulE<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, q.p, (const T*)E.p, ng);
    if (dbl) hipfftExecD2Z(fwd, (hipfftDoubleReal*)q.p, (hipfftDoubleComplex*)s.p);
    else hipfftExecR2C(fwd, (hipfftReal*)q.p, (hipfftComplex*)s.p);
    hipLaunchKernelGGL(k_cmp_mulK<T>, dim3(grid_for(nk)), dim3(kThreads), 0, c->stream, s.p,
                       (const T*)En.kf[0].p, (const T*)En.kf[1].p, (const T*)En.kf[2].p, nh, c->ny, c->nz);
    if (dbl) hipfftExecZ2D(inv, (hipfftDoubleComplex*)s.p, (hipfftDoubleReal*)q.p);
    else hipfftExecC2R(inv, (hipfftComplex*)s.p, (hipfftReal*)q.p);
    hipLaunchKernelGGL(k_cmp_mulE<T>, dim3(grid_for(ng)), dim3(kThreads), 0, c->stream, q.p, (const T*)E.p, ng);
  };
  for (int r = 0; r < 5; ++r) step();
  HIPCK(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; ++r) step();
  HIPCK(hipEventRecord(c->ev1, c->stream));
  HIPCK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  hipfftDestroy(fwd);
  hipfftDestroy(inv);
  return 1000.0 * ms / reps;
}

void reset_axes(adscf_ctx* c) {
  const int d[3] = {c->nx, c->ny, c->nz};
  for (int a = 0; a < 3; ++a) {
    c->ax[a].n = d[a];
    c->ax[a].plan = make_line_plan(d[a]);
    if (c->ax[a].plan.npass < 0) throw ArgError{"grid dimension has more than 12 prime factors"};
    c->ax[a].L = choose_tile(a, d[a], c->nx, c->ny, c->nz);
    default_schedule(c->ax[a], a, c->prec);
  }
}

void host_energy_parts(adscf_ctx* c, const double* d_phi, const double* d_w, std::vector<double>& out) {
  const int nb = grid_for(c->ng);
  hipLaunchKernelGGL(k_energy_parts, dim3(nb), dim3(kThreads), 0, c->stream, d_phi, d_w, c->ng, c->fh,
                     c->epart.p, 16);
  HIPCK(hipGetLastError());
  std::vector<double> part((size_t)nb * 16);
  HIPCK(hipMemcpyAsync(part.data(), c->epart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  const int nq = 2 + c->nt * (c->nt - 1) / 2;
  out.assign(nq, 0.0);
  for (int b = 0; b < nb; ++b)
    for (int q = 0; q < nq; ++q) out[q] += part[(size_t)b * 16 + q];
}

double interaction_from_parts(const adscf_ctx* c, const std::vector<double>& p) {
  double e = 0.5 * c->fh.kappa * c->dvol * p[1];
  int q = 2;
  for (int t = 0; t < c->nt; ++t)
    for (int t2 = t + 1; t2 < c->nt; ++t2) e = e + c->fh.chi[t + t2 * c->nt] * c->dvol * p[q++];
  return e;
}


// ---------------------------------------------------------------------------
// adjoint of one density evaluation (canonical ensemble)
// ---------------------------------------------------------------------------
__global__ void k_acc_partials(const double* __restrict__ part, int n, double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += part[i];
  *acc += s;
}

struct PosWeights {
  double w1 = 0.0, w2 = 0.0;
  int b1 = -1, b2 = -1;  // blocks whose contour integral covers point s (two at a junction)
};

// contour weights of point s in the blocks that cover it (mod_multiblock.f90:157-168 for
// the trapezoid; contour_weight for both rules).  Chains reaching the adjoint have no empty
// block, so at most two blocks meet at a point.
PosWeights pos_weights(const ChainDesc& cd, int rule, double ds, int s) {
  PosWeights pw;
  for (int b = 0; b < cd.nblocks; ++b) {
    const int s0 = cd.start[b], s1 = cd.start[b] + cd.seg[b];
    if (s < s0 || s > s1) continue;
    const double w = contour_weight(rule, s - s0, cd.seg[b], ds);
    if (pw.b1 < 0) {
      pw.b1 = b;
      pw.w1 = w;
    } else {
      pw.b2 = b;
      pw.w2 = w;
    }
  }
  return pw;
}

Seed seed_of(const ChainDesc& cd, const PosWeights& pw) {
  Seed sd{};
  sd.w1 = pw.w1;
  sd.t1 = cd.type[pw.b1];
  sd.w2 = pw.w2;
  sd.t2 = pw.b2 >= 0 ? cd.type[pw.b2] : -1;
  return sd;
}

template <typename T>
void upload_dtables(adscf_ctx* c) {
  Engine<T>& E = eng<T>(c);
  const int dims[3] = {c->nx, c->ny, c->nz};
  const double lens[3] = {c->lx, c->ly, c->lz};
  for (int a = 0; a < 3; ++a) {
    const int n = dims[a];
    std::vector<T> kd(n);
    for (int t = 0; t < n; ++t) {
      const double q = wavenumber(t, n, lens[a]);
      kd[t] = (T)(q * q / (double)n);
    }
    E.kd[a].ensure(n);
    HIPCK(hipMemcpyAsync(E.kd[a].p, kd.data(), sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
  }
}

template <typename T>
struct AdjointRun {
  adscf_ctx* c;
  const ChainDesc& cd;
  Engine<T>& En;
  long ng;
  std::vector<int> cks;

  int ck_slot(int s) const {
    for (size_t i = 0; i < cks.size(); ++i)
      if (cks[i] == s) return (int)i;
    return -1;
  }
  int ck_floor(int s) const {
    int best = 0;
    for (int v : cks)
      if (v <= s) best = v;
    return best;
  }
  const T* Eof(int chain, int s) const { return En.E.p + (long)step_type(cd, chain, s) * ng; }

  // one packed step: chain 1 at forward step s1, chain 2 at complementary step s2
  void step(const T* in1, const T* in2, int s1, int s2, T* out1, T* out2, bool post_E) {
    const std::vector<int> A = active_axes(c);
    for (size_t i = 0; i < A.size(); ++i) {
      PassArgs<T> p = base_pass<T>(c, A[i]);
      p.Ea1 = Eof(0, s1);
      p.Ea2 = Eof(1, s2);
      if (i == 0) {
        p.in_kind = IN_PLANAR;
        p.q1_in = in1;
        p.q2_in = in2;
        p.pre_E = 1;
      } else {
        p.in_kind = IN_TMP;
        p.tmp_in = En.tmp.p;
      }
      if (i + 1 == A.size()) {
        p.out_kind = OUT_PLANAR;
        p.h1 = out1;
        p.h2 = out2;
        p.post_E = post_E ? 1 : 0;
      } else {
        p.out_kind = OUT_TMP;
        p.tmp_out = En.tmp.p;
      }
      launch_pass<T>(c, p);
    }
  }

  // states of chain 1 from checkpoint a1 for n1 steps and chain 2 from a2 for n2 steps
  void recompute(int a1, int n1, T* b1, int a2, int n2, T* b2) {
    HIPCK(hipMemcpyAsync(b1, En.ck1.p + (long)ck_slot(a1) * ng, ng * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    HIPCK(hipMemcpyAsync(b2, En.ck2.p + (long)ck_slot(a2) * ng, ng * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    const int nmax = std::max(n1, n2);
    for (int i = 1; i <= nmax; ++i)
      step(b1 + (long)(i - 1) * ng, b2 + (long)(i - 1) * ng, a1 + i, a2 + i, b1 + (long)i * ng, b2 + (long)i * ng,
           true);
  }
};

template <typename T>
int adjoint_impl(adscf_ctx* c, const ChainDesc& cd, double rho, const double* field, int target_type,
                 const double* phit, double lam, int ckev, double* loss, double* Fout, double* grad_w,
                 double* grad_box, double* phi_out) {
  ensure_tables<T>(c);
  ensure_partials(c);
  Engine<T>& En = eng<T>(c);
  const long ng = c->ng;
  const int nt = c->nt, ns = cd.ns;
  for (int b = 0; b < cd.nblocks; ++b)
    if (cd.seg[b] < 1) throw ArgError{"adjoint: every block needs at least one segment"};
  const int ck = ckev > 0 ? ckev : std::max(1, (int)lround(sqrt((double)ns)));
  AdjointRun<T> R{c, cd, En, ng, {}};
  for (int s = 0; s <= ns; s += ck) R.cks.push_back(s);
  if (R.cks.back() != ns) R.cks.push_back(ns);
  const size_t nck = R.cks.size();
  const size_t seg = 2 * (size_t)ck + 2;
  En.E.ensure((size_t)nt * ng);
  En.tmp.ensure(ng);
  En.ck1.ensure(nck * ng);
  En.ck2.ensure(nck * ng);
  En.sa1.ensure(seg * ng);
  En.sa2.ensure(seg * ng);
  En.sb1.ensure(seg * ng);
  En.sb2.ensure(seg * ng);
  En.mu.ensure(2 * ng);
  En.nu.ensure(2 * ng);
  En.y.ensure(2 * ng);
  En.scr.ensure(4 * ng);
  const int nb = grid_for(ng);
  // fields, E, target
  c->io.ensure((size_t)nt * ng);
  HIPCK(hipMemcpyAsync(c->io.p, field, (size_t)nt * ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_expfield<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)c->io.p, En.E.p, ng, nt,
                     c->ds, (double*)nullptr, (const int*)nullptr);
  DevBuf<double> phit_d, dist, phi, g, gE;
  phit_d.ensure(ng);
  HIPCK(hipMemcpyAsync(phit_d.p, phit, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // checkpoint slot 0 holds q(0) = p(0) = 1
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kThreads), 0, c->stream, En.ck1.p, T(1), ng);
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kThreads), 0, c->stream, En.ck2.p, T(1), ng);
  // ---- forward sweep keeping only the checkpoints
  {
    const T* cur1 = En.ck1.p;
    const T* cur2 = En.ck2.p;
    for (int s = 1; s <= ns; ++s) {
      const int k = R.ck_slot(s);
      T* o1 = k >= 0 ? En.ck1.p + (long)k * ng : En.scr.p + (long)(s % 2) * ng;
      T* o2 = k >= 0 ? En.ck2.p + (long)k * ng : En.scr.p + (long)(2 + s % 2) * ng;
      R.step(cur1, cur2, s, s, o1, o2, true);
      cur1 = o1;
      cur2 = o2;
    }
  }
  // ---- contour integral by segments: q on [a,b], p on [ns-b, ns-a]
  dist.ensure((size_t)cd.nblocks * ng);
  hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(kThreads), 0, c->stream, dist.p, 0.0, (long)cd.nblocks * ng);
  for (size_t j = 0; j + 1 < nck; ++j) {
    const int a = R.cks[j], b = R.cks[j + 1];
    const int pb = R.ck_floor(ns - b);
    R.recompute(a, b - a, En.sa1.p, pb, ns - a - pb, En.sa2.p);
    const int last = (b == ns) ? b : b - 1;
    for (int s = a; s <= last; ++s) {
      const PosWeights pw = pos_weights(cd, c->contour_rule, c->ds, s);
      hipLaunchKernelGGL(k_adj_contour<T>, dim3(nb), dim3(kThreads), 0, c->stream,
                         (const T*)(En.sa1.p + (long)(s - a) * ng), (const T*)(En.sa2.p + (long)(ns - s - pb) * ng), ng,
                         pw.w1, pw.b1, pw.w2, pw.b2, dist.p);
    }
  }
  HIPCK(hipGetLastError());
  // ---- S, phi, F, loss and dL/dphi
  hipLaunchKernelGGL(k_qsum<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const T*)(En.ck1.p + (long)(nck - 1) * ng),
                     ng, c->qpart.p, (const int*)nullptr);
  std::vector<double> part(nb);
  HIPCK(hipMemcpyAsync(part.data(), c->qpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  double S = 0.0;
  for (double v : part) S += v;
  const double A = rho * (double)ng / S;  // rho V / Q with Q = dvol S
  phi.ensure((size_t)nt * ng);
  hipLaunchKernelGGL(k_adj_phi, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)dist.p, ng, cd, nt, A,
                     target_type, (const double*)phit_d.p, phi.p, c->qpart.p);
  HIPCK(hipMemcpyAsync(part.data(), c->qpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  double misfit = 0.0;
  for (double v : part) misfit += v;
  std::vector<double> ep;
  host_energy_parts(c, phi.p, c->io.p, ep);
  const double V = c->volume;
  const double mu = log(rho * V / (c->dvol * S));
  const double F = rho * V * (mu - 1.0) - c->dvol * ep[0] + interaction_from_parts(c, ep);
  if (loss) *loss = misfit + lam * F;
  if (Fout) *Fout = F;
  if (phi_out) HIPCK(hipMemcpy(phi_out, phi.p, (size_t)nt * ng * sizeof(double), hipMemcpyDeviceToHost));
  if (!grad_w && !grad_box) return ADSCF_OK;
  g.ensure((size_t)nt * ng);
  hipLaunchKernelGGL(k_adj_g, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)phi.p, (const double*)c->io.p, ng,
                     c->fh, lam, c->dvol, target_type, (const double*)phit_d.p, g.p, c->qpart.p);
  HIPCK(hipMemcpyAsync(part.data(), c->qpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  double gphi = 0.0;
  for (double v : part) gphi += v;
  const double dLdS = -gphi / S - lam * rho * V / S;
  // ---- backward sweep
  upload_dtables<T>(c);
  gE.ensure((size_t)nt * ng);
  hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(kThreads), 0, c->stream, gE.p, 0.0, (long)nt * ng);
  DevBuf<double> spart, sacc;
  spart.ensure((size_t)ng + 64);
  sacc.ensure(3);
  HIPCK(hipMemsetAsync(sacc.p, 0, 3 * sizeof(double), c->stream));
  const Seed sN = seed_of(cd, pos_weights(cd, c->contour_rule, c->ds, ns));
  const Seed s0 = seed_of(cd, pos_weights(cd, c->contour_rule, c->ds, 0));
  hipLaunchKernelGGL(k_adj_init<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)g.p, ng, A, sN, s0, dLdS,
                     En.mu.p, En.nu.p);
  int cur = 0;
  const std::vector<int> axes = active_axes(c);
  for (size_t j = nck - 1; j >= 1; --j) {
    const int lo = R.cks[j - 1], hi = R.cks[j];
    R.recompute(lo, hi - lo, En.sa1.p, lo, hi - lo, En.sa2.p);
    const int f0 = R.ck_floor(ns - hi + 1);
    R.recompute(f0, ns - lo - f0, En.sb1.p, f0, ns - lo - f0, En.sb2.p);
    for (int s = hi; s > lo; --s) {
      T* mu_s = En.mu.p + (long)cur * ng;
      T* nu_s = En.nu.p + (long)cur * ng;
      const T* q_s = En.sa1.p + (long)(s - lo) * ng;
      const T* p_s = En.sa2.p + (long)(s - lo) * ng;
      // y = L(E mu(s)), yp = L(E nu(s))
      R.step(mu_s, nu_s, s, s, En.y.p, En.y.p + ng, false);
      // box stress: <E mu(s), D_a q(s)/E> per axis
      if (grad_box) {
        for (int a : axes) {
          if (c->ax[a].n < 2) continue;
          PassArgs<T> p = base_pass<T>(c, a);
          p.kf = En.kd[a].p;
          p.in_kind = IN_PLANAR;
          p.q1_in = q_s;
          p.q2_in = p_s;
          p.pre_E = -1;
          p.Ea1 = R.Eof(0, s);
          p.Ea2 = R.Eof(1, s);
          p.out_kind = OUT_DOT;
          p.d1 = mu_s;
          p.d2 = nu_s;
          p.dot_out = spart.p;
          const int items = launch_pass<T>(c, p);
          hipLaunchKernelGGL(k_acc_partials, dim3(1), dim3(64), 0, c->stream, (const double*)spart.p, items,
                             sacc.p + a);
        }
      }
      AdjStepArgs<T> st{};
      st.ng = ng;
      st.A = A;
      st.seed_q = seed_of(cd, pos_weights(cd, c->contour_rule, c->ds, s - 1));
      st.seed_p = seed_of(cd, pos_weights(cd, c->contour_rule, c->ds, ns - s + 1));
      st.g = g.p;
      st.Eq = R.Eof(0, s);
      st.Ep = R.Eof(1, s);
      st.y = En.y.p;
      st.yp = En.y.p + ng;
      st.mu = mu_s;
      st.nu = nu_s;
      st.q_s = q_s;
      st.q_m = En.sa1.p + (long)(s - 1 - lo) * ng;
      st.p_s = p_s;
      st.p_m = En.sa2.p + (long)(s - 1 - lo) * ng;
      st.p_f = En.sb2.p + (long)(ns - s + 1 - f0) * ng;
      st.q_f = En.sb1.p + (long)(ns - s + 1 - f0) * ng;
      st.mu_out = En.mu.p + (long)(1 - cur) * ng;
      st.nu_out = En.nu.p + (long)(1 - cur) * ng;
      st.gEq = gE.p + (long)step_type(cd, 0, s) * ng;
      st.gEp = gE.p + (long)step_type(cd, 1, s) * ng;
      hipLaunchKernelGGL(k_adj_step<T>, dim3(nb), dim3(kThreads), 0, c->stream, st);
      HIPCK(hipGetLastError());
      cur = 1 - cur;
    }
  }
  if (grad_w) {
    DevBuf<double> gw;
    gw.ensure((size_t)nt * ng);
    hipLaunchKernelGGL(k_adj_gradw<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)gE.p,
                       (const T*)En.E.p, (const double*)phi.p, (long)nt * ng, c->ds, lam, c->dvol, gw.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(grad_w, gw.p, (size_t)nt * ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
  }
  if (grad_box) {
    double sa[3];
    HIPCK(hipMemcpy(sa, sacc.p, 3 * sizeof(double), hipMemcpyDeviceToHost));
    const double lens[3] = {c->lx, c->ly, c->lz};
    for (int a = 0; a < 3; ++a) grad_box[a] = 2.0 * c->ds / lens[a] * sa[a] + lam * F / lens[a];
  }
  return ADSCF_OK;
}

// ---------------------------------------------------------------------------
// batched instances: one launch per line pass for all instances
// ---------------------------------------------------------------------------
template <typename T>
struct BatchTables {
  // persistent sweep (k_batch_persist): one launch for every pass of every instance
  bool persist = false;
  DevBuf<PassArgs<T>> pargs;
  DevBuf<int> pshape, pneed, pdone;
  DevBuf<PersistItem> pitems;
  int pnitems = 0, npass_total = 0, pgrid = 0;
  size_t psmem = 0;
  // per-pass launches (lanes), the path for shapes the persistent kernel does not carry
  DevBuf<PassArgs<T>> args;
  DevBuf<int> blk;
  struct Launch {
    size_t arg_off, blk_off;
    int blocks, N1, N2;
    size_t smem;
    int lane;
    int plane = 0, mode = 0, threads = 64;  // plane passes: fixed variant and workgroup size
  };
  std::vector<Launch> launches;  // pass-major, one per lane and shape present in the pass
};

// instance slots are packed above kBatchItemBits in a signed 32-bit block code
constexpr int kMaxBatchInstances = 1 << (31 - kBatchItemBits);

bool batch_shape_ok(int N1, int N2) {
  for (const auto& s : kBatchShapes)
    if (s[0] == N1 && s[1] == N2) return true;
  return false;
}

// every axis of a batch instance needs a register shape the batched kernel knows (the batched
// line passes size their tiles by kBatchLineWaves, not by the axis's waves)
void batch_fix_schedule(adscf_ctx* c) {
  for (int a = 0; a < 3; ++a) {
    AxisTables& ax = c->ax[a];
    if (ax.n < 2) continue;
    ax.waves = 1;
    if (ax.N1 && batch_shape_ok(ax.N1, ax.N2)) continue;
    bool found = false;
    for (const RegShape& s : kRegShapes)
      if (s.n == ax.n && batch_shape_ok(s.N1, s.N2)) {
        ax.N1 = s.N1;
        ax.N2 = s.N2;
        found = true;
        break;
      }
    if (!found) throw ArgError{"batched instances need axis lengths 8, 16, 24, 32, 48, 64 or 128"};
  }
}

}  // namespace

// Independent lanes of a batched iteration: the captured graph forks into this many branches
// (the batch stream and kBatchLanes - 1 side streams, default priority), each carrying the
// batched passes and the post-sweep kernels (contour, Gram, solve, mix, schedule) of its own
// instances, and joins at the end of the iteration.  4 matches the hardware queues a process
// gets by default.
constexpr int kBatchLanes = 4;

struct adscf_batch {
  hipStream_t stream = nullptr;
  int plane_req = -1;  // xy-plane passes of the instances (adscf_batch_set_plane_passes)
  std::vector<int> lane_of;  // lane of each instance (batch_build_tables)
  hipStream_t side[kBatchLanes - 1] = {};
  hipEvent_t fork = nullptr, join[kBatchLanes - 1] = {};
  int prec = ADSCF_FP32;
  std::vector<adscf_ctx*> inst;
  BatchTables<double> t64;
  BatchTables<float> t32;
  bool tables_ok = false;
  hipGraph_t graph_src = nullptr;
  hipGraphExec_t graph = nullptr;
  GraphKeep keep;  // retired and timing graphs, destroyed at adscf_batch_destroy
  bool persist_req = false;  // adscf_batch_set_persistent
  bool use_graph = true;     // ADSCF_GRAPH=0 in the environment: stream launches (profiling aid)
  DevBuf<int> perr;  // set by a persistent sweep whose bounded wait expired
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms_iter = 0;
  std::string err;
};

namespace {

template <typename T>
BatchTables<T>& btab(adscf_batch* b);
template <>
BatchTables<double>& btab<double>(adscf_batch* b) { return b->t64; }
template <>
BatchTables<float>& btab<float>(adscf_batch* b) { return b->t32; }

int persist_shape_id(int N1, int N2) {
  for (int k = 0; k < kNumPersistShapes; ++k)
    if (kPersistShapes[k][0] == N1 && kPersistShapes[k][1] == N2) return k;
  return -1;
}

// the persistent sweep kernel of precision T over kPersistShapes
template <typename T>
constexpr auto persist_kernel() {
  return &k_batch_persist<T, 4, 8, 8, 4, 8, 6, 6, 8, 8, 8, 4, 0>;
}

// whether every pass of every instance has a shape the persistent kernel carries and a tile
// that fills the workgroup as the cube engine needs
template <typename T>
bool persist_shapes_ok(adscf_batch* b) {
  for (adscf_ctx* c : b->inst) {
    if (active_axes(c).size() == 3 && plane_sweep<T>(c)) return false;  // line passes only
    for (int a : active_axes(c)) {
      const AxisTables& ax = c->ax[a];
      if (persist_shape_id(ax.N1, ax.N2) < 0) return false;
      PassArgs<T> p = base_pass<T>(c, a);
      int threads = 0;
      prep_reg<T>(c, p, ax.N1, ax.N2, 4, &threads, nullptr);
      // every thread of the workgroup must sit on a line (the engines take t < threads per line)
      if (threads != kPersistThreads) return false;
    }
  }
  return true;
}

// Tables of the persistent sweep: every pass of every instance with a 256-thread tile
// geometry, the items in pass-major order, and the grid the occupancy allows.  Leaves
// tb.persist false (the per-pass launches run instead) unless adscf_batch_set_persistent
// asked for it, every shape is one the kernel carries and every tile fills its workgroup.
template <typename T>
void batch_build_persist(adscf_batch* b) {
  BatchTables<T>& tb = btab<T>(b);
  tb.persist = false;
  if (!b->persist_req || !persist_shapes_ok<T>(b)) return;
  std::vector<PassArgs<T>> args;
  std::vector<int> shape, need, tiles_of, base;
  size_t smem = 0;
  int maxpass = 0;
  for (adscf_ctx* c : b->inst) {
    Engine<T>& En = eng<T>(c);
    auto ps = build_sweep<T>(c, c->cd, En.E.p, En.h1.p, En.h2.p, c->d1.p, c->d2.p, &c->st.p->done);
    base.push_back((int)args.size());
    int prev_tiles = 0;
    for (size_t q = 0; q < ps.size(); ++q) {
      PassArgs<T> p = ps[q];
      if (p.axis == kAxisPlane) return;  // the persistent kernel carries line passes only
      const AxisTables& ax = c->ax[p.axis];
      const int id = persist_shape_id(ax.N1, ax.N2);
      if (id < 0) return;
      int threads = 0;
      size_t sm = 0;
      const int tiles = prep_reg<T>(c, p, ax.N1, ax.N2, 4, &threads, &sm);
      if (threads != kPersistThreads) return;  // every thread on a line, as the engines assume
      smem = std::max(smem, sm);
      args.push_back(p);
      shape.push_back(id);
      need.push_back(q == 0 ? 0 : prev_tiles);
      tiles_of.push_back(tiles);
      prev_tiles = tiles;
    }
    maxpass = std::max(maxpass, (int)ps.size());
  }
  std::vector<PersistItem> items;
  for (int q = 0; q < maxpass; ++q)
    for (size_t i = 0; i < b->inst.size(); ++i) {
      const int npass = (i + 1 < b->inst.size() ? base[i + 1] : (int)args.size()) - base[i];
      if (q >= npass) continue;
      const int f = base[i] + q;
      for (int t = 0; t < tiles_of[f]; ++t) items.push_back(PersistItem{f, t});
    }
  int per_cu = 0, ncu = 0, dev = 0;
  HIPCK(hipGetDevice(&dev));
  HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel<T>(), kPersistThreads, smem));
  if (per_cu < 1 || ncu < 1) return;
  tb.pargs.ensure(args.size());
  tb.pshape.ensure(shape.size());
  tb.pneed.ensure(need.size());
  tb.pdone.ensure(args.size());
  tb.pitems.ensure(items.size());
  b->perr.ensure(1);
  HIPCK(hipMemcpy(tb.pargs.p, args.data(), args.size() * sizeof(PassArgs<T>), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(tb.pshape.p, shape.data(), shape.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(tb.pneed.p, need.data(), need.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(tb.pitems.p, items.data(), items.size() * sizeof(PersistItem), hipMemcpyHostToDevice));
  const int zero = 0;
  HIPCK(hipMemcpy(b->perr.p, &zero, sizeof zero, hipMemcpyHostToDevice));
  tb.pnitems = (int)items.size();
  tb.npass_total = (int)args.size();
  tb.pgrid = std::min(tb.pnitems, per_cu * ncu);
  tb.psmem = smem;
  tb.persist = true;
}

template <typename T>
void launch_batch_persist(adscf_batch* b) {
  BatchTables<T>& tb = btab<T>(b);
  hipLaunchKernelGGL(k_fill<int>, dim3(grid_for(tb.npass_total)), dim3(kThreads), 0, b->stream, tb.pdone.p, 0,
                     (long)tb.npass_total);
  PersistArgs<T> pa{};
  pa.args = tb.pargs.p;
  pa.shape = tb.pshape.p;
  pa.need = tb.pneed.p;
  pa.items = tb.pitems.p;
  pa.nitems = tb.pnitems;
  pa.done = tb.pdone.p;
  pa.err = b->perr.p;
  pa.spin_limit = kPersistSpinLimit;
  hipLaunchKernelGGL(persist_kernel<T>(), dim3(tb.pgrid), dim3(kPersistThreads), tb.psmem, b->stream, pa);
  HIPCK(hipGetLastError());
}

// a persistent sweep that gave up waiting leaves *err set: report it (and clear it)
void batch_check_persist(adscf_batch* b) {
  if (!b->perr.p) return;
  int e = 0;
  HIPCK(hipMemcpy(&e, b->perr.p, sizeof e, hipMemcpyDeviceToHost));
  if (e) {
    const int zero = 0;
    HIPCK(hipMemcpy(b->perr.p, &zero, sizeof zero, hipMemcpyHostToDevice));
    throw ArgError{"the persistent batched sweep gave up waiting for a pass (not every workgroup was resident); "
                   "results of this call are invalid (adscf_batch_set_persistent(batch, 0) selects per-pass "
                   "launches)"};
  }
}

template <typename T>
void batch_build_tables(adscf_batch* b) {
  BatchTables<T>& tb = btab<T>(b);
  struct Entry {
    PassArgs<T> args;
    int items, N1, N2;
    size_t smem;
    int plane, mode, threads;  // plane pass: its fixed variant and workgroup size
    std::array<int, 4> key() const { return {plane, N1, N2, mode}; }
  };
  std::vector<std::vector<Entry>> per;
  size_t npass = 0;
  for (adscf_ctx* c : b->inst) {
    ensure_tables<T>(c);
    Engine<T>& En = eng<T>(c);
    auto ps = build_sweep<T>(c, c->cd, En.E.p, En.h1.p, En.h2.p, c->d1.p, c->d2.p, &c->st.p->done);
    std::vector<Entry> es;
    for (auto& p : ps) {
      if (p.axis == kAxisPlane) {
        Entry e{p, 16 * ((c->nz + 7) / 8), 0, 0, 0, 1, pass_mode(p), 0};
        plane_shape<T>(c, e.N1, e.N2);
        if (!plane_batch_geom<T>(c->nx, e.N1, e.N2, &e.args.lds_line, &e.args.lds_line2, &e.smem, &e.threads))
          throw ArgError{"no plane kernel for this shape"};
        es.push_back(e);
        continue;
      }
      const AxisTables& ax = c->ax[p.axis];
      Entry e{p, 0, ax.N1, ax.N2, 0, 0, PM_GENERIC, 64 * kBatchLineWaves};
      int threads = 0;
      e.items = prep_reg<T>(c, e.args, ax.N1, ax.N2, kBatchLineWaves, &threads, &e.smem);
      if (e.items >= (1 << kBatchItemBits)) throw ArgError{"batched pass has too many blocks per instance"};
      // the batched kernel runs 64 kBatchLineWaves threads, and the engines assume each of
      // them on a line
      if (ax.N1 > 1 && threads != 64 * kBatchLineWaves) throw ArgError{"batched pass shape does not fill its wavefronts"};
      es.push_back(e);
    }
    npass = std::max(npass, es.size());
    per.push_back(std::move(es));
  }
  // Lanes: a lane's sweep and post-sweep kernels depend on nothing in another lane, so the
  // captured iteration runs the lanes as parallel graph branches.  Within a lane, pass p of
  // all its instances with one kernel shape is one launch, and launches of different shapes
  // queue behind each other.  Instances form groups, largest work first, and an instance
  // goes to the least loaded lane that already holds its group and stays within 5% over the
  // even share of the work, else to the least loaded lane; work is points x passes.  With
  // plane passes a group is the set of kernel shapes of the sweep (on cubic grids, one per
  // axis length: 69.0 -> 71.7 batch iterations/s on the 64-instance sweep against the
  // grouping below); with line passes only it is the sequence of shapes, pass by pass (so
  // also the contour length), which measured 10% faster there than grouping by shape.
  bool planes = false;
  for (const auto& es : per)
    for (const Entry& e : es) planes = planes || e.plane;
  std::vector<std::vector<std::array<int, 3>>> sigs;
  std::vector<int> group_of(per.size());
  std::vector<double> gcost, icost(per.size());
  double total = 0.0;
  for (size_t i = 0; i < per.size(); ++i) {
    std::vector<std::array<int, 3>> sg;
    for (const Entry& e : per[i]) {
      const std::array<int, 3> k = {e.plane, e.N1, e.N2};
      if (!planes || std::find(sg.begin(), sg.end(), k) == sg.end()) sg.push_back(k);
    }
    if (planes) std::sort(sg.begin(), sg.end());
    auto it = std::find(sigs.begin(), sigs.end(), sg);
    const int g = (int)(it - sigs.begin());
    if (it == sigs.end()) {
      sigs.push_back(sg);
      gcost.push_back(0.0);
    }
    group_of[i] = g;
    icost[i] = (double)b->inst[i]->ng * (double)per[i].size();
    gcost[g] += icost[i];
    total += icost[i];
  }
  std::vector<int> order(per.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    if (group_of[x] != group_of[y]) {
      if (gcost[group_of[x]] != gcost[group_of[y]]) return gcost[group_of[x]] > gcost[group_of[y]];
      return group_of[x] < group_of[y];
    }
    if (icost[x] != icost[y]) return icost[x] > icost[y];
    return x < y;
  });
  const double share = total / kBatchLanes;
  double load[kBatchLanes] = {};
  std::vector<std::vector<int>> holds(kBatchLanes);  // shape groups in each lane
  b->lane_of.assign(per.size(), 0);
  for (int i : order) {
    const int g = group_of[i];
    int best = -1;
    for (int k = 0; k < kBatchLanes; ++k)
      if (std::find(holds[k].begin(), holds[k].end(), g) != holds[k].end() && load[k] + icost[i] <= 1.05 * share &&
          (best < 0 || load[k] < load[best]))
        best = k;
    if (best < 0) {
      best = 0;
      for (int k = 1; k < kBatchLanes; ++k)
        if (load[k] < load[best]) best = k;
    }
    b->lane_of[i] = best;
    load[best] += icost[i];
    if (std::find(holds[best].begin(), holds[best].end(), g) == holds[best].end()) holds[best].push_back(g);
  }

  std::vector<PassArgs<T>> args;
  std::vector<int> blk;
  tb.launches.clear();
  for (size_t p = 0; p < npass; ++p) {
    for (int lane = 0; lane < kBatchLanes; ++lane) {
      std::vector<std::array<int, 4>> shapes;  // distinct kernels of this pass in the lane
      for (size_t i = 0; i < per.size(); ++i) {
        if (p >= per[i].size() || b->lane_of[i] != lane) continue;
        const std::array<int, 4> sh = per[i][p].key();
        if (std::find(shapes.begin(), shapes.end(), sh) == shapes.end()) shapes.push_back(sh);
      }
      for (const auto& sh : shapes) {
        typename BatchTables<T>::Launch L{args.size(), blk.size(), 0, sh[1], sh[2], 0, lane};
        L.plane = sh[0];
        L.mode = sh[3];
        int slot = 0;
        std::vector<int> codes;
        for (size_t i = 0; i < per.size(); ++i) {
          if (p >= per[i].size() || b->lane_of[i] != lane || per[i][p].key() != sh) continue;
          const Entry& e = per[i][p];
          args.push_back(e.args);
          for (int it = 0; it < e.items; ++it) codes.push_back((slot << kBatchItemBits) | it);
          L.blocks += e.items;
          L.smem = std::max(L.smem, e.smem);
          L.threads = e.threads;
          ++slot;
        }
        // line passes: block b runs on XCD b % 8, and XCD k takes a contiguous range of the
        // tiles (xcd_tile), so neighbouring tiles that share cache lines meet in one L2.
        // Plane passes keep the order their kernel's own mapping assumes.
        for (int bk = 0; bk < (int)codes.size(); ++bk)
          blk.push_back(codes[L.plane ? bk : xcd_tile(bk, (int)codes.size())]);
        tb.launches.push_back(L);
      }
    }
  }
  tb.args.ensure(args.size());
  tb.blk.ensure(blk.size());
  HIPCK(hipMemcpy(tb.args.p, args.data(), args.size() * sizeof(PassArgs<T>), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(tb.blk.p, blk.data(), blk.size() * sizeof(int), hipMemcpyHostToDevice));
  batch_build_persist<T>(b);
}

template <typename T, int N1, int N2>
void launch_plane_batch(hipStream_t s, const typename BatchTables<T>::Launch& L, BatchGroup<T> g) {
  const dim3 grid(L.blocks), block(PlaneGeom<T, PassEngine<T, N1, N2>>::threads);
  switch (L.mode) {
    case PM_FIRST: hipLaunchKernelGGL((k_plane_pass_batch<T, N1, N2, PM_FIRST>), grid, block, L.smem, s, g); break;
    case PM_MERGED: hipLaunchKernelGGL((k_plane_pass_batch<T, N1, N2, PM_MERGED>), grid, block, L.smem, s, g); break;
    case PM_LAST: hipLaunchKernelGGL((k_plane_pass_batch<T, N1, N2, PM_LAST>), grid, block, L.smem, s, g); break;
    default: throw ArgError{"plane pass of a kind the sweep does not build"};
  }
}

template <typename T>
void launch_batch_group(hipStream_t s, const typename BatchTables<T>::Launch& L, BatchGroup<T> g) {
  if (L.plane) {
#define ADSCF_PLANE(a, b, cc)                    \
  if (L.N1 == b && L.N2 == cc) {                 \
    launch_plane_batch<T, b, cc>(s, L, g);       \
    return;                                      \
  }
    ADSCF_PLANE_SHAPES(ADSCF_PLANE)
#undef ADSCF_PLANE
    throw ArgError{"no batched plane kernel for this shape"};
  }
  const dim3 grid(L.blocks), block(64 * kBatchLineWaves);
#define ADSCF_BATCH_SHAPE(a, c)                                                             \
  if (L.N1 == a && L.N2 == c) {                                                             \
    hipLaunchKernelGGL((k_line_pass_batch<T, a, c>), grid, block, L.smem, s, g);             \
    return;                                                                                 \
  }
  ADSCF_BATCH_SHAPE(4, 8) ADSCF_BATCH_SHAPE(8, 4) ADSCF_BATCH_SHAPE(2, 16) ADSCF_BATCH_SHAPE(8, 6)
  ADSCF_BATCH_SHAPE(6, 8) ADSCF_BATCH_SHAPE(16, 3) ADSCF_BATCH_SHAPE(8, 8) ADSCF_BATCH_SHAPE(4, 16)
  ADSCF_BATCH_SHAPE(16, 4) ADSCF_BATCH_SHAPE(8, 16) ADSCF_BATCH_SHAPE(16, 8) ADSCF_BATCH_SHAPE(1, 16)
  ADSCF_BATCH_SHAPE(1, 8) ADSCF_BATCH_SHAPE(4, 0)
#undef ADSCF_BATCH_SHAPE
  throw ArgError{"no batched kernel for this shape"};
}

hipStream_t lane_stream(adscf_batch* b, int lane) { return lane == 0 ? b->stream : b->side[lane - 1]; }

// One batched iteration (post = true) or sweep (post = false) as kBatchLanes branches forked
// from and joined back into the batch stream (for capture into a graph).
template <typename T>
void batch_enqueue_lanes(adscf_batch* b, bool post) {
  BatchTables<T>& tb = btab<T>(b);
  // the persistent sweep runs alone on the batch stream (every workgroup resident), then the
  // lanes fork for the post-sweep kernels
  if (tb.persist) launch_batch_persist<T>(b);
  HIPCK(hipEventRecord(b->fork, b->stream));
  for (hipStream_t s : b->side) HIPCK(hipStreamWaitEvent(s, b->fork, 0));
  for (int lane = 0; lane < kBatchLanes; ++lane) {
    hipStream_t s = lane_stream(b, lane);
    if (!tb.persist)
      for (const auto& L : tb.launches)
        if (L.lane == lane) launch_batch_group<T>(s, L, BatchGroup<T>{tb.args.p + L.arg_off, tb.blk.p + L.blk_off});
    HIPCK(hipGetLastError());
    if (!post) continue;
    for (size_t i = 0; i < b->inst.size(); ++i) {
      if (b->lane_of[i] != lane) continue;
      adscf_ctx* c = b->inst[i];
      c->stream = s;
      try {
        enqueue_iteration<T>(c, false);
      } catch (...) {
        c->stream = b->stream;
        throw;
      }
      c->stream = b->stream;
    }
  }
  for (int k = 0; k < kBatchLanes - 1; ++k) {
    HIPCK(hipEventRecord(b->join[k], b->side[k]));
    HIPCK(hipStreamWaitEvent(b->stream, b->join[k], 0));
  }
}

void batch_drop_graph(adscf_batch* b) { b->keep.retire(b->graph_src, b->graph); }

template <typename T>
void batch_build_graph(adscf_batch* b) {
  if (b->graph) return;
  batch_build_tables<T>(b);
  HIPCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    batch_enqueue_lanes<T>(b, true);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCK(hipStreamEndCapture(b->stream, &b->graph_src));
  HIPCK(hipGraphInstantiate(&b->graph, b->graph_src, nullptr, nullptr, 0));
}

template <typename T>
double batch_sweep_timing(adscf_batch* b, int reps) {
  batch_build_tables<T>(b);
  // Captured afresh on every call: the graph bakes in the batch tables (per-block codes and
  // per-instance pass arguments), which batch_build_tables rewrites in place for another set
  // of schedules, so a sweep graph is valid only until the next rebuild.  It is retired (not
  // destroyed) after timing and released with the batch.
  GraphPair np;
  HIPCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    batch_enqueue_lanes<T>(b, false);
  } catch (...) {
    (void)hipStreamEndCapture(b->stream, &np.g);
    if (np.g) (void)hipGraphDestroy(np.g);
    throw;
  }
  HIPCK(hipStreamEndCapture(b->stream, &np.g));
  const hipError_t e = hipGraphInstantiate(&np.e, np.g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(np.g);
    throw HipError{e, "hipGraphInstantiate(batch sweep)", __LINE__};
  }
  hipGraphExec_t ge = np.e;
  b->keep.retire(np.g, np.e);
  HIPCK(hipGraphLaunch(ge, b->stream));
  HIPCK(hipEventRecord(b->ev0, b->stream));
  for (int r = 0; r < reps; ++r) HIPCK(hipGraphLaunch(ge, b->stream));
  HIPCK(hipEventRecord(b->ev1, b->stream));
  HIPCK(hipEventSynchronize(b->ev1));
  batch_check_persist(b);
  float ms = 0.f;
  HIPCK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  return ms / reps;
}

template <typename F>
int bguard(adscf_batch* b, F&& f) {
  if (!b) return ADSCF_ERR_ARG;
  try {
    return f();
  } catch (const HipError& e) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s [adscf.hip:%d]", (int)e.e, hipGetErrorString(e.e), e.what,
             e.line);
    b->err = buf;
    return ADSCF_ERR_HIP;
  } catch (const ArgError& e) {
    b->err = e.msg;
    return ADSCF_ERR_ARG;
  } catch (const std::bad_alloc&) {
    b->err = "host allocation failed";
    return ADSCF_ERR_HIP;
  }
}

// the first failing instance's message becomes the batch's
int bcheck(adscf_batch* b, adscf_ctx* c, int rc) {
  if (rc != ADSCF_OK) {
    b->err = c->err;
    throw ArgError{c->err};
  }
  return rc;
}

// ---------------------------------------------------------------------------
// no-allocation chain solve: every work array in the device scratch
// ---------------------------------------------------------------------------
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxScratchDevices] = {};

// The address of `scratch` on device dev (current), looked up once per device.
unsigned char* scratch_address(int dev) {
  if (dev < 0 || dev >= kMaxScratchDevices) throw ArgError{"device index out of range for the scratch table"};
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ArgError{"hipGetSymbolAddress returned no address for the scratch array"};
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_dev[dev];
}

// Carves 256-byte aligned pieces off the scratch (or only measures them when base is null).
struct ScratchCarve {
  unsigned char* base;
  size_t off = 0;
  template <typename V>
  V* take(size_t count) {
    const size_t at = off;
    off = (off + count * sizeof(V) + 255) & ~size_t(255);
    return base ? reinterpret_cast<V*>(base + at) : nullptr;
  }
};

template <typename T>
struct ChainScratch {
  ScfState* st;
  double *qpart, *spart, *rpart, *e2part;
  int *e1, *e2, *d1, *d2;
  C2<T>* tmp;
  T* E;  // FP32: E converted from the caller's FP64 array; FP64: unused
  T *h1, *h2;
  size_t bytes;
};

template <typename T>
ChainScratch<T> carve_chain(unsigned char* base, long ng, int nt, int ns) {
  ScratchCarve sc{base};
  ChainScratch<T> w{};
  w.st = sc.take<ScfState>(1);
  w.qpart = sc.take<double>(kRedBlocks);
  w.spart = sc.take<double>(kRedBlocks);
  w.rpart = sc.take<double>(kRedBlocks);
  w.e2part = sc.take<double>((size_t)kRedBlocks * kMaxTypes);
  w.e1 = sc.take<int>(ns + 1);
  w.e2 = sc.take<int>(ns + 1);
  w.d1 = sc.take<int>(ns + 1);
  w.d2 = sc.take<int>(ns + 1);
  w.tmp = sc.take<C2<T>>(ng);
  w.E = std::is_same<T, double>::value ? nullptr : sc.take<T>((size_t)nt * ng);
  w.h1 = sc.take<T>((size_t)(ns + 1) * ng);
  w.h2 = sc.take<T>((size_t)(ns + 1) * ng);
  w.bytes = sc.off;
  return w;
}

// Propagate_Multiblock (mod_multiblock.f90:112-188) of adscf_propagate_chain on device arrays,
// enqueued with the same kernels in the same order, every work array in the scratch.
template <typename T>
void propagate_chain_noalloc_impl(adscf_ctx* c, const ChainDesc& cd, double rho, double mu, const double* expfield,
                                  double* distribution, double* density, double* scalars) {
  const long ng = c->ng;
  const ChainScratch<T> need = carve_chain<T>(nullptr, ng, c->nt, cd.ns);
  if (need.bytes > kScratchBytes) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "the chain needs %zu bytes of scratch (grid %ld points, %d contour steps); the no-allocation entry "
             "point holds %zu",
             need.bytes, ng, cd.ns, kScratchBytes);
    throw ArgError{buf};
  }
  const bool ready = std::is_same<T, double>::value ? c->tables64 : c->tables32;
  if (!ready) throw ArgError{"the transform tables of this precision are not on the device"};
  const ChainScratch<T> w = carve_chain<T>(c->scratch_dev, ng, c->nt, cd.ns);
  const size_t nE = (size_t)c->nt * ng;
  const T* E;
  if constexpr (std::is_same<T, double>::value) {
    E = expfield;
  } else {
    hipLaunchKernelGGL(k_convert<T>, dim3(grid_for((long)nE)), dim3(kThreads), 0, c->stream, expfield, w.E, (long)nE);
    E = w.E;
  }
  const int nb = grid_for(ng);
  hipLaunchKernelGGL(k_e2<T>, dim3(nb), dim3(kThreads), 0, c->stream, E, ng, c->nt, w.e2part);
  hipLaunchKernelGGL(k_schedule, dim3(1), dim3(kThreads), 0, c->stream, cd, (const double*)w.e2part, nb, ng, c->nt,
                     w.e1, w.e2, w.d1, w.d2, (ScfState*)nullptr, 0, 0, (const int*)nullptr);
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kThreads), 0, c->stream, w.h1, T(1), ng);
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kThreads), 0, c->stream, w.h2, T(1), ng);
  HIPCK(hipGetLastError());
  for (auto& p : build_sweep<T>(c, cd, E, w.h1, w.h2, w.d1, w.d2, nullptr, w.tmp)) launch_pass<T>(c, p);
  hipLaunchKernelGGL(k_qsum<T>, dim3(nb), dim3(kThreads), 0, c->stream, (const T*)(w.h1 + (long)cd.ns * ng), ng,
                     w.qpart, (const int*)nullptr);
  hipLaunchKernelGGL(k_qfinal, dim3(1), dim3(kThreads), 0, c->stream, (const double*)w.qpart, nb,
                     (const int*)w.e1, cd.ns, c->dvol, c->volume, cd.ensemble, rho, mu, w.st, (const int*)nullptr);
  ContourArgs<T> a = contour_args<T>(c, cd, w.st);
  a.h1 = w.h1;
  a.h2 = w.h2;
  a.e1 = w.e1;
  a.e2 = w.e2;
  a.spart = w.spart;
  a.rpart = nullptr;
  a.mode = CONTOUR_BLOCKS;
  a.dist = distribution;
  a.dens = density;
  const int nbc = launch_contour<T>(c->stream, a);
  hipLaunchKernelGGL(k_chain_scalars, dim3(1), dim3(64), 0, c->stream, (const ScfState*)w.st, (const double*)w.spart,
                     nbc, cd.ensemble, rho, c->dvol, c->volume, c->ds, cd.ns, scalars);
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int adscf_init(adscf_ctx** out, int nx, int ny, int nz, int ntypes, double lx, double ly, double lz, double ds,
               int precision) {
  if (!out) return ADSCF_ERR_ARG;
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || ntypes < 1 || ntypes > kMaxTypes || !(lx > 0) || !(ly > 0) || !(lz > 0) ||
      !(ds > 0) || (precision != ADSCF_FP64 && precision != ADSCF_FP32))
    return ADSCF_ERR_ARG;
  // the line passes address one grid-sized array with 32-bit byte offsets
  if ((long)nx * ny * nz > kMaxGridPoints) return ADSCF_ERR_ARG;
  auto* c = new (std::nothrow) adscf_ctx();
  if (!c) return ADSCF_ERR_HIP;
  int rc = guarded(c, [&] {
    int ndev = 0;
    HIPCK(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw ArgError{"no HIP device"};
    HIPCK(hipGetDevice(&c->dev));
    if (const char* g = getenv("ADSCF_GRAPH")) c->use_graph = strcmp(g, "0") != 0;
    HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCK(hipEventCreate(&c->ev0));
    HIPCK(hipEventCreate(&c->ev1));
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->nt = ntypes;
    c->ng = (long)nx * ny * nz;
    c->lx = lx;
    c->ly = ly;
    c->lz = lz;
    c->ds = ds;
    c->volume = lx * ly * lz;                      // mod_fourier.f90:154
    c->dvol = c->volume / (double)(nx * ny * nz);  // :155
    c->prec = precision;
    c->fh.nt = ntypes;
    reset_axes(c);
    c->scratch_dev = scratch_address(c->dev);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
  if (rc != ADSCF_OK) {
    fprintf(stderr, "adscf_init: %s\n", c->err.c_str());
    adscf_finalize(c);
    return rc;
  }
  *out = c;
  return ADSCF_OK;
}

int adscf_finalize(adscf_ctx* c) {
  if (!c) return ADSCF_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->keep.retire(c->graph_src, c->graph);
  c->keep.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  hipStream_t s = c->stream;
  delete c;  // DevBuf destructors free device memory
  if (s) (void)hipStreamDestroy(s);
  return ADSCF_OK;
}

int adscf_set_box(adscf_ctx* c, double lx, double ly, double lz) {
  return guarded(c, [&] {
    if (!(lx > 0) || !(ly > 0) || !(lz > 0)) throw ArgError{"box lengths must be positive"};
    c->lx = lx;
    c->ly = ly;
    c->lz = lz;
    c->volume = lx * ly * lz;
    c->dvol = c->volume / (double)(c->nx * c->ny * c->nz);
    invalidate_tables(c);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
}

int adscf_set_ds(adscf_ctx* c, double ds) {
  return guarded(c, [&] {
    if (!(ds > 0)) throw ArgError{"ds must be positive"};
    c->ds = ds;
    invalidate_tables(c);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
}

int adscf_set_precision(adscf_ctx* c, int precision) {
  return guarded(c, [&] {
    if (precision != ADSCF_FP64 && precision != ADSCF_FP32) throw ArgError{"precision must be 0 (FP64) or 1 (FP32)"};
    if (precision == c->prec) return ADSCF_OK;
    c->prec = precision;
    ensure_current_tables(c);
    if (c->field_set) {  // continue the same SCF run (history kept) in the new precision
      if (precision == ADSCF_FP64) refresh_E<double>(c);
      else refresh_E<float>(c);
      c->phi_fresh = false;
    }
    return ADSCF_OK;
  });
}

int adscf_volume(const adscf_ctx* c, double* volume, double* dvol) {
  if (!c) return ADSCF_ERR_ARG;
  if (volume) *volume = c->volume;
  if (dvol) *dvol = c->dvol;
  return ADSCF_OK;
}

const char* adscf_last_error(const adscf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int adscf_real2fourier(adscf_ctx* c, const double* prop, double* pf) {
  return guarded(c, [&] {
    if (!prop || !pf) throw ArgError{"null array"};
    return c->prec == ADSCF_FP64 ? r2c_impl<double>(c, prop, pf) : r2c_impl<float>(c, prop, pf);
  });
}

int adscf_fourier2real(adscf_ctx* c, const double* pf, double* prop) {
  return guarded(c, [&] {
    if (!prop || !pf) throw ArgError{"null array"};
    return c->prec == ADSCF_FP64 ? c2r_impl<double>(c, pf, prop) : c2r_impl<float>(c, pf, prop);
  });
}

int adscf_laplace(adscf_ctx* c, double* laplace, double* explaplace, double* nabla) {
  // Host tables, as Calculate_Laplace fills module arrays (mod_fourier.f90:175-194).
  return guarded(c, [&] {
    const double pi = 3.141592653589793;
    const int nh = c->nx / 2 + 1;
    const long nk = (long)nh * c->ny * c->nz;
    const double dx = c->lx / c->nx, dy = c->ly / c->ny, dz = c->lz / c->nz;
    for (int z = 1; z <= c->nz; ++z)
      for (int y = 1; y <= c->ny; ++y)
        for (int x = 1; x <= nh; ++x) {
          double qx = 2.0 * pi * (double)(x - 1) / c->lx;
          double qy = 2.0 * pi * (double)(y - 1) / c->ly;
          double qz = 2.0 * pi * (double)(z - 1) / c->lz;
          if (y > c->ny / 2 + 1) qy = 2.0 * pi * ((double)(y - 1) - (double)c->ny) / c->ly;
          if (z > c->nz / 2 + 1) qz = 2.0 * pi * ((double)(z - 1) - (double)c->nz) / c->lz;
          const long i = ((long)(z - 1) * c->ny + (y - 1)) * nh + (x - 1);
          const double k2 = qx * qx + qy * qy + qz * qz;
          if (laplace) {
            laplace[2 * i] = -k2;
            laplace[2 * i + 1] = 0.0;
          }
          if (explaplace) {
            explaplace[2 * i] = exp(-c->ds * k2);
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
    return ADSCF_OK;
  });
}

int adscf_expfield(adscf_ctx* c, const double* field, double* expfield) {
  return guarded(c, [&] {
    if (!field || !expfield) throw ArgError{"null array"};
    const long n = (long)c->nt * c->ng;
    c->io.ensure(n);
    c->io2.ensure(n);
    HIPCK(hipMemcpyAsync(c->io.p, field, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_expfield<double>, dim3(grid_for(c->ng)), dim3(kThreads), 0, c->stream,
                       (const double*)c->io.p, c->io2.p, c->ng, c->nt, c->ds, (double*)nullptr, (const int*)nullptr);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(expfield, c->io2.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_propagate_step(adscf_ctx* c, double* prop, const double* E) {
  return guarded(c, [&] {
    if (!prop || !E) throw ArgError{"null array"};
    return c->prec == ADSCF_FP64 ? propagate_step_impl<double>(c, prop, E) : propagate_step_impl<float>(c, prop, E);
  });
}

int adscf_propagate_chain(adscf_ctx* c, int nblocks, const int* monomer_type, const int* block_segments, int ensemble,
                          double* rho, double* mu, const double* expfield, double* distribution, double* density,
                          double* qsum, double* energy) {
  return guarded(c, [&]() -> int {
    if (!monomer_type || !block_segments || !rho || !mu || !expfield || !qsum || !energy)
      throw ArgError{"null argument"};
    if (ensemble != 1 && ensemble != 2) {
      c->err = "Invalid ensemble; Cannot calculate Density";  // mod_multiblock.f90:182
      return ADSCF_ERR_ENSEMBLE;
    }
    ChainDesc cd = make_chain(c, nblocks, monomer_type, block_segments, ensemble);
    return c->prec == ADSCF_FP64
               ? propagate_chain_impl<double>(c, cd, rho, mu, expfield, distribution, density, qsum, energy)
               : propagate_chain_impl<float>(c, cd, rho, mu, expfield, distribution, density, qsum, energy);
  });
}

int adscf_propagate_chain_noalloc(adscf_ctx* c, int nblocks, const int* monomer_type, const int* block_segments,
                                  int ensemble, double rho, double mu, const double* expfield, double* distribution,
                                  double* density, double* scalars) {
  return guarded(c, [&]() -> int {
    if (!monomer_type || !block_segments || !expfield || !distribution || !density || !scalars)
      throw ArgError{"null argument"};
    if (ensemble != 1 && ensemble != 2) {
      c->err = "Invalid ensemble; Cannot calculate Density";  // mod_multiblock.f90:182
      return ADSCF_ERR_ENSEMBLE;
    }
    if (!c->scratch_dev) throw ArgError{"the device scratch was not resolved at adscf_init"};
    ChainDesc cd = make_chain(c, nblocks, monomer_type, block_segments, ensemble);
    if (c->prec == ADSCF_FP64)
      propagate_chain_noalloc_impl<double>(c, cd, rho, mu, expfield, distribution, density, scalars);
    else
      propagate_chain_noalloc_impl<float>(c, cd, rho, mu, expfield, distribution, density, scalars);
    return ADSCF_OK;
  });
}

int adscf_propagate_homopolymer(adscf_ctx* c, int monomer_type, int segments, int ensemble, double* rho, double* mu,
                                const double* expfield, double* distribution, double* density, double* qsum,
                                double* energy) {
  // mod_homopolymer.f90:90-141 is the one-block chain: q+ runs the same steps as q, so the
  // packed sweep's complementary half reproduces qq(mono - s) of :119-125
  return adscf_propagate_chain(c, 1, &monomer_type, &segments, ensemble, rho, mu, expfield, distribution, density,
                               qsum, energy);
}

int adscf_update_solvent(adscf_ctx* c, int ensemble, double* rho, double* mu, const double* field,
                         double* distribution, double* density, double* qsum, double* energy) {
  return guarded(c, [&]() -> int {
    if (!rho || !mu || !field || !qsum || !energy) throw ArgError{"null argument"};
    if (ensemble != 1 && ensemble != 2) {
      c->err = "Invalid ensemble; Cannot calculate Density";  // mod_solvent.f90:105
      return ADSCF_ERR_ENSEMBLE;
    }
    ensure_partials(c);
    const long ng = c->ng;
    c->io.ensure((size_t)3 * ng);
    double* w = c->io.p;
    double* dist = w + ng;
    double* den = w + 2 * ng;
    const int nb = grid_for(ng);
    HIPCK(hipMemcpyAsync(w, field, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_solvent_dist, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)w, ng, dist,
                       c->qpart.p);
    std::vector<double> part(nb);
    HIPCK(hipMemcpyAsync(part.data(), c->qpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    *qsum = c->dvol * s;
    const double f = ensemble == 1 ? *rho * c->volume / *qsum : exp(*mu);
    hipLaunchKernelGGL(k_scale_sum, dim3(nb), dim3(kThreads), 0, c->stream, (const double*)dist, ng, f, den,
                       c->spart.p);
    HIPCK(hipMemcpyAsync(part.data(), c->spart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (distribution) HIPCK(hipMemcpyAsync(distribution, dist, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (density) HIPCK(hipMemcpyAsync(density, den, ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (ensemble == 1) {
      *mu = log(*rho * c->volume / *qsum);
      *energy = *rho * c->volume * (*mu - 1.0);
    } else {
      double t = 0.0;
      for (int b = 0; b < nb; ++b) t += part[b];
      *rho = t * c->dvol / c->volume;
      *energy = -*rho * c->volume;
    }
    return ADSCF_OK;
  });
}

int adscf_set_interactions(adscf_ctx* c, double kappa, const double* chi, const double* vol) {
  return guarded(c, [&] {
    if (!chi || !vol) throw ArgError{"null array"};
    c->fh.nt = c->nt;
    c->fh.kappa = kappa;
    for (int i = 0; i < c->nt * c->nt; ++i) c->fh.chi[i] = chi[i];
    for (int i = 0; i < c->nt; ++i) c->fh.vol[i] = vol[i];
    c->fh_set = true;
    invalidate_tables(c);  // the captured iteration embeds the parameters
    c->tables64 = c->tables32 = false;
    return ADSCF_OK;
  });
}

int adscf_mean_fields(adscf_ctx* c, const double* density, double* field) {
  return guarded(c, [&] {
    if (!density || !field) throw ArgError{"null array"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    const long n = (long)c->nt * c->ng;
    c->io.ensure(n);
    c->io2.ensure(n);
    HIPCK(hipMemcpyAsync(c->io.p, density, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_mean_fields, dim3(grid_for(c->ng)), dim3(kThreads), 0, c->stream, (const double*)c->io.p,
                       c->io2.p, c->ng, c->fh);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(field, c->io2.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_interaction_energy(adscf_ctx* c, const double* density, double* energy) {
  return guarded(c, [&] {
    if (!density || !energy) throw ArgError{"null array"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    ensure_partials(c);
    const long n = (long)c->nt * c->ng;
    c->io.ensure(n);
    HIPCK(hipMemcpyAsync(c->io.p, density, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<double> p;
    host_energy_parts(c, c->io.p, nullptr, p);
    *energy = interaction_from_parts(c, p);
    return ADSCF_OK;
  });
}

int adscf_average_volume_fraction(adscf_ctx* c, const double* density, double* average) {
  return guarded(c, [&] {
    if (!density || !average) throw ArgError{"null array"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    ensure_partials(c);
    const long n = (long)c->nt * c->ng;
    c->io.ensure(n);
    HIPCK(hipMemcpyAsync(c->io.p, density, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const int nb = grid_for(c->ng);
    // sum(volume_fraction) / Ng with volume_fraction_t = density_t * v_t (mod_FloryHuggins.f90:147-151)
    double tot = 0.0;
    for (int t = 0; t < c->nt; ++t) {
      hipLaunchKernelGGL(k_qsum<double>, dim3(nb), dim3(kThreads), 0, c->stream,
                         (const double*)(c->io.p + (long)t * c->ng), c->ng, c->qpart.p, (const int*)nullptr);
      std::vector<double> part(nb);
      HIPCK(hipMemcpyAsync(part.data(), c->qpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCK(hipStreamSynchronize(c->stream));
      double s = 0.0;
      for (int b = 0; b < nb; ++b) s += part[b];
      tot += s * c->fh.vol[t];
    }
    *average = tot / (double)(c->nx * c->ny * c->nz);
    return ADSCF_OK;
  });
}

int adscf_mixing_simple(adscf_ctx* c, double* field, const double* dfield, long fieldsize, int mixing_dim,
                        double lambda) {
  return guarded(c, [&] {
    if (!field || !dfield || fieldsize < 1 || mixing_dim < 0) throw ArgError{"bad mixing arguments"};
    ensure_partials(c);
    c->io.ensure(fieldsize);
    c->io2.ensure(fieldsize);
    HIPCK(hipMemcpyAsync(c->io.p, field, fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->io2.p, dfield, fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // field(:,0) = field(:,0) + lambda*dfield(:,0): one mixing step with no history
    ScfState st{};
    st.mode = 0;
    st.lam = lambda;
    HIPCK(hipMemcpyAsync(c->cst.p, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    c->e32.cmixE.ensure(fieldsize);
    // with m = 0 the ring has one slot and the new field overwrites it in place
    hipLaunchKernelGGL(k_mix<float>, dim3(grid_for(fieldsize)), dim3(kThreads), 0, c->stream, c->io.p,
                       (const double*)c->io2.p, fieldsize, 1, 0, (const ScfState*)c->cst.p, 0.0, c->e32.cmixE.p,
                       c->e2part.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(field, c->io.p, fieldsize * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_mixing_lambda(adscf_ctx* c, double* field, double* dfield, long fieldsize, int mixing_dim, double lammax,
                        int iter) {
  return guarded(c, [&] {
    if (!field || !dfield || fieldsize < 1 || mixing_dim < 1) throw ArgError{"bad mixing arguments"};
    ensure_partials(c);
    // Mixing_Lambda (mod_iterate.f90:46-84): lambda from two norms, then shift column 0 into 1
    const double lammin = 0.00001;
    double lambda = lammin;
    if (iter != 1) {
      c->io.ensure(2 * fieldsize);
      c->io2.ensure(2 * fieldsize);
      HIPCK(hipMemcpyAsync(c->io.p, field, 2 * fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCK(hipMemcpyAsync(c->io2.p, dfield, 2 * fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
      // the m = 1 Gram kernel over the two-slot rings gives <D0-D1, D0-D1>; the field
      // differences use the same kernel on the field ring
      c->gpart.ensure((size_t)kGramBlocks * 3);
      c->gsum.ensure(3);
      ScfState st{};
      HIPCK(hipMemcpyAsync(c->cst.p, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
      double sums[2];
      for (int which = 0; which < 2; ++which) {
        launch_gram(c->stream, (const double*)(which == 0 ? c->io.p : c->io2.p), fieldsize, 1,
                    (const ScfState*)c->cst.p, 1, c->gpart.p, c->gsum.p, nullptr);
        HIPCK(hipGetLastError());
        double g3[3];
        HIPCK(hipMemcpyAsync(g3, c->gsum.p, sizeof g3, hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        sums[which] = g3[2];  // pair (1,1)
      }
      const double num = sums[0], nom = sums[1];
      if (nom > 1.0e-50) {
        lambda = sqrt(num / nom);
        if (lambda < lammin) lambda = lammin;
        else if (lambda > lammax) lambda = lammax;
      }
    }
    // shift column 0 into 1 and update column 0 (host-array semantics, done on device)
    c->io.ensure(fieldsize);
    c->io2.ensure(fieldsize);
    HIPCK(hipMemcpyAsync(c->io.p, field, fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->io2.p, dfield, fieldsize * sizeof(double), hipMemcpyHostToDevice, c->stream));
    ScfState st{};
    st.mode = 0;
    st.lam = lambda;
    HIPCK(hipMemcpyAsync(c->cst.p, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    c->e32.cmixE.ensure(fieldsize);
    hipLaunchKernelGGL(k_mix<float>, dim3(grid_for(fieldsize)), dim3(kThreads), 0, c->stream, c->io.p,
                       (const double*)c->io2.p, fieldsize, 1, 0, (const ScfState*)c->cst.p, 0.0, c->e32.cmixE.p,
                       c->e2part.p);
    HIPCK(hipGetLastError());
    std::memcpy(field + fieldsize, field, fieldsize * sizeof(double));
    std::memcpy(dfield + fieldsize, dfield, fieldsize * sizeof(double));
    HIPCK(hipMemcpyAsync(field, c->io.p, fieldsize * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_mixing_anderson(adscf_ctx* c, double* field, double* dfield, long fieldsize, int m, double lam_mix,
                          int iter) {
  return guarded(c, [&] {
    if (!field || !dfield || fieldsize < 1 || m < 1 || m > kMaxHist) throw ArgError{"bad mixing arguments"};
    ensure_partials(c);
    const size_t tot = (size_t)(m + 1) * fieldsize;
    c->io.ensure(tot);
    c->io2.ensure(tot);
    HIPCK(hipMemcpyAsync(c->io.p, field, tot * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->io2.p, dfield, tot * sizeof(double), hipMemcpyHostToDevice, c->stream));
    ScfState st{};
    HIPCK(hipMemcpyAsync(c->cst.p, &st, sizeof st, hipMemcpyHostToDevice, c->stream));
    const int np = (m + 1) * (m + 2) / 2;
    c->gpart.ensure((size_t)kGramBlocks * np);
    c->gsum.ensure(np);
    const int status = iter < 1000 ? iter : 1000;
    if (status > m)
      launch_gram(c->stream, (const double*)c->io2.p, fieldsize, m, (const ScfState*)c->cst.p, 1, c->gpart.p,
                  c->gsum.p, nullptr);
    SolveArgs s{};
    s.m = m;
    s.lam_mix = lam_mix;
    s.gsum = c->gsum.p;
    s.iter_override = iter;
    s.st = c->cst.p;
    hipLaunchKernelGGL(k_solve, dim3(1), dim3(kThreads), 0, c->stream, s);
    c->e32.cmixE.ensure(fieldsize);
    hipLaunchKernelGGL(k_mix<float>, dim3(grid_for(fieldsize)), dim3(kThreads), 0, c->stream, c->io.p,
                       (const double*)c->io2.p, fieldsize, 1, m, (const ScfState*)c->cst.p, 0.0, c->e32.cmixE.p,
                       c->e2part.p);
    HIPCK(hipGetLastError());
    // ring with head 0: the new field sits in slot m; the shift (:143-146) maps slot i-1 to column i
    const size_t fb = fieldsize * sizeof(double);
    HIPCK(hipMemcpyAsync(field, c->io.p + (size_t)m * fieldsize, fb, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(field + fieldsize, c->io.p, (size_t)m * fb, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(dfield + fieldsize, c->io2.p, (size_t)m * fb, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_device_lu_solve(adscf_ctx* c, const double* a, int n, double* b, double* lu, int* indx, int* singular) {
  return guarded(c, [&] {
    if (!a || !b || n < 1 || n > kMaxHist) throw ArgError{"adscf_device_lu_solve: need a, b and 1 <= n <= 40"};
    c->io.ensure((size_t)n * n + n);
    c->io2.ensure((size_t)n * n + n);
    DevBuf<int> piv;
    piv.ensure(n + 1);
    HIPCK(hipMemcpyAsync(c->io.p, a, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(hipMemcpyAsync(c->io.p + (size_t)n * n, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_lu_check, dim3(1), dim3(kThreads), 0, c->stream, (const double*)c->io.p,
                       (const double*)c->io.p + (size_t)n * n, n, c->io2.p, piv.p);
    HIPCK(hipGetLastError());
    std::vector<int> hp(n + 1);
    HIPCK(hipMemcpyAsync(b, c->io2.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lu)
      HIPCK(hipMemcpyAsync(lu, c->io2.p + n, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipMemcpyAsync(hp.data(), piv.p, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    if (indx) std::copy(hp.begin(), hp.begin() + n, indx);
    if (singular) *singular = hp[n];
    return ADSCF_OK;
  });
}

int adscf_ludcmp(double* a, int n, int* indx, double* d) {
  if (!a || !indx || !d || n < 1) return ADSCF_ERR_ARG;
  std::vector<double> vv(n + 1);
  lu_decompose(a, n, indx, d, vv.data());
  return ADSCF_OK;
}

int adscf_lubksb(const double* a, int n, const int* indx, double* b) {
  if (!a || !indx || !b || n < 1) return ADSCF_ERR_ARG;
  lu_backsubstitute(a, n, indx, b);
  return ADSCF_OK;
}

int adscf_solve_lu(const double* a, const double* b, double* x, int n) {
  if (!a || !b || !x || n < 1) return ADSCF_ERR_ARG;
  std::vector<double> aa(a, a + (size_t)n * n), vv(n + 1);
  std::vector<int> indx(n);
  double d;
  for (int i = 0; i < n; ++i) x[i] = b[i];
  lu_decompose(aa.data(), n, indx.data(), &d, vv.data());
  lu_backsubstitute(aa.data(), n, indx.data(), x);
  return ADSCF_OK;
}

int adscf_solve_cholesky(const double* A, const double* B, double* X, int n) {
  // Solve_LinearSystem_Cholesky (mod_iterate.f90:160-195), A = L L^T
  if (!A || !B || !X || n < 1) return ADSCF_ERR_ARG;
  std::vector<double> L((size_t)n * n, 0.0), Y(n);
  auto a = [&](int i, int j) { return A[(i - 1) + (size_t)(j - 1) * n]; };
  auto l = [&](int i, int j) -> double& { return L[(i - 1) + (size_t)(j - 1) * n]; };
  l(1, 1) = sqrt(a(1, 1));
  for (int i = 2; i <= n; ++i) {
    for (int j = 1; j <= i - 1; ++j) {
      double s = 0.0;
      for (int k = 1; k <= j - 1; ++k) s += l(i, k) * l(j, k);
      l(i, j) = (a(i, j) - s) / l(j, j);
    }
    double s = 0.0;
    for (int k = 1; k <= i - 1; ++k) s += l(i, k) * l(i, k);
    l(i, i) = sqrt(a(i, i) - s);
  }
  for (int i = 1; i <= n; ++i) {
    double s = 0.0;
    for (int k = 1; k <= i - 1; ++k) s += l(i, k) * Y[k - 1];
    Y[i - 1] = (B[i - 1] - s) / l(i, i);
  }
  for (int j = n; j >= 1; --j) {
    double s = 0.0;
    for (int k = j + 1; k <= n; ++k) s += l(k, j) * X[k - 1];
    X[j - 1] = (Y[j - 1] - s) / l(j, j);
  }
  return ADSCF_OK;
}

int adscf_scf_setup(adscf_ctx* c, int nblocks, const int* monomer_type, const int* block_segments, int ensemble,
                    double rho, double mu, int mixing_dim, double lam_mix) {
  return guarded(c, [&]() -> int {
    if (!monomer_type || !block_segments) throw ArgError{"null array"};
    if (ensemble != 1 && ensemble != 2) {
      c->err = "Invalid ensemble; Cannot calculate Density";
      return ADSCF_ERR_ENSEMBLE;
    }
    if (mixing_dim < 1 || mixing_dim > kMaxHist) throw ArgError{"mixing_dim must be in [1, 40]"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    c->cd = make_chain(c, nblocks, monomer_type, block_segments, ensemble);
    c->bl_cd.clear();
    c->bl_rho.clear();
    c->m = mixing_dim;
    c->lam_mix = lam_mix;
    c->rho = rho;
    c->mu = mu;
    c->scf_ready = true;
    c->field_set = false;
    invalidate_tables(c);
    return ADSCF_OK;
  });
}

int adscf_scf_set_field(adscf_ctx* c, const double* field) {
  return guarded(c, [&]() -> int {
    if (!field) throw ArgError{"null array"};
    if (!c->scf_ready) {
      c->err = "adscf_scf_setup has not been called";
      return ADSCF_ERR_STATE;
    }
    if (c->prec == ADSCF_FP64) scf_set_field_impl<double>(c, field);
    else scf_set_field_impl<float>(c, field);
    c->field_set = true;
    c->phi_fresh = false;
    c->last_loop = -1;
    return ADSCF_OK;
  });
}

int adscf_scf_iterate(adscf_ctx* c, int niter, double tol, int* iters, double* accuracy, int* converged) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (niter < 0) throw ArgError{"niter must be >= 0"};
    if (c->user_measure == 2 && c->bl_cd.size() > 1)
      throw ArgError{"measure 2 (change of the densities) is not available for a blend of several polymers"};
    c->resid_kind = c->user_kind;
    c->measure = c->user_measure;
    c->mixing = c->user_mixing;
    return c->prec == ADSCF_FP64 ? scf_iterate_impl<double>(c, niter, tol, iters, accuracy, converged)
                                 : scf_iterate_impl<float>(c, niter, tol, iters, accuracy, converged);
  });
}

int adscf_scf_evaluate(adscf_ctx* c) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (c->prec == ADSCF_FP64) enqueue_evaluate<double>(c);
    else enqueue_evaluate<float>(c);
    HIPCK(hipStreamSynchronize(c->stream));
    c->phi_fresh = true;
    return ADSCF_OK;
  });
}

int adscf_scf_get(adscf_ctx* c, double* field, double* density) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    ScfState st{};
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    const long nvar = c->ng * c->nt;
    if (field)
      HIPCK(hipMemcpyAsync(field, c->wring.p + (size_t)st.head * nvar, nvar * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    if (density) HIPCK(hipMemcpyAsync(density, c->phi.p, nvar * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return ADSCF_OK;
  });
}

int adscf_scf_energy(adscf_ctx* c, double* F, double* Fc, double* qsum, double* mu, double* rho) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (!c->phi_fresh) {
      if (c->prec == ADSCF_FP64) enqueue_evaluate<double>(c);
      else enqueue_evaluate<float>(c);
      c->phi_fresh = true;
    }
    ScfState st{};
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double rho_v = st.rho, energy = st.energy;
    if (c->cd.ensemble == 2) {
      // grand canonical rho from the densities of this evaluation
      std::vector<double> sp(grid_for(c->ng));
      HIPCK(hipMemcpyAsync(sp.data(), c->spart.p, sp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCK(hipStreamSynchronize(c->stream));
      double s = 0.0;
      for (double v : sp) s += v;
      rho_v = s * c->dvol / c->volume / (c->ds * c->cd.ns);
      energy = -rho_v * c->volume;
    }
    if (!c->bl_cd.empty()) {  // Calculate_Energy of the blend (main_blend.f90:356-366)
      std::vector<ScfState> ps(c->bl_cd.size());
      HIPCK(hipMemcpy(ps.data(), c->pst.p, ps.size() * sizeof(ScfState), hipMemcpyDeviceToHost));
      energy = 0.0;
      rho_v = 0.0;
      for (size_t i = 0; i < ps.size(); ++i) {
        energy += ps[i].energy;
        rho_v += c->bl_rho[i];
      }
      st.qsum = ps[0].qsum;
      st.mu = ps[0].mu;
    }
    std::vector<double> p;
    host_energy_parts(c, c->phi.p, c->wring.p + (size_t)st.head * c->ng * c->nt, p);
    const double chain_energy = energy - c->dvol * p[0];
    const double f = chain_energy + interaction_from_parts(c, p);
    if (F) *F = f;
    if (Fc) *Fc = f / (rho_v * c->volume);
    if (qsum) *qsum = st.qsum;
    if (mu) *mu = st.mu;
    if (rho) *rho = rho_v;
    return ADSCF_OK;
  });
}

int adscf_blend_setup(adscf_ctx* c, int npol, const int* monomer_type, const int* segments, const double* rho,
                      int mixing_dim, double lam_mix) {
  return guarded(c, [&]() -> int {
    if (!monomer_type || !segments || !rho) throw ArgError{"null array"};
    if (npol < 1 || npol > 16) throw ArgError{"npol must be in [1, 16]"};
    if (mixing_dim < 1 || mixing_dim > kMaxHist) throw ArgError{"mixing_dim must be in [1, 40]"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    std::vector<ChainDesc> cds;
    for (int i = 0; i < npol; ++i) cds.push_back(make_chain(c, 1, monomer_type + i, segments + i, 1));
    c->bl_cd = cds;
    c->bl_rho.assign(rho, rho + npol);
    c->cd = cds[0];
    c->rho = rho[0];
    c->mu = 0.0;
    c->m = mixing_dim;
    c->lam_mix = lam_mix;
    c->pst.ensure(npol);
    c->scf_ready = true;
    c->field_set = false;
    invalidate_tables(c);
    return ADSCF_OK;
  });
}

int adscf_blend_scalars(adscf_ctx* c, double* qsum, double* mu, double* energy) {
  return guarded(c, [&]() -> int {
    if (c->bl_cd.empty()) {
      c->err = "adscf_blend_setup has not been called";
      return ADSCF_ERR_STATE;
    }
    std::vector<ScfState> ps(c->bl_cd.size());
    HIPCK(hipStreamSynchronize(c->stream));
    HIPCK(hipMemcpy(ps.data(), c->pst.p, ps.size() * sizeof(ScfState), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ps.size(); ++i) {
      if (qsum) qsum[i] = ps[i].qsum;
      if (mu) mu[i] = ps[i].mu;
      if (energy) energy[i] = ps[i].energy;
    }
    return ADSCF_OK;
  });
}

int adscf_epd_step(adscf_ctx* c, double timestep, double accuracy_scf, int itermax, int* iterations) {
  return guarded(c, [&]() -> int {
    if (!c->field_set || !c->target_set) {
      c->err = "adscf_epd_step needs adscf_scf_set_field and adscf_ddft_set_density";
      return ADSCF_ERR_STATE;
    }
    if (itermax < 1) throw ArgError{"itermax must be >= 1"};
    return c->prec == ADSCF_FP64 ? epd_step_impl<double>(c, timestep, accuracy_scf, itermax, iterations)
                                 : epd_step_impl<float>(c, timestep, accuracy_scf, itermax, iterations);
  });
}

int adscf_field_derivative(adscf_ctx* c, double* dfield) {
  return guarded(c, [&]() -> int {
    if (!dfield) throw ArgError{"null array"};
    if (!c->field_set || !c->fh_set) {
      c->err = "adscf_field_derivative needs adscf_set_interactions and adscf_scf_set_field";
      return ADSCF_ERR_STATE;
    }
    if (c->prec == ADSCF_FP64) field_derivative_impl<double>(c, dfield);
    else field_derivative_impl<float>(c, dfield);
    return ADSCF_OK;
  });
}

int adscf_epd_field_step(adscf_ctx* c, double timestep, int method) {
  return guarded(c, [&]() -> int {
    if (method != 0 && method != 1) throw ArgError{"method must be 0 (Euler) or 1 (Runge-Kutta)"};
    if (!c->field_set || !c->fh_set) {
      c->err = "adscf_epd_field_step needs adscf_set_interactions and adscf_scf_set_field";
      return ADSCF_ERR_STATE;
    }
    if (c->prec == ADSCF_FP64) epd_field_step_impl<double>(c, timestep, method);
    else epd_field_step_impl<float>(c, timestep, method);
    c->phi_fresh = false;
    c->last_loop = -1;
    return ADSCF_OK;
  });
}

int adscf_scf_set_residual(adscf_ctx* c, int kind, int measure, const double* target) {
  return guarded(c, [&]() -> int {
    if (kind != 0 && kind != 1) throw ArgError{"residual kind must be 0 (mean field) or 1 (density)"};
    if (measure < 0 || measure > 2)
      throw ArgError{"measure must be 0 (max), 1 (sum * dvol) or 2 (sum of squared density changes)"};
    const long nvar = c->ng * c->nt;
    if (target) {
      c->target.ensure(nvar);
      HIPCK(hipMemcpy(c->target.p, target, nvar * sizeof(double), hipMemcpyHostToDevice));
      c->target_set = true;
    }
    if (kind == 1 && !c->target_set) throw ArgError{"a density residual needs target densities"};
    c->user_kind = kind;
    c->user_measure = measure;
    return ADSCF_OK;
  });
}

int adscf_scf_set_density(adscf_ctx* c, const double* density) {
  return guarded(c, [&]() -> int {
    if (!density) throw ArgError{"null array"};
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    HIPCK(hipMemcpy(c->phi.p, density, c->ng * c->nt * sizeof(double), hipMemcpyHostToDevice));
    c->phi_fresh = false;
    return ADSCF_OK;
  });
}

int adscf_set_contour_rule(adscf_ctx* c, int rule) {
  return guarded(c, [&]() -> int {
    if (rule != RULE_TRAPEZOID && rule != RULE_SIMPSON) throw ArgError{"rule must be 0 (trapezoid) or 1 (Simpson)"};
    if (rule != c->contour_rule) {
      c->contour_rule = rule;
      c->keep.retire(c->graph_src, c->graph);  // the captured iteration bakes the rule in
      c->phi_fresh = false;
    }
    return ADSCF_OK;
  });
}

int adscf_scf_set_mixing(adscf_ctx* c, int mixing_type) {
  return guarded(c, [&]() -> int {
    if (mixing_type < 1 || mixing_type > 3) throw ArgError{"mixing_type must be 1 (simple), 2 (lambda) or 3 (Anderson)"};
    c->user_mixing = mixing_type;
    return ADSCF_OK;
  });
}

int adscf_conjugate_fields(adscf_ctx* c, const double* density, double thresh, int maxit, int measure, int* iterations,
                           double* accuracy) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (maxit < 0) throw ArgError{"maxit must be >= 0"};
    if (measure != 0 && measure != 1) throw ArgError{"measure must be 0 (max) or 1 (sum * dvol)"};
    const long nvar = c->ng * c->nt;
    if (density) {
      c->target.ensure(nvar);
      HIPCK(hipMemcpy(c->target.p, density, nvar * sizeof(double), hipMemcpyHostToDevice));
      c->target_set = true;
    }
    if (!c->target_set) throw ArgError{"no target densities"};
    return c->prec == ADSCF_FP64 ? conjugate_impl<double>(c, thresh, maxit, measure, iterations, accuracy)
                                 : conjugate_impl<float>(c, thresh, maxit, measure, iterations, accuracy);
  });
}

int adscf_mobility_qq(adscf_ctx* c, int nq, double dq, const double* table, double wcutoff, double* mobility_qq) {
  // Read_Mobilities, table part (main_multiblock.f90:486-504), on the host at set-up
  return guarded(c, [&]() -> int {
    if (!table || nq < 1 || !(dq > 0.0)) throw ArgError{"bad mobility table"};
    const int nt = c->nt;
    const long nk = (long)(c->nx / 2 + 1) * c->ny * c->nz;
    std::vector<double> lap(2 * nk), mob((size_t)nk * nt * nt);
    const int rc = adscf_laplace(c, lap.data(), nullptr, nullptr);
    if (rc != ADSCF_OK) return rc;
    for (long i = 0; i < nk; ++i) {
      const double qq = -lap[2 * i];
      const double q = sqrt(qq);
      const int iq = (int)(q / dq);
      for (int t2 = 0; t2 < nt; ++t2)
        for (int t1 = 0; t1 < nt; ++t1) {
          const double* col = table + (long)(t1 + t2 * nt) * nq;  // mobility_table(:, t1, t2)
          double mq;
          if (iq == 0) mq = col[0] * q / dq;
          else if (iq >= nq) mq = col[nq - 1];
          else mq = (col[iq - 1] * (q - (double)iq * dq) + col[iq] * ((double)(iq + 1) * dq - q)) / dq;
          const double v = qq * mq;
          mob[(size_t)(t1 + t2 * nt) * nk + i] = v < wcutoff ? v : wcutoff;
        }
    }
    if (mobility_qq) std::copy(mob.begin(), mob.end(), mobility_qq);
    return adscf_ddft_set_mobility(c, mob.data());
  });
}

int adscf_ddft_set_mobility(adscf_ctx* c, const double* mobility_qq) {
  return guarded(c, [&]() -> int {
    if (!mobility_qq) throw ArgError{"null array"};
    const long n = (long)(c->nx / 2 + 1) * c->ny * c->nz * c->nt * c->nt;
    c->mob.ensure(n);
    HIPCK(hipMemcpy(c->mob.p, mobility_qq, n * sizeof(double), hipMemcpyHostToDevice));
    c->mob_set = true;
    return ADSCF_OK;
  });
}

int adscf_ddft_set_density(adscf_ctx* c, const double* density) {
  return guarded(c, [&]() -> int {
    if (!density) throw ArgError{"null array"};
    const long nvar = c->ng * c->nt;
    c->target.ensure(nvar);
    HIPCK(hipMemcpy(c->target.p, density, nvar * sizeof(double), hipMemcpyHostToDevice));
    c->target_set = true;
    return ADSCF_OK;
  });
}

int adscf_ddft_get_density(adscf_ctx* c, double* density) {
  return guarded(c, [&]() -> int {
    if (!density) throw ArgError{"null array"};
    if (!c->target_set) {
      c->err = "no DDFT density has been set";
      return ADSCF_ERR_STATE;
    }
    HIPCK(hipStreamSynchronize(c->stream));
    HIPCK(hipMemcpy(density, c->target.p, c->ng * c->nt * sizeof(double), hipMemcpyDeviceToHost));
    return ADSCF_OK;
  });
}

int adscf_ddft_step(adscf_ctx* c, double timestep, int qdmin, double accuracy_scf, int itermax_scf, int* gcfstat,
                    int* iterations) {
  return guarded(c, [&]() -> int {
    if (!c->field_set || !c->target_set || !c->mob_set) {
      c->err = "adscf_ddft_step needs adscf_scf_set_field, adscf_ddft_set_density and a mobility table";
      return ADSCF_ERR_STATE;
    }
    if (itermax_scf < 1) throw ArgError{"itermax_scf must be >= 1"};
    return c->prec == ADSCF_FP64 ? ddft_step_impl<double>(c, timestep, qdmin, accuracy_scf, itermax_scf, gcfstat, iterations)
                                 : ddft_step_impl<float>(c, timestep, qdmin, accuracy_scf, itermax_scf, gcfstat, iterations);
  });
}

int adscf_ddft_energy(adscf_ctx* c, double* F, double* mu, double* rho) {
  // Calculate_Energy (main_multiblock.f90:548-557) with the DDFT density: the chain energy of
  // the last density evaluation - dvol sum(density*field) + Interaction_Energy(density)
  return guarded(c, [&]() -> int {
    if (!c->field_set || !c->target_set) {
      c->err = "no DDFT state";
      return ADSCF_ERR_STATE;
    }
    ScfState st{};
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double chain = st.energy;
    if (!c->bl_cd.empty()) {  // the polymers' energies (main_blend.f90:359-362)
      std::vector<ScfState> ps(c->bl_cd.size());
      HIPCK(hipMemcpy(ps.data(), c->pst.p, ps.size() * sizeof(ScfState), hipMemcpyDeviceToHost));
      chain = 0.0;
      for (const ScfState& q : ps) chain += q.energy;
      st.mu = ps[0].mu;
      st.rho = c->bl_rho[0];
    }
    std::vector<double> p;
    host_energy_parts(c, c->target.p, c->wring.p + (size_t)st.head * c->ng * c->nt, p);
    if (F) *F = chain - c->dvol * p[0] + interaction_from_parts(c, p);
    if (mu) *mu = st.mu;
    if (rho) *rho = st.rho;
    return ADSCF_OK;
  });
}

int adscf_scf_residuals(adscf_ctx* c, double* out, int cap, int* n) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (!out || !n || cap < 0) throw ArgError{"bad residual buffer"};
    ScfState st{};
    HIPCK(hipMemcpyAsync(&st, c->st.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    const int avail = std::min<int>(st.iter, (int)c->resid.n);
    const int k = std::min(avail, cap);
    if (k > 0) HIPCK(hipMemcpy(out, c->resid.p, k * sizeof(double), hipMemcpyDeviceToHost));
    *n = avail;
    return ADSCF_OK;
  });
}

int adscf_scf_timing(adscf_ctx* c, double* ms_per_iter, double* ms_sweep) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_scf_set_field has not been called";
      return ADSCF_ERR_STATE;
    }
    if (ms_per_iter) *ms_per_iter = c->last_ms_iter;
    if (ms_sweep) {
      c->last_ms_sweep = c->prec == ADSCF_FP64 ? sweep_timing<double>(c, 5) : sweep_timing<float>(c, 5);
      *ms_sweep = c->last_ms_sweep;
    }
    return ADSCF_OK;
  });
}

int adscf_adjoint(adscf_ctx* c, int nblocks, const int* monomer_type, const int* block_segments, double rho,
                  const double* field, int target_type, const double* phi_target, double lam, int checkpoint_every,
                  double* loss, double* free_energy, double* grad_field, double* grad_box, double* phi) {
  return guarded(c, [&]() -> int {
    if (!monomer_type || !block_segments || !field || !phi_target) throw ArgError{"null argument"};
    if (!c->fh_set) throw ArgError{"adscf_set_interactions has not been called"};
    if (target_type < 1 || target_type > c->nt) throw ArgError{"target_type out of range 1..ntypes"};
    ChainDesc cd = make_chain(c, nblocks, monomer_type, block_segments, 1);
    for (int b = 0; b < cd.nblocks; ++b)
      if (cd.seg[b] < 1) throw ArgError{"adscf_adjoint: every block needs at least one contour segment"};
    return c->prec == ADSCF_FP64
               ? adjoint_impl<double>(c, cd, rho, field, target_type - 1, phi_target, lam, checkpoint_every, loss,
                                      free_energy, grad_field, grad_box, phi)
               : adjoint_impl<float>(c, cd, rho, field, target_type - 1, phi_target, lam, checkpoint_every, loss,
                                     free_energy, grad_field, grad_box, phi);
  });
}

int adscf_autotune(adscf_ctx* c, int reps, char* report, int report_cap) {
  return guarded(c, [&]() -> int {
    if (!c->field_set) {
      c->err = "adscf_autotune needs adscf_scf_setup and adscf_scf_set_field";
      return ADSCF_ERR_STATE;
    }
    const std::string r = c->prec == ADSCF_FP64 ? autotune_impl<double>(c, std::max(reps, 1))
                                                : autotune_impl<float>(c, std::max(reps, 1));
    if (report && report_cap > 0) {
      const size_t k = std::min(r.size(), (size_t)report_cap - 1);
      memcpy(report, r.data(), k);
      report[k] = 0;
    }
    return ADSCF_OK;
  });
}

int adscf_set_middle_axis(adscf_ctx* c, int axis) {
  return guarded(c, [&]() -> int {
    if (axis < -1 || axis > 2) throw ArgError{"axis must be -1, 0, 1 or 2"};
    c->mid_axis = axis;
    invalidate_tables(c);
    return ADSCF_OK;
  });
}

int adscf_get_middle_axis(const adscf_ctx* c, int* axis) {
  if (!c || !axis) return ADSCF_ERR_ARG;
  *axis = c->mid_axis < 0 ? 1 : c->mid_axis;
  return ADSCF_OK;
}

int adscf_set_plane_passes(adscf_ctx* c, int mode, int n1, int n2) {
  return guarded(c, [&]() -> int {
    if (mode < -1 || mode > 1) throw ArgError{"mode must be -1, 0 or 1"};
    if (n1 != 0 && !(c->prec == ADSCF_FP64 ? (bool)plane_launcher<double>(c->nx, n1, n2, nullptr)
                                            : (bool)plane_launcher<float>(c->nx, n1, n2, nullptr)))
      throw ArgError{"no plane kernel for this shape in this precision"};
    const int old_mode = c->plane, old_n1 = c->plane_N1, old_n2 = c->plane_N2;
    const bool old_set = c->plane_shape_set;
    c->plane = mode;
    if (n1 != 0) {
      c->plane_shape_set = true;
      c->plane_N1 = n1;
      c->plane_N2 = n2;
    }
    if (mode == 1 && !(c->prec == ADSCF_FP64 ? plane_sweep<double>(c) : plane_sweep<float>(c))) {
      c->plane = old_mode;
      c->plane_shape_set = old_set;
      c->plane_N1 = old_n1;
      c->plane_N2 = old_n2;
      throw ArgError{"plane passes need nx == ny, nz > 1 and a plane shape whose LDS fits"};
    }
    invalidate_tables(c);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
}

int adscf_get_plane_passes(const adscf_ctx* c, int* active, int* n1, int* n2) {
  if (!c) return ADSCF_ERR_ARG;
  int N1, N2;
  if (c->prec == ADSCF_FP64) plane_shape<double>(c, N1, N2);
  else plane_shape<float>(c, N1, N2);
  if (active) *active = c->prec == ADSCF_FP64 ? plane_sweep<double>(c) : plane_sweep<float>(c);
  if (n1) *n1 = N1;
  if (n2) *n2 = N2;
  return ADSCF_OK;
}

int adscf_get_schedule(const adscf_ctx* c, int axis, int* n, int* n1, int* n2, int* waves) {
  if (!c || axis < 0 || axis > 2) return ADSCF_ERR_ARG;
  if (n) *n = c->ax[axis].n;
  if (n1) *n1 = c->ax[axis].N1;
  if (n2) *n2 = c->ax[axis].N2;
  if (waves) *waves = c->ax[axis].waves;
  return ADSCF_OK;
}

int adscf_set_pass_tiles(adscf_ctx* c, int axis, int tiles) {
  return guarded(c, [&]() -> int {
    if (axis < 0 || axis > 2 || (tiles != 1 && tiles != 2 && tiles != 4)) throw ArgError{"tiles must be 1, 2 or 4"};
    c->ax[axis].tiles = tiles;
    invalidate_tables(c);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
}

int adscf_get_pass_tiles(const adscf_ctx* c, int axis, int* tiles) {
  if (!c || axis < 0 || axis > 2 || !tiles) return ADSCF_ERR_ARG;
  *tiles = c->ax[axis].tiles;
  return ADSCF_OK;
}

int adscf_set_schedule(adscf_ctx* c, int axis, int n1, int n2, int waves) {
  return guarded(c, [&]() -> int {
    if (axis < 0 || axis > 2 || waves < 1 || waves > 4) throw ArgError{"bad schedule"};
    AxisTables& ax = c->ax[axis];
    if (n1 != 0 && !reg_launcher<double>(ax.n, n1, n2)) throw ArgError{"no register kernel for this shape"};
    ax.N1 = n1;
    ax.N2 = n1 ? n2 : 0;
    ax.waves = waves;
    invalidate_tables(c);
    ensure_current_tables(c);
    return ADSCF_OK;
  });
}

int adscf_batch_create(adscf_batch** out, int ninst, const int* grid, const double* box, const double* ds,
                       const int* nblocks, const int* monomer_type, const int* block_segments, const double* kappa,
                       const double* chi, const double* vol, int ensemble, double rho, int mixing_dim,
                       double lam_mix, int precision) {
  if (!out || ninst < 1 || ninst > kMaxBatchInstances || !grid || !box || !ds || !nblocks || !monomer_type ||
      !block_segments || !kappa || !chi || !vol)
    return ADSCF_ERR_ARG;
  *out = nullptr;
  auto* b = new (std::nothrow) adscf_batch();
  if (!b) return ADSCF_ERR_HIP;
  b->prec = precision;
  if (const char* g = getenv("ADSCF_GRAPH")) b->use_graph = strcmp(g, "0") != 0;
  int rc = bguard(b, [&]() -> int {
    HIPCK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (hipStream_t& s : b->side) HIPCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (hipEvent_t& e : b->join) HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCK(hipEventCreate(&b->ev0));
    HIPCK(hipEventCreate(&b->ev1));
    int boff = 0;
    for (int i = 0; i < ninst; ++i) {
      adscf_ctx* c = nullptr;
      int r = adscf_init(&c, grid[3 * i], grid[3 * i + 1], grid[3 * i + 2], 2, box[3 * i], box[3 * i + 1],
                         box[3 * i + 2], ds[i], precision);
      if (r != ADSCF_OK) throw ArgError{"batch instance init failed"};
      (void)hipStreamSynchronize(c->stream);
      (void)hipStreamDestroy(c->stream);
      c->stream = b->stream;  // every instance runs on the batch stream
      c->plane = b->plane_req;
      b->inst.push_back(c);
      bcheck(b, c, adscf_set_interactions(c, kappa[i], chi + 4 * i, vol + 2 * i));
      bcheck(b, c, adscf_scf_setup(c, nblocks[i], monomer_type + boff, block_segments + boff, ensemble, rho, 0.0,
                                   mixing_dim, lam_mix));
      boff += nblocks[i];
      batch_fix_schedule(c);
    }
    return ADSCF_OK;
  });
  if (rc != ADSCF_OK) {
    fprintf(stderr, "adscf_batch_create: %s\n", b->err.c_str());
    adscf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return ADSCF_OK;
}

int adscf_batch_set_persistent(adscf_batch* b, int enable, int* active) {
  return bguard(b, [&]() -> int {
    if (enable != 0 && enable != 1) throw ArgError{"enable must be 0 or 1"};
    b->persist_req = enable != 0;
    batch_drop_graph(b);
    if (active)
      *active = b->persist_req && (b->prec == ADSCF_FP64 ? persist_shapes_ok<double>(b) : persist_shapes_ok<float>(b));
    return ADSCF_OK;
  });
}

int adscf_batch_set_plane_passes(adscf_batch* b, int mode, int* active) {
  return bguard(b, [&]() -> int {
    if (mode < -1 || mode > 1) throw ArgError{"mode must be -1, 0 or 1"};
    const bool fp64 = b->prec == ADSCF_FP64;
    int n_active = 0;
    for (adscf_ctx* c : b->inst) {
      c->plane = mode;
      n_active += fp64 ? plane_sweep<double>(c) : plane_sweep<float>(c);
    }
    if (mode == 1 && n_active != (int)b->inst.size()) {
      for (adscf_ctx* c : b->inst) c->plane = b->plane_req;
      throw ArgError{"plane passes need nx == ny, nz > 1 and a compiled plane shape on every instance"};
    }
    b->plane_req = mode;
    batch_drop_graph(b);
    if (active) *active = n_active;
    return ADSCF_OK;
  });
}

int adscf_batch_set_schedule(adscf_batch* b, int n, int n1, int n2) {
  return bguard(b, [&]() -> int {
    bool known = false;
    for (const RegShape& s : kRegShapes) known = known || (s.n == n && s.N1 == n1 && s.N2 == n2);
    if (!known || !batch_shape_ok(n1, n2)) throw ArgError{"no batched register shape n1 x n2 of this length"};
    int hit = 0;
    for (adscf_ctx* c : b->inst)
      for (int a = 0; a < 3; ++a)
        if (c->ax[a].n == n) {
          c->ax[a].N1 = n1;
          c->ax[a].N2 = n2;
          c->ax[a].waves = 1;
          ++hit;
        }
    if (!hit) throw ArgError{"no instance has an axis of this length"};
    batch_drop_graph(b);
    return ADSCF_OK;
  });
}

int adscf_batch_destroy(adscf_batch* b) {
  if (!b) return ADSCF_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  b->keep.release();
  for (adscf_ctx* c : b->inst) {
    c->stream = nullptr;
    adscf_finalize(c);
  }
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (hipEvent_t e : b->join)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : b->side)
    if (s) (void)hipStreamDestroy(s);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return ADSCF_OK;
}

const char* adscf_batch_last_error(const adscf_batch* b) { return b ? b->err.c_str() : "null batch"; }

int adscf_batch_set_fields(adscf_batch* b, const double* fields) {
  return bguard(b, [&]() -> int {
    if (!fields) throw ArgError{"null fields"};
    size_t off = 0;
    for (adscf_ctx* c : b->inst) {
      bcheck(b, c, adscf_scf_set_field(c, fields + off));
      off += (size_t)c->nt * c->ng;
    }
    batch_drop_graph(b);
    return ADSCF_OK;
  });
}

int adscf_batch_set_precision(adscf_batch* b, int precision) {
  return bguard(b, [&]() -> int {
    for (adscf_ctx* c : b->inst) {
      bcheck(b, c, adscf_set_precision(c, precision));
      batch_fix_schedule(c);
    }
    b->prec = precision;
    batch_drop_graph(b);
    return ADSCF_OK;
  });
}

int adscf_batch_iterate(adscf_batch* b, int niter, double tol, int* iters, double* accuracy, int* converged) {
  return bguard(b, [&]() -> int {
    for (adscf_ctx* c : b->inst)
      if (!c->field_set) throw ArgError{"adscf_batch_set_fields has not been called"};
    if (!b->use_graph) {
      if (b->prec == ADSCF_FP64) batch_build_tables<double>(b);
      else batch_build_tables<float>(b);
    } else if (b->prec == ADSCF_FP64) {
      batch_build_graph<double>(b);
    } else {
      batch_build_graph<float>(b);
    }
    for (adscf_ctx* c : b->inst)
      hipLaunchKernelGGL(k_state_update, dim3(1), dim3(64), 0, b->stream, c->st.p, tol, 1, 0);
    HIPCK(hipGetLastError());
    const int n = (int)b->inst.size();
    std::vector<ScfState> st(n);
    for (int i = 0; i < n; ++i)
      HIPCK(hipMemcpyAsync(&st[i], b->inst[i]->st.p, sizeof(ScfState), hipMemcpyDeviceToHost, b->stream));
    HIPCK(hipStreamSynchronize(b->stream));
    // instances that stopped at convergence in an earlier call continue their loops under the
    // current tolerance from the mixing step of their last evaluation (as adscf_scf_iterate)
    for (int i = 0; i < n && niter > 0; ++i) {
      if (!(st[i].done && st[i].iter > 0)) continue;
      adscf_ctx* c = b->inst[i];
      hipLaunchKernelGGL(k_state_update, dim3(1), dim3(64), 0, b->stream, c->st.p, 0.0, 0, 1);
      HIPCK(hipGetLastError());
      hipStream_t own = c->stream;
      c->stream = b->stream;
      try {
        if (b->prec == ADSCF_FP64) enqueue_mix_tail<double>(c, 0, true);
        else enqueue_mix_tail<float>(c, 0, true);
      } catch (...) {
        c->stream = own;
        throw;
      }
      c->stream = own;
    }
    HIPCK(hipStreamSynchronize(b->stream));
    int done_iters = 0;
    float total = 0.f;
    while (done_iters < niter) {
      const int chunk = std::min(niter - done_iters, tol > 0 ? 25 : niter);
      HIPCK(hipEventRecord(b->ev0, b->stream));
      for (int i = 0; i < chunk; ++i) {
        if (b->use_graph) {
          HIPCK(hipGraphLaunch(b->graph, b->stream));
        } else if (b->prec == ADSCF_FP64) {
          batch_enqueue_lanes<double>(b, true);
        } else {
          batch_enqueue_lanes<float>(b, true);
        }
      }
      HIPCK(hipEventRecord(b->ev1, b->stream));
      for (int i = 0; i < n; ++i)
        HIPCK(hipMemcpyAsync(&st[i], b->inst[i]->st.p, sizeof(ScfState), hipMemcpyDeviceToHost, b->stream));
      HIPCK(hipStreamSynchronize(b->stream));
      batch_check_persist(b);
      float ms = 0.f;
      HIPCK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      total += ms;
      done_iters += chunk;
      bool all = true;
      for (auto& s : st) all = all && s.done;
      if (all) break;
    }
    b->last_ms_iter = done_iters ? total / done_iters : 0.0;
    for (int i = 0; i < n; ++i) {
      b->inst[i]->phi_fresh = st[i].done != 0;
      if (iters) iters[i] = st[i].iter;
      if (accuracy) accuracy[i] = st[i].accuracy;
      if (converged) converged[i] = st[i].done;
    }
    return ADSCF_OK;
  });
}

int adscf_batch_energy(adscf_batch* b, double* F_per_chain) {
  return bguard(b, [&]() -> int {
    for (size_t i = 0; i < b->inst.size(); ++i) {
      double fc = 0.0;
      bcheck(b, b->inst[i], adscf_scf_energy(b->inst[i], nullptr, &fc, nullptr, nullptr, nullptr));
      if (F_per_chain) F_per_chain[i] = fc;
    }
    return ADSCF_OK;
  });
}

int adscf_batch_get(adscf_batch* b, int i, double* field, double* density) {
  return bguard(b, [&]() -> int {
    if (i < 0 || i >= (int)b->inst.size()) throw ArgError{"instance index out of range"};
    return bcheck(b, b->inst[i], adscf_scf_get(b->inst[i], field, density));
  });
}

int adscf_batch_timing(adscf_batch* b, double* ms_per_iter, double* ms_sweep) {
  return bguard(b, [&]() -> int {
    if (ms_per_iter) *ms_per_iter = b->last_ms_iter;
    if (ms_sweep) *ms_sweep = b->prec == ADSCF_FP64 ? batch_sweep_timing<double>(b, 3) : batch_sweep_timing<float>(b, 3);
    return ADSCF_OK;
  });
}

// Autotuner for the batch: for every distinct axis length, try each batched register
// shape on all instances together and keep the fastest batched sweep (the shape in place
// unless another is at least 2% faster).
int adscf_batch_autotune(adscf_batch* b, int reps, char* report, int report_cap) {
  return bguard(b, [&]() -> int {
    std::string rep;
    char line[256];
    std::vector<int> lengths;
    for (adscf_ctx* c : b->inst)
      for (int a = 0; a < 3; ++a)
        if (c->ax[a].n > 1 && std::find(lengths.begin(), lengths.end(), c->ax[a].n) == lengths.end())
          lengths.push_back(c->ax[a].n);
    for (int n : lengths) {
      // the shape in place (the default, or an earlier choice) is timed first and a
      // candidate replaces it only when its sweep is at least 2% faster: the shapes of one
      // length change only part of a batched sweep, and timings of equal schedules differ
      // by about 1% from run to run
      std::pair<int, int> cur = {0, 0};
      for (adscf_ctx* c : b->inst)
        for (int a = 0; a < 3; ++a)
          if (c->ax[a].n == n && cur.first == 0) cur = {c->ax[a].N1, c->ax[a].N2};
      std::vector<std::pair<int, int>> cands;
      if (batch_shape_ok(cur.first, cur.second)) cands.push_back(cur);
      for (const RegShape& s : kRegShapes)
        if (s.n == n && batch_shape_ok(s.N1, s.N2) && std::make_pair(s.N1, s.N2) != cur) cands.push_back({s.N1, s.N2});
      double best = 1e300;
      std::pair<int, int> bestsh = cands.empty() ? std::make_pair(0, 0) : cands[0];
      for (auto sh : cands) {
        for (adscf_ctx* c : b->inst)
          for (int a = 0; a < 3; ++a)
            if (c->ax[a].n == n) {
              c->ax[a].N1 = sh.first;
              c->ax[a].N2 = sh.second;
              c->ax[a].waves = 1;
            }
        const double ms = b->prec == ADSCF_FP64 ? batch_sweep_timing<double>(b, std::max(reps, 1))
                                                : batch_sweep_timing<float>(b, std::max(reps, 1));
        snprintf(line, sizeof line, "n=%d shape=%s batched sweep=%.4f ms\n", n, shape_name(sh.first, sh.second).c_str(),
                 ms);
        rep += line;
        if (ms < (best < 1e300 ? 0.98 * best : best)) {
          best = ms;
          bestsh = sh;
        }
      }
      for (adscf_ctx* c : b->inst)
        for (int a = 0; a < 3; ++a)
          if (c->ax[a].n == n) {
            c->ax[a].N1 = bestsh.first;
            c->ax[a].N2 = bestsh.second;
          }
      snprintf(line, sizeof line, "n=%d chosen %s (%.4f ms)\n", n, shape_name(bestsh.first, bestsh.second).c_str(),
               best);
      rep += line;
    }
    batch_drop_graph(b);
    if (report && report_cap > 0) {
      const size_t k = std::min(rep.size(), (size_t)report_cap - 1);
      memcpy(report, rep.data(), k);
      report[k] = 0;
    }
    return ADSCF_OK;
  });
}

int adscf_hbm_bandwidth(adscf_ctx* c, long bytes, int reps, double* copy_GBs, double* read_GBs) {
  return guarded(c, [&]() -> int {
    if (bytes < (1L << 20) || reps < 1) throw ArgError{"adscf_hbm_bandwidth: bytes >= 1 MiB and reps >= 1"};
    const long n = bytes / (long)sizeof(double);
    DevBuf<double> a, b, part;
    a.ensure(n);
    b.ensure(n);
    part.ensure(kRedBlocks * 4);
    hipLaunchKernelGGL(k_fill<double>, dim3(kRedBlocks * 4), dim3(kThreads), 0, c->stream, a.p, 1.0, n);
    HIPCK(hipGetLastError());
    auto median_ms = [&](const std::function<void()>& f) {
      f();
      std::vector<float> ms(reps);
      for (int r = 0; r < reps; ++r) {
        HIPCK(hipEventRecord(c->ev0, c->stream));
        f();
        HIPCK(hipEventRecord(c->ev1, c->stream));
        HIPCK(hipEventSynchronize(c->ev1));
        HIPCK(hipEventElapsedTime(&ms[r], c->ev0, c->ev1));
      }
      std::sort(ms.begin(), ms.end());
      return (double)ms[reps / 2];
    };
    const double nb = (double)n * sizeof(double);
    const double tc = median_ms([&] {
      HIPCK(hipMemcpyAsync(b.p, a.p, (size_t)nb, hipMemcpyDeviceToDevice, c->stream));
    });
    const double tr = median_ms([&] {
      hipLaunchKernelGGL(k_read_sum, dim3(kRedBlocks * 4), dim3(kThreads), 0, c->stream, (const double*)a.p, n,
                         part.p);
      HIPCK(hipGetLastError());
    });
    HIPCK(hipStreamSynchronize(c->stream));
    if (copy_GBs) *copy_GBs = 2.0 * nb / (tc * 1e-3) / 1e9;
    if (read_GBs) *read_GBs = nb / (tr * 1e-3) / 1e9;
    return ADSCF_OK;
  });
}

int adscf_hipfft_step_time(adscf_ctx* c, int reps, double* us_per_step) {
  return guarded(c, [&]() -> int {
    if (!us_per_step || reps < 1) throw ArgError{"bad arguments"};
    *us_per_step = c->prec == ADSCF_FP64 ? hipfft_step_us<double>(c, reps) : hipfft_step_us<float>(c, reps);
    return ADSCF_OK;
  });
}

}  // extern "C"
