// Host emulation of the R^3 register line operator (adscf_amd/csrc/fft_cube.hpp): the
// R^2 "threads" of a line run each phase in turn (the barrier is a no-op between
// phases run to completion), and L x = IDFT(K * DFT(x)) (unnormalised inverse) is
// compared with a direct long-double evaluation.  Exits nonzero on any error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../adscf_amd/csrc/fft_cube.hpp"

using namespace adscf;

static const long double PI = 3.14159265358979323846264338327950288L;

template <typename T, int R>
static double check() {
  using CB = Cube<R>;
  const int n = CB::n;
  std::vector<C2<T>> tw(n), x(n);
  std::vector<T> kf(n);
  srand(R * 977);
  for (int t = 0; t < n; ++t) {
    long double a = -2 * PI * t / n;
    tw[t] = {(T)cosl(a), (T)sinl(a)};
    x[t] = {(T)(rand() / (double)RAND_MAX - 0.5), (T)(rand() / (double)RAND_MAX - 0.5)};
    kf[t] = (T)(rand() / (double)RAND_MAX);
  }
  // reference
  std::vector<long double> fr(n), fi(n), re(n, 0), im(n, 0);
  for (int k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    for (int j = 0; j < n; ++j) {
      long double a = -2 * PI * ((long)j * k % n) / n;
      sr += x[j].x * cosl(a) - x[j].y * sinl(a);
      si += x[j].x * sinl(a) + x[j].y * cosl(a);
    }
    fr[k] = sr * kf[k];
    fi[k] = si * kf[k];
  }
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      long double a = 2 * PI * ((long)j * k % n) / n;
      re[j] += fr[k] * cosl(a) - fi[k] * sinl(a);
      im[j] += fr[k] * sinl(a) + fi[k] * cosl(a);
    }
  // emulation: every phase runs for all R^2 threads before the next (the device
  // puts a block barrier between phases)
  std::vector<C2<T>> lds(CB::lds_elems);
  const int nt = CB::tpl;
  std::vector<std::vector<C2<T>>> out(nt, std::vector<C2<T>>(R));
  for (int t = 0; t < nt; ++t)
    for (int m = 0; m < R; ++m) out[t][m] = x[t + nt * m];
  std::vector<CubeTables<T, R>> W(nt);
  for (int t = 0; t < nt; ++t) W[t].load(t, tw.data(), kf.data());
  for (int t = 0; t < nt; ++t) cube_a<T, R>(out[t].data(), lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) cube_b<T, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) cube_c<T, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) cube_bi<T, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) cube_ai<T, R>(out[t].data(), lds.data(), t, W[t]);
  double worst = 0;
  for (int t = 0; t < nt; ++t)
    for (int m = 0; m < R; ++m) {
      const int j = t + nt * m;
      worst = std::max(worst, std::fabs((double)(out[t][m].x - re[j])) + std::fabs((double)(out[t][m].y - im[j])));
    }
  return worst / n;
}

// the same check for the A R^2 engine (tri_* of fft_cube.hpp)
template <typename T, int A, int R>
static double check_tri() {
  using TR = Tri<A, R>;
  const int n = TR::n;
  std::vector<C2<T>> tw(n), x(n);
  std::vector<T> kf(n);
  srand(A * 131 + R * 977);
  for (int t = 0; t < n; ++t) {
    long double a = -2 * PI * t / n;
    tw[t] = {(T)cosl(a), (T)sinl(a)};
    x[t] = {(T)(rand() / (double)RAND_MAX - 0.5), (T)(rand() / (double)RAND_MAX - 0.5)};
    kf[t] = (T)(rand() / (double)RAND_MAX);
  }
  std::vector<long double> fr(n), fi(n), re(n, 0), im(n, 0);
  for (int k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    for (int j = 0; j < n; ++j) {
      long double a = -2 * PI * ((long)j * k % n) / n;
      sr += x[j].x * cosl(a) - x[j].y * sinl(a);
      si += x[j].x * sinl(a) + x[j].y * cosl(a);
    }
    fr[k] = sr * kf[k];
    fi[k] = si * kf[k];
  }
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      long double a = 2 * PI * ((long)j * k % n) / n;
      re[j] += fr[k] * cosl(a) - fi[k] * sinl(a);
      im[j] += fr[k] * sinl(a) + fi[k] * cosl(a);
    }
  std::vector<C2<T>> lds(TR::lds_elems);
  const int nt = TR::tpl;
  std::vector<std::vector<C2<T>>> out(nt, std::vector<C2<T>>(A));
  for (int t = 0; t < nt; ++t)
    for (int m = 0; m < A; ++m) out[t][m] = x[t + nt * m];
  std::vector<TriTables<T, A, R>> W(nt);
  for (int t = 0; t < nt; ++t) W[t].load(t, tw.data(), kf.data());
  for (int t = 0; t < nt; ++t) tri_a<T, A, R>(out[t].data(), lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) tri_b<T, A, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) tri_c<T, A, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) tri_bi<T, A, R>(lds.data(), t, W[t]);
  for (int t = 0; t < nt; ++t) tri_ai<T, A, R>(out[t].data(), lds.data(), t, W[t]);
  double worst = 0;
  for (int t = 0; t < nt; ++t)
    for (int m = 0; m < A; ++m) {
      const int j = t + nt * m;
      worst = std::max(worst, std::fabs((double)(out[t][m].x - re[j])) + std::fabs((double)(out[t][m].y - im[j])));
    }
  return worst / n;
}

int main() {
  int bad = 0;
#define CC(r)                                                               \
  {                                                                         \
    double e64 = check<double, r>(), e32 = check<float, r>();               \
    bool ok = e64 < 1e-15 && e32 < 3e-7;                                    \
    printf("cube %d^3 n=%d %.3e %.3e %s\n", r, r * r * r, e64, e32, ok ? "ok" : "BAD"); \
    bad |= !ok;                                                             \
  }
  CC(2) CC(3) CC(4) CC(5) CC(8)
#define CT(a, r)                                                                         \
  {                                                                                      \
    double e64 = check_tri<double, a, r>(), e32 = check_tri<float, a, r>();              \
    bool ok = e64 < 1e-15 && e32 < 3e-7;                                                 \
    printf("tri %dx%dx%d n=%d %.3e %.3e %s\n", a, r, r, a * r * r, e64, e32, ok ? "ok" : "BAD"); \
    bad |= !ok;                                                                          \
  }
  CT(2, 2) CT(4, 2) CT(8, 2) CT(3, 3) CT(6, 3) CT(4, 4) CT(8, 4) CT(16, 4) CT(5, 5) CT(8, 8)
  return bad;
}
