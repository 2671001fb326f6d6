// Host emulation of the register-resident line operator (adscf_amd/csrc/fft_reg.hpp):
// runs phases A, B, C for every "thread" in order and compares L x = IDFT(K * DFT(x))
// (unnormalised inverse) with a direct long-double evaluation, for each four-step shape
// and each single-thread length.  Exits nonzero on any error above the bound.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../adscf_amd/csrc/fft_reg.hpp"

using namespace adscf;

static const long double PI = 3.14159265358979323846264338327950288L;

template <typename T>
static void reference(int n, const std::vector<C2<T>>& x, const std::vector<T>& kf, std::vector<long double>& re,
                      std::vector<long double>& im) {
  std::vector<long double> fr(n), fi(n);
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
  re.assign(n, 0);
  im.assign(n, 0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      long double a = 2 * PI * ((long)j * k % n) / n;
      re[j] += fr[k] * cosl(a) - fi[k] * sinl(a);
      im[j] += fr[k] * sinl(a) + fi[k] * cosl(a);
    }
}

template <typename T, int N1, int N2>
static double check4() {
  using FS = FourStep<N1, N2>;
  const int n = FS::n;
  std::vector<C2<T>> tw(n), x(n);
  std::vector<T> kf(n);
  srand(N1 * 131 + N2);
  for (int t = 0; t < n; ++t) {
    long double a = -2 * PI * t / n;
    tw[t] = {(T)cosl(a), (T)sinl(a)};
    x[t] = {(T)(rand() / (double)RAND_MAX - 0.5), (T)(rand() / (double)RAND_MAX - 0.5)};
    kf[t] = (T)(rand() / (double)RAND_MAX);
  }
  std::vector<long double> re, im;
  reference<T>(n, x, kf, re, im);
  std::vector<C2<T>> lds(FS::lds_elems);
  C2<T> v[N1][N2];
  for (int t = 0; t < N1; ++t)
    for (int j2 = 0; j2 < N2; ++j2) v[t][j2] = x[t + N1 * j2];
  const int tpl = N1 > N2 ? N1 : N2;
  std::vector<RegTables<T, N1, N2>> R(tpl);
  for (int t = 0; t < tpl; ++t) R[t].load(t, tw.data(), kf.data());
  for (int t = 0; t < N1; ++t) fs_phase_a<T, N1, N2>(v[t], lds.data(), t, R[t]);
  for (int t = 0; t < N2; ++t) fs_phase_b<T, N1, N2>(lds.data(), t, R[t]);
  for (int t = 0; t < N1; ++t) fs_phase_c<T, N1, N2>(v[t], lds.data(), t);
  double worst = 0;
  for (int t = 0; t < N1; ++t)
    for (int j2 = 0; j2 < N2; ++j2) {
      int j = t + N1 * j2;
      double e = std::fabs((double)(v[t][j2].x - re[j])) + std::fabs((double)(v[t][j2].y - im[j]));
      worst = std::max(worst, e);
    }
  return worst / n;
}

template <typename T, int N>
static double check1() {
  std::vector<C2<T>> x(N);
  std::vector<T> kf(N);
  srand(N * 17);
  for (int t = 0; t < N; ++t) {
    x[t] = {(T)(rand() / (double)RAND_MAX - 0.5), (T)(rand() / (double)RAND_MAX - 0.5)};
    kf[t] = (T)(rand() / (double)RAND_MAX);
  }
  std::vector<long double> re, im;
  reference<T>(N, x, kf, re, im);
  C2<T> v[N];
  for (int j = 0; j < N; ++j) v[j] = x[j];
  RegTables<T, 1, N> R;
  R.load(0, nullptr, kf.data());
  line_op_single<T, N>(v, R);
  double worst = 0;
  for (int j = 0; j < N; ++j)
    worst = std::max(worst, std::fabs((double)(v[j].x - re[j])) + std::fabs((double)(v[j].y - im[j])));
  return worst / N;
}

#define C4(a, b)                                                                       \
  {                                                                                    \
    double e64 = check4<double, a, b>(), e32 = check4<float, a, b>();                  \
    bool ok = e64 < 1e-15 && e32 < 3e-7;                                               \
    printf("four-step %dx%d n=%d %.3e %.3e %s\n", a, b, a * b, e64, e32, ok ? "ok" : "BAD"); \
    bad |= !ok;                                                                        \
  }
#define C1(a)                                                                \
  {                                                                          \
    double e64 = check1<double, a>(), e32 = check1<float, a>();              \
    bool ok = e64 < 1e-15 && e32 < 3e-7;                                     \
    printf("single n=%d %.3e %.3e %s\n", a, e64, e32, ok ? "ok" : "BAD");    \
    bad |= !ok;                                                              \
  }

int main() {
  int bad = 0;
  C1(2) C1(3) C1(4) C1(5) C1(6) C1(8) C1(16)
  C4(4, 8) C4(8, 4) C4(8, 6) C4(6, 8) C4(8, 8) C4(8, 16) C4(16, 8) C4(16, 16) C4(2, 3) C4(4, 4) C4(8, 5)
  C4(2, 5) C4(4, 3) C4(4, 5) C4(8, 3) C4(4, 6) C4(2, 16) C4(6, 6) C4(16, 3) C4(4, 16) C4(16, 4) C4(16, 5) C4(16, 6)
  return bad;
}
