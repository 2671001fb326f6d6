// Host emulation of the device line-FFT engine (adscf_amd/csrc/fft_line.hpp):
// runs the same Stockham passes sequentially and compares forward and inverse
// transforms of L random lines against a direct O(n^2) DFT in long double.
// Prints "n maxerr" per size; exits nonzero when any error exceeds the bound.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../adscf_amd/csrc/fft_line.hpp"

using namespace adscf;

template <typename T>
static double check(int n, int L, double tol) {
  LinePlan pl = make_line_plan(n);
  if (pl.npass < 0) return 1e9;
  const int ldn = n + 1;
  std::vector<C2<T>> tw(n), b0(L * ldn), b1(L * ldn);
  for (int t = 0; t < n; ++t) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * t / n;
    tw[t] = {(T)cosl(a), (T)sinl(a)};
  }
  std::vector<long double> re(L * n), im(L * n);
  srand(n * 7 + L);
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < n; ++j) {
      re[l * n + j] = (long double)rand() / RAND_MAX - 0.5L;
      im[l * n + j] = (long double)rand() / RAND_MAX - 0.5L;
      b0[l * ldn + j] = {(T)re[l * n + j], (T)im[l * n + j]};
    }
  double worst = 0;
  for (int dir : {-1, +1}) {
    std::vector<C2<T>> a0 = b0, a1 = b1;
    int nthr = 7;  // odd thread count exercises the strided loops
    C2<T>* out = nullptr;
    // emulate threads: every pass runs all "threads" before the next pass (the sync)
    {
      C2<T>* src = a0.data();
      C2<T>* dst = a1.data();
      int Ns = 1;
      for (int p = 0; p < pl.npass; ++p) {
        int R = pl.radix[p];
        for (int tid = 0; tid < nthr; ++tid) {
          switch (R) {
            case 2: stockham_pass<T, 2>(src, dst, n, ldn, L, Ns, tw.data(), dir, tid, nthr); break;
            case 3: stockham_pass<T, 3>(src, dst, n, ldn, L, Ns, tw.data(), dir, tid, nthr); break;
            case 4: stockham_pass<T, 4>(src, dst, n, ldn, L, Ns, tw.data(), dir, tid, nthr); break;
            case 5: stockham_pass<T, 5>(src, dst, n, ldn, L, Ns, tw.data(), dir, tid, nthr); break;
            case 8: stockham_pass<T, 8>(src, dst, n, ldn, L, Ns, tw.data(), dir, tid, nthr); break;
            default: stockham_pass_generic<T>(src, dst, n, ldn, L, Ns, R, tw.data(), dir, tid, nthr); break;
          }
        }
        std::swap(src, dst);
        Ns *= R;
      }
      out = src;
    }
    // also the convenience driver must agree bit-for-bit
    std::vector<C2<T>> c0 = b0, c1 = b1;
    C2<T>* out2 = line_fft<T>(c0.data(), c1.data(), pl, ldn, L, tw.data(), dir, 0, 1, [] {});
    for (int l = 0; l < L; ++l)
      for (int k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
          long double a = dir * 2.0L * 3.14159265358979323846264338327950288L * ((long)j * k % n) / n;
          sr += re[l * n + j] * cosl(a) - im[l * n + j] * sinl(a);
          si += re[l * n + j] * sinl(a) + im[l * n + j] * cosl(a);
        }
        C2<T> g = out[l * ldn + k];
        C2<T> g2 = out2[l * ldn + k];
        if (g.x != g2.x || g.y != g2.y) return 1e9;
        double e = std::fabs((double)(g.x - sr)) + std::fabs((double)(g.y - si));
        if (e > worst) worst = e;
      }
  }
  return worst / std::sqrt((double)n);
}

int main() {
  // 49, 77, 91, 121, 143, 169, 221: two generic-radix stages in a row (stage twiddles on a
  // generic pass that follows another)
  int sizes[] = {2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 15, 16, 17, 24, 25, 27, 30, 32, 34, 36, 38, 40, 43, 48, 49, 57,
                 60, 64, 75, 77, 86, 91, 96, 97, 100, 121, 128, 143, 169, 221, 256, 300};
  int bad = 0;
  for (int n : sizes) {
    double e64 = check<double>(n, 3, 0);
    double e32 = check<float>(n, 2, 0);
    int ok = e64 < 1e-14 && e32 < 2e-6;
    printf("%d %.3e %.3e %s\n", n, e64, e32, ok ? "ok" : "BAD");
    bad |= !ok;
  }
  return bad;
}
