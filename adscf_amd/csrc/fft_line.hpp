// Line FFT engine: mixed-radix Stockham passes over a tile of L lines held in
// LDS (device) or plain memory (host emulation, used by the CPU unit test).
//
// A tile holds L complex lines of length n with leading dimension ldn.  The
// forward transform uses tw[t] = exp(-2*pi*i*t/n); the inverse uses conj(tw).
// Stockham auto-sort keeps the output in natural bin order, so the Fourier
// multiplier is applied by bin index k directly.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADSCF_HD __host__ __device__ __forceinline__
#else
#define ADSCF_HD inline
#endif

namespace adscf {

// Aligned to its own size, so that one complex value moves with one 8-byte (FP32) or 16-byte
// (FP64) LDS or global access whose offset folds into the instruction.
template <typename T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};

template <typename T>
ADSCF_HD C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
ADSCF_HD C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
ADSCF_HD C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * conj(b)
template <typename T>
ADSCF_HD C2<T> cmulc(C2<T> a, C2<T> b) { return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }
// dir * i * a  (dir = -1 forward, +1 inverse)
template <typename T>
ADSCF_HD C2<T> crot(C2<T> a, int dir) { return dir < 0 ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x}; }
template <typename T>
ADSCF_HD C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }

template <typename T>
struct KC;
template <>
struct KC<double> {
  static constexpr double s2 = 0.70710678118654752440;   // sqrt(1/2)
  static constexpr double s3 = 0.86602540378443864676;   // sqrt(3)/2
  static constexpr double c5a = 0.30901699437494742410;  // cos(2pi/5)
  static constexpr double s5a = 0.95105651629515357212;  // sin(2pi/5)
  static constexpr double c5b = -0.80901699437494742410; // cos(4pi/5)
  static constexpr double s5b = 0.58778525229247312917;  // sin(4pi/5)
};
template <>
struct KC<float> {
  static constexpr float s2 = 0.70710678118654752440f;
  static constexpr float s3 = 0.86602540378443864676f;
  static constexpr float c5a = 0.30901699437494742410f;
  static constexpr float s5a = 0.95105651629515357212f;
  static constexpr float c5b = -0.80901699437494742410f;
  static constexpr float s5b = 0.58778525229247312917f;
};

// In-register R-point DFTs, X_k = sum_j v_j exp(dir*2*pi*i*j*k/R).
template <typename T>
ADSCF_HD void dft2(C2<T>* v) {
  C2<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T>
ADSCF_HD void dft3(C2<T>* v, int dir) {
  C2<T> t = cadd(v[1], v[2]);
  C2<T> u = csub(v[1], v[2]);
  C2<T> a = v[0];
  C2<T> m = {a.x - T(0.5) * t.x, a.y - T(0.5) * t.y};
  C2<T> r = crot(cscale(u, KC<T>::s3), dir);
  v[0] = cadd(a, t);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <typename T>
ADSCF_HD void dft4(C2<T>* v, int dir) {
  C2<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  C2<T> t2 = cadd(v[1], v[3]), t3 = crot(csub(v[1], v[3]), dir);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <typename T>
ADSCF_HD void dft5(C2<T>* v, int dir) {
  C2<T> a = v[0];
  C2<T> s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
  C2<T> s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
  C2<T> m1 = {a.x + KC<T>::c5a * s1.x + KC<T>::c5b * s2.x, a.y + KC<T>::c5a * s1.y + KC<T>::c5b * s2.y};
  C2<T> m2 = {a.x + KC<T>::c5b * s1.x + KC<T>::c5a * s2.x, a.y + KC<T>::c5b * s1.y + KC<T>::c5a * s2.y};
  C2<T> n1 = {KC<T>::s5a * d1.x + KC<T>::s5b * d2.x, KC<T>::s5a * d1.y + KC<T>::s5b * d2.y};
  C2<T> n2 = {KC<T>::s5b * d1.x - KC<T>::s5a * d2.x, KC<T>::s5b * d1.y - KC<T>::s5a * d2.y};
  C2<T> r1 = crot(n1, dir), r2 = crot(n2, dir);
  v[0] = cadd(a, cadd(s1, s2));
  v[1] = cadd(m1, r1);
  v[4] = csub(m1, r1);
  v[2] = cadd(m2, r2);
  v[3] = csub(m2, r2);
}

template <typename T>
ADSCF_HD void dft8(C2<T>* v, int dir) {
  // two radix-4 DFTs over even/odd samples, then radix-2 with W8^k
  C2<T> e[4] = {v[0], v[2], v[4], v[6]};
  C2<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e, dir);
  dft4(o, dir);
  const T h = KC<T>::s2;
  // W8^1 = h*(1 + dir*i), W8^2 = dir*i, W8^3 = h*(-1 + dir*i)
  C2<T> o1 = {h * (o[1].x - dir * o[1].y), h * (o[1].y + dir * o[1].x)};
  C2<T> o2 = crot(o[2], dir);
  C2<T> o3 = {h * (-o[3].x - dir * o[3].y), h * (-o[3].y + dir * o[3].x)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// Twiddle tw^(dir) at index t: forward table entry or its conjugate.
template <typename T>
ADSCF_HD C2<T> twid(const C2<T>* tw, int t, int dir) {
  C2<T> w = tw[t];
  return dir < 0 ? w : C2<T>{w.x, -w.y};
}

// One Stockham pass of radix R (2,3,4,5,8) over L lines: Ns is the product of
// the radices already applied.  Butterfly b = (line, j), j < n/R.
template <typename T, int R>
ADSCF_HD void stockham_pass(const C2<T>* src, C2<T>* dst, int n, int ldn, int L, int Ns, const C2<T>* tw, int dir,
                            int tid, int nthr) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  const int total = L * nb;
  for (int b = tid; b < total; b += nthr) {
    const int line = b / nb;
    const int j = b - line * nb;
    const C2<T>* s = src + line * ldn;
    C2<T>* d = dst + line * ldn;
    const int k = j % Ns;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid(tw, r * k * tstep, dir));
    }
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v, dir);
    else if constexpr (R == 4) dft4(v, dir);
    else if constexpr (R == 5) dft5(v, dir);
    else dft8(v, dir);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[base + r * Ns] = v[r];
  }
}

// Generic radix R (an odd prime above 5): one output per work item.  Inputs r and R - r
// meet conjugate kernel factors, w^(rq) and conj(w^(rq)), so each output folds them into
// one pair: a w + b conj(w) = (a + b) Re w + i (a - b) Im w, (R - 1)/2 steps of four
// multiply-adds.  The stage twiddles of the inputs (Ns > 1) advance by addition; (r q) mod R
// wraps once per step.  Leftover primes come first in the plan, where Ns = 1 and the inputs
// carry no stage twiddle.
template <typename T>
ADSCF_HD void stockham_pass_generic(const C2<T>* src, C2<T>* dst, int n, int ldn, int L, int Ns, int R, const C2<T>* tw,
                                    int dir, int tid, int nthr) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  const int rstep = n / R;
  const int total = L * n;
  const int half = (R - 1) / 2;
  for (int o = tid; o < total; o += nthr) {
    const int line = o / n;
    const int idx = o - line * n;
    const int k = idx % Ns;
    const int q = (idx / Ns) % R;
    const int j = (idx / (Ns * R)) * Ns + k;
    const C2<T>* s = src + line * ldn;
    const int step = k * tstep;  // stage twiddle index step per input
    T ax = s[j].x, ay = s[j].y;
    int rq = 0;
    auto fold = [&](C2<T> u, C2<T> v, C2<T> w) {
      const T sx = u.x + v.x, sy = u.y + v.y, dx = u.x - v.x, dy = u.y - v.y;
      ax += sx * w.x - dy * w.y;
      ay += sy * w.x + dx * w.y;
    };
    int r = 1;
    if (Ns == 1) {
      // groups of four steps: the twelve LDS reads of a group are issued before its folds
      for (; r + 3 <= half; r += 4) {
        C2<T> u[4], v[4], w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          rq += q;
          if (rq >= R) rq -= R;
          u[t] = s[j + (r + t) * nb];
          v[t] = s[j + (R - r - t) * nb];
          w[t] = twid(tw, rq * rstep, dir);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) fold(u[t], v[t], w[t]);
      }
    }
    for (; r <= half; ++r) {
      rq += q;
      if (rq >= R) rq -= R;
      C2<T> u = s[j + r * nb], v = s[j + (R - r) * nb];
      if (Ns > 1) {
        u = cmul(u, twid(tw, r * step, dir));
        v = cmul(v, twid(tw, (R - r) * step, dir));
      }
      fold(u, v, twid(tw, rq * rstep, dir));
    }
    dst[line * ldn + idx] = C2<T>{ax, ay};
  }
}

constexpr int kMaxPasses = 12;

struct LinePlan {
  int n;                  // line length (> 1)
  int npass;              // number of Stockham passes
  int radix[kMaxPasses];  // radix of each pass, product = n
};

// Factor n into radices: leftover primes above 5 first (generic passes, which then run
// without stage twiddles), then 8s, 4, 2, 3, 5.
inline LinePlan make_line_plan(int n) {
  LinePlan p{};
  p.n = n;
  int m = n;
  int np = 0;
  {
    int r = m;
    for (int f : {2, 3, 5})
      while (r % f == 0) r /= f;
    for (int f = 7; r > 1 && np < kMaxPasses; f += 2) {
      while (r % f == 0 && np < kMaxPasses) { p.radix[np++] = f; r /= f; m /= f; }
    }
  }
  while (m % 8 == 0 && np < kMaxPasses) { p.radix[np++] = 8; m /= 8; }
  while (m % 4 == 0 && np < kMaxPasses) { p.radix[np++] = 4; m /= 4; }
  while (m % 2 == 0 && np < kMaxPasses) { p.radix[np++] = 2; m /= 2; }
  while (m % 3 == 0 && np < kMaxPasses) { p.radix[np++] = 3; m /= 3; }
  while (m % 5 == 0 && np < kMaxPasses) { p.radix[np++] = 5; m /= 5; }
  p.npass = (m == 1) ? np : -1;  // -1: too many factors
  return p;
}

// Run the passes of one transform direction over the tile.  src/dst ping-pong;
// returns the buffer holding the result.  `sync` is called between passes.
template <typename T, typename Sync>
ADSCF_HD C2<T>* line_fft(C2<T>* b0, C2<T>* b1, const LinePlan& pl, int ldn, int L, const C2<T>* tw, int dir, int tid,
                         int nthr, Sync sync) {
  C2<T>* src = b0;
  C2<T>* dst = b1;
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    switch (R) {
      case 2: stockham_pass<T, 2>(src, dst, pl.n, ldn, L, Ns, tw, dir, tid, nthr); break;
      case 3: stockham_pass<T, 3>(src, dst, pl.n, ldn, L, Ns, tw, dir, tid, nthr); break;
      case 4: stockham_pass<T, 4>(src, dst, pl.n, ldn, L, Ns, tw, dir, tid, nthr); break;
      case 5: stockham_pass<T, 5>(src, dst, pl.n, ldn, L, Ns, tw, dir, tid, nthr); break;
      case 8: stockham_pass<T, 8>(src, dst, pl.n, ldn, L, Ns, tw, dir, tid, nthr); break;
      default: stockham_pass_generic<T>(src, dst, pl.n, ldn, L, Ns, R, tw, dir, tid, nthr); break;
    }
    sync();
    C2<T>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

}  // namespace adscf
