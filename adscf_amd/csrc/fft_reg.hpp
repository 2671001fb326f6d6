// Register-resident line operator L = F^-1 diag(K) F for lines of length
// n = N1 * N2 (four-step decomposition, index j = j1 + N1*j2, bin
// k = k2 + N2*k1):
//   phase A  thread j1 holds x[j1 + N1*j2] (j2 < N2): DFT_N2 forward, twiddle
//            W_n^(j1*k2), write the N2 results to LDS
//   phase B  thread k2 reads Y[j1][k2] (j1 < N1): DFT_N1 forward -> bins
//            k2 + N2*k1, multiply by K, DFT_N1 inverse, conj twiddle, write back
//   phase C  thread j1 reads its N2 values: DFT_N2 inverse -> x[j1 + N1*j2]
// Two LDS transposes per operator; every DFT runs in registers.  With N1 == 1
// the whole line lives in one thread and no LDS is used.
#pragma once

#include "fft_line.hpp"

namespace adscf {

// cos(pi*m/8), sin(pi*m/8), m = 0..15
template <typename T>
struct W16 {
  static ADSCF_HD T c(int m) {
    constexpr double tc[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                               0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
    return (T)tc[m & 15];
  }
  static ADSCF_HD T s(int m) { return c(m + 12); }  // sin(x) = cos(x - pi/2) = cos(pi*(m-4)/8)
};

// exp(DIR * i * pi * m / 8)
template <typename T, int DIR>
ADSCF_HD C2<T> w16(int m) {
  return {W16<T>::c(m), (T)DIR * W16<T>::s(m)};
}

template <typename T, int N, int DIR>
ADSCF_HD void dft_reg(C2<T>* v);

template <typename T, int N, int DIR>
struct DftReg;

template <typename T, int DIR>
struct DftReg<T, 1, DIR> {
  static ADSCF_HD void run(C2<T>*) {}
};
template <typename T, int DIR>
struct DftReg<T, 2, DIR> {
  static ADSCF_HD void run(C2<T>* v) { dft2(v); }
};
template <typename T, int DIR>
struct DftReg<T, 3, DIR> {
  static ADSCF_HD void run(C2<T>* v) { dft3(v, DIR); }
};
template <typename T, int DIR>
struct DftReg<T, 4, DIR> {
  static ADSCF_HD void run(C2<T>* v) { dft4(v, DIR); }
};
template <typename T, int DIR>
struct DftReg<T, 5, DIR> {
  static ADSCF_HD void run(C2<T>* v) { dft5(v, DIR); }
};
template <typename T, int DIR>
struct DftReg<T, 8, DIR> {
  static ADSCF_HD void run(C2<T>* v) { dft8(v, DIR); }
};
// 6 = 2 x 3: X[k2 + 3*k1] = sum_j1 W2^(j1 k1) W6^(j1 k2) sum_j2 x[j1 + 2 j2] W3^(j2 k2)
template <typename T, int DIR>
struct DftReg<T, 6, DIR> {
  static ADSCF_HD void run(C2<T>* v) {
    C2<T> a[3] = {v[0], v[2], v[4]};
    C2<T> b[3] = {v[1], v[3], v[5]};
    dft3(a, DIR);
    dft3(b, DIR);
    // W6^1 = (1/2, DIR*sqrt3/2), W6^2 = (-1/2, DIR*sqrt3/2)
    const T h = T(0.5), s = KC<T>::s3 * (T)DIR;
    b[1] = cmul(b[1], C2<T>{h, s});
    b[2] = cmul(b[2], C2<T>{-h, s});
    for (int k = 0; k < 3; ++k) {
      v[k] = cadd(a[k], b[k]);
      v[k + 3] = csub(a[k], b[k]);
    }
  }
};
// 16 = 4 x 4
template <typename T, int DIR>
struct DftReg<T, 16, DIR> {
  static ADSCF_HD void run(C2<T>* v) {
    C2<T> y[4][4];
#pragma unroll
    for (int j1 = 0; j1 < 4; ++j1) {
      y[j1][0] = v[j1];
      y[j1][1] = v[j1 + 4];
      y[j1][2] = v[j1 + 8];
      y[j1][3] = v[j1 + 12];
      dft4(y[j1], DIR);
#pragma unroll
      for (int k2 = 1; k2 < 4; ++k2)
        if (j1 > 0) y[j1][k2] = cmul(y[j1][k2], w16<T, DIR>(j1 * k2));
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      C2<T> z[4] = {y[0][k2], y[1][k2], y[2][k2], y[3][k2]};
      dft4(z, DIR);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = z[k1];
    }
  }
};

template <typename T, int N, int DIR>
ADSCF_HD void dft_reg(C2<T>* v) {
  DftReg<T, N, DIR>::run(v);
}

template <int N1, int N2>
struct FourStep {
  static constexpr int n = N1 * N2;
  static constexpr int tpl = N1 > N2 ? N1 : N2;  // threads per line
  static constexpr int n1p = N1 + 1;             // padded LDS row (bank spread)
  static constexpr int lds_elems = N2 * n1p;     // per line
};

// Per-thread twiddles and multipliers of one line operator, loaded once per
// pass so that their memory latency overlaps the loads of the data.  Callers run
// exactly max(N1, N2) threads per line (t < max(N1, N2)), so a guard against the
// larger of the two is always true.
//   ta[k2] = W_n^(t*k2)       phase A, thread t < N1
//   tb[j1] = W_n^(j1*t)       phase B, thread t < N2
//   k[k1]  = K[t + N2*k1]     phase B; with N1 == 1, k[k] = K[k] for the whole line
template <typename T, int N1, int N2>
struct RegTables {
  static constexpr int NK = N1 > 1 ? N1 : N2;
  C2<T> ta[N1 > 1 ? N2 : 1];
  C2<T> tb[N1];
  T k[NK];
  ADSCF_HD void load(int t, const C2<T>* tw, const T* kf) {
    if constexpr (N1 == 1) {
#pragma unroll
      for (int i = 0; i < N2; ++i) k[i] = kf[i];
    } else {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) ta[k2] = (N1 >= N2 || t < N1) ? tw[t * k2] : C2<T>{T(1), T(0)};
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) tb[j1] = (N2 >= N1 || t < N2) ? tw[j1 * t] : C2<T>{T(1), T(0)};
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) k[k1] = (N2 >= N1 || t < N2) ? kf[t + N2 * k1] : T(0);
    }
  }
};

// phase A: thread t < N1 holds v[j2] = x[t + N1*j2]
template <typename T, int N1, int N2>
ADSCF_HD void fs_phase_a(C2<T>* v, C2<T>* lds, int t, const RegTables<T, N1, N2>& R) {
  using FS = FourStep<N1, N2>;
  dft_reg<T, N2, -1>(v);
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) {
    C2<T> y = v[k2];
    if (k2 > 0) y = cmul(y, R.ta[k2]);  // lane 0: W^0 = (1, -0), an exact multiply
    lds[t + FS::n1p * k2] = y;
  }
}

// phase B: thread t < N2 (as k2)
template <typename T, int N1, int N2>
ADSCF_HD void fs_phase_b(C2<T>* lds, int t, const RegTables<T, N1, N2>& R) {
  using FS = FourStep<N1, N2>;
  C2<T> u[N1];
#pragma unroll
  for (int j1 = 0; j1 < N1; ++j1) u[j1] = lds[j1 + FS::n1p * t];
  dft_reg<T, N1, -1>(u);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) u[k1] = cscale(u[k1], R.k[k1]);
  dft_reg<T, N1, +1>(u);
#pragma unroll
  for (int j1 = 0; j1 < N1; ++j1) {
    C2<T> z = u[j1];
    if (j1 > 0) z = cmulc(z, R.tb[j1]);  // lane 0: W^0, exact (no per-lane branch)
    lds[j1 + FS::n1p * t] = z;
  }
}

// phase C: thread t < N1 ends with v[j2] = (L x)[t + N1*j2]
template <typename T, int N1, int N2>
ADSCF_HD void fs_phase_c(C2<T>* v, const C2<T>* lds, int t) {
  using FS = FourStep<N1, N2>;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) v[k2] = lds[t + FS::n1p * k2];
  dft_reg<T, N2, +1>(v);
}

// whole line in one thread (N1 == 1): forward DFT, K, inverse DFT
template <typename T, int N>
ADSCF_HD void line_op_single(C2<T>* v, const RegTables<T, 1, N>& R) {
  dft_reg<T, N, -1>(v);
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = cscale(v[k], R.k[k]);
  dft_reg<T, N, +1>(v);
}

}  // namespace adscf
