// Register-resident line operator L = F^-1 diag(K) F for n = R^3 held by R^2
// threads of one line, R points each (index j = t + R^2 m).  With every thread
// busy in every phase it gives a 64-point line 16 threads (four per thread in
// registers) where the 8 x 8 four-step of fft_reg.hpp gives 8: twice the waves
// at a small grid, and radix-4 butterflies need no multiplications.
//   forward  A  thread t:        DFT_R over m, twiddle W_n^(t k2)        -> S[t][k2]
//            B  thread (k2, a):  DFT_R over b of S[a + R b][k2], twiddle W_(R^2)^(a c) -> S2[k2][c][a]
//            C  thread (k2, c):  DFT_R over a -> bin k = k2 + R (c + R d); multiply by K[k]
//   inverse  C, B, A with conjugate DFTs and twiddles in reverse order
// (n = N1 N2 four-step with N1 = R^2, N2 = R, its inner DFT_(R^2) split again as R x R).
#pragma once

#include "fft_reg.hpp"

namespace adscf {

// Twiddle multiplies are unconditional in the lane index: the twiddle of lane 0 is
// W^0 = (1, -0), and z (1, -0) = z exactly (z.x 1 - z.y (-0) = z.x, z.x (-0) + z.y 1 = z.y, up
// to the sign of an exact zero), so skipping it for lane 0 would only add a per-lane branch,
// which serialises the phase's LDS reads behind it.

template <int R>
struct Cube {
  static constexpr int n = R * R * R;
  static constexpr int tpl = R * R;           // threads per line
  static constexpr int np = R;                // points per thread
  static constexpr int sp = R + 1;            // padded row of S and S2
  static constexpr int lds_elems = 2 * R * R * sp;  // S then S2, per line
};

// twiddles and multipliers of thread t of a line
template <typename T, int R>
struct CubeTables {
  C2<T> wa[R];  // W_n^(t k2)
  C2<T> wb[R];  // W_n^(R a c), a = t % R, c = 0..R-1
  T k[R];       // K[k2 + R (c + R d)], (k2, c) = (t / R, t % R), d = 0..R-1
  ADSCF_HD void load(int t, const C2<T>* tw, const T* kf) {
    const int a = t % R, k2 = t / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      wa[q] = tw[t * q];
      wb[q] = tw[R * a * q];
      k[q] = kf[k2 + R * (a + R * q)];
    }
  }
};

// The five phases; thread t of the line, lds = S (R^2 padded rows) then S2.
template <typename T, int R>
ADSCF_HD void cube_a(C2<T>* v, C2<T>* lds, int t, const CubeTables<T, R>& W) {
  constexpr int sp = Cube<R>::sp;
  dft_reg<T, R, -1>(v);
#pragma unroll
  for (int q = 0; q < R; ++q) lds[t * sp + q] = q > 0 ? cmul(v[q], W.wa[q]) : v[q];
}
template <typename T, int R>
ADSCF_HD void cube_b(C2<T>* lds, int t, const CubeTables<T, R>& W) {
  constexpr int sp = Cube<R>::sp;
  C2<T>* S2 = lds + R * R * sp;
  const int a = t % R, k2 = t / R;
  C2<T> u[R];
#pragma unroll
  for (int b = 0; b < R; ++b) u[b] = lds[(a + R * b) * sp + k2];
  dft_reg<T, R, -1>(u);
#pragma unroll
  for (int c = 0; c < R; ++c) S2[(k2 * R + c) * sp + a] = c > 0 ? cmul(u[c], W.wb[c]) : u[c];
}
template <typename T, int R>
ADSCF_HD void cube_c(C2<T>* lds, int t, const CubeTables<T, R>& W) {
  constexpr int sp = Cube<R>::sp;
  C2<T>* row = lds + R * R * sp + t * sp;  // (k2, c) = (t / R, t % R): row k2 R + c = t
  C2<T> u[R];
#pragma unroll
  for (int q = 0; q < R; ++q) u[q] = row[q];
  dft_reg<T, R, -1>(u);
#pragma unroll
  for (int d = 0; d < R; ++d) u[d] = cscale(u[d], W.k[d]);
  dft_reg<T, R, +1>(u);
#pragma unroll
  for (int q = 0; q < R; ++q) row[q] = u[q];
}
template <typename T, int R>
ADSCF_HD void cube_bi(C2<T>* lds, int t, const CubeTables<T, R>& W) {
  constexpr int sp = Cube<R>::sp;
  const C2<T>* S2 = lds + R * R * sp;
  const int a = t % R, k2 = t / R;
  C2<T> u[R];
#pragma unroll
  for (int c = 0; c < R; ++c) {
    const C2<T> z = S2[(k2 * R + c) * sp + a];
    u[c] = c > 0 ? cmulc(z, W.wb[c]) : z;
  }
  dft_reg<T, R, +1>(u);
#pragma unroll
  for (int b = 0; b < R; ++b) lds[(a + R * b) * sp + k2] = u[b];
}
template <typename T, int R>
ADSCF_HD void cube_ai(C2<T>* v, const C2<T>* lds, int t, const CubeTables<T, R>& W) {
  constexpr int sp = Cube<R>::sp;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const C2<T> z = lds[t * sp + q];
    v[q] = q > 0 ? cmulc(z, W.wa[q]) : z;
  }
  dft_reg<T, R, +1>(v);
}

// ---------------------------------------------------------------------------
// n = A R^2 on R^2 threads, A points each (A a multiple of R): the cube engine with a
// longer first stage.  128 = 8 x 4 x 4 on 16 threads keeps every thread busy in every
// phase, where the 16 x 8 four-step of fft_reg.hpp leaves half of its 16 threads idle in
// its 16-point phase.  Index j = t + R^2 m (t < R^2, m < A); bin k = k2 + A (c + R d).
//   A   thread t:             DFT_A over m, twiddle W_n^(t k2)            -> S[t][k2]
//   B   pair (k2, a), P each:  DFT_R over b of S[a + R b][k2], W_n^(A a c) -> S2[k2 R + c][a]
//   C   row (k2, c), P each:   DFT_R over a -> bin k2 + A (c + R d), times K, inverse DFT_R
//   then B and A inverted (conjugate twiddles, inverse DFTs).  P = A / R pairs and rows
//   per thread, pair/row index p = t + R^2 i.
// ---------------------------------------------------------------------------
template <int A, int R>
struct Tri {
  static_assert(A % R == 0, "A must be a multiple of R");
  static constexpr int n = A * R * R;
  static constexpr int tpl = R * R;
  static constexpr int np = A;
  static constexpr int P = A / R;
  static constexpr int sa = A + 1;  // padded row of S
  static constexpr int sb = R + 1;  // padded row of S2
  static constexpr int lds_elems = R * R * sa + A * R * sb;
};

template <typename T, int A, int R>
struct TriTables {
  static constexpr int P = Tri<A, R>::P;
  C2<T> wa[A];     // W_n^(t k2)
  C2<T> wb[P][R];  // W_n^(A a c) of pair i
  T k[P][R];       // K[k2 + A (c + R d)] of row i
  ADSCF_HD void load(int t, const C2<T>* tw, const T* kf) {
    constexpr int TT = R * R;
#pragma unroll
    for (int q = 0; q < A; ++q) wa[q] = tw[t * q];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int p = t + TT * i;
      const int a = p % R, k2 = p / R;  // pair (k2, a); row (k2, c) with c = p % R
#pragma unroll
      for (int c = 0; c < R; ++c) wb[i][c] = tw[A * a * c];
#pragma unroll
      for (int d = 0; d < R; ++d) k[i][d] = kf[k2 + A * (a + R * d)];
    }
  }
};

template <typename T, int A, int R>
ADSCF_HD void tri_a(C2<T>* v, C2<T>* lds, int t, const TriTables<T, A, R>& W) {
  constexpr int sa = Tri<A, R>::sa;
  dft_reg<T, A, -1>(v);
#pragma unroll
  for (int q = 0; q < A; ++q) lds[t * sa + q] = q > 0 ? cmul(v[q], W.wa[q]) : v[q];
}
template <typename T, int A, int R>
ADSCF_HD void tri_b(C2<T>* lds, int t, const TriTables<T, A, R>& W) {
  constexpr int sa = Tri<A, R>::sa, sb = Tri<A, R>::sb, TT = R * R;
  C2<T>* S2 = lds + TT * sa;
#pragma unroll
  for (int i = 0; i < Tri<A, R>::P; ++i) {
    const int p = t + TT * i;
    const int a = p % R, k2 = p / R;
    C2<T> u[R];
#pragma unroll
    for (int b = 0; b < R; ++b) u[b] = lds[(a + R * b) * sa + k2];
    dft_reg<T, R, -1>(u);
#pragma unroll
    for (int c = 0; c < R; ++c) S2[(k2 * R + c) * sb + a] = c > 0 ? cmul(u[c], W.wb[i][c]) : u[c];
  }
}
template <typename T, int A, int R>
ADSCF_HD void tri_c(C2<T>* lds, int t, const TriTables<T, A, R>& W) {
  constexpr int sa = Tri<A, R>::sa, sb = Tri<A, R>::sb, TT = R * R;
#pragma unroll
  for (int i = 0; i < Tri<A, R>::P; ++i) {
    C2<T>* row = lds + TT * sa + (t + TT * i) * sb;  // row (k2, c) = ((t + TT i) / R, (t + TT i) % R)
    C2<T> u[R];
#pragma unroll
    for (int q = 0; q < R; ++q) u[q] = row[q];
    dft_reg<T, R, -1>(u);
#pragma unroll
    for (int d = 0; d < R; ++d) u[d] = cscale(u[d], W.k[i][d]);
    dft_reg<T, R, +1>(u);
#pragma unroll
    for (int q = 0; q < R; ++q) row[q] = u[q];
  }
}
template <typename T, int A, int R>
ADSCF_HD void tri_bi(C2<T>* lds, int t, const TriTables<T, A, R>& W) {
  constexpr int sa = Tri<A, R>::sa, sb = Tri<A, R>::sb, TT = R * R;
  const C2<T>* S2 = lds + TT * sa;
#pragma unroll
  for (int i = 0; i < Tri<A, R>::P; ++i) {
    const int p = t + TT * i;
    const int a = p % R, k2 = p / R;
    C2<T> u[R];
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const C2<T> z = S2[(k2 * R + c) * sb + a];
      u[c] = c > 0 ? cmulc(z, W.wb[i][c]) : z;
    }
    dft_reg<T, R, +1>(u);
#pragma unroll
    for (int b = 0; b < R; ++b) lds[(a + R * b) * sa + k2] = u[b];
  }
}
template <typename T, int A, int R>
ADSCF_HD void tri_ai(C2<T>* v, const C2<T>* lds, int t, const TriTables<T, A, R>& W) {
  constexpr int sa = Tri<A, R>::sa;
#pragma unroll
  for (int q = 0; q < A; ++q) {
    const C2<T> z = lds[t * sa + q];
    v[q] = q > 0 ? cmulc(z, W.wa[q]) : z;
  }
  dft_reg<T, A, +1>(v);
}

}  // namespace adscf
