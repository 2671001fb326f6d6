// xy-plane pass: the x and y line operators of a contour step in one kernel.
//
// A contour step is E Lx Ly Lz E.  The line passes of kernels.hpp apply one axis per
// pass, so a step costs two dependent launches (the outer axes of consecutive steps
// merge).  A plane pass holds one z-plane of one chain in a workgroup and applies Lx and
// Ly there, so that the sweep alternates plane passes and z passes and a step costs one
// launch: the plane pass of a step boundary applies  Ly Lx | E store E | Lx Ly  and the z
// pass  Lz | E store E | Lz.
//
// The chain's plane is real, and every line operator is real (L = F^-1 diag(K) F with K
// real and even), so two real lines travel as the real and imaginary parts of one complex
// line with no cross-talk: the x operator runs on the complex lines (row j, row j + n/2)
// and the y operator on (column j, column j + n/2).  The two chains of the sweep get
// separate workgroups (one per chain and plane), which doubles the workgroups of a pass
// against a packed (q, q+) plane.  The complex line operator itself is the register
// engine of the line passes (fft_reg.hpp, fft_cube.hpp), unchanged.
//
// Layouts inside the workgroup (nx = ny = n, NL = n / 2 complex lines per operator, each
// thread carrying LPT of them, lines j and j + NL / LPT; NLT = NL / LPT):
//   x layout  thread id = j TPL + t: line j (rows j, j + NL), points x = t + STRIDE m;
//             a line's threads are consecutive lanes, so its exchanges stay in one wave
//   y layout  thread id = t NLT + j: line j (columns j, j + NL), points y = t + STRIDE m;
//             consecutive lanes hold consecutive x, so the E loads and the q(s) stores
//             of the step boundary are coalesced rows
// The plane moves between the layouts through an LDS array of n padded rows.  The exchange
// areas of the lines are a.lds_line (x layout) or a.lds_line2 (y layout) elements apart, so
// that each layout's lines start in banks the host's model finds free of conflicts.  The packed
// line buffer between passes keeps the (q, q+) pairs of the z passes; a plane workgroup
// reads and writes its chain's component.
#pragma once

#include "kernels.hpp"

namespace adscf {

constexpr int kAxisPlane = 3;  // PassArgs::axis of a plane pass

template <typename T, class ENG>
struct PlaneGeom {
  static constexpr int n = ENG::NP * ENG::STRIDE;  // points per line of the engine
  static constexpr int NL = n / 2;                 // complex lines per operator
  // lines per thread: a 128-point plane (64 complex lines of 16 threads) would need 1024
  // threads with both operators' tables in registers; two lines per thread halve that
  static constexpr int LPT = ENG::TPL * NL > 512 ? 2 : 1;
  static constexpr int NLT = NL / LPT;  // exchange areas (lines in flight)
  static constexpr int threads = ENG::TPL * NLT;
  static constexpr int ld = n + 1;  // padded row of the real plane in LDS
  static constexpr size_t plane_bytes = ((size_t)n * ld * sizeof(T) + 15) / 16 * 16;
  // LDS of a workgroup whose lines' exchange areas are `stride` elements apart
  static constexpr size_t lds_bytes(int stride) { return plane_bytes + (size_t)NLT * stride * sizeof(C2<T>); }
};

// Workgroup b takes chain (b / 8) % 2 of plane 8 (b / 16) + b % 8: both chains of a plane
// land on one XCD (workgroups are dealt to the 8 XCDs round-robin), so the two halves of
// each (q, q+) pair of the line buffer meet in one L2.
__device__ __forceinline__ void plane_of_block(int b, int& chain, int& z) {
  chain = (b >> 3) & 1;
  z = (b >> 4) * 8 + (b & 7);
}

// one z-plane of one chain; every thread of the workgroup calls it (it has block barriers)
template <typename T, int N1, int N2, int MODE>
__device__ __forceinline__ void plane_body(const PassArgs<T>& a, int chain, int z, unsigned char* smem_raw) {
  using ENG = PassEngine<T, N1, N2>;
  using G = PlaneGeom<T, ENG>;
  constexpr int n = G::n, NL = G::NL, NP = ENG::NP, STRIDE = ENG::STRIDE, TPL = ENG::TPL, LD = G::ld;
  constexpr int LPT = G::LPT, NLT = G::NLT;
  constexpr bool FIRST = MODE == PM_FIRST, MERGED = MODE == PM_MERGED, LAST = MODE == PM_LAST;
  // with many points per thread the step-boundary E factors are loaded after the first
  // transpose, so that they do not hold registers through the first operators
  constexpr bool EARLY_E = NP * LPT <= 8;
  static_assert(FIRST || MERGED || LAST, "plane passes come in the first, merged and last variants");
  // the transposes rely on the block barriers inside the y operator (engines with N1 > 1)
  static_assert(ENG::LDS > 0, "plane passes need an engine that exchanges through LDS");
  T* P = reinterpret_cast<T*>(smem_raw);
  C2<T>* scr = reinterpret_cast<C2<T>*>(smem_raw + G::plane_bytes);
  const int tid = threadIdx.x;
  const int jx = tid / TPL, tx = tid % TPL;  // x layout: lines jx + r NLT
  const int jy = tid % NLT, ty = tid / NLT;  // y layout
  const bool actx = tx < ENG::NACT, acty = ty < ENG::NACT;
  const unsigned zb = (unsigned)z * (unsigned)(n * n);
  const T* Ea = chain ? a.Ea2 : a.Ea1;
  const T* Eb = chain ? a.Eb2 : a.Eb1;
  T* h = chain ? a.h2 : a.h1;
  const int stop = *(a.done ? a.done : &g_no_stop);
  // point index of element m of line r in either layout; the partner row / column of the
  // complex line is NL n / NL away
  auto px = [&](int r, int m) { return zb + (unsigned)((jx + r * NLT) * n + tx + STRIDE * m); };
  auto py = [&](int r, int m) { return zb + (unsigned)((ty + STRIDE * m) * n + jy + r * NLT); };
  constexpr unsigned kRow = (unsigned)(NL * n);  // x layout: second row of the pair
  constexpr unsigned kCol = (unsigned)NL;        // y layout: second column of the pair
  constexpr unsigned B = (unsigned)sizeof(T);
  const unsigned ch = (unsigned)chain * B;
  auto pair_off = [&](unsigned g) { return 2u * g * B + ch; };  // this chain's half

  C2<T> v[LPT][NP];
  T e0[LPT][NP], e1[LPT][NP], f0[LPT][NP], f1[LPT][NP];  // E factors of the two points
  auto load_E_y = [&] {  // E of the step boundary, in the y layout where it is applied
    if (acty) {
#pragma unroll
      for (int r = 0; r < LPT; ++r)
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          e0[r][m] = ldo(Ea, py(r, m) * B);
          e1[r][m] = ldo(Ea, (py(r, m) + kCol) * B);
          if (MERGED) {
            f0[r][m] = ldo(Eb, py(r, m) * B);
            f1[r][m] = ldo(Eb, (py(r, m) + kCol) * B);
          }
        }
    }
  };
#pragma unroll
  for (int r = 0; r < LPT; ++r)
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      v[r][m] = {T(0), T(0)};
      e0[r][m] = e1[r][m] = f0[r][m] = f1[r][m] = T(1);
      if (actx) {
        if (FIRST) {
          v[r][m] = {T(1), T(1)};
          e0[r][m] = ldo(Ea, px(r, m) * B);
          e1[r][m] = ldo(Ea, (px(r, m) + kRow) * B);
        } else {
          v[r][m].x = ldo(reinterpret_cast<const T*>(a.tmp_in), pair_off(px(r, m)));
          v[r][m].y = ldo(reinterpret_cast<const T*>(a.tmp_in), pair_off(px(r, m) + kRow));
        }
      }
    }
  if (!FIRST && EARLY_E) load_E_y();
  typename ENG::Tables Rx, Ry;
  Rx.load(tx, a.tw, a.kf);
  Ry.load(ty, a.tw2, a.kf2);
  const T sc = (MERGED || LAST) ? exp2i<T>((chain ? a.dexp2 : a.dexp1)[a.sidx]) : T(1);
  const bool wlx = lines_in_wave<TPL>(0);
  C2<T>* sx = scr + jx * a.lds_line;
  C2<T>* sy = scr + jy * a.lds_line2;
  if (FIRST) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        v[r][m].x *= e0[r][m];
        v[r][m].y *= e1[r][m];
      }
  }
#pragma unroll
  for (int r = 0; r < LPT; ++r) ENG::op(v[r], sx, tx, Rx, wlx);
  if (stop) return;  // uniform over the block; nothing stored yet
  // x layout -> y layout
  if (actx) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        P[(jx + r * NLT) * LD + tx + STRIDE * m] = v[r][m].x;
        P[(jx + r * NLT + NL) * LD + tx + STRIDE * m] = v[r][m].y;
      }
  }
  __syncthreads();
  if (acty) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        v[r][m].x = P[(ty + STRIDE * m) * LD + jy + r * NLT];
        v[r][m].y = P[(ty + STRIDE * m) * LD + jy + r * NLT + NL];
      }
  }
  if (!FIRST && !EARLY_E) load_E_y();
#pragma unroll
  for (int r = 0; r < LPT; ++r) ENG::op(v[r], sy, ty, Ry, false);
  if (FIRST) {
    if (acty) {
#pragma unroll
      for (int r = 0; r < LPT; ++r)
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          sto(reinterpret_cast<T*>(a.tmp_out), pair_off(py(r, m)), v[r][m].x);
          sto(reinterpret_cast<T*>(a.tmp_out), pair_off(py(r, m) + kCol), v[r][m].y);
        }
    }
    return;
  }
  if (LAST) {
    if (acty) {
#pragma unroll
      for (int r = 0; r < LPT; ++r)
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          sto_nt(h, py(r, m) * B, v[r][m].x * e0[r][m] * sc);
          sto_nt(h, (py(r, m) + kCol) * B, v[r][m].y * e1[r][m] * sc);
        }
    }
    return;
  }
  // merged: E of the ending step and the rescale, store q(s), E of the next step
  if (acty) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        C2<T> x = v[r][m];
        x.x = x.x * e0[r][m] * sc;
        x.y = x.y * e1[r][m] * sc;
        sto_nt(h, py(r, m) * B, x.x);
        sto_nt(h, (py(r, m) + kCol) * B, x.y);
        x.x *= f0[r][m];
        x.y *= f1[r][m];
        v[r][m] = x;
      }
  }
#pragma unroll
  for (int r = 0; r < LPT; ++r) ENG::op(v[r], sy, ty, Ry, false);
  // y layout -> x layout (the y operators' barriers have retired every read of the
  // first transpose)
  if (acty) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        P[(ty + STRIDE * m) * LD + jy + r * NLT] = v[r][m].x;
        P[(ty + STRIDE * m) * LD + jy + r * NLT + NL] = v[r][m].y;
      }
  }
  __syncthreads();
  if (actx) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        v[r][m].x = P[(jx + r * NLT) * LD + tx + STRIDE * m];
        v[r][m].y = P[(jx + r * NLT + NL) * LD + tx + STRIDE * m];
      }
  }
#pragma unroll
  for (int r = 0; r < LPT; ++r) ENG::op(v[r], sx, tx, Rx, wlx);
  if (actx) {
#pragma unroll
    for (int r = 0; r < LPT; ++r)
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        sto(reinterpret_cast<T*>(a.tmp_out), pair_off(px(r, m)), v[r][m].x);
        sto(reinterpret_cast<T*>(a.tmp_out), pair_off(px(r, m) + kRow), v[r][m].y);
      }
  }
}

template <typename T, int N1, int N2, int MODE>
__global__ void __launch_bounds__((PlaneGeom<T, PassEngine<T, N1, N2>>::threads)) k_plane_pass(PassArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int chain, z;
  plane_of_block(blockIdx.x, chain, z);
  if (z >= a.nz) return;  // uniform over the block
  plane_body<T, N1, N2, MODE>(a, chain, z, smem_raw);
}

// Batched independent instances: block b runs item (code & mask) of the instance in slot
// (code >> kBatchItemBits) of the launch, an item being one (chain, z-plane) as above.
template <typename T, int N1, int N2, int MODE>
__global__ void __launch_bounds__((PlaneGeom<T, PassEngine<T, N1, N2>>::threads))
    k_plane_pass_batch(BatchGroup<T> g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int code = g.blk[blockIdx.x];
  const PassArgs<T> a = g.args[code >> kBatchItemBits];
  int chain, z;
  plane_of_block(code & ((1 << kBatchItemBits) - 1), chain, z);
  if (z >= a.nz) return;  // uniform over the block
  plane_body<T, N1, N2, MODE>(a, chain, z, smem_raw);
}

}  // namespace adscf
