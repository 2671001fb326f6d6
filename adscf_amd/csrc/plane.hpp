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
// Layouts inside the workgroup (nx = ny = n, NL = n / 2 complex lines per operator):
//   x layout  thread id = j TPL + t: line j (rows j, j + NL), points x = t + STRIDE m;
//             a line's threads are consecutive lanes, so its exchanges stay in one wave
//   y layout  thread id = t NL + j: line j (columns j, j + NL), points y = t + STRIDE m;
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
  static constexpr int threads = ENG::TPL * NL;
  static constexpr int ld = n + 1;  // padded row of the real plane in LDS
  static constexpr size_t plane_bytes = ((size_t)n * ld * sizeof(T) + 15) / 16 * 16;
  // LDS of a workgroup whose lines' exchange areas are `stride` elements apart
  static constexpr size_t lds_bytes(int stride) { return plane_bytes + (size_t)NL * stride * sizeof(C2<T>); }
};

// Workgroup b takes chain (b / 8) % 2 of plane 8 (b / 16) + b % 8: both chains of a plane
// land on one XCD (workgroups are dealt to the 8 XCDs round-robin), so the two halves of
// each (q, q+) pair of the line buffer meet in one L2.
__device__ __forceinline__ void plane_of_block(int b, int& chain, int& z) {
  chain = (b >> 3) & 1;
  z = (b >> 4) * 8 + (b & 7);
}

template <typename T, int N1, int N2, int MODE>
__global__ void __launch_bounds__((PlaneGeom<T, PassEngine<T, N1, N2>>::threads)) k_plane_pass(PassArgs<T> a) {
  using ENG = PassEngine<T, N1, N2>;
  using G = PlaneGeom<T, ENG>;
  constexpr int n = G::n, NL = G::NL, NP = ENG::NP, STRIDE = ENG::STRIDE, TPL = ENG::TPL, LD = G::ld;
  constexpr bool FIRST = MODE == PM_FIRST, MERGED = MODE == PM_MERGED, LAST = MODE == PM_LAST;
  static_assert(FIRST || MERGED || LAST, "plane passes come in the first, merged and last variants");
  // the transposes rely on the block barriers inside the y operator (engines with N1 > 1)
  static_assert(ENG::LDS > 0, "plane passes need an engine that exchanges through LDS");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* P = reinterpret_cast<T*>(smem_raw);
  C2<T>* scr = reinterpret_cast<C2<T>*>(smem_raw + G::plane_bytes);
  int chain, z;
  plane_of_block(blockIdx.x, chain, z);
  if (z >= a.nz) return;  // uniform over the block
  const int tid = threadIdx.x;
  const int jx = tid / TPL, tx = tid % TPL;  // x layout
  const int jy = tid % NL, ty = tid / NL;    // y layout
  const bool actx = tx < ENG::NACT, acty = ty < ENG::NACT;
  const unsigned zb = (unsigned)z * (unsigned)(n * n);
  const T* Ea = chain ? a.Ea2 : a.Ea1;
  const T* Eb = chain ? a.Eb2 : a.Eb1;
  T* h = chain ? a.h2 : a.h1;
  const int stop = *(a.done ? a.done : &g_no_stop);
  // point index of element m in either layout; the partner row / column is NL n / NL away
  auto px = [&](int m) { return zb + (unsigned)(jx * n + tx + STRIDE * m); };
  auto py = [&](int m) { return zb + (unsigned)((ty + STRIDE * m) * n + jy); };
  constexpr unsigned kRow = (unsigned)(NL * n);  // x layout: second row of the pair
  constexpr unsigned kCol = (unsigned)NL;        // y layout: second column of the pair
  const unsigned ch = (unsigned)chain * (unsigned)sizeof(T);
  auto pair_off = [&](unsigned g) { return 2u * g * (unsigned)sizeof(T) + ch; };  // this chain's half

  C2<T> v[NP];
  T e0[NP], e1[NP], f0[NP], f1[NP];  // E factors of the two points of each element
#pragma unroll
  for (int m = 0; m < NP; ++m) {
    v[m] = {T(0), T(0)};
    e0[m] = e1[m] = f0[m] = f1[m] = T(1);
    if (FIRST) {
      if (actx) {
        v[m] = {T(1), T(1)};
        e0[m] = ldo(Ea, px(m) * (unsigned)sizeof(T));
        e1[m] = ldo(Ea, (px(m) + kRow) * (unsigned)sizeof(T));
      }
    } else {
      if (actx) {
        v[m].x = ldo(reinterpret_cast<const T*>(a.tmp_in), pair_off(px(m)));
        v[m].y = ldo(reinterpret_cast<const T*>(a.tmp_in), pair_off(px(m) + kRow));
      }
      if (acty) {  // E of the step boundary, in the y layout where it is applied
        e0[m] = ldo(Ea, py(m) * (unsigned)sizeof(T));
        e1[m] = ldo(Ea, (py(m) + kCol) * (unsigned)sizeof(T));
        if (MERGED) {
          f0[m] = ldo(Eb, py(m) * (unsigned)sizeof(T));
          f1[m] = ldo(Eb, (py(m) + kCol) * (unsigned)sizeof(T));
        }
      }
    }
  }
  typename ENG::Tables Rx, Ry;
  Rx.load(tx, a.tw, a.kf);
  Ry.load(ty, a.tw2, a.kf2);
  const T sc = (MERGED || LAST) ? exp2i<T>((chain ? a.dexp2 : a.dexp1)[a.sidx]) : T(1);
  const bool wlx = lines_in_wave<TPL>(0);
  if (FIRST) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      v[m].x *= e0[m];
      v[m].y *= e1[m];
    }
  }
  ENG::op(v, scr + jx * a.lds_line, tx, Rx, wlx);
  if (stop) return;  // uniform over the block; nothing stored yet
  // x layout -> y layout
  if (actx) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      P[jx * LD + tx + STRIDE * m] = v[m].x;
      P[(jx + NL) * LD + tx + STRIDE * m] = v[m].y;
    }
  }
  __syncthreads();
  if (acty) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      v[m].x = P[(ty + STRIDE * m) * LD + jy];
      v[m].y = P[(ty + STRIDE * m) * LD + jy + NL];
    }
  }
  ENG::op(v, scr + jy * a.lds_line2, ty, Ry, false);
  if (FIRST) {
    if (acty) {
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        sto(reinterpret_cast<T*>(a.tmp_out), pair_off(py(m)), v[m].x);
        sto(reinterpret_cast<T*>(a.tmp_out), pair_off(py(m) + kCol), v[m].y);
      }
    }
    return;
  }
  if (LAST) {
    if (acty) {
#pragma unroll
      for (int m = 0; m < NP; ++m) {
        sto(h, py(m) * (unsigned)sizeof(T), v[m].x * e0[m] * sc);
        sto(h, (py(m) + kCol) * (unsigned)sizeof(T), v[m].y * e1[m] * sc);
      }
    }
    return;
  }
  // merged: E of the ending step and the rescale, store q(s), E of the next step
  if (acty) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      C2<T> x = v[m];
      x.x = x.x * e0[m] * sc;
      x.y = x.y * e1[m] * sc;
      sto(h, py(m) * (unsigned)sizeof(T), x.x);
      sto(h, (py(m) + kCol) * (unsigned)sizeof(T), x.y);
      x.x *= f0[m];
      x.y *= f1[m];
      v[m] = x;
    }
  }
  ENG::op(v, scr + jy * a.lds_line2, ty, Ry, false);
  // y layout -> x layout (the y operators' barriers have retired every read of the
  // first transpose)
  if (acty) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      P[(ty + STRIDE * m) * LD + jy] = v[m].x;
      P[(ty + STRIDE * m) * LD + jy + NL] = v[m].y;
    }
  }
  __syncthreads();
  if (actx) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      v[m].x = P[jx * LD + tx + STRIDE * m];
      v[m].y = P[(jx + NL) * LD + tx + STRIDE * m];
    }
  }
  ENG::op(v, scr + jx * a.lds_line, tx, Rx, wlx);
  if (actx) {
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      sto(reinterpret_cast<T*>(a.tmp_out), pair_off(px(m)), v[m].x);
      sto(reinterpret_cast<T*>(a.tmp_out), pair_off(px(m) + kRow), v[m].y);
    }
  }
}

}  // namespace adscf

