// Crout LU with implicit-scaling partial pivoting and its back substitution,
// with the semantics of ludcmp / lubksb in General_Routines/mod_iterate.f90
// (232-298, 302-337): column-major n x n matrix, 1-based pivot indices in
// indx, a zero pivot replaced by 1e-20, the last maximal scaled pivot chosen
// ('.ge.' comparison).  Shared by the host C-ABI and the device solve; FMA
// contraction is off so both give the same bits as a plain sequential loop.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADSCF_LU_HD __host__ __device__
#else
#define ADSCF_LU_HD
#endif

#include <cmath>

namespace adscf {

// returns 1 when a row of a is entirely zero ('singular matrix in ludcmp', :251)
// vv: workspace of n + 1 doubles
ADSCF_LU_HD inline int lu_decompose(double* a, int n, int* indx, double* d, double* vv) {
#pragma clang fp contract(off)
  int singular = 0, imax = 1;
  *d = 1.0;
  for (int i = 1; i <= n; ++i) {
    double aamax = 0.0;
    for (int j = 1; j <= n; ++j) {
      const double v = fabs(a[(i - 1) + (j - 1) * n]);
      if (v > aamax) aamax = v;
    }
    if (aamax == 0.0) singular = 1;
    vv[i] = 1.0 / aamax;
  }
  for (int j = 1; j <= n; ++j) {
    for (int i = 1; i <= j - 1; ++i) {
      double s = a[(i - 1) + (j - 1) * n];
      for (int k = 1; k <= i - 1; ++k) s = s - a[(i - 1) + (k - 1) * n] * a[(k - 1) + (j - 1) * n];
      a[(i - 1) + (j - 1) * n] = s;
    }
    double aamax = 0.0;
    for (int i = j; i <= n; ++i) {
      double s = a[(i - 1) + (j - 1) * n];
      for (int k = 1; k <= j - 1; ++k) s = s - a[(i - 1) + (k - 1) * n] * a[(k - 1) + (j - 1) * n];
      a[(i - 1) + (j - 1) * n] = s;
      const double dum = vv[i] * fabs(s);
      if (dum >= aamax) {
        imax = i;
        aamax = dum;
      }
    }
    if (j != imax) {
      for (int k = 1; k <= n; ++k) {
        const double dum = a[(imax - 1) + (k - 1) * n];
        a[(imax - 1) + (k - 1) * n] = a[(j - 1) + (k - 1) * n];
        a[(j - 1) + (k - 1) * n] = dum;
      }
      *d = -*d;
      vv[imax] = vv[j];
    }
    indx[j - 1] = imax;
    if (a[(j - 1) + (j - 1) * n] == 0.0) a[(j - 1) + (j - 1) * n] = 1.0e-20;
    if (j != n) {
      const double dum = 1.0 / a[(j - 1) + (j - 1) * n];
      for (int i = j + 1; i <= n; ++i) a[(i - 1) + (j - 1) * n] = a[(i - 1) + (j - 1) * n] * dum;
    }
  }
  return singular;
}

ADSCF_LU_HD inline void lu_backsubstitute(const double* a, int n, const int* indx, double* b) {
#pragma clang fp contract(off)
  int ii = 0;
  for (int i = 1; i <= n; ++i) {
    const int ll = indx[i - 1];
    double s = b[ll - 1];
    b[ll - 1] = b[i - 1];
    if (ii != 0) {
      for (int j = ii; j <= i - 1; ++j) s = s - a[(i - 1) + (j - 1) * n] * b[j - 1];
    } else if (s != 0.0) {
      ii = i;
    }
    b[i - 1] = s;
  }
  for (int i = n; i >= 1; --i) {
    double s = b[i - 1];
    for (int j = i + 1; j <= n; ++j) s = s - a[(i - 1) + (j - 1) * n] * b[j - 1];
    b[i - 1] = s / a[(i - 1) + (i - 1) * n];
  }
}

}  // namespace adscf
