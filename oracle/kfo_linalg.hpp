// kfo_linalg.hpp -- fixed-size dense matrices for the oracle.
//
// TEST INFRASTRUCTURE ONLY (see kfo_scalar.hpp).
//
// The reference does its algebra with Eigen fixed-size matrices
// (include/Utilities/Definitions.hpp:58-151).  Eigen is not available here and
// its summation order is not pinned by any reference test, so the oracle uses
// the textbook order: c(i,j) = a(i,0)b(0,j) + a(i,1)b(1,j) + ... left to right.
#pragma once

#include "kfo_scalar.hpp"

namespace kfo {

template <int R, int C>
struct Mat {
  Real a[R][C];
  Real &operator()(int i, int j) { return a[i][j]; }
  const Real &operator()(int i, int j) const { return a[i][j]; }
  // vector-style access for column vectors
  Real &operator[](int i) { return a[i][0]; }
  const Real &operator[](int i) const { return a[i][0]; }
  static Mat zero() {
    Mat m;
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < C; ++j) m.a[i][j] = 0.0;
    return m;
  }
  static Mat identity() {
    Mat m = zero();
    for (int i = 0; i < R && i < C; ++i) m.a[i][i] = 1.0;
    return m;
  }
};

template <int N>
using Vec = Mat<N, 1>;
using V2 = Vec<2>;
using V3 = Vec<3>;
using V6 = Vec<6>;
using V8 = Vec<8>;
using M22 = Mat<2, 2>;
using M33 = Mat<3, 3>;
using M66 = Mat<6, 6>;
using M86 = Mat<8, 6>;
using M68 = Mat<6, 8>;
using M88 = Mat<8, 8>;
using R6 = Mat<1, 6>;

inline V3 v3(Real x, Real y, Real z) {
  V3 v;
  v[0] = x;
  v[1] = y;
  v[2] = z;
  return v;
}

template <int R, int K, int C>
inline Mat<R, C> operator*(const Mat<R, K> &a, const Mat<K, C> &b) {
  Mat<R, C> r;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) {
      Real s = a(i, 0) * b(0, j);
      for (int k = 1; k < K; ++k) s = s + a(i, k) * b(k, j);
      r(i, j) = s;
    }
  return r;
}

template <int R, int C>
inline Mat<R, C> operator+(const Mat<R, C> &a, const Mat<R, C> &b) {
  Mat<R, C> r;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) r(i, j) = a(i, j) + b(i, j);
  return r;
}

template <int R, int C>
inline Mat<R, C> operator-(const Mat<R, C> &a, const Mat<R, C> &b) {
  Mat<R, C> r;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) r(i, j) = a(i, j) - b(i, j);
  return r;
}

template <int R, int C>
inline Mat<R, C> operator*(Real s, const Mat<R, C> &a) {
  Mat<R, C> r;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) r(i, j) = s * a(i, j);
  return r;
}

template <int R, int C>
inline Mat<C, R> transpose(const Mat<R, C> &a) {
  Mat<C, R> r;
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j) r(j, i) = a(i, j);
  return r;
}

inline Real dot(const V3 &a, const V3 &b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

inline V3 cross(const V3 &a, const V3 &b) {
  return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
            a[0] * b[1] - a[1] * b[0]);
}

inline Real norm(const V3 &a) { return ksqrt(dot(a, a)); }

inline V3 normalized(const V3 &a) {
  Real n = norm(a);
  return v3(a[0] / n, a[1] / n, a[2] / n);
}

inline V3 col(const M33 &m, int j) { return v3(m(0, j), m(1, j), m(2, j)); }

// Column-wise cross product of a 3x3 matrix with a vector,
// include/Utilities/Helpers.hpp:168-176.
inline M33 cross_cols(const M33 &m, const V3 &v) {
  M33 r;
  for (int j = 0; j < 3; ++j) {
    V3 c = cross(col(m, j), v);
    r(0, j) = c[0];
    r(1, j) = c[1];
    r(2, j) = c[2];
  }
  return r;
}

template <int R, int C>
inline bool has_nan(const Mat<R, C> &a) {
  for (int i = 0; i < R; ++i)
    for (int j = 0; j < C; ++j)
      if (kisnan(a(i, j))) return true;
  return false;
}

// Inverse by LU decomposition with partial pivoting, the method of the
// reference's host path (Eigen's fixed-size inverse() used by
// calculateInverse, include/Utilities/Math.hpp:103-112).
template <int N>
inline Mat<N, N> inverse(const Mat<N, N> &m) {
  Mat<N, N> lu = m;
  int perm[N];
  for (int i = 0; i < N; ++i) perm[i] = i;
  for (int k = 0; k < N; ++k) {
    int p = k;
    Real best = kabs(lu(k, k));
    for (int i = k + 1; i < N; ++i) {
      Real v = kabs(lu(i, k));
      if (best < v) {
        best = v;
        p = i;
      }
    }
    if (p != k) {
      for (int j = 0; j < N; ++j) {
        Real t = lu(k, j);
        lu(k, j) = lu(p, j);
        lu(p, j) = t;
      }
      int t = perm[k];
      perm[k] = perm[p];
      perm[p] = t;
    }
    for (int i = k + 1; i < N; ++i) {
      lu(i, k) = lu(i, k) / lu(k, k);
      for (int j = k + 1; j < N; ++j) lu(i, j) = lu(i, j) - lu(i, k) * lu(k, j);
    }
  }
  Mat<N, N> inv;
  for (int c = 0; c < N; ++c) {
    Real y[N];
    for (int i = 0; i < N; ++i) {
      Real s = (perm[i] == c) ? Real(1.0) : Real(0.0);
      for (int k = 0; k < i; ++k) s = s - lu(i, k) * y[k];
      y[i] = s;
    }
    for (int i = N - 1; i >= 0; --i) {
      Real s = y[i];
      for (int k = i + 1; k < N; ++k) s = s - lu(i, k) * inv(k, c);
      inv(i, c) = s / lu(i, i);
    }
  }
  return inv;
}

}  // namespace kfo
