// The per-particle 2x2 log-conformation update of predictor_mod.F90:141-331, FP64.
#pragma once

#include "vb_common.h"

namespace vb {

// A value with one directional derivative (forward-mode dual number), for the log-conformation
// map's Jacobian.  Comparisons and branches act on the value, so a derivative is that of the branch
// the value takes; the clamped square root (argument set to exactly 0) has derivative 0.  The
// overloads are hidden friends: the double instantiations below call the ordinary sqrt/log/exp.
struct Dual {
  double v, d;
  __host__ __device__ Dual(double v_ = 0.0, double d_ = 0.0) : v(v_), d(d_) {}
  friend __host__ __device__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
  friend __host__ __device__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
  friend __host__ __device__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
  friend __host__ __device__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
  friend __host__ __device__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
  }
  friend __host__ __device__ bool operator<=(Dual a, double b) { return a.v <= b; }
  friend __host__ __device__ bool operator!=(Dual a, double b) { return a.v != b; }
  friend __host__ __device__ Dual sqrt(Dual a) {
    const double s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
  }
  friend __host__ __device__ Dual log(Dual a) { return Dual(::log(a.v), a.d / a.v); }
  friend __host__ __device__ Dual exp(Dual a) {
    const double e = ::exp(a.v);
    return Dual(e, e * a.d);
  }
};

// Which branch of calc_eigens a matrix takes.  The reference takes the eigenvector branch when
// A21 != 0 and otherwise fixes R = I, L = (A11, A22) (predictor_mod.F90:280-288).  Where A21 = 0 and
// A11 > A22 with the discriminant unclamped (gap above ~3e-8), the map is smooth across A21 = 0 and
// the eigenvector branch gives the same values there (r11 = L1 - A22 = A11 - A22 > 0, r21 = 0, so
// R = I; L differs from (A11, A22) only by the rounding of its formula), but the A21 = 0 branch
// drops the derivative of R in A21 (dR = dA21 / (A11 - A22)).  So the derivative pass (Dual) takes
// the eigenvector branch there; the step's FP64 map keeps the reference's test.  Left non-smooth:
// A21 = 0 with A11 <= A22 (the map's two sides differ by the 1e-10 of the rotation-rate denominator)
// and the clamped discriminant at A21 = 0.
__host__ __device__ inline bool eigenvector_branch(const double* A, double) { return A[2] != 0.0; }
__host__ __device__ inline bool eigenvector_branch(const Dual* A, Dual tmp) {
  return A[2].v != 0.0 || (A[0].v > A[3].v && tmp.v > 0.0);
}

// calc_eigens, predictor_mod.F90:253-290 (A, R as a11,a12,a21,a22)
template <class T>
__host__ __device__ inline void calc_eigens(const T* A, T* L, T* R) {
  T tr = A[0] + A[3];
  T det = A[0] * A[3] - A[2] * A[1];
  T tmp = 0.25 * tr * tr - det;
  if (tmp <= kEps) tmp = 0.0;
  T sq = sqrt(tmp);
  L[0] = 0.5 * tr + sq;
  L[1] = 0.5 * tr - sq;
  if (eigenvector_branch(A, tmp)) {
    T r11 = L[0] - A[3], r21 = A[2];
    T nrm = sqrt(r11 * r11 + r21 * r21);
    r11 = r11 / nrm; r21 = r21 / nrm;
    R[0] = r11; R[2] = r21; R[1] = -r21; R[3] = r11;
  } else {
    L[0] = A[0]; L[1] = A[3];
    R[0] = 1.0; R[3] = 1.0; R[1] = 0.0; R[2] = 0.0;
  }
}

template <class T>
__host__ __device__ inline void mm2(const T* A, const T* B, T* C) {
  T c11 = A[0] * B[0] + A[1] * B[2], c12 = A[0] * B[1] + A[1] * B[3];
  T c21 = A[2] * B[0] + A[3] * B[2], c22 = A[2] * B[1] + A[3] * B[3];
  C[0] = c11; C[1] = c12; C[2] = c21; C[3] = c22;
}

// R diag(d1,d2) R^T, summed as matmul(R, matmul(D, transpose(R))) with the zero products dropped
template <class T>
__host__ __device__ inline void rdrt(const T* R, T d1, T d2, T* o) {
  T m11 = d1 * R[0], m12 = d1 * R[2], m21 = d2 * R[1], m22 = d2 * R[3];
  o[0] = R[0] * m11 + R[1] * m21; o[1] = R[0] * m12 + R[1] * m22;
  o[2] = R[2] * m11 + R[3] * m21; o[3] = R[2] * m12 + R[3] * m22;
}

// constitutive_eqn_evolve for one particle (:165-245) with f_relaxation_integrate (:318-331); T is
// double on the step's path and Dual for the map's Jacobian (the same expressions either way)
template <class T>
__host__ __device__ inline void logconf(const T* tau, T trc_prev, const T* G, T dt, T wi, double eps_ptt,
                                        double kick, bool use_kick, T* to, T* trc_out) {
  T A[4] = {tau[0] / 1.0 + 1.0, tau[1] / 1.0 + 0.0, tau[2] / 1.0 + 0.0, tau[3] / 1.0 + 1.0};
  T L[2], R[4], Psi0[4];
  calc_eigens(A, L, R);
  rdrt(R, log(L[0]), log(L[1]), Psi0);
  T RT[4] = {R[0], R[2], R[1], R[3]}, GR[4], M[4];
  mm2(G, R, GR);
  mm2(RT, GR, M);
  T om = (L[1] * M[1] + L[0] * M[2]) / (L[1] - L[0] + 1.0e-10);
  T B[4];
  rdrt(R, M[0], M[3], B);
  if (use_kick) { B[0] = B[0] + kick; B[1] = B[1] + kick; B[2] = B[2] + kick; B[3] = B[3] + kick; }
  T w11 = om * R[1], w12 = om * R[3], w21 = -om * R[0], w22 = -om * R[2];
  T Om[4] = {R[0] * w11 + R[1] * w21, R[0] * w12 + R[1] * w22, R[2] * w11 + R[3] * w21,
             R[2] * w12 + R[3] * w22};
  T OP[4], PO[4], Psi[4];
  mm2(Om, Psi0, OP);
  mm2(Psi0, Om, PO);
  for (int q = 0; q < 4; q++) Psi[q] = Psi0[q] + dt * ((2.0 * B[q] + OP[q]) - PO[q]);
  calc_eigens(Psi, L, R);
  T Aa[4];
  rdrt(R, exp(L[0]), exp(L[1]), Aa);
  T aR = 1.0 + eps_ptt * (trc_prev - 2.0);
  T t1 = exp(-aR * 1.0 * dt / wi), t2 = (1.0 - t1) / 1.0;
  Aa[0] = Aa[0] * t1 + t2; Aa[1] = Aa[1] * t1; Aa[2] = Aa[2] * t1; Aa[3] = Aa[3] * t1 + t2;
  *trc_out = Aa[0] + Aa[3];
  to[0] = Aa[0] - 1.0; to[1] = Aa[1] - 0.0; to[2] = Aa[2] - 0.0; to[3] = Aa[3] - 1.0;
}

}  // namespace vb
