// FP64 CPU oracle for the linear-tetmesh hyperelastic hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls this file: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as the
// checker (or as the timed CPU baseline), never as the thing measured on the GPU.
//
// It is a plain C++17 restatement (no Eigen) of the reference algorithm, written from the
// reference's source read as text.  Each function cites the reference file:line it follows.
// Conventions (reference include/flatten.h:59-72, include/flatten_multiply.h:52-75):
//   vec(F) is column-major: F(i,j) at i + 3j.
//   dphidX row e is the column-major flattening of the 4x3 matrix Phi: Phi(m,j) at e + (m+4j)*nT.
//   B(i+3j, l+3m) = delta_il * Phi(m,j), so vec(F) = B * qe, qe = [x0;x1;x2;x3].
//   E is nT x 4 column-major int32, params nT x 2 column-major, q interleaved xyz.
//
// What pins each part (the tests do the pinning):
//   pattern + assembly fold + signed volume : reference fixtures M_coarse_bunny.txt /
//                                             H_neohookean_coarser_bunny.txt (tests/test_oracle_fixtures.py)
//   legacy isochoric NH                     : H_neohookean_coarser_bunny.txt
//   active NH, StVK, corot, ARAP            : torch-FP64 autograd golden vectors (tests/golden/),
//                                             finite differences, invariants
//   dphidX                                  : F = diag(2,4,3) known answer (tests/test_constitutive_model.cpp:42-66)
//   PSD fix                                 : numpy eigh spectral definition
//   line search / Newton                    : closed-form toy problems
//   SVD perturbation at near-repeated singular values, Eigen SparseLU / SelfAdjointEigenSolver
//   rounding                                : parity unpinned (tolerances only)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#include <limits>

namespace {

enum Model { NEOHOOKEAN = 0, STVK = 1, COROTATIONAL = 2, ARAP = 3, NEOHOOKEAN_LEGACY = 4, NEOHOOKEAN_MUSCLE = 5 };

inline double det3(const double* F) {
  return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
         F[6] * (F[1] * F[5] - F[4] * F[2]);
}

// cofactor matrix cof(F) = det(F) F^{-T}, column-major
inline void cofactor(const double* F, double* c) {
  auto f = [&](int i, int j) { return F[i + 3 * j]; };
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      c[i + 3 * j] = f(i1, j1) * f(i2, j2) - f(i1, j2) * f(i2, j1);
    }
}

// Levi-Civita
inline int eps3(int a, int b, int c) {
  if (a == b || b == c || a == c) return 0;
  return ((a + 1) % 3 == b) ? 1 : -1;
}

// ---------------------------------------------------------------------------------------------
// hyper-dual numbers: exact first and mixed second derivatives (used for the legacy NH form,
// whose Hessian the H_neohookean_coarser_bunny.txt fixture pins)
struct HD {
  double a, b, c, d;  // f, df/dx_i, df/dx_j, d2f/dx_i dx_j
};
inline HD operator+(HD x, HD y) { return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
inline HD operator-(HD x, HD y) { return {x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
inline HD operator*(HD x, HD y) {
  return {x.a * y.a, x.a * y.b + x.b * y.a, x.a * y.c + x.c * y.a,
          x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a};
}
inline HD operator*(double s, HD x) { return {s * x.a, s * x.b, s * x.c, s * x.d}; }
inline HD cst(double v) { return {v, 0, 0, 0}; }
template <class Fn, class D1, class D2>
inline HD chain(HD x, Fn f, D1 df, D2 d2f) {
  double f1 = df(x.a), f2 = d2f(x.a);
  return {f(x.a), f1 * x.b, f1 * x.c, f1 * x.d + f2 * x.b * x.c};
}
inline HD hd_log(HD x) {
  return chain(x, [](double v) { return std::log(v); }, [](double v) { return 1.0 / v; },
               [](double v) { return -1.0 / (v * v); });
}
inline HD hd_recip(HD x) {
  return chain(x, [](double v) { return 1.0 / v; }, [](double v) { return -1.0 / (v * v); },
               [](double v) { return 2.0 / (v * v * v); });
}
inline HD hd_cbrt(HD x) {
  return chain(x, [](double v) { return std::cbrt(v); },
               [](double v) { double c = std::cbrt(v); return 1.0 / (3.0 * c * c); },
               [](double v) { double c = std::cbrt(v); return -2.0 / (9.0 * c * c * c * c * c); });
}

template <class T>
T det3t(const T* F) {
  return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
         F[6] * (F[1] * F[5] - F[4] * F[2]);
}

// legacy isochoric Neo-Hookean: psi = C (J^{-2/3} tr(F^T F) - 3) + D (J-1)^2.
// Reference: src/psi_neohookean_F.cpp:8-14 (commented form), stablePow include/EigenTypes.h:100-102.
HD psi_nh_legacy_hd(const HD* F, double C, double D) {
  HD J = det3t(F);
  HD c = hd_cbrt(J);
  HD J23 = hd_recip(c * c);
  HD I1 = cst(0);
  for (int k = 0; k < 9; ++k) I1 = I1 + F[k] * F[k];
  HD Jm1 = J - cst(1.0);
  return C * (J23 * I1 - cst(3.0)) + D * (Jm1 * Jm1);
}

// ---------------------------------------------------------------------------------------------
// 3x3 SVD (FP64) in the rotation convention of igl::svd3x3 (McAdams et al.): U, V rotations,
// singular values sorted descending by magnitude, the smallest carries the sign of det F.
// Jacobi eigen-decomposition of F^T F, then U = F V / sigma with a cross product for the last
// column. The reference runs this in float (src/psi_stretch_F.cpp:13-18); the oracle is FP64.
void jacobi_sym3(double* A /*3x3 sym, col-major, destroyed*/, double* V /*col-major*/, double* w) {
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = A[3] * A[3] + A[6] * A[6] + A[7] * A[7];
    double dg = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-36 * dg || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double apq = A[p + 3 * q];
        if (apq == 0.0) continue;
        double app = A[p + 3 * p], aqq = A[q + 3 * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A = A J
          double akp = A[k + 3 * p], akq = A[k + 3 * q];
          A[k + 3 * p] = c * akp - s * akq;
          A[k + 3 * q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A = J^T A
          double apk = A[p + 3 * k], aqk = A[q + 3 * k];
          A[p + 3 * k] = c * apk - s * aqk;
          A[q + 3 * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k + 3 * p], vkq = V[k + 3 * q];
          V[k + 3 * p] = c * vkp - s * vkq;
          V[k + 3 * q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = A[i + 3 * i];
}

void svd3(const double* F, double* U, double* S, double* V) {
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += F[k + 3 * i] * F[k + 3 * j];
      A[i + 3 * j] = s;
    }
  double Vt[9], w[3];
  jacobi_sym3(A, Vt, w);
  int idx[3] = {0, 1, 2};
  std::sort(idx, idx + 3, [&](int a, int b) { return w[a] > w[b]; });
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) V[r + 3 * c] = Vt[r + 3 * idx[c]];
  // make V a rotation
  if (det3(V) < 0)
    for (int r = 0; r < 3; ++r) V[r + 6] = -V[r + 6];
  double fv[3][3];  // columns F v_c
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += F[r + 3 * k] * V[k + 3 * c];
      fv[c][r] = s;
    }
  double n0 = std::sqrt(fv[0][0] * fv[0][0] + fv[0][1] * fv[0][1] + fv[0][2] * fv[0][2]);
  double u0[3], u1[3], u2[3];
  if (n0 > 0) {
    for (int r = 0; r < 3; ++r) u0[r] = fv[0][r] / n0;
  } else {
    u0[0] = 1; u0[1] = 0; u0[2] = 0;
  }
  double d01 = u0[0] * fv[1][0] + u0[1] * fv[1][1] + u0[2] * fv[1][2];
  for (int r = 0; r < 3; ++r) u1[r] = fv[1][r] - d01 * u0[r];
  double n1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (n1 > 1e-300) {
    for (int r = 0; r < 3; ++r) u1[r] /= n1;
  } else {  // any unit vector orthogonal to u0
    double e[3] = {0, 0, 0};
    int k = (std::fabs(u0[0]) < 0.6) ? 0 : ((std::fabs(u0[1]) < 0.6) ? 1 : 2);
    e[k] = 1;
    double dd = u0[k];
    for (int r = 0; r < 3; ++r) u1[r] = e[r] - dd * u0[r];
    double nn = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    for (int r = 0; r < 3; ++r) u1[r] /= nn;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  for (int r = 0; r < 3; ++r) {
    U[r] = u0[r];
    U[r + 3] = u1[r];
    U[r + 6] = u2[r];
  }
  S[0] = n0;
  S[1] = u1[0] * fv[1][0] + u1[1] * fv[1][1] + u1[2] * fv[1][2];
  S[2] = u2[0] * fv[2][0] + u2[1] * fv[2][1] + u2[2] * fv[2][2];
}

// stretch energies in singular values.
// Reference: src/psi_corotational_S.cpp:5-9, src/dpsi_corotational_dS.cpp:6-13,
// src/d2psi_corotational_dS2.cpp:6-14 (params [lambda, mu]); src/psi_arap_S.cpp:5-11,
// src/dpsi_arap_dS.cpp:6-12, src/d2psi_arap_dS2.cpp:6-12 (k = params(0)).
double psiS(int model, const double* S, const double* p) {
  if (model == COROTATIONAL) {
    double t = S[0] + S[1] + S[2] - 3.0;
    return p[1] * ((S[0] - 1) * (S[0] - 1) + (S[1] - 1) * (S[1] - 1) + (S[2] - 1) * (S[2] - 1)) +
           0.5 * p[0] * t * t;
  }
  return p[0] * ((S[0] - 1) * (S[0] - 1) + (S[1] - 1) * (S[1] - 1) + (S[2] - 1) * (S[2] - 1));
}
void dpsiS(int model, const double* S, const double* p, double* d) {
  if (model == COROTATIONAL) {
    double t = S[0] + S[1] + S[2] - 3.0;
    for (int i = 0; i < 3; ++i) d[i] = 2.0 * p[1] * (S[i] - 1.0) + p[0] * t;
  } else {
    for (int i = 0; i < 3; ++i) d[i] = 2.0 * p[0] * (S[i] - 1.0);
  }
}
void d2psiS(int model, const double* S, const double* p, double* H /*3x3 col-major*/) {
  (void)S;
  for (int i = 0; i < 9; ++i) H[i] = 0;
  if (model == COROTATIONAL) {
    for (int i = 0; i < 9; ++i) H[i] = p[0];
    for (int i = 0; i < 3; ++i) H[i + 3 * i] += 2.0 * p[1];
  } else {
    for (int i = 0; i < 3; ++i) H[i + 3 * i] = 2.0 * p[0];
  }
}

// deterministic stand-in for Eigen::Matrix3f::Random() (uniform in [-1,1]); the reference draws
// from std::rand (src/d2psi_stretch_dF2.cpp:25-28), which no restatement can reproduce.
void perturbation_matrix(double* R) {
  uint64_t s = 0x5eedULL;
  for (int i = 0; i < 9; ++i) {
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    R[i] = 2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

// SVD derivatives. Reference: src/dsvd.cpp:5-57 (conservative: |d| < 1e-8 -> no rotation).
// dU[r][s], dV[r][s]: 3x3 col-major; dS[r][s]: 3-vector.
void dsvd(double dU[3][3][9], double dS[3][3][3], double dV[3][3][9], const double* U, const double* S,
          const double* V) {
  const double tol = 1e-8;
  double d01 = S[1] * S[1] - S[0] * S[0];
  double d02 = S[2] * S[2] - S[0] * S[0];
  double d12 = S[2] * S[2] - S[1] * S[1];
  const double inf = std::numeric_limits<double>::infinity();
  d01 = 1.0 / (std::fabs(d01) < tol ? inf : d01);
  d02 = 1.0 / (std::fabs(d02) < tol ? inf : d02);
  d12 = 1.0 / (std::fabs(d12) < tol ? inf : d12);
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) {
      double UVT[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) UVT[i + 3 * j] = U[r + 3 * i] * V[s + 3 * j];
      for (int i = 0; i < 3; ++i) dS[r][s][i] = UVT[4 * i];
      for (int i = 0; i < 3; ++i) UVT[4 * i] = 0.0;
      double tmp[9];
      // tmp = S.asDiagonal()*UVT + UVT^T*S.asDiagonal()
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tmp[i + 3 * j] = S[i] * UVT[i + 3 * j] + UVT[j + 3 * i] * S[j];
      double w01 = tmp[0 + 3 * 1] * d01, w02 = tmp[0 + 3 * 2] * d02, w12 = tmp[1 + 3 * 2] * d12;
      double W[9] = {0, -w01, -w02, w01, 0, -w12, w02, w12, 0};  // col-major of [[0,w01,w02],[-w01,0,w12],[-w02,-w12,0]]
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double a = 0;
          for (int k = 0; k < 3; ++k) a += V[i + 3 * k] * W[k + 3 * j];
          dV[r][s][i + 3 * j] = a;
        }
      // tmp = UVT*S.asDiagonal() + S.asDiagonal()*UVT^T
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) tmp[i + 3 * j] = UVT[i + 3 * j] * S[j] + S[i] * UVT[j + 3 * i];
      w01 = tmp[0 + 3 * 1] * d01;
      w02 = tmp[0 + 3 * 2] * d02;
      w12 = tmp[1 + 3 * 2] * d12;
      double W2[9] = {0, -w01, -w02, w01, 0, -w12, w02, w12, 0};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double a = 0;
          for (int k = 0; k < 3; ++k) a += U[i + 3 * k] * W2[k + 3 * j];
          dU[r][s][i + 3 * j] = a;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// constitutive models on F

// Reference: src/psi_neohookean_F.cpp:16-31 (active), src/psi_stvk_F.cpp:5-13,
// src/psi_stretch_F.cpp:7-22.
// muscle fibre of the reference's MATLAB neohookean_muscle gateways (matlab/linear_tetmesh_neohookean_muscle_q.cpp,
// _dq.cpp, _dq2.cpp:53-77): Neo-Hookean with params (p[0], p[1]) plus a fibre with stiffness p[2] along the
// direction (p[3], p[4], p[5]): psi_muscle_fiber_F = 0.5 a d^T F^T F d (src/psi_muscle_fiber_F.cpp:5-10),
// dpsi = a D^T D vec F (src/dpsi_muscle_fiber_dF.cpp), d2psi = a D^T D (src/d2psi_muscle_fiber_dF2.cpp), with
// D = flatten_multiply_right(d) the 3x9 matrix of vec(F) -> F d (include/flatten_multiply.h:44-88).
void fibre_D(const double* d, double* D /*3x9 col-major*/) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 9; ++c) D[r + 3 * c] = (c % 3 == r) ? d[c / 3] : 0.0;
}
double psi_fibre(const double* F, const double* p) {
  double D[27], Fd[3] = {0, 0, 0};
  fibre_D(p + 3, D);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 9; ++c) Fd[r] += D[r + 3 * c] * F[c];
  return 0.5 * p[2] * (Fd[0] * Fd[0] + Fd[1] * Fd[1] + Fd[2] * Fd[2]);
}
void d2psi_fibre(const double* p, double* H /*9x9 col-major*/) {
  double D[27];
  fibre_D(p + 3, D);
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b) {
      double s = 0;
      for (int r = 0; r < 3; ++r) s += D[r + 3 * a] * D[r + 3 * b];
      H[a + 9 * b] = p[2] * s;
    }
}

double psi_F(int model, const double* F, const double* p) {
  switch (model) {
    case NEOHOOKEAN_MUSCLE:
      return psi_F(NEOHOOKEAN, F, p) + psi_fibre(F, p);
    case NEOHOOKEAN: {
      double J = det3(F), I1 = 0;
      for (int k = 0; k < 9; ++k) I1 += F[k] * F[k];
      return p[0] * (I1 - 3.0 - 2.0 * std::log(J)) + p[1] * (J - 1.0) * (J - 1.0);
    }
    case NEOHOOKEAN_LEGACY: {
      HD Fh[9];
      for (int k = 0; k < 9; ++k) Fh[k] = cst(F[k]);
      return psi_nh_legacy_hd(Fh, p[0], p[1]).a;
    }
    case STVK: {
      double E[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += F[k + 3 * i] * F[k + 3 * j];
          E[i + 3 * j] = 0.5 * (s - (i == j ? 1.0 : 0.0));
        }
      double trEE = 0, trE = E[0] + E[4] + E[8];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) trEE += E[i + 3 * j] * E[j + 3 * i];
      return p[1] * trEE + 0.5 * p[0] * trE * trE;
    }
    default: {
      double U[9], S[3], V[9];
      svd3(F, U, S, V);
      return psiS(model, S, p);
    }
  }
}

// Reference: src/dpsi_neohookean_dF.cpp:43-66, src/dpsi_stvk_dF.cpp:5-35,
// src/dpsi_stretch_dF.cpp:7-47 (inversion flip of negative singular values when det <= -1e-10).
void dpsi_F(int model, const double* F, const double* p, double* P) {
  switch (model) {
    case NEOHOOKEAN_MUSCLE: {
      dpsi_F(NEOHOOKEAN, F, p, P);
      double H[81];
      d2psi_fibre(p, H);  // the fibre energy is quadratic in vec F: its gradient is (a D^T D) vec F
      for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 9; ++b) P[a] += H[a + 9 * b] * F[b];
      return;
    }
    case NEOHOOKEAN: {
      double c[9];
      cofactor(F, c);
      double J = det3(F);
      for (int k = 0; k < 9; ++k) P[k] = p[0] * (2.0 * F[k] - 2.0 * c[k] / J) + 2.0 * p[1] * (J - 1.0) * c[k];
      return;
    }
    case NEOHOOKEAN_LEGACY: {
      for (int k = 0; k < 9; ++k) {
        HD Fh[9];
        for (int l = 0; l < 9; ++l) Fh[l] = cst(F[l]);
        Fh[k].b = 1.0;
        P[k] = psi_nh_legacy_hd(Fh, p[0], p[1]).b;
      }
      return;
    }
    case STVK: {
      double E[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += F[k + 3 * i] * F[k + 3 * j];
          E[i + 3 * j] = 0.5 * (s - (i == j ? 1.0 : 0.0));
        }
      double trE = E[0] + E[4] + E[8];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += F[i + 3 * k] * E[k + 3 * j];
          P[i + 3 * j] = 2.0 * p[1] * s + p[0] * trE * F[i + 3 * j];
        }
      return;
    }
    default: {
      double U[9], S[3], V[9];
      svd3(F, U, S, V);
      double det = S[0] * S[1] * S[2];
      if (det <= -1e-10)
        for (int i = 0; i < 3; ++i)
          if (S[i] < 0) S[i] = -S[i];
      if (det3(U) <= 0)
        for (int r = 0; r < 3; ++r) U[r + 6] = -U[r + 6];
      if (det3(V) <= 0)
        for (int r = 0; r < 3; ++r) V[r + 6] = -V[r + 6];
      double d[3];
      dpsiS(model, S, p, d);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += U[i + 3 * k] * d[k] * V[j + 3 * k];
          P[i + 3 * j] = s;
        }
      return;
    }
  }
}

// 9x9 Hessian, col-major H[r + 9c].
// Reference: src/d2psi_neohookean_dF2.cpp:185-283 (active), :8-183 (legacy),
// src/d2psi_stvk_dF2.cpp:5-106, src/d2psi_stretch_dF2.cpp:7-47 (+ src/dsvd.cpp).
// ref_layout (stretch models only): reproduce the reference's row-major write
// (src/d2psi_stretch_dF2.cpp:36-45), i.e. the Hessian evaluated at F^T.
void d2psi_F(int model, const double* F, const double* p, int ref_layout, double* H) {
  switch (model) {
    case NEOHOOKEAN_MUSCLE: {
      d2psi_F(NEOHOOKEAN, F, p, ref_layout, H);
      double Hf[81];
      d2psi_fibre(p, Hf);
      for (int k = 0; k < 81; ++k) H[k] += Hf[k];
      return;
    }
    case NEOHOOKEAN: {
      // d2psi = 2C I + (2C/J^2 + 2D) g g^T + (2D(J-1) - 2C/J) d2J,  g = vec(cof F)
      double c[9];
      cofactor(F, c);
      double J = det3(F);
      double a = 2.0 * p[0] / (J * J) + 2.0 * p[1];
      double b = 2.0 * p[1] * (J - 1.0) - 2.0 * p[0] / J;
      for (int r = 0; r < 9; ++r)
        for (int s = 0; s < 9; ++s) {
          int i = r % 3, j = r / 3, k = s % 3, l = s / 3;
          double d2J = 0;
          for (int e = 0; e < 3; ++e)
            for (int f = 0; f < 3; ++f) {
              int s1 = eps3(i, k, e), s2 = eps3(j, l, f);
              if (s1 && s2) d2J += s1 * s2 * F[e + 3 * f];
            }
          H[r + 9 * s] = (r == s ? 2.0 * p[0] : 0.0) + a * c[r] * c[s] + b * d2J;
        }
      return;
    }
    case NEOHOOKEAN_LEGACY: {
      for (int r = 0; r < 9; ++r)
        for (int s = r; s < 9; ++s) {
          HD Fh[9];
          for (int l = 0; l < 9; ++l) Fh[l] = cst(F[l]);
          Fh[r].b = 1.0;
          Fh[s].c = 1.0;
          double v = psi_nh_legacy_hd(Fh, p[0], p[1]).d;
          H[r + 9 * s] = v;
          H[s + 9 * r] = v;
        }
      return;
    }
    case STVK: {
      // dP_ij/dF_kl = 2mu d_ik E_lj + mu (F_il F_kj + (FF^T)_ik d_jl) + lambda (F_kl F_ij + trE d_ik d_jl)
      double E[9], FFt[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0, t = 0;
          for (int k = 0; k < 3; ++k) {
            s += F[k + 3 * i] * F[k + 3 * j];
            t += F[i + 3 * k] * F[j + 3 * k];
          }
          E[i + 3 * j] = 0.5 * (s - (i == j ? 1.0 : 0.0));
          FFt[i + 3 * j] = t;
        }
      double trE = E[0] + E[4] + E[8];
      const double lam = p[0], mu = p[1];
      for (int r = 0; r < 9; ++r)
        for (int s = 0; s < 9; ++s) {
          int i = r % 3, j = r / 3, k = s % 3, l = s / 3;
          double v = 0;
          if (i == k) v += 2.0 * mu * E[l + 3 * j];
          v += mu * F[i + 3 * l] * F[k + 3 * j];
          if (j == l) v += mu * FFt[i + 3 * k];
          v += lam * F[k + 3 * l] * F[i + 3 * j];
          if (i == k && j == l) v += lam * trE;
          H[r + 9 * s] = v;
        }
      return;
    }
    default: {
      double Fp[9];
      std::memcpy(Fp, F, sizeof(Fp));
      double U[9], S[3], V[9];
      svd3(Fp, U, S, V);
      if (std::fabs(S[0] - S[1]) < 1e-5 || std::fabs(S[1] - S[2]) < 1e-5 || std::fabs(S[0] - S[2]) < 1e-5) {
        double R[9];
        perturbation_matrix(R);
        for (int i = 0; i < 9; ++i) Fp[i] += R[i] * 1e-5;
        svd3(Fp, U, S, V);
      }
      double Plam[3], Plam2[9];
      dpsiS(model, S, p, Plam);
      d2psiS(model, S, p, Plam2);
      static thread_local double dU[3][3][9], dS[3][3][3], dV[3][3][9];
      dsvd(dU, dS, dV, U, S, V);
      for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
          double pv[3];
          for (int i = 0; i < 3; ++i) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += Plam2[i + 3 * k] * dS[r][s][k];
            pv[i] = a;
          }
          double M[9];  // dP/dF_rs
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              double a = 0;
              for (int k = 0; k < 3; ++k)
                a += dU[r][s][i + 3 * k] * Plam[k] * V[j + 3 * k] + U[i + 3 * k] * Plam[k] * dV[r][s][j + 3 * k] +
                     U[i + 3 * k] * pv[k] * V[j + 3 * k];
              M[i + 3 * j] = a;
            }
          if (ref_layout) {
            // ddw.row(3r+s) = row-major flatten of M: ddw(3r+s, 3i+j) = M(i,j)
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) H[(3 * r + s) + 9 * (3 * i + j)] = M[i + 3 * j];
          } else {
            // true Hessian: H(vec(F)_{i+3j}, vec(F)_{r+3s}) = dP_ij / dF_rs
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) H[(i + 3 * j) + 9 * (r + 3 * s)] = M[i + 3 * j];
          }
        }
      return;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// element level.  Reference: src/linear_tet_neohookean_{q,dq,dq2}.cpp and siblings; B from
// include/flatten_multiply.h:52-75 (index functor).

inline double Bcoef(const double* phi /*4x3 col-major: phi[m+4j]*/, int row, int col) {
  // B(i+3j, l+3m) = delta_il * Phi(m,j)
  int i = row % 3, j = row / 3, l = col % 3, m = col / 3;
  return (l != i) ? 0.0 : phi[m + 4 * j];
}

struct Tet {
  double phi[12];
  double qe[12];
  double F[9];
  double vol;
  double p[6];
};

// parameter columns: 6 for the muscle model with one model for all tets, else 2
inline int cols_of(int model, const unsigned char* ids) { return (!ids && model == NEOHOOKEAN_MUSCLE) ? 6 : 2; }

void load_tet(Tet& t, int e, int nT, const int* E, const double* q, const double* dphidX, const double* vol,
              const double* params, int cols = 2) {
  for (int k = 0; k < 12; ++k) t.phi[k] = dphidX[e + (int64_t)k * nT];
  for (int a = 0; a < 4; ++a) {
    int v = E[e + (int64_t)a * nT];
    for (int c = 0; c < 3; ++c) t.qe[3 * a + c] = q[3 * (int64_t)v + c];
  }
  // F = unflatten<3,3>(B * qe)
  for (int r = 0; r < 9; ++r) {
    double s = 0;
    for (int c = 0; c < 12; ++c) s += Bcoef(t.phi, r, c) * t.qe[c];
    t.F[r] = s;
  }
  t.vol = vol[e];
  for (int c = 0; c < 6; ++c) t.p[c] = (c < cols) ? params[e + (int64_t)c * nT] : 0.0;
}

// K = B^T d2psi B * vol  (12x12 col-major); B materialised once per element, as the reference's
// Matrix<9,12> B = flatten_multiply_right<...>(dphidX) is (src/linear_tet_neohookean_dq2.cpp:12-14)
void element_hessian(const Tet& t, int model, int ref_layout, double* K) {
  double H[81], HB[108], B[108];
  for (int c = 0; c < 12; ++c)
    for (int r = 0; r < 9; ++r) B[r + 9 * c] = Bcoef(t.phi, r, c);
  d2psi_F(model, t.F, t.p, ref_layout, H);
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < 12; ++c) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += H[r + 9 * k] * B[k + 9 * c];
      HB[r + 9 * c] = s;
    }
  for (int r = 0; r < 12; ++r)
    for (int c = 0; c < 12; ++c) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += B[k + 9 * r] * HB[k + 9 * c];
      K[r + 12 * c] = s * t.vol;
    }
}

void element_gradient(const Tet& t, int model, double* ge) {
  double P[9];
  dpsi_F(model, t.F, t.p, P);
  for (int c = 0; c < 12; ++c) {
    double s = 0;
    for (int k = 0; k < 9; ++k) s += Bcoef(t.phi, k, c) * P[k];
    ge[c] = s * t.vol;
  }
}

// simple_psd_fix: eigenvalues < tol -> tol, then V diag V^T.  Reference: src/simple_psd_fix.cpp:5-21.
// Cyclic Jacobi eigen-decomposition of the n x n symmetric matrix (n = 12 for a tet).
void psd_fix(double* A, int n, double tol) {
  std::vector<double> a(A, A + n * n), V(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + n * i] = 1.0;
  double scale = 0;
  for (int i = 0; i < n * n; ++i) scale += a[i] * a[i];
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += a[p + n * q] * a[p + n * q];
    if (off <= 1e-32 * scale || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[p + n * q];
        if (apq == 0.0) continue;
        double theta = (a[q + n * q] - a[p + n * p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[k + n * p], akq = a[k + n * q];
          a[k + n * p] = c * akp - s * akq;
          a[k + n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[p + n * k], aqk = a[q + n * k];
          a[p + n * k] = c * apk - s * aqk;
          a[q + n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + n * p], vkq = V[k + n * q];
          V[k + n * p] = c * vkp - s * vkq;
          V[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<double> lam(n);
  for (int i = 0; i < n; ++i) lam[i] = (a[i + n * i] < tol) ? tol : a[i + n * i];
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += V[r + n * k] * lam[k] * V[c + n * k];
      A[r + n * c] = s;
    }
}

// The same spectral function by the method of Eigen's SelfAdjointEigenSolver (which the reference
// calls, src/simple_psd_fix.cpp:9): Householder tridiagonalisation with the reflections accumulated,
// then implicit symmetric QR steps with the Wilkinson shift (Golub & Van Loan, Matrix Computations,
// 4th ed., Algorithms 8.3.1-8.3.3), eigenvectors accumulated by the Givens rotations.  Fixed-size
// arrays, no allocation: the CPU-baseline timing build uses it; tests pin it against psd_fix and
// numpy's eigh.
template <int N>
void sym_eigen_ql(const double* A /*N x N col-major, symmetric*/, double* lam, double* Vout /*col-major*/) {
  double a[N][N], Q[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      a[i][j] = A[i + N * j];
      Q[i][j] = (i == j) ? 1.0 : 0.0;
    }
  // tridiagonalisation: for column k, the reflection H = I - 2 v v^T / v^T v (v zero above k+1) maps
  // a(k+1:, k) onto a multiple of e_{k+1}; A <- H A H by the symmetric rank-2 update, Q <- Q H
  for (int k = 0; k < N - 2; ++k) {
    double nrm = 0.0;
    for (int i = k + 1; i < N; ++i) nrm += a[i][k] * a[i][k];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    const double alpha = (a[k + 1][k] > 0.0) ? -nrm : nrm;
    double v[N] = {};
    for (int i = k + 1; i < N; ++i) v[i] = a[i][k];
    v[k + 1] -= alpha;
    double vv = 0.0;
    for (int i = k + 1; i < N; ++i) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    double pvec[N], K = 0.0;
    for (int i = 0; i < N; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < N; ++j) t += a[i][j] * v[j];
      pvec[i] = 2.0 * t / vv;
    }
    for (int i = k + 1; i < N; ++i) K += v[i] * pvec[i];
    K /= vv;
    double w[N];
    for (int i = 0; i < N; ++i) w[i] = pvec[i] - K * v[i];
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) a[i][j] -= v[i] * w[j] + w[i] * v[j];
    for (int i = 0; i < N; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < N; ++j) t += Q[i][j] * v[j];
      t *= 2.0 / vv;
      for (int j = k + 1; j < N; ++j) Q[i][j] -= t * v[j];
    }
  }
  double d[N], e[N];
  for (int i = 0; i < N; ++i) d[i] = a[i][i];
  for (int i = 0; i + 1 < N; ++i) e[i] = a[i + 1][i];
  e[N - 1] = 0.0;
  // implicit symmetric QR on the unreduced trailing block [l, m]
  const double eps = std::numeric_limits<double>::epsilon();
  int guard = 0;
  for (int m = N - 1; m > 0 && guard < 60 * N;) {
    if (std::fabs(e[m - 1]) <= eps * (std::fabs(d[m - 1]) + std::fabs(d[m]))) {
      e[m - 1] = 0.0;
      --m;
      continue;
    }
    int l = m - 1;
    while (l > 0 && std::fabs(e[l - 1]) > eps * (std::fabs(d[l - 1]) + std::fabs(d[l]))) --l;
    ++guard;
    const double dd = 0.5 * (d[m - 1] - d[m]), b = e[m - 1];
    const double den = dd + ((dd >= 0.0) ? 1.0 : -1.0) * std::hypot(dd, b);
    const double mu = d[m] - ((den != 0.0) ? b * b / den : 0.0);
    double x = d[l] - mu, z = e[l], bulge = 0.0;
    for (int k = l; k < m; ++k) {
      const double r = std::hypot(x, z);
      const double c = (r != 0.0) ? x / r : 1.0, sn = (r != 0.0) ? z / r : 0.0;
      if (k > l) e[k - 1] = r;
      const double ak = d[k], bk = e[k], ck = d[k + 1];
      d[k] = c * c * ak + 2.0 * c * sn * bk + sn * sn * ck;
      d[k + 1] = sn * sn * ak - 2.0 * c * sn * bk + c * c * ck;
      e[k] = (c * c - sn * sn) * bk + c * sn * (ck - ak);
      if (k + 1 < m) {
        bulge = sn * e[k + 1];
        e[k + 1] *= c;
      }
      for (int i = 0; i < N; ++i) {
        const double qk = Q[i][k], qk1 = Q[i][k + 1];
        Q[i][k] = c * qk + sn * qk1;
        Q[i][k + 1] = -sn * qk + c * qk1;
      }
      x = e[k];
      z = bulge;
    }
  }
  for (int i = 0; i < N; ++i) {
    lam[i] = d[i];
    for (int j = 0; j < N; ++j) Vout[i + N * j] = Q[i][j];
  }
}

template <int N>
void psd_fix_ql(double* A, double tol) {
  double lam[N], V[N * N];
  sym_eigen_ql<N>(A, lam, V);
  for (int k = 0; k < N; ++k) lam[k] = (lam[k] < tol) ? tol : lam[k];
  for (int c = 0; c < N; ++c)
    for (int r = c; r < N; ++r) {
      double sum = 0.0;
      for (int k = 0; k < N; ++k) sum += V[r + N * k] * lam[k] * V[c + N * k];
      A[r + N * c] = sum;
      A[c + N * r] = sum;
    }
}

// ---------------------------------------------------------------------------------------------
// assembly: 144 triplets per tet, then setFromTriplets semantics (duplicates summed left to right
// in insertion order = ascending element order, explicit zeros kept, column-major compressed,
// sorted inner indices).  Reference: src/assemble.cpp:36-97.
struct Triplet {
  int r, c;
  double v;
};

struct CSC {
  int64_t n = 0, nnz = 0;
  std::vector<int32_t> colptr, rowind;
  std::vector<double> vals;
};

void set_from_triplets(CSC& M, int64_t n, const std::vector<Triplet>& trips) {
  M.n = n;
  const int64_t nt = (int64_t)trips.size();
  // two stable counting sorts (by row, then by column): within a column the entries end up sorted
  // by row, and entries with the same (row, column) stay in insertion order -- so the fold below
  // sums duplicates left to right in ascending element order, as setFromTriplets does
  std::vector<int64_t> cnt(n + 1, 0);
  for (auto& t : trips) cnt[t.r + 1]++;
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> byrow(nt);
  for (int64_t k = 0; k < nt; ++k) byrow[cnt[trips[k].r]++] = k;
  std::fill(cnt.begin(), cnt.end(), 0);
  for (auto& t : trips) cnt[t.c + 1]++;
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> colstart(cnt.begin(), cnt.end());
  std::vector<int64_t> order(nt);
  for (int64_t k = 0; k < nt; ++k) order[cnt[trips[byrow[k]].c]++] = byrow[k];
  M.colptr.assign(n + 1, 0);
  M.rowind.clear();
  M.vals.clear();
  M.rowind.reserve(nt / 3);
  M.vals.reserve(nt / 3);
  for (int64_t c = 0; c < n; ++c) {
    int64_t k = colstart[c];
    const int64_t kend = colstart[c + 1];
    while (k < kend) {
      const int r = trips[order[k]].r;
      double s = trips[order[k]].v;
      int64_t k2 = k + 1;
      while (k2 < kend && trips[order[k2]].r == r) s += trips[order[k2++]].v;
      M.rowind.push_back(r);
      M.vals.push_back(s);
      k = k2;
    }
    M.colptr[c + 1] = (int32_t)M.rowind.size();
  }
  M.nnz = (int64_t)M.rowind.size();
}

void push_element(std::vector<Triplet>& trips, const int* E, int nT, int e, const double* K) {
  // reference loop order: for jj (col), for ii (row) -- src/assemble.cpp:68-74
  for (int jj = 0; jj < 12; ++jj)
    for (int ii = 0; ii < 12; ++ii)
      trips.push_back({3 * E[e + (int64_t)(ii / 3) * nT] + ii % 3, 3 * E[e + (int64_t)(jj / 3) * nT] + jj % 3,
                       K[ii + 12 * jj]});
}

inline int model_of(int model, const unsigned char* ids, int e) { return ids ? (int)ids[e] : model; }

}  // namespace

// =============================================================================================
// C ABI for the ctypes harness (tests/, bench.py cpu_baseline leg)
extern "C" {

// dphidX (nT x 12 col-major) and signed volume.  Reference: src/linear_tet_dphi_dX.cpp:5-18,
// include/EigenTypes.h:105-120 (inverse33), volume as igl::volume: -(a-d).((b-d)x(c-d))/6.
int or_dphidX_volume(int nV, const double* V /*nV x 3 col-major*/, int nT, const int* E, double* dphidX,
                     double* vol) {
  for (int e = 0; e < nT; ++e) {
    double x[4][3];
    for (int a = 0; a < 4; ++a) {
      int v = E[e + (int64_t)a * nT];
      if (v < 0 || v >= nV) return -1;
      for (int c = 0; c < 3; ++c) x[a][c] = V[v + (int64_t)c * nV];
    }
    // A: rows are edges x_m - x_0, then transposed -> columns are edges (Dm)
    double A[9];
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 3; ++c) A[c + 3 * m] = x[m + 1][c] - x[0][c];
    // inverse33 of A (cofactor / determinant)
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
                 A[6] * (A[1] * A[5] - A[4] * A[2]);
    double inv = 1.0 / det;
    double c[9];
    cofactor(A, c);  // c = det A^{-T}
    double Ainv[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Ainv[i + 3 * j] = c[j + 3 * i] * inv;
    // rows 1..3 of Phi = Ainv rows; row 0 = -sum
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int m = 0; m < 3; ++m) {
        dphidX[e + (int64_t)((m + 1) + 4 * j) * nT] = Ainv[m + 3 * j];
        s += Ainv[m + 3 * j];
      }
      dphidX[e + (int64_t)(0 + 4 * j) * nT] = -s;
    }
    double ad[3], bd[3], cd[3];
    for (int k = 0; k < 3; ++k) {
      ad[k] = x[0][k] - x[3][k];
      bd[k] = x[1][k] - x[3][k];
      cd[k] = x[2][k] - x[3][k];
    }
    double cr[3] = {bd[1] * cd[2] - bd[2] * cd[1], bd[2] * cd[0] - bd[0] * cd[2], bd[0] * cd[1] - bd[1] * cd[0]};
    vol[e] = -(ad[0] * cr[0] + ad[1] * cr[1] + ad[2] * cr[2]) / 6.0;
  }
  return 0;
}

double or_psi(int model, const double* F, const double* p) { return psi_F(model, F, p); }
void or_dpsi(int model, const double* F, const double* p, double* P) { dpsi_F(model, F, p, P); }
void or_d2psi(int model, const double* F, const double* p, int ref_layout, double* H) {
  d2psi_F(model, F, p, ref_layout, H);
}
void or_svd3(const double* F, double* U, double* S, double* V) { svd3(F, U, S, V); }
void or_psd_fix(double* A, int n, double tol) { psd_fix(A, n, tol); }

// per-element 12x12 (col-major) for tet e, optional PSD
void or_element_hessian(int model, int nT, int e, const int* E, const double* q, const double* dphidX,
                        const double* vol, const double* params, int ref_layout, int psd, double psd_tol,
                        double* K) {
  Tet t;
  load_tet(t, e, nT, E, q, dphidX, vol, params, cols_of(model, nullptr));
  element_hessian(t, model, ref_layout, K);
  if (psd) psd_fix(K, 12, psd_tol);
}

// Energy: fold from element 0.  Reference: src/linear_tetmesh_neohookean_q.cpp:8-24,
// src/assemble.cpp:134-150 (undefined for nT == 0; the oracle returns 0).
double or_tetmesh_q(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                    const double* vol, const double* params, const unsigned char* model_ids) {
  (void)nV;
  double energy = 0.0;
  for (int e = 0; e < nT; ++e) {
    Tet t;
    load_tet(t, e, nT, E, q, dphidX, vol, params, cols_of(model, model_ids));
    double v = psi_F(model_of(model, model_ids, e), t.F, t.p) * t.vol;
    energy = (e == 0) ? v : energy + v;
  }
  return energy;
}

// Gradient: zero-init, += in ascending element order.  Reference: src/linear_tetmesh_neohookean_dq.cpp:5-22,
// src/assemble.cpp:99-132, src/linear_tet_neohookean_dq.cpp:5-22.
void or_tetmesh_dq(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                   const double* vol, const double* params, const unsigned char* model_ids, double* g) {
  std::memset(g, 0, sizeof(double) * 3 * (size_t)nV);
  for (int e = 0; e < nT; ++e) {
    Tet t;
    load_tet(t, e, nT, E, q, dphidX, vol, params, cols_of(model, model_ids));
    double ge[12];
    element_gradient(t, model_of(model, model_ids, e), ge);
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) g[3 * (int64_t)E[e + (int64_t)a * nT] + c] += ge[3 * a + c];
  }
}

struct or_csc {
  int64_t n, nnz;
  int32_t* colptr;
  int32_t* rowind;
  double* vals;
};

static or_csc* export_csc(CSC& M) {
  or_csc* o = (or_csc*)std::malloc(sizeof(or_csc));
  o->n = M.n;
  o->nnz = M.nnz;
  o->colptr = (int32_t*)std::malloc(sizeof(int32_t) * (M.n + 1));
  o->rowind = (int32_t*)std::malloc(sizeof(int32_t) * (M.nnz > 0 ? M.nnz : 1));
  o->vals = (double*)std::malloc(sizeof(double) * (M.nnz > 0 ? M.nnz : 1));
  std::memcpy(o->colptr, M.colptr.data(), sizeof(int32_t) * (M.n + 1));
  if (M.nnz) {
    std::memcpy(o->rowind, M.rowind.data(), sizeof(int32_t) * M.nnz);
    std::memcpy(o->vals, M.vals.data(), sizeof(double) * M.nnz);
  }
  return o;
}

void or_csc_free(or_csc* o) {
  if (!o) return;
  std::free(o->colptr);
  std::free(o->rowind);
  std::free(o->vals);
  std::free(o);
}

// Hessian with the optional simple_psd_fix callback.  Reference: src/linear_tetmesh_neohookean_dq2.cpp:8-28,
// src/assemble.cpp:36-97, src/simple_psd_fix.cpp:5-21.
or_csc* or_tetmesh_dq2(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                       const double* vol, const double* params, const unsigned char* model_ids, int ref_layout,
                       int psd, double psd_tol) {
  std::vector<Triplet> trips;
  trips.reserve((size_t)nT * 144);
  double K[144];
  for (int e = 0; e < nT; ++e) {
    Tet t;
    load_tet(t, e, nT, E, q, dphidX, vol, params, cols_of(model, model_ids));
    element_hessian(t, model_of(model, model_ids, e), ref_layout, K);
    if (psd) psd_fix(K, 12, psd_tol);
    push_element(trips, E, nT, e, K);
  }
  CSC M;
  set_from_triplets(M, 3 * (int64_t)nV, trips);
  return export_csc(M);
}

// The CPU baseline's timing mode of the Hessian above: the same per-element algorithm (index-functor
// B, dense B^T d2psi B, 144 triplets per tet) with simple_psd_fix by tridiagonalisation + implicit QR
// (fixed-size, no allocation), elements split over nthreads OpenMP threads in contiguous ranges.
// Each tet writes its 144 triplets at offset 144 e of one preallocated list, so the list is in
// element order whatever the thread count, then one serial setFromTriplets fold (Eigen's is serial).
// times (optional): [element and triplet seconds, setFromTriplets seconds].
or_csc* or_tetmesh_dq2_timed(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                             const double* vol, const double* params, int psd, double psd_tol, int nthreads,
                             double* times) {
  std::vector<Triplet> trips((size_t)nT * 144);
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int e = 0; e < nT; ++e) {
    Tet t;
    double K[144];
    load_tet(t, e, nT, E, q, dphidX, vol, params, cols_of(model, nullptr));
    element_hessian(t, model, 1, K);
    if (psd) psd_fix_ql<12>(K, psd_tol);
    Triplet* out = trips.data() + (size_t)e * 144;
    for (int jj = 0; jj < 12; ++jj)
      for (int ii = 0; ii < 12; ++ii)
        *out++ = {3 * E[e + (int64_t)(ii / 3) * nT] + ii % 3, 3 * E[e + (int64_t)(jj / 3) * nT] + jj % 3,
                  K[ii + 12 * jj]};
  }
  const auto t1 = std::chrono::steady_clock::now();
  CSC M;
  set_from_triplets(M, 3 * (int64_t)nV, trips);
  const auto t2 = std::chrono::steady_clock::now();
  if (times) {
    times[0] = std::chrono::duration<double>(t1 - t0).count();
    times[1] = std::chrono::duration<double>(t2 - t1).count();
  }
  return export_csc(M);
}

// 12x12 simple_psd_fix by the tridiagonalisation + QR eigensolver (tests compare it with or_psd_fix)
void or_psd_fix_ql(double* A, double tol) { psd_fix_ql<12>(A, tol); }

// Consistent mass matrix (m/10 diagonal, m/20 off-diagonal identity blocks).
// Reference: src/linear_tet_mass_matrix.cpp:8-44, src/linear_tetmesh_mass_matrix.cpp:7-24.
or_csc* or_tetmesh_mass_matrix(int nV, int nT, const int* E, const double* density, const double* vol) {
  std::vector<Triplet> trips;
  trips.reserve((size_t)nT * 144);
  double K[144];
  for (int e = 0; e < nT; ++e) {
    double m = density[e] * vol[e];
    double c0 = (1.0 / 10.0) * m, c1 = (1.0 / 20.0) * m;
    for (int i = 0; i < 144; ++i) K[i] = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int d = 0; d < 3; ++d) K[(3 * a + d) + 12 * (3 * b + d)] = (a == b) ? c0 : c1;
    push_element(trips, E, nT, e, K);
  }
  CSC M;
  set_from_triplets(M, 3 * (int64_t)nV, trips);
  return export_csc(M);
}

// f = sum over tets of traction*vol/4 on each vertex.  Reference: src/linear_tet_const_ext_force.cpp:5-10,
// src/linear_tetmesh_const_ext_force.cpp:5-18.
void or_tetmesh_const_ext_force(int nV, int nT, const int* E, const double* vol, const double* traction /*3*/,
                                double* f) {
  std::memset(f, 0, sizeof(double) * 3 * (size_t)nV);
  for (int e = 0; e < nT; ++e)
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) f[3 * (int64_t)E[e + (int64_t)a * nT] + c] += traction[c] * (0.25 * vol[e]);
}

}  // extern "C"

// =============================================================================================
// Newton (reference semantics) with a banded direct solve.
// Reference: include/optimization_gradient_based.h:33-110 (newtons_method_bisection),
// src/search_direction_newton_sparse.cpp:5-26, src/linesearch_backtracking_bisection.cpp:5-31.
namespace {

// banded LU without pivoting on a symmetric-pattern CSC (used on PSD / SPD systems)
struct Band {
  int64_t n = 0, b = 0;
  std::vector<double> a;  // a[(i - j + b) + (2b+1) * j] holds A(i,j), |i-j| <= b
  double& at(int64_t i, int64_t j) { return a[(i - j + b) + (2 * b + 1) * j]; }
};

bool band_factor_solve(const CSC& M, const std::vector<double>& rhs, std::vector<double>& x) {
  Band B;
  B.n = M.n;
  for (int64_t c = 0; c < M.n; ++c)
    for (int64_t k = M.colptr[c]; k < M.colptr[c + 1]; ++k) B.b = std::max<int64_t>(B.b, std::llabs(M.rowind[k] - c));
  const int64_t n = B.n, b = B.b;
  B.a.assign((size_t)(2 * b + 1) * n, 0.0);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t k = M.colptr[c]; k < M.colptr[c + 1]; ++k) B.at(M.rowind[k], c) += M.vals[k];
  // LU (Doolittle), in place
  for (int64_t k = 0; k < n; ++k) {
    double piv = B.at(k, k);
    if (!(std::fabs(piv) > 0) || !std::isfinite(piv)) return false;
    int64_t iend = std::min(n - 1, k + b);
    for (int64_t i = k + 1; i <= iend; ++i) B.at(i, k) /= piv;
    for (int64_t j = k + 1; j <= iend; ++j) {
      double ukj = B.at(k, j);
      if (ukj == 0.0) continue;
      for (int64_t i = k + 1; i <= iend; ++i) B.at(i, j) -= B.at(i, k) * ukj;
    }
  }
  x = rhs;
  for (int64_t i = 0; i < n; ++i) {
    int64_t j0 = std::max<int64_t>(0, i - b);
    double s = x[i];
    for (int64_t j = j0; j < i; ++j) s -= B.at(i, j) * x[j];
    x[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    int64_t j1 = std::min(n - 1, i + b);
    double s = x[i];
    for (int64_t j = i + 1; j <= j1; ++j) s -= B.at(i, j) * x[j];
    x[i] = s / B.at(i, i);
    if (!std::isfinite(x[i])) return false;
  }
  return true;
}

}  // namespace

extern "C" {

// Static equilibrium: minimise  Psi(q) - f^T q  over the free DOFs (fixed DOFs held at their
// input values), with the reference Newton / line-search semantics:
//   * convergence when g^T g < tol^2 (absolute), loop runs up to max_it + 1 iterations
//   * Armijo test f(x + a d) > f(x) + c g^T d (no alpha in the threshold), a *= p, step applied
//     even when the line search hits its cap
// The fixed DOFs are removed as P H P^T / P g (reference: src/fixed_point_constraint_matrix.cpp:7-38).
// trace (optional, length 4*(max_it+1)): per iteration [energy, |g|, alpha, ls_iterations].
// Returns the number of iterations run (>=0), or -1 when the direct solve fails.
static int newton_impl(int model, int nV, int nT, const int* E, const double* dphidX, const double* vol,
                       const double* params, const unsigned char* model_ids, const unsigned char* fixed_dof,
                       const double* f_ext, int psd, double psd_tol, double tol, int max_it, int max_ls, double c,
                       double p, double* q, double* trace, int* converged, const or_csc* mass, const double* q_pred,
                       double inertia);

int or_newton_static(int model, int nV, int nT, const int* E, const double* dphidX, const double* vol,
                     const double* params, const unsigned char* model_ids, const unsigned char* fixed_dof,
                     const double* f_ext, int psd, double psd_tol, double tol, int max_it, int max_ls, double c,
                     double p, double* q /*in/out, 3 nV*/, double* trace, int* converged) {
  return newton_impl(model, nV, nT, E, dphidX, vol, params, model_ids, fixed_dof, f_ext, psd, psd_tol, tol, max_it,
                     max_ls, c, p, q, trace, converged, nullptr, nullptr, 0.0);
}

// One backward-Euler step in the position form of the reference's velocity objective
// (matlab/examples/test_tet_materials.m:97-112): with q = q_t + dt v and q_pred = q_t + dt v_t,
//   1/2 v^T M v - v^T M v_t + Psi(q_t + dt v) - dt v^T f
//     = 1/(2 dt^2) (q - q_pred)^T M (q - q_pred) + Psi(q) - f^T q + const,
// so the minimiser is the same and the Newton iterates correspond (d_q = dt d_v).  Same Newton /
// line-search semantics as or_newton_static; M = consistent mass with the given per-tet density.
int or_newton_be(int model, int nV, int nT, const int* E, const double* dphidX, const double* vol,
                 const double* params, const unsigned char* model_ids, const unsigned char* fixed_dof,
                 const double* f_ext, const double* density, const double* q_pred, double dt, int psd,
                 double psd_tol, double tol, int max_it, int max_ls, double c, double p, double* q, double* trace,
                 int* converged) {
  or_csc* M = or_tetmesh_mass_matrix(nV, nT, E, density, vol);
  const int r = newton_impl(model, nV, nT, E, dphidX, vol, params, model_ids, fixed_dof, f_ext, psd, psd_tol, tol,
                            max_it, max_ls, c, p, q, trace, converged, M, q_pred, 1.0 / (dt * dt));
  or_csc_free(M);
  return r;
}

static int newton_impl(int model, int nV, int nT, const int* E, const double* dphidX, const double* vol,
                       const double* params, const unsigned char* model_ids, const unsigned char* fixed_dof,
                       const double* f_ext, int psd, double psd_tol, double tol, int max_it, int max_ls, double c,
                       double p, double* q, double* trace, int* converged, const or_csc* mass, const double* q_pred,
                       double inertia) {
  const int64_t n = 3 * (int64_t)nV;
  // M (q - q_pred), the inertia term's gradient
  auto mass_times = [&](const std::vector<double>& qq, std::vector<double>& out) {
    out.assign(n, 0.0);
    for (int64_t cc = 0; cc < n; ++cc) {
      const double dc = qq[cc] - q_pred[cc];
      for (int64_t k = mass->colptr[cc]; k < mass->colptr[cc + 1]; ++k) out[mass->rowind[k]] += mass->vals[k] * dc;
    }
  };
  std::vector<int64_t> free_of(n, -1);
  std::vector<int64_t> dofs;
  for (int64_t i = 0; i < n; ++i)
    if (!fixed_dof || !fixed_dof[i]) {
      free_of[i] = (int64_t)dofs.size();
      dofs.push_back(i);
    }
  const int64_t nf = (int64_t)dofs.size();
  std::vector<double> Md;
  auto energy = [&](const std::vector<double>& qq) {
    double e = or_tetmesh_q(model, nV, nT, E, qq.data(), dphidX, vol, params, model_ids);
    if (f_ext)
      for (int64_t i = 0; i < n; ++i) e -= f_ext[i] * qq[i];
    if (mass) {
      mass_times(qq, Md);
      double a = 0;
      for (int64_t i = 0; i < n; ++i) a += (qq[i] - q_pred[i]) * Md[i];
      e += 0.5 * inertia * a;
    }
    return e;
  };
  std::vector<double> qv(q, q + n), gfull(n), g(nf), d(nf);
  *converged = 0;
  int it = 0;
  do {
    or_tetmesh_dq(model, nV, nT, E, qv.data(), dphidX, vol, params, model_ids, gfull.data());
    if (mass) {
      mass_times(qv, Md);
      for (int64_t i = 0; i < n; ++i) gfull[i] += inertia * Md[i];
    }
    double gg = 0;
    for (int64_t k = 0; k < nf; ++k) {
      g[k] = gfull[dofs[k]] - (f_ext ? f_ext[dofs[k]] : 0.0);
      gg += g[k] * g[k];
    }
    if (trace) {
      trace[4 * it + 0] = energy(qv);
      trace[4 * it + 1] = std::sqrt(gg);
      trace[4 * it + 2] = 0;
      trace[4 * it + 3] = 0;
    }
    if (gg < tol * tol) {
      *converged = 1;
      break;
    }
    or_csc* H = or_tetmesh_dq2(model, nV, nT, E, qv.data(), dphidX, vol, params, model_ids, 1, psd, psd_tol);
    if (mass) {  // same pattern (full blocks of every vertex pair that shares a tet)
      if (mass->nnz != H->nnz) return -1;
      for (int64_t k = 0; k < H->nnz; ++k) H->vals[k] += inertia * mass->vals[k];
    }
    CSC R;  // reduced P H P^T
    R.n = nf;
    R.colptr.assign(nf + 1, 0);
    for (int64_t cc = 0; cc < n; ++cc) {
      if (free_of[cc] < 0) continue;
      for (int64_t k = H->colptr[cc]; k < H->colptr[cc + 1]; ++k)
        if (free_of[H->rowind[k]] >= 0) {
          R.rowind.push_back((int32_t)free_of[H->rowind[k]]);
          R.vals.push_back(H->vals[k]);
        }
      R.colptr[free_of[cc] + 1] = (int32_t)R.rowind.size();
    }
    R.nnz = (int64_t)R.rowind.size();
    or_csc_free(H);
    std::vector<double> sol;
    if (!band_factor_solve(R, g, sol)) return -1;
    for (int64_t k = 0; k < nf; ++k) d[k] = -sol[k];
    // line search
    double gd = 0;
    for (int64_t k = 0; k < nf; ++k) gd += g[k] * d[k];
    double alpha = 1.0;
    int ls = 0;
    std::vector<double> trial(qv);
    auto eval_at = [&](double a) {
      for (int64_t k = 0; k < nf; ++k) trial[dofs[k]] = qv[dofs[k]] + a * d[k];
      return energy(trial);
    };
    while ((eval_at(alpha) > energy(qv) + c * gd) && ls < max_ls) {
      alpha *= p;
      ++ls;
    }
    for (int64_t k = 0; k < nf; ++k) qv[dofs[k]] += alpha * d[k];
    if (trace) {
      trace[4 * it + 2] = alpha;
      trace[4 * it + 3] = ls;
    }
    ++it;
  } while (it <= max_it);
  std::memcpy(q, qv.data(), sizeof(double) * n);
  return it;
}

// reference line search on an arbitrary 1-D-parametrised callback (closed-form toy checks)
typedef double (*or_energy_cb)(const double* x, int n, void* user);
int or_linesearch_backtracking(double* x, const double* d, const double* g, int n, or_energy_cb f, void* user,
                               int max_it, double alpha, double c, double p, double* alpha_out) {
  std::vector<double> xt(n);
  double gd = 0;
  for (int i = 0; i < n; ++i) gd += g[i] * d[i];
  int it = 0;
  auto fa = [&](double a) {
    for (int i = 0; i < n; ++i) xt[i] = x[i] + a * d[i];
    return f(xt.data(), n, user);
  };
  while ((fa(alpha) > f(x, n, user) + c * gd) && it < max_it) {
    alpha *= p;
    ++it;
  }
  for (int i = 0; i < n; ++i) x[i] += alpha * d[i];
  *alpha_out = alpha;
  return (it == max_it) ? 1 : 0;  // 1 = MAX_ITERATIONS_REACHED, 0 = CONVERGED (include/BartelsTypes.h:6-9)
}

// direct solve helper for tests: solve A x = b with the banded LU (A given as CSC)
int or_band_solve(int64_t n, const int32_t* colptr, const int32_t* rowind, const double* vals, const double* b,
                  double* x) {
  CSC M;
  M.n = n;
  M.colptr.assign(colptr, colptr + n + 1);
  M.nnz = colptr[n];
  M.rowind.assign(rowind, rowind + M.nnz);
  M.vals.assign(vals, vals + M.nnz);
  std::vector<double> rhs(b, b + n), sol;
  if (!band_factor_solve(M, rhs, sol)) return -1;
  std::memcpy(x, sol.data(), sizeof(double) * n);
  return 0;
}

}  // extern "C"

// =============================================================================================
// Triangle siblings (the tri2d plane-strain elements and the tri3d StVK shell element): the
// reference's dense-B algorithm, element by element, then the same setFromTriplets fold.
namespace {

// tri2d: F3 = unflatten<3,3>(B qe + e_9) with B (9 x 6) = flatten_multiply_right<3x3>([dphidX | 0]) P,
// P embedding (x, y) -> (x, y, 0): B(i + 3j, 2a + c) = [i == c][j < 2] Phi(a, j).
// Reference: src/linear_tri2d_neohookean_q.cpp:5-13 (and its dq / dq2 and model siblings).
struct Tri2 {
  double B[54];  // 9 x 6 column-major
  double F[9];
  double vol;
  double p[2];
};

void load_tri2(Tri2& t, int e, int nT, const int* E, const double* q, const double* dphidX, const double* vol,
               const double* params) {
  double phi[6], qe[6];
  for (int k = 0; k < 6; ++k) phi[k] = dphidX[e + (int64_t)k * nT];  // Phi(m, j) at m + 3j
  for (int a = 0; a < 3; ++a) {
    const int v = E[e + (int64_t)a * nT];
    for (int c = 0; c < 2; ++c) qe[2 * a + c] = q[2 * (int64_t)v + c];
  }
  for (int col = 0; col < 6; ++col) {
    const int a = col / 2, c = col % 2;
    for (int r = 0; r < 9; ++r) {
      const int i = r % 3, j = r / 3;
      t.B[r + 9 * col] = (i == c && j < 2) ? phi[a + 3 * j] : 0.0;
    }
  }
  for (int r = 0; r < 9; ++r) {
    double s = (r == 8) ? 1.0 : 0.0;
    for (int col = 0; col < 6; ++col) s += t.B[r + 9 * col] * qe[col];
    t.F[r] = s;
  }
  t.vol = vol[e];
  t.p[0] = params[e];
  t.p[1] = params[e + (int64_t)nT];
}

// tri3d StVK shell: B (9 x 9) = flatten_multiply_right<3x3>(dphidX) + flatten_multiply_right<3x1>(N^T)
// unflatten<3,9>(dndq), F3 = unflatten<3,3>(B qe + vec(n N^T)).  Reference: src/linear_tri3d_stvk_dq2.cpp:6-40.
struct Tri3 {
  double B[81];  // 9 x 9 column-major
  double F[9];
  double vol;
  double p[2];
};

void load_tri3(Tri3& t, int e, int nT, const int* E, const double* q, const double* dphidX, const double* N,
               const double* n, const double* dndq, const double* vol, const double* params) {
  double phi[9], qe[9], Ne[3], ne[3], D[27];
  for (int k = 0; k < 9; ++k) phi[k] = dphidX[e + (int64_t)k * nT];  // Phi(m, j) at m + 3j
  for (int k = 0; k < 3; ++k) {
    Ne[k] = N[e + (int64_t)k * nT];
    ne[k] = n[e + (int64_t)k * nT];
  }
  for (int k = 0; k < 27; ++k) D[k] = dndq[e + (int64_t)k * nT];  // D(i, c) at i + 3c
  for (int a = 0; a < 3; ++a) {
    const int v = E[e + (int64_t)a * nT];
    for (int c = 0; c < 3; ++c) qe[3 * a + c] = q[3 * (int64_t)v + c];
  }
  for (int col = 0; col < 9; ++col) {
    const int a = col / 3, c = col % 3;
    for (int r = 0; r < 9; ++r) {
      const int i = r % 3, j = r / 3;
      t.B[r + 9 * col] = ((i == c) ? phi[a + 3 * j] : 0.0) + Ne[j] * D[i + 3 * col];
    }
  }
  for (int r = 0; r < 9; ++r) {
    const int i = r % 3, j = r / 3;
    double s = ne[i] * Ne[j];
    for (int col = 0; col < 9; ++col) s += t.B[r + 9 * col] * qe[col];
    t.F[r] = s;
  }
  t.vol = vol[e];
  t.p[0] = params[e];
  t.p[1] = params[e + (int64_t)nT];
}

// K = B^T H B vol (ne x ne, column-major) for B 9 x ne
void btHb(const double* B, int ne, const double* H, double vol, double* K) {
  double HB[81];
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < ne; ++c) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += H[r + 9 * k] * B[k + 9 * c];
      HB[r + 9 * c] = s;
    }
  for (int r = 0; r < ne; ++r)
    for (int c = 0; c < ne; ++c) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += B[k + 9 * r] * HB[k + 9 * c];
      K[r + ne * c] = s * vol;
    }
}

// triplets of an ne x ne element matrix of a dim-DOF mesh with 3-vertex elements, in the reference's
// loop order (src/assemble.cpp:68-74)
void push_tri(std::vector<Triplet>& trips, const int* E, int nT, int e, int dim, const double* K) {
  const int ne = 3 * dim;
  for (int jj = 0; jj < ne; ++jj)
    for (int ii = 0; ii < ne; ++ii)
      trips.push_back({dim * E[e + (int64_t)(ii / dim) * nT] + ii % dim, dim * E[e + (int64_t)(jj / dim) * nT] + jj % dim,
                       K[ii + ne * jj]});
}

}  // namespace

extern "C" {

// dphidX (nT x 6 col-major, Phi(m, j) at m + 3j) of 2-D triangles.  Reference:
// src/linear_tri2dmesh_dphi_dX.cpp:5-16, src/linear_tri2d_dphi_dX.cpp:5-17.
int or_tri2d_dphidX(int nV, const double* V /*nV x 2*/, int nT, const int* E, double* dphidX) {
  for (int e = 0; e < nT; ++e) {
    double x[3][2];
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
      if (v < 0 || v >= nV) return -1;
      for (int c = 0; c < 2; ++c) x[a][c] = V[v + (int64_t)c * nV];
    }
    // A = [x1 - x0, x2 - x0] (columns), rows 1..2 of Phi = A^-1, row 0 = -(row 1 + row 2)
    const double a00 = x[1][0] - x[0][0], a10 = x[1][1] - x[0][1], a01 = x[2][0] - x[0][0], a11 = x[2][1] - x[0][1];
    const double det = a00 * a11 - a01 * a10;
    const double inv[2][2] = {{a11 / det, -a01 / det}, {-a10 / det, a00 / det}};
    for (int j = 0; j < 2; ++j) {
      dphidX[e + (int64_t)(1 + 3 * j) * nT] = inv[0][j];
      dphidX[e + (int64_t)(2 + 3 * j) * nT] = inv[1][j];
      dphidX[e + (int64_t)(0 + 3 * j) * nT] = -(inv[0][j] + inv[1][j]);
    }
  }
  return 0;
}

// dphidX (nT x 9, Phi(m, j) at m + 3j) of triangles in 3-D: rows 1..2 of Phi = (T^T T)^-1 T^T for
// T = [x1 - x0, x2 - x0], row 0 = -(row 1 + row 2).  Reference: src/linear_tri3dmesh_dphi_dX.cpp:5-19,
// src/linear_tri3d_dphi_dX.cpp:5-19.
int or_tri3d_dphidX(int nV, const double* V /*nV x 3*/, int nT, const int* E, double* dphidX) {
  for (int e = 0; e < nT; ++e) {
    double x[3][3];
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
      if (v < 0 || v >= nV) return -1;
      for (int c = 0; c < 3; ++c) x[a][c] = V[v + (int64_t)c * nV];
    }
    double T[3][2];
    for (int c = 0; c < 3; ++c) {
      T[c][0] = x[1][c] - x[0][c];
      T[c][1] = x[2][c] - x[0][c];
    }
    double G[2][2] = {{0, 0}, {0, 0}};
    for (int r = 0; r < 2; ++r)
      for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 3; ++c) G[r][s] += T[c][r] * T[c][s];
    const double det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    const double Gi[2][2] = {{G[1][1] / det, -G[0][1] / det}, {-G[1][0] / det, G[0][0] / det}};
    for (int j = 0; j < 3; ++j) {
      double rows[2];
      for (int m = 0; m < 2; ++m) rows[m] = Gi[m][0] * T[j][0] + Gi[m][1] * T[j][1];
      dphidX[e + (int64_t)(1 + 3 * j) * nT] = rows[0];
      dphidX[e + (int64_t)(2 + 3 * j) * nT] = rows[1];
      dphidX[e + (int64_t)(0 + 3 * j) * nT] = -(rows[0] + rows[1]);
    }
  }
  return 0;
}

// sim::linear_tri2dmesh_<model>_q: fold from element 0 (src/linear_tri2dmesh_neohookean_q.cpp, src/assemble.cpp:134-150)
double or_tri2dmesh_q(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                      const double* vol, const double* params) {
  (void)nV;
  double energy = 0.0;
  for (int e = 0; e < nT; ++e) {
    Tri2 t;
    load_tri2(t, e, nT, E, q, dphidX, vol, params);
    const double v = psi_F(model, t.F, t.p) * t.vol;
    energy = (e == 0) ? v : energy + v;
  }
  return energy;
}

// sim::linear_tri2dmesh_<model>_dq: B^T vec(dpsi/dF) vol, += in ascending element order (src/linear_tri2d_neohookean_dq.cpp)
void or_tri2dmesh_dq(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                     const double* vol, const double* params, double* g) {
  std::memset(g, 0, sizeof(double) * 2 * (size_t)nV);
  for (int e = 0; e < nT; ++e) {
    Tri2 t;
    load_tri2(t, e, nT, E, q, dphidX, vol, params);
    double P[9];
    dpsi_F(model, t.F, t.p, P);
    for (int col = 0; col < 6; ++col) {
      double s = 0;
      for (int k = 0; k < 9; ++k) s += t.B[k + 9 * col] * P[k];
      g[2 * (int64_t)E[e + (int64_t)(col / 2) * nT] + col % 2] += s * t.vol;
    }
  }
}

// per-element 6x6 of tri2d element e (B^T d2psi B vol), optional simple_psd_fix
void or_tri2d_element_hessian(int model, int nT, int e, const int* E, const double* q, const double* dphidX,
                              const double* vol, const double* params, int psd, double psd_tol, double* K) {
  Tri2 t;
  load_tri2(t, e, nT, E, q, dphidX, vol, params);
  double H[81];
  d2psi_F(model, t.F, t.p, 1, H);
  btHb(t.B, 6, H, t.vol, K);
  if (psd) psd_fix(K, 6, psd_tol);
}

// sim::linear_tri2dmesh_<model>_dq2 with the optional simple_psd_fix callback (src/linear_tri2dmesh_neohookean_dq2.cpp:5-24)
or_csc* or_tri2dmesh_dq2(int model, int nV, int nT, const int* E, const double* q, const double* dphidX,
                         const double* vol, const double* params, int psd, double psd_tol) {
  std::vector<Triplet> trips;
  trips.reserve((size_t)nT * 36);
  double K[36];
  for (int e = 0; e < nT; ++e) {
    or_tri2d_element_hessian(model, nT, e, E, q, dphidX, vol, params, psd, psd_tol, K);
    push_tri(trips, E, nT, e, 2, K);
  }
  CSC M;
  set_from_triplets(M, 2 * (int64_t)nV, trips);
  return export_csc(M);
}

// sim::linear_tri2dmesh_mass_matrix: blocks (rho vol / 6) I2 on the diagonal, (rho vol / 12) I2 off it
// (src/linear_tri2d_mass_matrix.cpp:5-29, src/linear_tri2dmesh_mass_matrix.cpp)
or_csc* or_tri2dmesh_mass(int nV, int nT, const int* E, const double* density, const double* vol) {
  std::vector<Triplet> trips;
  trips.reserve((size_t)nT * 36);
  for (int e = 0; e < nT; ++e) {
    const double mass = density[e] * vol[e], c0 = (1.0 / 6.0) * mass, c1 = (1.0 / 12.0) * mass;
    double K[36];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) K[r + 6 * c] = (r % 2 == c % 2) ? ((r / 2 == c / 2) ? c0 : c1) : 0.0;
    push_tri(trips, E, nT, e, 2, K);
  }
  CSC M;
  set_from_triplets(M, 2 * (int64_t)nV, trips);
  return export_csc(M);
}

// the tri3d shell's element gradient B^T vec(dpsi_stvk/dF) vol (test helper: no reference counterpart;
// with the per-element N, n, dndq held fixed, F3 is affine in q and this is the derivative of
// psi(F3) vol, whose derivative in turn is the reference's B^T H B vol)
void or_tri3d_stvk_element_gradient(int nT, int e, const int* E, const double* q, const double* dphidX,
                                    const double* N, const double* n, const double* dndq, const double* vol,
                                    const double* params, double* ge) {
  Tri3 t;
  load_tri3(t, e, nT, E, q, dphidX, N, n, dndq, vol, params);
  double P[9];
  dpsi_F(STVK, t.F, t.p, P);
  for (int col = 0; col < 9; ++col) {
    double s = 0;
    for (int k = 0; k < 9; ++k) s += t.B[k + 9 * col] * P[k];
    ge[col] = s * t.vol;
  }
}

void or_tri3d_stvk_element_hessian(int nT, int e, const int* E, const double* q, const double* dphidX,
                                   const double* N, const double* n, const double* dndq, const double* vol,
                                   const double* params, int psd, double psd_tol, double* K) {
  Tri3 t;
  load_tri3(t, e, nT, E, q, dphidX, N, n, dndq, vol, params);
  double H[81];
  d2psi_F(STVK, t.F, t.p, 1, H);
  btHb(t.B, 9, H, t.vol, K);
  if (psd) psd_fix(K, 9, psd_tol);
}

// sim::linear_tri3dmesh_stvk_dq2 (src/linear_tri3dmesh_stvk_dq2.cpp:5-30, src/linear_tri3d_stvk_dq2.cpp:6-40)
or_csc* or_tri3dmesh_stvk_dq2(int nV, int nT, const int* E, const double* q, const double* dphidX, const double* N,
                              const double* n, const double* dndq, const double* vol, const double* params, int psd,
                              double psd_tol) {
  std::vector<Triplet> trips;
  trips.reserve((size_t)nT * 81);
  double K[81];
  for (int e = 0; e < nT; ++e) {
    or_tri3d_stvk_element_hessian(nT, e, E, q, dphidX, N, n, dndq, vol, params, psd, psd_tol, K);
    push_tri(trips, E, nT, e, 3, K);
  }
  CSC M;
  set_from_triplets(M, 3 * (int64_t)nV, trips);
  return export_csc(M);
}

}  // extern "C"
