// Per-tet constitutive math for the gfx950 kernels: energy density, first Piola stress and the
// 3x3 blocks of the element Hessian, written in closed block form instead of the reference's
// dense 12x9 * 9x9 * 9x12 product.  Templated on the element arithmetic type T:
//   T = double : the FP64 path (parity with the reference to 1e-12);
//   T = float  : the precision form -- F, stresses and Hessian blocks in FP32, while J = det F,
//                1/J, log J and the energy invariants are evaluated in FP64 from the FP32 F.
//
// Conventions (reference include/flatten.h:59-72, include/flatten_multiply.h:52-75):
//   F(i,j) at F[i + 3j];  phi_m = grad of shape function m (3-vector);  F = sum_m x_m phi_m^T.
//   Element Hessian block (a,m) (3x3, row = component l of vertex a, col = component k of m):
//     K_am[l][k] = vol * sum_{j,j'} phi_a[j] phi_m[j'] d2psi/dF_lj dF_kj'
//
// Closed forms used per model (all derived from the reference energies; the derivation is
// checked numerically against the FP64 oracle's dense path and torch autograd):
//   Neo-Hookean (src/psi_neohookean_F.cpp:31, src/d2psi_neohookean_dF2.cpp:200-280):
//     d2psi = 2C I + (2C/J^2 + 2D) g g^T + (2D(J-1) - 2C/J) d2J,   g = vec(cof F)
//     K_am  = vol [2C (phi_a.phi_m) I + alpha c_a c_m^T + beta S(F (phi_a x phi_m))]
//     with c = cof(F) phi and S(w)[l][k] = eps_lkp w_p.
//   legacy isochoric NH (src/d2psi_neohookean_dF2.cpp:8-183, commented): s = J^{-2/3}
//     d2psi = 2Cs I - 4/3 (Cs/J)(F g^T + g F^T) + (10/9 Cs I1/J^2 + 2D) g g^T + (2D(J-1) - 2/3 Cs I1/J) d2J
//   StVK (src/d2psi_stvk_dF2.cpp:5-106), S = 2mu E + lambda tr(E) I, u = F phi:
//     K_am  = vol [(phi_a^T S phi_m) I + mu u_m u_a^T + mu (phi_a.phi_m) F F^T + lambda u_a u_m^T]
//   stretch models (corotational, ARAP; src/d2psi_stretch_dF2.cpp:7-47, src/dsvd.cpp:5-57):
//     in the singular frame of G = U diag(s) V^T, with p = V^T phi:
//       Khat_ii = p_a[i] psi_ii p_m[i] + sum_{j!=i} A_ij p_a[j] p_m[j]
//       Khat_ij = p_a[i] psi_ij p_m[j] + B_ij p_a[j] p_m[i]          (i != j)
//       A = (s_ij + t_ij)/2, B = (s_ij - t_ij)/2,
//       s_ij = (psi_i - psi_j)/(s_i - s_j)  (exact per model: 2 mu or 2k),
//       t_ij = (psi_i + psi_j)/(s_i + s_j)
//     K_am = vol U Khat U^T.  With the reference layout G = F^T (U and V swap roles).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace bh {

enum : int {
  M_NH = 0,
  M_STVK = 1,
  M_COROT = 2,
  M_ARAP = 3,
  M_NH_LEGACY = 4,
  M_NH_MUSCLE = 5,  // Neo-Hookean plus a muscle fibre; params nT x 6: C, D, a, d0, d1, d2
  M_MASS = 100,
  M_PER_TET = 255
};
// parameter columns a model reads (nT x cols, column-major)
__host__ __device__ constexpr int param_cols(int model) { return model == M_NH_MUSCLE ? 6 : 2; }

template <typename T>
__host__ __device__ __forceinline__ T det3(const T* F) {
  return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
         F[6] * (F[1] * F[5] - F[4] * F[2]);
}

// cof(F) = det(F) F^{-T}, column-major: c[i + 3j] = (-1)^{i+j} minor_ij(F)
template <typename T>
__host__ __device__ __forceinline__ void cof3(const T* F, T* c) {
  c[0] = F[4] * F[8] - F[7] * F[5];
  c[1] = F[6] * F[5] - F[3] * F[8];
  c[2] = F[3] * F[7] - F[6] * F[4];
  c[3] = F[7] * F[2] - F[1] * F[8];
  c[4] = F[0] * F[8] - F[6] * F[2];
  c[5] = F[6] * F[1] - F[0] * F[7];
  c[6] = F[1] * F[5] - F[4] * F[2];
  c[7] = F[3] * F[2] - F[0] * F[5];
  c[8] = F[0] * F[4] - F[3] * F[1];
}

// y = A x, A col-major 3x3
template <typename T>
__host__ __device__ __forceinline__ void mv3(const T* A, const T* x, T* y) {
  y[0] = A[0] * x[0] + A[3] * x[1] + A[6] * x[2];
  y[1] = A[1] * x[0] + A[4] * x[1] + A[7] * x[2];
  y[2] = A[2] * x[0] + A[5] * x[1] + A[8] * x[2];
}
// y = A^T x
template <typename T>
__host__ __device__ __forceinline__ void mtv3(const T* A, const T* x, T* y) {
  y[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
  y[1] = A[3] * x[0] + A[4] * x[1] + A[5] * x[2];
  y[2] = A[6] * x[0] + A[7] * x[1] + A[8] * x[2];
}
template <typename T>
__host__ __device__ __forceinline__ T dot3(const T* a, const T* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <typename T>
__host__ __device__ __forceinline__ void cross3(const T* a, const T* b, T* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// determinant and cofactor in FP64 from an F of any precision (the precision form keeps J in FP64)
template <typename T>
__host__ __device__ __forceinline__ double det_cof_fp64(const T* F, double* c) {
  double Fd[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Fd[i] = (double)F[i];
  cof3(Fd, c);
  return Fd[0] * c[0] + Fd[3] * c[3] + Fd[6] * c[6];
}

// ---- 3x3 SVD, rotation convention (U, V rotations, s sorted descending, s[2] signed) --------
// Cyclic Jacobi on F^T F (bounded sweeps), then U from F V.  Same algorithm as the oracle.
template <typename T>
__host__ __device__ inline void jacobi_rot(T* A, T* V, int p, int q) {
  const T apq = A[p + 3 * q];
  if (apq == T(0)) return;
  const T app = A[p + 3 * p], aqq = A[q + 3 * q];
  const T theta = (aqq - app) / (T(2) * apq);
  const T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
  const T c = T(1) / sqrt(t * t + T(1)), s = t * c;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T akp = A[k + 3 * p], akq = A[k + 3 * q];
    A[k + 3 * p] = c * akp - s * akq;
    A[k + 3 * q] = s * akp + c * akq;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T apk = A[p + 3 * k], aqk = A[q + 3 * k];
    A[p + 3 * k] = c * apk - s * aqk;
    A[q + 3 * k] = s * apk + c * aqk;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T vkp = V[k + 3 * p], vkq = V[k + 3 * q];
    V[k + 3 * p] = c * vkp - s * vkq;
    V[k + 3 * q] = s * vkp + c * vkq;
  }
}

template <typename T>
__host__ __device__ inline void svd3(const T* F, T* U, T* s, T* V) {
  T A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      A[i + 3 * j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? T(1) : T(0);
  const T eps2 = sizeof(T) == 8 ? T(1e-36) : T(1e-16);
  for (int sweep = 0; sweep < 12; ++sweep) {
    const T off = A[3] * A[3] + A[6] * A[6] + A[7] * A[7];
    const T dg = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= eps2 * dg) break;
    jacobi_rot(A, V, 0, 1);
    jacobi_rot(A, V, 0, 2);
    jacobi_rot(A, V, 1, 2);
  }
  T w0 = A[0], w1 = A[4], w2 = A[8];
  // sort descending with column swaps (3-element network)
#define BH_SWAPCOL(a, b, wa, wb)                                       \
  if (wa < wb) {                                                       \
    T tw = wa; wa = wb; wb = tw;                                       \
    for (int r = 0; r < 3; ++r) {                                      \
      T tv = V[r + 3 * a]; V[r + 3 * a] = V[r + 3 * b]; V[r + 3 * b] = tv; \
    }                                                                  \
  }
  BH_SWAPCOL(0, 1, w0, w1)
  BH_SWAPCOL(0, 2, w0, w2)
  BH_SWAPCOL(1, 2, w1, w2)
#undef BH_SWAPCOL
  if (det3(V) < T(0)) {
    V[6] = -V[6]; V[7] = -V[7]; V[8] = -V[8];
  }
  T f0[3], f1[3], f2[3];
  mv3(F, V, f0);
  mv3(F, V + 3, f1);
  mv3(F, V + 6, f2);
  const T n0 = sqrt(dot3(f0, f0));
  T u0[3], u1[3], u2[3];
  if (n0 > T(0)) {
    u0[0] = f0[0] / n0; u0[1] = f0[1] / n0; u0[2] = f0[2] / n0;
  } else {
    u0[0] = T(1); u0[1] = T(0); u0[2] = T(0);
  }
  const T d01 = dot3(u0, f1);
  u1[0] = f1[0] - d01 * u0[0]; u1[1] = f1[1] - d01 * u0[1]; u1[2] = f1[2] - d01 * u0[2];
  const T n1 = sqrt(dot3(u1, u1));
  const T tiny = sizeof(T) == 8 ? T(1e-300) : T(1e-30);
  if (n1 > tiny) {
    u1[0] /= n1; u1[1] /= n1; u1[2] /= n1;
  } else {
    const int k = (fabs(u0[0]) < T(0.6)) ? 0 : ((fabs(u0[1]) < T(0.6)) ? 1 : 2);
    const T dd = u0[k];
    u1[0] = (k == 0 ? T(1) : T(0)) - dd * u0[0];
    u1[1] = (k == 1 ? T(1) : T(0)) - dd * u0[1];
    u1[2] = (k == 2 ? T(1) : T(0)) - dd * u0[2];
    const T nn = sqrt(dot3(u1, u1));
    u1[0] /= nn; u1[1] /= nn; u1[2] /= nn;
  }
  cross3(u0, u1, u2);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    U[r] = u0[r];
    U[r + 3] = u1[r];
    U[r + 6] = u2[r];
  }
  s[0] = n0;
  s[1] = dot3(u1, f1);
  s[2] = dot3(u2, f2);
}

// ---- per-tet precomputation -----------------------------------------------------------------
template <typename T>
struct Pre {
  T F[9];
  T vol;
  // NH / legacy / StVK
  T cof[9];
  T a0, a1, a2, a3;  // model scalars (pre-multiplied by vol)
  T S[9];            // StVK stress-like matrix (vol * S)
  T FFt[9];          // StVK (vol * mu * F F^T)
  // stretch
  T U[9], V[9];      // frame of G (reference layout: G = F^T)
  T psi2[9];         // vol * d2psi/ds2
  T A[3], B[3];      // vol * A_ij, vol * B_ij for pairs (0,1), (0,2), (1,2)
  // muscle fibre (M_NH_MUSCLE)
  T fa;              // vol * a
  T fd[3];           // fibre direction d
};

// ---- muscle fibre: psi = a/2 |F d|^2 (src/psi_muscle_fiber_F.cpp:5-10), P = a (F d) d^T
// (src/dpsi_muscle_fiber_dF.cpp, a D^T D vec F), d2psi = a D^T D (src/d2psi_muscle_fiber_dF2.cpp), with
// D vec(F) = F d.  fib = {a, d0, d1, d2}.
template <typename T>
__host__ __device__ __forceinline__ double fibre_energy(const T* F, const T* fib) {
  double Fd[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    Fd[i] = (double)F[i] * fib[1] + (double)F[i + 3] * fib[2] + (double)F[i + 6] * fib[3];
  return 0.5 * (double)fib[0] * (Fd[0] * Fd[0] + Fd[1] * Fd[1] + Fd[2] * Fd[2]);
}
template <typename T>
__host__ __device__ __forceinline__ void fibre_piola_add(const T* F, const T* fib, T* P) {
  T Fd[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) Fd[i] = F[i] * fib[1] + F[i + 3] * fib[2] + F[i + 6] * fib[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) P[i + 3 * j] += fib[0] * Fd[i] * fib[1 + j];
}
template <typename T>
__host__ __device__ __forceinline__ void set_fibre(Pre<T>& P, const T* fib) {
  P.fa = P.vol * fib[0];
  P.fd[0] = fib[1];
  P.fd[1] = fib[2];
  P.fd[2] = fib[3];
}

// stretch-energy derivatives in singular values. Corotational: params [lambda, mu]
// (src/dpsi_corotational_dS.cpp:6-13); ARAP: k = params(0) (src/dpsi_arap_dS.cpp:6-12).
template <int MODEL, typename T>
__host__ __device__ __forceinline__ void dpsi_s(const T* s, T p0, T p1, T* d) {
  if (MODEL == M_COROT) {
    const T tr = s[0] + s[1] + s[2] - T(3);
    d[0] = T(2) * p1 * (s[0] - T(1)) + p0 * tr;
    d[1] = T(2) * p1 * (s[1] - T(1)) + p0 * tr;
    d[2] = T(2) * p1 * (s[2] - T(1)) + p0 * tr;
  } else {
    d[0] = T(2) * p0 * (s[0] - T(1));
    d[1] = T(2) * p0 * (s[1] - T(1));
    d[2] = T(2) * p0 * (s[2] - T(1));
  }
}

// Hessian precomputation.  stretch_ref: reference (F^T) layout for corotational / ARAP.
template <int MODEL, typename T>
__host__ __device__ __forceinline__ void pre_hessian(Pre<T>& P, T p0, T p1, bool stretch_ref) {
  const T* F = P.F;
  const T vol = P.vol;
  if (MODEL == M_NH || MODEL == M_NH_LEGACY || MODEL == M_NH_MUSCLE) {
    double J;
    if constexpr (sizeof(T) == 8) {
      cof3(F, P.cof);
      J = (double)(F[0] * P.cof[0] + F[3] * P.cof[3] + F[6] * P.cof[6]);
    } else {
      double c[9];
      J = det_cof_fp64(F, c);
#pragma unroll
      for (int i = 0; i < 9; ++i) P.cof[i] = (T)c[i];
    }
    const double iJ = 1.0 / J;
    const double C = p0, D = p1, vd = vol;
    if (MODEL == M_NH || MODEL == M_NH_MUSCLE) {
      P.a0 = (T)(vd * 2.0 * C);                              // 2C
      P.a1 = (T)(vd * (2.0 * C * iJ * iJ + 2.0 * D));        // alpha
      P.a2 = (T)(vd * (2.0 * D * (J - 1.0) - 2.0 * C * iJ)); // beta
      P.a3 = T(0);
    } else {
      const double cb = cbrt(J);
      const double s = 1.0 / (cb * cb);
      double I1 = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) I1 += (double)F[k] * (double)F[k];
      const double Cs = C * s;
      P.a0 = (T)(vd * 2.0 * Cs);
      P.a1 = (T)(vd * ((10.0 / 9.0) * Cs * I1 * iJ * iJ + 2.0 * D));
      P.a2 = (T)(vd * (2.0 * D * (J - 1.0) - (2.0 / 3.0) * Cs * I1 * iJ));
      P.a3 = (T)(vd * (-(4.0 / 3.0) * Cs * iJ));
    }
  } else if (MODEL == M_STVK) {
    const T lam = p0, mu = p1;
    T C[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        C[i + 3 * j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
        P.FFt[i + 3 * j] = vol * mu * (F[i] * F[j] + F[i + 3] * F[j + 3] + F[i + 6] * F[j + 6]);
      }
    const T trE = T(0.5) * (C[0] + C[4] + C[8] - T(3));
#pragma unroll
    for (int i = 0; i < 9; ++i) P.S[i] = vol * (mu * (C[i] - ((i % 4 == 0) ? T(1) : T(0))));
    P.S[0] += vol * lam * trE;
    P.S[4] += vol * lam * trE;
    P.S[8] += vol * lam * trE;
    P.a0 = vol * mu;
    P.a1 = vol * lam;
  } else if (MODEL == M_COROT || MODEL == M_ARAP) {
    T U[9], V[9], s[3];
    svd3(F, U, s, V);
    // G = F^T = V diag(s) U^T in the reference layout
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      P.U[i] = stretch_ref ? V[i] : U[i];
      P.V[i] = stretch_ref ? U[i] : V[i];
    }
    T d[3];
    dpsi_s<MODEL>(s, p0, p1, d);
    T sij;
    if (MODEL == M_COROT) {
      sij = T(2) * p1;
#pragma unroll
      for (int i = 0; i < 9; ++i) P.psi2[i] = vol * (p0 + ((i % 4 == 0) ? T(2) * p1 : T(0)));
    } else {
      sij = T(2) * p0;
#pragma unroll
      for (int i = 0; i < 9; ++i) P.psi2[i] = vol * ((i % 4 == 0) ? T(2) * p0 : T(0));
    }
    const int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T den = s[pi[k]] + s[pj[k]];
      const T t = (fabs(den) < T(1e-8)) ? T(0) : (d[pi[k]] + d[pj[k]]) / den;
      P.A[k] = vol * T(0.5) * (sij + t);
      P.B[k] = vol * T(0.5) * (sij - t);
    }
  }
}

// 3x3 block K_am (row-major out[3l + k]).
template <int MODEL, typename T>
__host__ __device__ __forceinline__ void hess_block(const Pre<T>& P, const T* pa, const T* pm, T* out) {
  if (MODEL == M_NH || MODEL == M_NH_LEGACY || MODEL == M_NH_MUSCLE) {
    T ca[3], cm[3], x[3], w[3];
    mv3(P.cof, pa, ca);
    mv3(P.cof, pm, cm);
    cross3(pa, pm, x);
    mv3(P.F, x, w);
    const T d = P.a0 * dot3(pa, pm);
    const T b = P.a2;
    out[0] = d + P.a1 * ca[0] * cm[0];
    out[1] = P.a1 * ca[0] * cm[1] + b * w[2];
    out[2] = P.a1 * ca[0] * cm[2] - b * w[1];
    out[3] = P.a1 * ca[1] * cm[0] - b * w[2];
    out[4] = d + P.a1 * ca[1] * cm[1];
    out[5] = P.a1 * ca[1] * cm[2] + b * w[0];
    out[6] = P.a1 * ca[2] * cm[0] + b * w[1];
    out[7] = P.a1 * ca[2] * cm[1] - b * w[0];
    out[8] = d + P.a1 * ca[2] * cm[2];
    if (MODEL == M_NH_LEGACY) {
      T ua[3], um[3];
      mv3(P.F, pa, ua);
      mv3(P.F, pm, um);
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int k = 0; k < 3; ++k) out[3 * l + k] += P.a3 * (ua[l] * cm[k] + ca[l] * um[k]);
    }
    if (MODEL == M_NH_MUSCLE) {  // + vol a (pa.d)(pm.d) I
      const T f = P.fa * dot3(pa, P.fd) * dot3(pm, P.fd);
      out[0] += f;
      out[4] += f;
      out[8] += f;
    }
  } else if (MODEL == M_STVK) {
    T ua[3], um[3], Sm[3];
    mv3(P.F, pa, ua);
    mv3(P.F, pm, um);
    mv3(P.S, pm, Sm);
    const T sd = dot3(pa, Sm);
    const T dd = dot3(pa, pm);
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        out[3 * l + k] = ((l == k) ? sd : T(0)) + P.a0 * um[l] * ua[k] + dd * P.FFt[l + 3 * k] + P.a1 * ua[l] * um[k];
  } else if (MODEL == M_COROT || MODEL == M_ARAP) {
    T qa[3], qm[3];
    mtv3(P.V, pa, qa);
    mtv3(P.V, pm, qm);
    T K[9];  // Khat col-major
    // diagonal entries
    K[0] = qa[0] * P.psi2[0] * qm[0] + P.A[0] * qa[1] * qm[1] + P.A[1] * qa[2] * qm[2];
    K[4] = qa[1] * P.psi2[4] * qm[1] + P.A[0] * qa[0] * qm[0] + P.A[2] * qa[2] * qm[2];
    K[8] = qa[2] * P.psi2[8] * qm[2] + P.A[1] * qa[0] * qm[0] + P.A[2] * qa[1] * qm[1];
    // off-diagonal (i,j): qa[i] psi_ij qm[j] + B_ij qa[j] qm[i]
    K[3] = qa[0] * P.psi2[3] * qm[1] + P.B[0] * qa[1] * qm[0];  // (0,1)
    K[1] = qa[1] * P.psi2[1] * qm[0] + P.B[0] * qa[0] * qm[1];  // (1,0)
    K[6] = qa[0] * P.psi2[6] * qm[2] + P.B[1] * qa[2] * qm[0];  // (0,2)
    K[2] = qa[2] * P.psi2[2] * qm[0] + P.B[1] * qa[0] * qm[2];  // (2,0)
    K[7] = qa[1] * P.psi2[7] * qm[2] + P.B[2] * qa[2] * qm[1];  // (1,2)
    K[5] = qa[2] * P.psi2[5] * qm[1] + P.B[2] * qa[1] * qm[2];  // (2,1)
    // out = U K U^T
    T Tm[9];  // K U^T  (col-major)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Tm[i + 3 * j] = K[i] * P.U[j] + K[i + 3] * P.U[j + 3] + K[i + 6] * P.U[j + 6];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        out[3 * l + k] = P.U[l] * Tm[3 * k] + P.U[l + 3] * Tm[1 + 3 * k] + P.U[l + 6] * Tm[2 + 3 * k];
  }
}

// energy density psi(F) in FP64 (not volume scaled).  With an FP32 F the invariants are formed
// from the displacement gradient G = F - I (exact in FP32 near I) and J - 1 in FP64, so the
// energy keeps its relative accuracy at small strains (the line search compares energy sums).
template <int MODEL, typename T>
__host__ __device__ __forceinline__ double psi(const T* F, double p0, double p1, const T* Gin = nullptr) {
  if (MODEL == M_NH) {
    if (sizeof(T) == 8 && !Gin) {
      double I1 = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) I1 += (double)F[k] * (double)F[k];
      const double J = det3((const double*)F);
      return p0 * (I1 - 3.0 - 2.0 * log(J)) + p1 * (J - 1.0) * (J - 1.0);
    } else {
      double G[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) G[k] = Gin ? (double)Gin[k] : (double)(F[k] - ((k % 4 == 0) ? T(1) : T(0)));
      const double trG = G[0] + G[4] + G[8];
      double GG = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) GG += G[k] * G[k];
      // J - 1 = tr G + (sum of principal 2x2 minors of G) + det G
      const double m2 = (G[0] * G[4] - G[1] * G[3]) + (G[0] * G[8] - G[2] * G[6]) + (G[4] * G[8] - G[5] * G[7]);
      const double Jm1 = trG + m2 + det3(G);
      return p0 * (2.0 * trG + GG - 2.0 * log1p(Jm1)) + p1 * Jm1 * Jm1;
    }
  } else if (MODEL == M_NH_LEGACY) {
    double I1 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) I1 += (double)F[k] * (double)F[k];
    double c[9];
    const double J = det_cof_fp64(F, c);
    const double cb = cbrt(J);
    return p0 * (I1 / (cb * cb) - 3.0) + p1 * (J - 1.0) * (J - 1.0);
  } else if (MODEL == M_STVK) {
    double E[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (Gin)  // E = (G + G^T + G^T G) / 2: no cancellation at small strain
          E[i + 3 * j] = 0.5 * ((double)Gin[3 * i + j] + (double)Gin[3 * j + i] +
                                (double)Gin[3 * i] * Gin[3 * j] + (double)Gin[3 * i + 1] * Gin[3 * j + 1] +
                                (double)Gin[3 * i + 2] * Gin[3 * j + 2]);
        else
          E[i + 3 * j] = 0.5 * ((double)F[3 * i] * F[3 * j] + (double)F[3 * i + 1] * F[3 * j + 1] +
                                (double)F[3 * i + 2] * F[3 * j + 2] - ((i == j) ? 1.0 : 0.0));
      }
    double trEE = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) trEE += E[k] * E[k];  // E symmetric: tr(E E) = sum E_ij^2
    const double trE = E[0] + E[4] + E[8];
    return p1 * trEE + 0.5 * p0 * trE * trE;
  } else {
    T U[9], V[9], s[3];
    svd3(F, U, s, V);
    const double s0 = s[0], s1 = s[1], s2 = s[2];
    if (MODEL == M_COROT) {
      const double t = s0 + s1 + s2 - 3.0;
      return p1 * ((s0 - 1.0) * (s0 - 1.0) + (s1 - 1.0) * (s1 - 1.0) + (s2 - 1.0) * (s2 - 1.0)) + 0.5 * p0 * t * t;
    }
    return p0 * ((s0 - 1.0) * (s0 - 1.0) + (s1 - 1.0) * (s1 - 1.0) + (s2 - 1.0) * (s2 - 1.0));
  }
}

// first Piola stress P (col-major), not volume scaled.
// Stretch models follow src/dpsi_stretch_dF.cpp:20-42: when s0 s1 s2 <= -1e-10 the negative
// singular values are flipped positive (U, V unchanged).
template <int MODEL, typename T>
__host__ __device__ __forceinline__ void piola(const T* F, T p0, T p1, T* P, const T* G = nullptr) {
  if (MODEL == M_NH && G && sizeof(T) == 4) {
    // from G = F - I (cancellation-free at small strain):
    //   cof F = (1 + tr G) I - G^T + cof G,  J = 1 + tr G + e2 + det G (e2 = tr cof G),
    //   J F - cof F = (e2 + det G) I + (1 + tr G + e2 + det G) G + G^T - cof G,
    //   P = 2C (J F - cof F) / J + 2D (J - 1) cof F
    double Gd[9], cg[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Gd[k] = (double)G[k];
    cof3(Gd, cg);
    const double trG = Gd[0] + Gd[4] + Gd[8];
    const double e2 = cg[0] + cg[4] + cg[8];
    const double dG = Gd[0] * cg[0] + Gd[3] * cg[3] + Gd[6] * cg[6];
    const double Jm1 = trG + e2 + dG, J = 1.0 + Jm1, iJ = 1.0 / J;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int k = i + 3 * j, kt = j + 3 * i;
        const double cofF = ((i == j) ? 1.0 + trG : 0.0) - Gd[kt] + cg[k];
        const double dev = ((i == j) ? e2 + dG : 0.0) + (1.0 + Jm1) * Gd[k] + Gd[kt] - cg[k];
        P[k] = (T)(2.0 * p0 * dev * iJ + 2.0 * p1 * Jm1 * cofF);
      }
  } else if (MODEL == M_NH) {
    T c[9];
    double J;
    if constexpr (sizeof(T) == 8) {
      cof3(F, c);
      J = (double)(F[0] * c[0] + F[3] * c[3] + F[6] * c[6]);
    } else {
      double cd[9];
      J = det_cof_fp64(F, cd);
#pragma unroll
      for (int i = 0; i < 9; ++i) c[i] = (T)cd[i];
    }
    const T iJ = (T)(1.0 / J), Jm1 = (T)(J - 1.0);
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = T(2) * p0 * (F[k] - c[k] * iJ) + T(2) * p1 * Jm1 * c[k];
  } else if (MODEL == M_NH_LEGACY) {
    double cd[9];
    const double J = det_cof_fp64(F, cd);
    double I1 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) I1 += (double)F[k] * (double)F[k];
    const double cb = cbrt(J);
    const double s = 1.0 / (cb * cb);
#pragma unroll
    for (int k = 0; k < 9; ++k)
      P[k] = (T)(p0 * (2.0 * s * F[k] - (2.0 / 3.0) * (s / J) * I1 * cd[k]) + 2.0 * p1 * (J - 1.0) * cd[k]);
  } else if (MODEL == M_STVK) {
    T S[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        S[i + 3 * j] = G ? G[3 * i + j] + G[3 * j + i] + G[3 * i] * G[3 * j] + G[3 * i + 1] * G[3 * j + 1] +
                               G[3 * i + 2] * G[3 * j + 2]
                         : F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2] -
                               ((i == j) ? T(1) : T(0));  // 2E
    const T trE = T(0.5) * (S[0] + S[4] + S[8]);
#pragma unroll
    for (int k = 0; k < 9; ++k) S[k] *= p1;  // 2 mu E
    S[0] += p0 * trE;
    S[4] += p0 * trE;
    S[8] += p0 * trE;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        P[i + 3 * j] = F[i] * S[3 * j] + F[i + 3] * S[3 * j + 1] + F[i + 6] * S[3 * j + 2];
  } else {
    T U[9], V[9], s[3];
    svd3(F, U, s, V);
    if (s[0] * s[1] * s[2] <= T(-1e-10)) {
      s[0] = fabs(s[0]);
      s[1] = fabs(s[1]);
      s[2] = fabs(s[2]);
    }
    T d[3];
    dpsi_s<MODEL>(s, p0, p1, d);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        P[i + 3 * j] = U[i] * d[0] * V[j] + U[i + 3] * d[1] * V[j + 3] + U[i + 6] * d[2] * V[j + 6];
  }
}

// ---- PSD projection helpers (simple_psd_fix, src/simple_psd_fix.cpp:5-21) ---------------------------
// K = vol B^T H B, B = Phi^T (x) I3.  With the thin QR Phi = Q R, K = (Q (x) I3) M (Q (x) I3)^T where
// M = (R (x) I3) vol H (R (x) I3)^T is the block formula evaluated at the rows of R.  The three
// translations (the complement of range Q (x) I3) are eigenvectors of K with eigenvalue 0, so
//   simple_psd_fix(K, tol) = K + (Q (x) I3) dM (Q (x) I3)^T + max(tol, 0)/4 (1 1^T (x) I3),
//   dM = W diag(max(lambda, tol) - lambda) W^T for M = W diag(lambda) W^T.

// Q (4x3, orthonormal columns, orthogonal to the ones vector) and R (3x3 upper) of Phi (4x3)
template <typename T>
__host__ __device__ __forceinline__ void phi_qr(const T phi[4][3], T Q[4][3], T R[3][3]) {
  T c0[4], c1[4], c2[4];
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) {
    c0[mm] = phi[mm][0];
    c1[mm] = phi[mm][1];
    c2[mm] = phi[mm][2];
  }
  auto d4 = [](const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; };
  const T r00 = sqrt(d4(c0, c0));
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c0[mm] /= r00;
  const T r01 = d4(c0, c1);
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c1[mm] -= r01 * c0[mm];
  const T r11 = sqrt(d4(c1, c1));
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c1[mm] /= r11;
  const T r02 = d4(c0, c2);
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c2[mm] -= r02 * c0[mm];
  const T r12 = d4(c1, c2);
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c2[mm] -= r12 * c1[mm];
  const T r22 = sqrt(d4(c2, c2));
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) c2[mm] /= r22;
#pragma unroll
  for (int mm = 0; mm < 4; ++mm) {
    Q[mm][0] = c0[mm];
    Q[mm][1] = c1[mm];
    Q[mm][2] = c2[mm];
  }
  R[0][0] = r00; R[0][1] = r01; R[0][2] = r02;
  R[1][0] = T(0); R[1][1] = r11; R[1][2] = r12;
  R[2][0] = T(0); R[2][1] = T(0); R[2][2] = r22;
}

// 1/sqrt(x): the hardware reciprocal square root with its library refinement on the device
template <typename T>
__host__ __device__ __forceinline__ T inv_sqrt(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rsqrt(x);
#else
  return T(1) / std::sqrt(x);
#endif
}

// packed upper-triangular index of a symmetric 9x9
__host__ __device__ __forceinline__ int sym9(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a * 9 - (a * (a - 1)) / 2 + (b - a);
}

// M (symmetric 9x9, index 3p + l) from the Hessian precomputation P and R, as its packed upper
// triangle a[sym9(i, j)]: the six blocks (p, r), p <= r (the lower blocks are their transposes up
// to rounding)
template <int MODEL, typename T>
__host__ __device__ __forceinline__ void reduced_packed(const Pre<T>& P, const T R[3][3], T a[45]) {
#pragma unroll
  for (int pp = 0; pp < 3; ++pp)
#pragma unroll
    for (int rr = pp; rr < 3; ++rr) {
      T b[9];
      hess_block<MODEL>(P, R[pp], R[rr], b);
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
          if (rr > pp || kk >= l) a[sym9(3 * pp + l, 3 * rr + kk)] = b[3 * l + kk];
    }
}

// true when M - tol I is positive definite (Cholesky of a packed copy)
template <typename T>
__host__ __device__ __forceinline__ bool shifted_pd(const T a[45], T tol) {
  T L[45];
  bool pd = true;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    T d = a[sym9(j, j)] - tol;
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[sym9(k, j)] * L[sym9(k, j)];
    pd = pd && (d > T(0));
    const T ljj = sqrt(fmax(d, sizeof(T) == 8 ? T(1e-300) : T(1e-30)));
    const T inv = T(1) / ljj;
#pragma unroll
    for (int i = j + 1; i < 9; ++i) {
      T s = a[sym9(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[sym9(k, i)] * L[sym9(k, j)];
      L[sym9(j, i)] = s * inv;
    }
    L[sym9(j, j)] = ljj;
  }
  return pd;
}

// Implicit QL on the tridiagonal (d, e) (e[i] couples d[i] and d[i+1]; e[8] = 0): on return d holds
// the eigenvalues (unsorted).  With WANTZ the rotations are also applied to the columns of Z.
template <bool WANTZ, typename T>
__host__ __device__ __forceinline__ void tridiag_ql(T d[9], T e[9], T* Z) {
  constexpr T kEps = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07);
  // Rotations at or below the split point m are evaluated and discarded (select), which keeps the
  // sweep free of lane-divergent branches.
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    for (int iter = 0; iter < 40; ++iter) {
      int m = 8;
#pragma unroll
      for (int j = 7; j >= l; --j)
        if (fabs(e[j]) <= kEps * (fabs(d[j]) + fabs(d[j + 1]))) m = j;
      if (m == l) break;
      T dm = d[8];
#pragma unroll
      for (int j = l + 1; j < 8; ++j) dm = (j == m) ? d[j] : dm;
      T g = (d[l + 1] - d[l]) / (T(2) * e[l]);
      const T r0 = sqrt(g * g + T(1));
      g = dm - d[l] + e[l] / (g + (g >= T(0) ? r0 : -r0));
      T s = T(1), c = T(1), pacc = T(0);
      // No split below l (m = 8) in any active lane of the wave -- the common case: every rotation
      // is active, so the sweep needs none of the selects that park the rotations below a split.
      // Same operations as the general sweep with every rotation active: same bits.
#if defined(__HIP_DEVICE_COMPILE__)
      const bool nosplit = __all(m == 8);
#else
      const bool nosplit = m == 8;
#endif
      if (!WANTZ && nosplit) {
#pragma unroll
        for (int i = 7; i >= l; --i) {
          const T f = s * e[i], b = c * e[i];
          const T r2 = f * f + g * g;
          const T inv = r2 > T(0) ? inv_sqrt(r2) : T(0);
          const T r = r2 * inv;
          const T sn = f * inv;  // f = 0 when r2 = 0
          const T cn = r2 > T(0) ? g * inv : T(1);
          const T gi = d[i + 1] - pacc;
          const T ri = (d[i] - gi) * sn + T(2) * cn * b;
          const T pi = sn * ri;
          e[i + 1] = r;
          d[i + 1] = gi + pi;
          s = sn;
          c = cn;
          pacc = pi;
          g = cn * ri - b;
        }
        d[l] -= pacc;
        e[l] = g;
        e[8] = T(0);
        continue;
      }
#pragma unroll
      for (int i = 7; i >= l; --i) {
        const bool act = i < m;
        const T f = s * e[i], b = c * e[i];
        const T r2 = f * f + g * g;
        const T inv = r2 > T(0) ? inv_sqrt(r2) : T(0);  // one reciprocal square root per rotation
        const T r = r2 * inv;
        const T sn = r2 > T(0) ? f * inv : T(0);
        const T cn = r2 > T(0) ? g * inv : T(1);
        const T gi = d[i + 1] - pacc;
        const T ri = (d[i] - gi) * sn + T(2) * cn * b;
        const T pi = sn * ri;
        e[i + 1] = act ? r : e[i + 1];
        d[i + 1] = act ? gi + pi : d[i + 1];
        s = act ? sn : s;
        c = act ? cn : c;
        pacc = act ? pi : pacc;
        g = act ? cn * ri - b : g;
        if constexpr (WANTZ) {
          const T zs = act ? sn : T(0), zc = act ? cn : T(1);
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            const T zf = Z[k + 9 * (i + 1)], zi = Z[k + 9 * i];
            Z[k + 9 * (i + 1)] = zs * zi + zc * zf;
            Z[k + 9 * i] = zc * zi - zs * zf;
          }
        }
      }
      d[l] -= pacc;
      e[l] = g;
#pragma unroll
      for (int j = l + 1; j < 9; ++j) e[j] = (j == m) ? T(0) : e[j];
    }
  }
}

// Number of eigenvalues of the tridiagonal (d, e) below x: the negative pivots of the LDL^T
// factorisation of T - x I (Sturm count; a zero pivot is replaced by -tiny and counted negative).
template <typename T>
__host__ __device__ __forceinline__ int sturm_below(const T d[9], const T e[9], T x) {
  constexpr T kTiny = sizeof(T) == 8 ? T(1e-300) : T(1e-30);
  T q = d[0] - x;
  q = (q == T(0)) ? -kTiny : q;
  int c = q < T(0) ? 1 : 0;
#pragma unroll
  for (int i = 1; i < 9; ++i) {
    q = (d[i] - x) - e[i - 1] * (e[i - 1] / q);
    q = (q == T(0)) ? -kTiny : q;
    c += q < T(0) ? 1 : 0;
  }
  return c;
}

// The eigenvalues of the tridiagonal (d, e) below x0, by implicit QL that hunts them first: each
// deflation at the top starts with three zero-shift sweeps (the top then converges toward the
// eigenvalue nearest zero -- below a small x0 these are the few smallest in magnitude, well separated
// from the stiff rest) and continues with the Wilkinson shift of tridiag_ql; it stops as soon as the
// deflated values below x0 number what the Sturm count of T - x0 I says.  Returns ldone: d[0 .. ldone)
// are eigenvalues, and every eigenvalue below x0 is among them; with ldone = 9 all of d are.  The
// remaining block's diagonal entries d[ldone ..] are not eigenvalues (its eigenvalues are all >= x0).
// Compared with the full QL on the 64x64x256 beam's reduced matrices (two eigenvalues below 1e-3 in
// every tet): 51 rotations and 9 sweeps per matrix instead of 103 and 28, the same count in every
// tet of a wave (two zero-shift sweeps: 51 on average but up to 55 in a wave), and the same
// eigenvalues to 1e-15 of |M|.
template <typename T>
__host__ __device__ __forceinline__ int tridiag_ql_below(T d[9], T e[9], T x0) {
  constexpr T kEps = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07);
  const int want = sturm_below(d, e, x0);
  int found = 0, ldone = 9;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    if (found == want) {
      ldone = l;
      break;
    }
    for (int iter = 0; iter < 60; ++iter) {
      int m = 8;
#pragma unroll
      for (int j = 7; j >= l; --j)
        if (fabs(e[j]) <= kEps * (fabs(d[j]) + fabs(d[j + 1]))) m = j;
      if (m == l) break;
      T dm = d[8];
#pragma unroll
      for (int j = l + 1; j < 8; ++j) dm = (j == m) ? d[j] : dm;
      T g = dm;  // zero shift
      if (iter >= 3) {
        const T gg = (d[l + 1] - d[l]) / (T(2) * e[l]);
        const T r0 = sqrt(gg * gg + T(1));
        g = dm - d[l] + e[l] / (gg + (gg >= T(0) ? r0 : -r0));
      }
      T s = T(1), c = T(1), pacc = T(0);
#if defined(__HIP_DEVICE_COMPILE__)
      const bool nosplit = __all(m == 8);
#else
      const bool nosplit = m == 8;
#endif
#pragma unroll
      for (int i = 7; i >= l; --i) {
        const bool act = nosplit || i < m;
        const T f = s * e[i], b = c * e[i];
        const T r2 = f * f + g * g;
        const T inv = r2 > T(0) ? inv_sqrt(r2) : T(0);
        const T r = r2 * inv;
        const T sn = f * inv;  // f = 0 when r2 = 0
        const T cn = r2 > T(0) ? g * inv : T(1);
        const T gi = d[i + 1] - pacc;
        const T ri = (d[i] - gi) * sn + T(2) * cn * b;
        const T pi = sn * ri;
        e[i + 1] = act ? r : e[i + 1];
        d[i + 1] = act ? gi + pi : d[i + 1];
        s = act ? sn : s;
        c = act ? cn : c;
        pacc = act ? pi : pacc;
        g = act ? cn * ri - b : g;
      }
      d[l] -= pacc;
      e[l] = g;
#pragma unroll
      for (int j = l + 1; j < 9; ++j) e[j] = (j == m) ? T(0) : e[j];
    }
    found += d[l] < x0 ? 1 : 0;
  }
  return ldone;
}

// Householder reduction of a symmetric 9x9 (packed upper triangle a[sym9(i, j)]) to tridiagonal form
// (d, e), e[i] coupling d[i] and d[i+1]: reflector k (k = 0..6) maps row k's tail a(k, k+1..8) onto its
// first entry and is kept in that row of a with its scale beta[k], so Z = H_0 H_1 ... H_6 (the
// eigenvectors' back-transformation) can be applied later.
template <typename T>
__host__ __device__ __forceinline__ void sym9_tridiag(T a[45], T d[9], T e[9], T beta[7]) {
  constexpr T kTiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38);
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    T sig2 = T(0);
#pragma unroll
    for (int j = k + 2; j < 9; ++j) sig2 += a[sym9(k, j)] * a[sym9(k, j)];
    const T x0 = a[sym9(k, k + 1)];
    const T nrm = sqrt(sig2 + x0 * x0);
    const T alpha = x0 >= T(0) ? -nrm : nrm;
    const T vtv = T(2) * nrm * (nrm + fabs(x0));
    const bool act = sig2 > T(0) && vtv > kTiny;
    const T bt = act ? T(2) / vtv : T(0);
    e[k] = act ? alpha : x0;
    T v[9];
#pragma unroll
    for (int j = k + 1; j < 9; ++j) v[j] = (j == k + 1) ? x0 - alpha : a[sym9(k, j)];
    // A22 <- H A22 H = A22 - v w^T - w v^T with p = bt A22 v, w = p - (bt/2)(v.p) v
    T p[9];
    T vp = T(0);
#pragma unroll
    for (int i = k + 1; i < 9; ++i) {
      T s = T(0);
#pragma unroll
      for (int j = k + 1; j < 9; ++j) s += a[sym9(i, j)] * v[j];
      p[i] = bt * s;
      vp += v[i] * p[i];
    }
    const T kk = T(0.5) * bt * vp;
#pragma unroll
    for (int i = k + 1; i < 9; ++i) p[i] -= kk * v[i];
#pragma unroll
    for (int i = k + 1; i < 9; ++i)
#pragma unroll
      for (int j = i; j < 9; ++j) a[sym9(i, j)] -= v[i] * p[j] + p[i] * v[j];
#pragma unroll
    for (int j = k + 1; j < 9; ++j) a[sym9(k, j)] = v[j];
    beta[k] = bt;
  }
  e[7] = a[sym9(7, 8)];
  e[8] = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = a[sym9(i, i)];
}

// Eigen-decomposition of a symmetric 9x9 (packed upper triangle a[sym9(i, j)], destroyed):
// Householder reduction to tridiagonal form, the reflectors accumulated into Z, then implicit QL
// sweeps with the Wilkinson-type shift from the leading 2x2 (Golub & Van Loan, Matrix
// Computations, 4th ed., sections 8.3.1 and 8.3.5, in the QL orientation that deflates from the
// top).  On return d holds the eigenvalues (unsorted) and column c of Z (Z[r + 9 c]) the unit
// eigenvector of d[c].  Every array index is a compile-time constant after unrolling, so a, d, e
// and Z stay in registers.
template <typename T>
__host__ __device__ inline void sym9_eigen(T a[45], T d[9], T Z[81]) {
  T e[9], beta[7];
  sym9_tridiag(a, d, e, beta);
  // Z = H_0 H_1 ... H_6, accumulated from the right end (H_k touches rows and columns k+1..8)
#pragma unroll
  for (int i = 0; i < 81; ++i) Z[i] = (i % 10 == 0) ? T(1) : T(0);
#pragma unroll
  for (int k = 6; k >= 0; --k) {
#pragma unroll
    for (int c = k + 1; c < 9; ++c) {
      T s = T(0);
#pragma unroll
      for (int r = k + 1; r < 9; ++r) s += a[sym9(k, r)] * Z[r + 9 * c];
      s *= beta[k];
#pragma unroll
      for (int r = k + 1; r < 9; ++r) Z[r + 9 * c] -= s * a[sym9(k, r)];
    }
  }
  tridiag_ql<true>(d, e, Z);
}

// Clamp correction of a reduced matrix M (packed upper triangle, destroyed): with
// M = sum_i lambda_i v_i v_i^T, dM = M' - M = sum_{lambda_i < tol} (tol - lambda_i) v_i v_i^T.
// Returns the number k of eigenvalues below tol and stores dM in the form the assembly reads:
//   k <= 3: k columns w_j = sqrt(tol - lambda_j) v_j at out[9 j .. 9 j + 8] (dM = W W^T);
//   k >  3: the packed upper triangle, out[sym9(r, c)] (45 values).
template <typename T>
__host__ __device__ inline int psd_correction(T a[45], T tol, T* out) {
  T d[9], Z[81];
  sym9_eigen(a, d, Z);
  int k = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) k += (d[i] < tol) ? 1 : 0;
  if (k <= 3) {
    // column j of out for the j-th eigenvalue below tol; the column is chosen by predication over
    // compile-time indices (a run-time index would put out in scratch memory)
    int j = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      if (d[i] < tol) {
        const T w = sqrt(tol - d[i]);
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
          if (j == jj)
#pragma unroll
            for (int r = 0; r < 9; ++r) out[9 * jj + r] = w * Z[r + 9 * i];
        ++j;
      }
    }
  } else {
    T lift[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) lift[i] = (d[i] < tol) ? (tol - d[i]) : T(0);
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
      for (int c = r; c < 9; ++c) {
        T sacc = T(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) sacc += Z[r + 9 * i] * lift[i] * Z[c + 9 * i];
        out[sym9(r, c)] = sacc;
      }
  }
  return k;
}

// One step of inverse iteration on a symmetric tridiagonal T (d, e): x <- (T - mu I)^{-1} x by
// Gaussian elimination with row interchanges (at step i the pivot row is whichever of the active
// row i and row i+1 has the larger entry in column i, so U gains a second superdiagonal), fused with
// the forward substitution; pivots smaller than pmin in magnitude are replaced by +-pmin, as
// inverse iteration at a computed eigenvalue needs the nearly singular solve (Golub & Van Loan,
// section 8.2.2; Wilkinson, The Algebraic Eigenvalue Problem, ch. 5).
template <typename T>
__host__ __device__ __forceinline__ void tridiag_shift_solve(const T d[9], const T e[9], T mu, T pmin, T x[9]) {
  T r0[9], u1[8], u2[7];
  T al = d[0] - mu, be = e[0], act = x[0];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const T c = e[i], dn = d[i + 1] - mu, en = (i < 7) ? e[i + 1] : T(0);
    const bool sw = fabs(c) > fabs(al);
    T p0 = sw ? c : al;
    p0 = fabs(p0) < pmin ? (p0 < T(0) ? -pmin : pmin) : p0;
    const T rp = T(1) / p0;
    const T p1 = sw ? dn : be, p2 = sw ? en : T(0);
    const T ml = (sw ? al : c) * rp;
    al = (sw ? be : dn) - ml * p1;
    be = (sw ? T(0) : en) - ml * p2;
    const T nx = x[i + 1];
    const T piv = sw ? nx : act;
    act = (sw ? act : nx) - ml * piv;
    x[i] = piv;
    r0[i] = rp;
    u1[i] = p1;
    if (i < 7) u2[i] = p2;
  }
  al = fabs(al) < pmin ? (al < T(0) ? -pmin : pmin) : al;
  x[8] = act / al;
  x[7] = (x[7] - u1[7] * x[8]) * r0[7];
#pragma unroll
  for (int i = 6; i >= 0; --i) x[i] = (x[i] - u1[i] * x[i + 1] - u2[i] * x[i + 2]) * r0[i];
}

// x <- x / |x|, scaled by its largest entry first so the squares cannot overflow
template <typename T>
__host__ __device__ __forceinline__ void normalize9(T x[9]) {
  T mx = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) mx = fmax(mx, fabs(x[i]));
  const T rs = mx > T(0) ? T(1) / mx : T(1);
  T n2 = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    x[i] *= rs;
    n2 += x[i] * x[i];
  }
  const T rn = n2 > T(0) ? inv_sqrt(n2) : T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] *= rn;
}

// x <- x - (x.y) y for unit y
template <typename T>
__host__ __device__ __forceinline__ void deflate9(T x[9], const T y[9]) {
  T s = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) s += x[i] * y[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) x[i] -= s * y[i];
}

// Clamp correction of a reduced matrix M (packed upper triangle, destroyed) when at most three
// eigenvalues lie below tol, without forming the eigenvector matrix: Householder tridiagonalisation
// (reflectors kept), the eigenvalues below tol by tridiag_ql_below on a copy of (d, e) (no rotations
// applied to vectors), then for each eigenvalue mu_j < tol kInvIt steps of inverse iteration on the tridiagonal
// form from a fixed start vector, each followed by Gram-Schmidt against the vectors already found
// (members of a cluster then span its invariant subspace; the clamp only needs that subspace, and
// the error any rotation within a cluster leaves in dM is bounded by the cluster's width), and the
// result mapped back through the reflectors.  Hands each column w_j = sqrt(tol - mu_j) v_j to
// emit(j, w_j) as soon as it is formed and returns k; for k > 3 it emits nothing (the caller runs
// psd_correction on the matrix again).
// Inverse-iteration steps per vector.  On the 64x64x256 beam's reduced matrices the clamp is within
// 8e-16 of |M| of the full eigensolver's after two steps (1.9e-15 after one); three change nothing.
constexpr int kInvIt = 2;

template <typename T, typename Emit>
__host__ __device__ inline int psd_correction_low(T a[45], T tol, Emit&& emit) {
  constexpr T kEps = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07);
  constexpr T kFloor = sizeof(T) == 8 ? T(1e-290) : T(1e-30);
  T d[9], e[9], beta[7];
  sym9_tridiag(a, d, e, beta);
  T lam[9], le[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    lam[i] = d[i];
    le[i] = e[i];
  }
  // FP64 (the device's FP64 pass): only the eigenvalues below tol; FP32 (host checks only): all of them
  if constexpr (sizeof(T) == 8) {
    const int ldone = tridiag_ql_below(lam, le, tol);
#pragma unroll
    for (int i = 0; i < 9; ++i) lam[i] = (i >= ldone) ? tol : lam[i];  // not eigenvalues; none below tol
  } else {
    tridiag_ql<false>(lam, le, (T*)nullptr);
  }
  int k = 0;
#pragma unroll
  for (int i = 0; i < 9; ++i) k += (lam[i] < tol) ? 1 : 0;
  if (k > 3) return k;
  T nrm = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) nrm = fmax(nrm, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : T(0)));
  const T pmin = fmax(kEps * nrm, kFloor);
  // fixed start vectors, entries of magnitude 0.25..1 with mixed signs
  const T start[27] = {-0.878349, 0.700260, -0.389749, 0.521367, 0.650314, -0.272610, 0.266693, -0.944151, 0.636744,
                       0.754404,  0.638914, -0.455676, -0.561953, 0.431300, 0.811135, 0.623978, 0.274232, -0.751266,
                       -0.314536, 0.953812, 0.275041, -0.520199, 0.844775, -0.664799, 0.968093, -0.325816, -0.891424};
  // mus[j] = the j-th eigenvalue below tol (in QL order)
  T mus[3] = {T(0), T(0), T(0)};
  {
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const bool sel = lam[i] < tol;
#pragma unroll
      for (int j = 0; j < 3; ++j) mus[j] = (sel && cnt == j) ? lam[i] : mus[j];
      cnt += sel ? 1 : 0;
    }
  }
  T y[3][9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < k) {
      const T mu = mus[j];
      T x[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) x[i] = start[9 * j + i];
#pragma unroll
      for (int it = 0; it < kInvIt; ++it) {
        tridiag_shift_solve(d, e, mu, pmin, x);
#pragma unroll
        for (int q = 0; q < j; ++q) deflate9(x, y[q]);
        if (it == kInvIt - 1) {
#pragma unroll
          for (int q = 0; q < j; ++q) deflate9(x, y[q]);
        }
        normalize9(x);
      }
#pragma unroll
      for (int i = 0; i < 9; ++i) y[j][i] = x[i];
      // v = H_0 H_1 ... H_6 x
#pragma unroll
      for (int r = 6; r >= 0; --r) {
        T sdot = T(0);
#pragma unroll
        for (int c = r + 1; c < 9; ++c) sdot += a[sym9(r, c)] * x[c];
        sdot *= beta[r];
#pragma unroll
        for (int c = r + 1; c < 9; ++c) x[c] -= sdot * a[sym9(r, c)];
      }
      const T w = sqrt(tol - mu);
#pragma unroll
      for (int r = 0; r < 9; ++r) x[r] *= w;
      emit(j, x);
    }
  }
  return k;
}

// class of a tet's clamp: 0 none, 1..3 eigenvalues below tol, 4 more
__host__ __device__ __forceinline__ int psd_format(int k) { return k == 0 ? 0 : (k <= 3 ? k : 4); }

// Tm[r] = sum_p qa[p] S_{p r} (3x3 sub-blocks of the packed symmetric 9x9 S, row-major l, k): the
// row factor of the corner whose row of Q is qa
template <typename T>
__host__ __device__ __forceinline__ void row_factor(const T qa[3], const T* S, T Tm[3][9]) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        T s = T(0);
#pragma unroll
        for (int p = 0; p < 3; ++p) s += qa[p] * S[sym9(3 * p + l, 3 * r + kk)];
        Tm[r][3 * l + kk] = s;
      }
}

// blk = the (a, m) block of (Q (x) I3) S (Q (x) I3)^T + tolPos I3 for the packed symmetric 9x9 S,
// from the rows ql = Q[lo] and qh = Q[hi] of the corners lo = min(a, m), hi = max(a, m): computed as
// the (lo, hi) block (row factor of lo, then hi's combination) and transposed when a > m, so the
// (m, a) block of a tet is bit for bit the transpose of its (a, m) block; a diagonal block is
// symmetrized as (B + B^T) / 2.
template <typename T>
__host__ __device__ __forceinline__ void rec_block(const T ql[3], const T qh[3], int a, int m, const T* S, T tolPos,
                                                   T blk[9]) {
  T Tm[3][9];
  row_factor(ql, S, Tm);
  T b[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) b[i] = qh[0] * Tm[0][i] + qh[1] * Tm[1][i] + qh[2] * Tm[2][i];
  const bool tr = a > m, sy = a == m;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T x = b[3 * l + k], y = b[3 * k + l];
      const T v = sy ? T(0.5) * (x + y) : (tr ? y : x);
      blk[3 * l + k] = (l == k) ? v + tolPos : v;
    }
}

}  // namespace bh
