// Host build of the device element math (bartels_amd/csrc/elements.h) so the CPU test suite can
// check every closed-form block against the oracle's dense B^T H B.  Test infrastructure only.
#include "../bartels_amd/csrc/elements.h"

using namespace bh;

template <typename T>
static void make_F(const double* phi12, const double* x12, T phi[4][3], T* F) {
  for (int m = 0; m < 4; ++m)
    for (int j = 0; j < 3; ++j) phi[m][j] = (T)phi12[m + 4 * j];
  if (sizeof(T) == 8) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        F[i + 3 * j] = x12[i] * phi[0][j] + x12[3 + i] * phi[1][j] + x12[6 + i] * phi[2][j] + x12[9 + i] * phi[3][j];
  } else {
    T ed[3][3];
    for (int m = 0; m < 3; ++m)
      for (int i = 0; i < 3; ++i) ed[m][i] = (T)(x12[3 * (m + 1) + i] - x12[i]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) F[i + 3 * j] = ed[0][i] * phi[1][j] + ed[1][i] * phi[2][j] + ed[2][i] * phi[3][j];
  }
}

template <int MODEL, typename T>
static void blocks(const double* phi12, const double* x12, double vol, double p0, double p1, int ref, double* K) {
  T phi[4][3];
  Pre<T> P;
  make_F(phi12, x12, phi, P.F);
  P.vol = (T)vol;
  pre_hessian<MODEL>(P, (T)p0, (T)p1, ref != 0);
  for (int a = 0; a < 4; ++a)
    for (int m = 0; m < 4; ++m) {
      T b[9];
      hess_block<MODEL>(P, phi[a], phi[m], b);
      for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 3; ++k) K[(3 * a + l) + 12 * (3 * m + k)] = (double)b[3 * l + k];
    }
}

template <int MODEL, typename T>
static void grad(const double* phi12, const double* x12, double vol, double p0, double p1, double* g) {
  T phi[4][3], F[9], P[9];
  make_F(phi12, x12, phi, F);
  piola<MODEL>(F, (T)p0, (T)p1, P);
  for (int a = 0; a < 4; ++a) {
    T f[3];
    mv3(P, phi[a], f);
    for (int l = 0; l < 3; ++l) g[3 * a + l] = (double)((T)vol * f[l]);
  }
}

// FP32 stress / energy from G = F - I formed from displacement edges (rest positions X12)
template <int MODEL, typename T>
static void grad_g(const double* phi12, const double* x12, const double* X12, double vol, double p0, double p1,
                   double* g) {
  T phi[4][3], F[9], G[9], P[9];
  for (int m = 0; m < 4; ++m)
    for (int j = 0; j < 3; ++j) phi[m][j] = (T)phi12[m + 4 * j];
  T ed[3][3];
  for (int m = 0; m < 3; ++m)
    for (int i = 0; i < 3; ++i)
      ed[m][i] = (T)((x12[3 * (m + 1) + i] - X12[3 * (m + 1) + i]) - (x12[i] - X12[i]));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      G[i + 3 * j] = ed[0][i] * phi[1][j] + ed[1][i] * phi[2][j] + ed[2][i] * phi[3][j];
      F[i + 3 * j] = G[i + 3 * j] + ((i == j) ? T(1) : T(0));
    }
  piola<MODEL>(F, (T)p0, (T)p1, P, G);
  for (int a = 0; a < 4; ++a) {
    T f[3];
    mv3(P, phi[a], f);
    for (int l = 0; l < 3; ++l) g[3 * a + l] = (double)((T)vol * f[l]);
  }
}

template <int MODEL, typename T>
static double energy(const double* phi12, const double* x12, double vol, double p0, double p1) {
  T phi[4][3], F[9];
  make_F(phi12, x12, phi, F);
  return psi<MODEL>(F, p0, p1) * vol;
}


// simple_psd_fix through the reduced 9x9 as the device computes it: the record M' = M (+ dM when M -
// tol I fails the Cholesky test: inverse iteration for k <= 3 eigenvalues below tol, the full
// eigensolver on a fresh M for k > 3), then every block from M' and Q.  Returns the storage format
// code of k (0 none, 1..3 low rank, 4 more)
template <int MODEL, typename T>
static int psd_blocks(const double* phi12, const double* x12, double vol, double p0, double p1, int ref, double tol,
                      double* K) {
  T phi[4][3];
  Pre<T> P;
  make_F(phi12, x12, phi, P.F);
  P.vol = (T)vol;
  pre_hessian<MODEL>(P, (T)p0, (T)p1, ref != 0);
  T Q[4][3], R[3][3], M[45], rec[45];
  phi_qr(phi, Q, R);
  reduced_packed<MODEL>(P, R, M);
  for (int i = 0; i < 45; ++i) rec[i] = M[i];
  int k = 0;
  // FP64: the QL eigenvalues decide, inverse iteration for k <= 3, the full eigensolver for k > 3;
  // FP32: a Cholesky test of M - tol I, then the full eigensolver
  if (sizeof(T) == 8 || !shifted_pd(M, (T)tol)) {
    T Mc[45];
    for (int i = 0; i < 45; ++i) Mc[i] = M[i];
    k = sizeof(T) == 8 ? psd_correction_low(Mc, (T)tol, [&](int, const T* w) {
      for (int i = 0; i < 9; ++i)
        for (int j = i; j < 9; ++j) rec[sym9(i, j)] += w[i] * w[j];
    }) : 4;
    if (k > 3) {
      T dM[45];
      k = psd_correction(M, (T)tol, dM);
      for (int i = 0; i < 9; ++i)
        for (int j = i; j < 9; ++j) {
          T acc = rec[sym9(i, j)];
          if (k > 3)
            acc += dM[sym9(i, j)];
          else
            for (int c = 0; c < k; ++c) acc += dM[9 * c + i] * dM[9 * c + j];
          rec[sym9(i, j)] = acc;
        }
    }
  }
  const T tp = (T)((tol > 0 ? tol : 0.0) * 0.25);
  for (int a = 0; a < 4; ++a) {
    for (int m = 0; m < 4; ++m) {
      T b[9];
      rec_block(Q[a < m ? a : m], Q[a < m ? m : a], a, m, rec, tp, b);
      for (int l = 0; l < 3; ++l)
        for (int kk = 0; kk < 3; ++kk) K[(3 * a + l) + 12 * (3 * m + kk)] = (double)b[3 * l + kk];
    }
  }
  return psd_format(k);
}

// eigen-decomposition of a symmetric 9x9 (col-major; upper triangle read): eigenvalues d, vectors Z
template <typename T>
static void eig9(const double* A, double* d, double* Z) {
  T a[45], dd[9], ZZ[81];
  for (int r = 0; r < 9; ++r)
    for (int c = r; c < 9; ++c) a[sym9(r, c)] = (T)A[r + 9 * c];
  sym9_eigen(a, dd, ZZ);
  for (int i = 0; i < 9; ++i) d[i] = (double)dd[i];
  for (int i = 0; i < 81; ++i) Z[i] = (double)ZZ[i];
}

#define DISPATCH(fn, ...)                                              \
  switch (model) {                                                     \
    case 0: return fp32 ? fn<0, float>(__VA_ARGS__) : fn<0, double>(__VA_ARGS__); \
    case 1: return fp32 ? fn<1, float>(__VA_ARGS__) : fn<1, double>(__VA_ARGS__); \
    case 2: return fp32 ? fn<2, float>(__VA_ARGS__) : fn<2, double>(__VA_ARGS__); \
    case 3: return fp32 ? fn<3, float>(__VA_ARGS__) : fn<3, double>(__VA_ARGS__); \
    default: return fp32 ? fn<4, float>(__VA_ARGS__) : fn<4, double>(__VA_ARGS__); \
  }

// clamp correction of a symmetric 9x9 (col-major; upper triangle read) by the inverse-iteration
// path: returns k and, for k <= 3, the columns w_j in W[9 j + r]
template <typename T>
static int psd_low9(const double* A, double tol, double* W) {
  T a[45], w[27];
  for (int r = 0; r < 9; ++r)
    for (int c = r; c < 9; ++c) a[sym9(r, c)] = (T)A[r + 9 * c];
  const int k = psd_correction_low(a, (T)tol, [&](int j, const T* col) {
    for (int r = 0; r < 9; ++r) w[9 * j + r] = col[r];
  });
  for (int i = 0; i < 9 * (k <= 3 ? k : 0); ++i) W[i] = (double)w[i];
  return k;
}
extern "C" {
void eh_blocks(int model, int fp32, const double* phi12, const double* x12, double vol, double p0, double p1, int ref,
               double* K) {
  DISPATCH(blocks, phi12, x12, vol, p0, p1, ref, K)
}
void eh_grad(int model, int fp32, const double* phi12, const double* x12, double vol, double p0, double p1, double* g) {
  DISPATCH(grad, phi12, x12, vol, p0, p1, g)
}
int eh_psd_blocks(int model, int fp32, const double* phi12, const double* x12, double vol, double p0, double p1,
                  int ref, double tol, double* K) {
  DISPATCH(psd_blocks, phi12, x12, vol, p0, p1, ref, tol, K)
}
int eh_psd_low9(int fp32, const double* A, double tol, double* W) {
  return fp32 ? psd_low9<float>(A, tol, W) : psd_low9<double>(A, tol, W);
}
void eh_eig9(int fp32, const double* A, double* d, double* Z) {
  if (fp32)
    eig9<float>(A, d, Z);
  else
    eig9<double>(A, d, Z);
}
void eh_grad_g(int model, const double* phi12, const double* x12, const double* X12, double vol, double p0,
               double p1, double* g) {
  if (model == 0)
    grad_g<0, float>(phi12, x12, X12, vol, p0, p1, g);
  else
    grad_g<1, float>(phi12, x12, X12, vol, p0, p1, g);
}
double eh_energy(int model, int fp32, const double* phi12, const double* x12, double vol, double p0, double p1) {
  DISPATCH(energy, phi12, x12, vol, p0, p1)
}
}
