// Triangle siblings of the tet path (included at the end of bartels_hip.hip; same translation unit):
// the 2-D plane-strain triangles of sim::linear_tri2dmesh_{neohookean,stvk,corotational,arap}_{q,dq,dq2}
// and linear_tri2dmesh_mass_matrix, and the StVK shell triangle of sim::linear_tri3dmesh_stvk_dq2.
//
// tri2d (src/linear_tri2d_neohookean_dq2.cpp:5-35 and siblings): the reference embeds the triangle
// in 3-D, F3 = [[F2, 0], [0, 1]] with F2 = sum_a x_a phi_a^T (plane strain), and evaluates the 3-D
// energies.  So psi, the stress and the element Hessian are the tet path's closed forms at F3, with
// the shape gradients padded to 3-vectors (phi_a, 0); the Hessian's 2x2 block (a, m) is the leading
// 2x2 of the 3-D block.
// tri3d (src/linear_tri3d_stvk_dq2.cpp:6-40): F3 = unflatten(B q_e + vec(n N^T)) with
// B(i + 3j, 3a + c) = [i == c] phi(a, j) + N[j] dndq(i, 3a + c); the element Hessian is the dense
// B^T d2psi B vol with d2psi assembled from the closed-form blocks at unit vectors.
//
// Layout: the element arrays are the reference's (E nT x 3, dphidX nT x 6 or nT x 9 with Phi(m, j) at
// column m + 3j, N and n nT x 3, dndq nT x 27, all column-major).  The mesh handle holds only the
// pattern: block rows of dim x dim blocks (sorted neighbours, the vertex itself included), CSR row
// dim v + l at dim^2 bptr[v] + dim l nb(v), the block of rank r at + dim r (the reference's CSC by
// symmetry), and per-vertex incidence lists in ascending element order.  Assembly: one lane per
// element forms its element matrix (with simple_psd_fix when asked: the full 9x9 eigensolver, a 6x6
// padded with zeros) and the workgroup's matrices leave through LDS as one contiguous block; then 64
// vertices per workgroup fold the blocks of their rows in LDS in ascending element order (the order
// of Eigen's setFromTriplets on the reference's triplet list) and write them out contiguously.

struct bh_trimesh {
  int dim = 2, nV = 0, nT = 0;
  int64_t nnz = 0, nblocks = 0;
  std::vector<int32_t> h_bptr, h_bcol;
  int32_t* E = nullptr;     // nT x 3, column-major
  int32_t* bptr = nullptr;  // nV + 1
  int32_t* bcol = nullptr;
  int32_t* iptr = nullptr;  // nV + 1: incidence list of v at [iptr[v], iptr[v + 1])
  int32_t* inc = nullptr;   // (e << 2) | corner, ascending e
  uint32_t* irank = nullptr;  // ranks of the element's 3 vertices in v's neighbour list, 8 bits each
  double* work = nullptr;     // element results between the two passes (grown on demand, kept)
  size_t workN = 0;
  int maxSeg64 = 0;           // most pattern blocks in the rows of 64 consecutive vertices
};

namespace {

constexpr int kTriModels = 4;  // neohookean, stvk, corotational, arap

void free_trimesh(bh_trimesh* m) {
  void* ptrs[] = {m->E, m->bptr, m->bcol, m->iptr, m->inc, m->irank, m->work};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// the mesh's workspace of at least n doubles (calls on one mesh share it; each call synchronises its
// stream before it returns, so calls one after another are safe, calls overlapping on one mesh are not)
int tri_work(bh_trimesh* m, size_t n, double** out) {
  n = std::max<size_t>(1, n);
  if (m->workN < n) {
    if (m->work) BH_CHECK(hipFree(m->work));
    m->work = nullptr;
    m->workN = 0;
    BH_CHECK(hipMalloc(&m->work, sizeof(double) * n));
    m->workN = n;
  }
  *out = m->work;
  return BH_OK;
}

// device array freed on scope exit
template <typename T>
struct DevArr {
  T* p = nullptr;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    BH_CHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(1, n)));
    return BH_OK;
  }
  int upload(const T* h, size_t n, hipStream_t s) {
    int st = alloc(n);
    if (st) return st;
    if (n) BH_CHECK(hipMemcpyAsync(p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return BH_OK;
  }
};

int build_trimesh(bh_trimesh* m, const int* E) {
  const int nV = m->nV, nT = m->nT;
  // vertex pairs of every element (3 x 3, the vertex itself included), sorted and unique per vertex
  std::vector<int64_t> keys;
  keys.reserve(9 * (size_t)nT);
  for (int e = 0; e < nT; ++e)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        keys.push_back((int64_t)E[e + (int64_t)a * nT] * nV + E[e + (int64_t)b * nT]);
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  m->h_bptr.assign(nV + 1, 0);
  m->h_bcol.resize(keys.size());
  for (size_t k = 0; k < keys.size(); ++k) {
    m->h_bptr[keys[k] / nV + 1]++;
    m->h_bcol[k] = (int32_t)(keys[k] % nV);
  }
  for (int v = 0; v < nV; ++v) m->h_bptr[v + 1] += m->h_bptr[v];
  for (int v0 = 0; v0 < nV; v0 += 64)
    m->maxSeg64 = std::max(m->maxSeg64, m->h_bptr[std::min(v0 + 64, nV)] - m->h_bptr[v0]);
  m->nblocks = (int64_t)keys.size();
  m->nnz = m->nblocks * m->dim * m->dim;
  if (m->nnz >= (1ll << 31)) return BH_UNSUPPORTED;
  // incidence lists in ascending element order, with the ranks of the element's vertices
  std::vector<int32_t> iptr(nV + 1, 0), inc(3 * (size_t)nT);
  std::vector<uint32_t> irank(3 * (size_t)nT);
  for (int e = 0; e < nT; ++e)
    for (int a = 0; a < 3; ++a) iptr[E[e + (int64_t)a * nT] + 1]++;
  for (int v = 0; v < nV; ++v) iptr[v + 1] += iptr[v];
  std::vector<int32_t> fill(iptr.begin(), iptr.end() - 1);
  for (int e = 0; e < nT; ++e)
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
      const int32_t* b0 = m->h_bcol.data() + m->h_bptr[v];
      const int32_t* b1 = m->h_bcol.data() + m->h_bptr[v + 1];
      if (b1 - b0 > 256) return BH_UNSUPPORTED;  // ranks are packed in 8 bits
      uint32_t rk = 0;
      for (int b = 0; b < 3; ++b) {
        const int n = E[e + (int64_t)b * nT];
        rk |= (uint32_t)(std::lower_bound(b0, b1, n) - b0) << (8 * b);
      }
      const int k = fill[v]++;
      inc[k] = (e << 2) | a;
      irank[k] = rk;
    }
  hipStream_t s = nullptr;
  BH_CHECK(hipMalloc(&m->E, sizeof(int32_t) * std::max<size_t>(1, 3 * (size_t)nT)));
  BH_CHECK(hipMalloc(&m->bptr, sizeof(int32_t) * (nV + 1)));
  BH_CHECK(hipMalloc(&m->bcol, sizeof(int32_t) * std::max<size_t>(1, keys.size())));
  BH_CHECK(hipMalloc(&m->iptr, sizeof(int32_t) * (nV + 1)));
  BH_CHECK(hipMalloc(&m->inc, sizeof(int32_t) * std::max<size_t>(1, inc.size())));
  BH_CHECK(hipMalloc(&m->irank, sizeof(uint32_t) * std::max<size_t>(1, irank.size())));
  if (nT) BH_CHECK(hipMemcpyAsync(m->E, E, sizeof(int32_t) * 3 * (size_t)nT, hipMemcpyHostToDevice, s));
  BH_CHECK(hipMemcpyAsync(m->bptr, m->h_bptr.data(), sizeof(int32_t) * (nV + 1), hipMemcpyHostToDevice, s));
  if (!keys.empty())
    BH_CHECK(hipMemcpyAsync(m->bcol, m->h_bcol.data(), sizeof(int32_t) * keys.size(), hipMemcpyHostToDevice, s));
  BH_CHECK(hipMemcpyAsync(m->iptr, iptr.data(), sizeof(int32_t) * (nV + 1), hipMemcpyHostToDevice, s));
  if (nT) {
    BH_CHECK(hipMemcpyAsync(m->inc, inc.data(), sizeof(int32_t) * inc.size(), hipMemcpyHostToDevice, s));
    BH_CHECK(hipMemcpyAsync(m->irank, irank.data(), sizeof(uint32_t) * irank.size(), hipMemcpyHostToDevice, s));
  }
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

// the clamped spectral function of a symmetric n x n (n <= 9, column-major, in place): eigenvalues
// below tol raised to tol (src/simple_psd_fix.cpp:5-21); a matrix smaller than 9 x 9 is padded with
// zeros, whose block is decoupled from it, so the clamp of the padded matrix restricted to the leading
// n x n is the clamp of the matrix
template <int N>
__device__ __forceinline__ void tri_psd(double* K, double tol) {
  double a[45], d[9], Z[81];
#pragma unroll
  for (int i = 0; i < 9; ++i)
#pragma unroll
    for (int j = i; j < 9; ++j) a[sym9(i, j)] = (i < N && j < N) ? K[i + N * j] : 0.0;
  sym9_eigen(a, d, Z);
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = d[i] < tol ? tol : d[i];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) s += Z[r + 9 * i] * d[i] * Z[c + 9 * i];
      K[r + N * c] = s;
    }
}

// F3 of tri2d element e (column-major, F3[i + 3 j]) and its padded shape gradients
__device__ __forceinline__ void tri2_load(int nT, const int32_t* __restrict__ E, const double* __restrict__ q,
                                          const double* __restrict__ dX, int e, double phi[3][3], double F[9]) {
  double x[3][2];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int v = E[e + (int64_t)a * nT];
    x[a][0] = q[2ll * v];
    x[a][1] = q[2ll * v + 1];
    phi[a][0] = dX[e + (int64_t)a * nT];
    phi[a][1] = dX[e + (int64_t)(a + 3) * nT];
    phi[a][2] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) F[i + 3 * j] = x[0][i] * phi[0][j] + x[1][i] * phi[1][j] + x[2][i] * phi[2][j];
  F[8] = 1.0;
}

template <int MODEL>
__device__ double tri2_energy(const double* F, double p0, double p1) {
  return psi<MODEL>(F, p0, p1);
}

template <int MODEL>
__device__ void tri2_forces(const double* F, const double phi[3][3], double vol, double p0, double p1, double* f) {
  double P[9];
  piola<MODEL>(F, p0, p1, P);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) f[2 * a + c] = vol * (P[c] * phi[a][0] + P[c + 3] * phi[a][1]);
}

template <int MODEL>
__device__ void tri2_hessian(const double* F, const double phi[3][3], double vol, double p0, double p1, double* K) {
  Pre<double> P;
#pragma unroll
  for (int i = 0; i < 9; ++i) P.F[i] = F[i];
  P.vol = vol;
  pre_hessian<MODEL>(P, p0, p1, true);
  for (int a = 0; a < 3; ++a)
    for (int mm = 0; mm < 3; ++mm) {
      double blk[9];
      hess_block<MODEL>(P, phi[a], phi[mm], blk);
#pragma unroll
      for (int l = 0; l < 2; ++l)
#pragma unroll
        for (int k = 0; k < 2; ++k) K[(2 * a + l) + 6 * (2 * mm + k)] = blk[3 * l + k];
    }
}

// Cyclic Jacobi on a symmetric N x N (A destroyed): eigenvalues d, eigenvectors in the columns of V
// (Golub & Van Loan, Matrix Computations, 4th ed., section 8.5.2); every index is a compile-time
// constant after unrolling, so A and V stay in registers.
template <int N, int SWEEPS>
__device__ __forceinline__ void jacobi_sym(double A[N][N], double d[N], double V[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < SWEEPS; ++sweep) {
    double off = 0.0, dia = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
      dia += A[p][p] * A[p][p];
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += A[p][q] * A[p][q];
    }
    if (off <= 1e-34 * dia) break;  // converged (quadratically: a few sweeps)
#pragma unroll
    for (int p = 0; p < N - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = A[p][q];
        if (fabs(apq) <= 1e-300) continue;
        const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {  // A <- A J (columns p, q)
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {  // A <- J^T A (rows p, q)
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) d[i] = A[i][i];
}

// The tri2d element Hessian with simple_psd_fix: as for the tets, the three shape gradients sum to
// zero, so with the thin QR Phi = Q R (Q 3 x 2) the 6x6 K = (Q x I2) M (Q x I2)^T with the 4x4
// M(l + 2p, k + 2r) = the leading 2x2 of the closed-form block at (R_p, R_r) (rows of R), and the two
// translations in K's null space.  The clamp is then K' = (Q x I2) M' (Q x I2)^T + max(tol, 0)/3 (1 1^T x I2)
// with M' = M with its eigenvalues raised to tol: the eigenvalues of K are those of M and the two zeros
// of the translations (eigenvectors (1, 1, 1) / sqrt 3 x e_i), which the clamp lifts to tol.
template <int MODEL>
__device__ void tri2_hessian_psd(const double* F, const double phi[3][3], double vol, double p0, double p1,
                                 double tol, double* K) {
  double Q[3][2], R[2][2];
  {
    double r00 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) r00 += phi[a][0] * phi[a][0];
    r00 = sqrt(r00);
    double r01 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Q[a][0] = phi[a][0] / r00;
      r01 += Q[a][0] * phi[a][1];
    }
    double u[3], r11 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      u[a] = phi[a][1] - r01 * Q[a][0];
      r11 += u[a] * u[a];
    }
    r11 = sqrt(r11);
#pragma unroll
    for (int a = 0; a < 3; ++a) Q[a][1] = u[a] / r11;
    R[0][0] = r00;
    R[0][1] = r01;
    R[1][0] = 0.0;
    R[1][1] = r11;
  }
  Pre<double> P;
#pragma unroll
  for (int i = 0; i < 9; ++i) P.F[i] = F[i];
  P.vol = vol;
  pre_hessian<MODEL>(P, p0, p1, true);
  double M[4][4];
#pragma unroll
  for (int pp = 0; pp < 2; ++pp)
#pragma unroll
    for (int rr = pp; rr < 2; ++rr) {
      const double rp[3] = {R[pp][0], R[pp][1], 0.0}, rq[3] = {R[rr][0], R[rr][1], 0.0};
      double blk[9];
      hess_block<MODEL>(P, rp, rq, blk);
#pragma unroll
      for (int l = 0; l < 2; ++l)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          M[l + 2 * pp][k + 2 * rr] = blk[3 * l + k];
          M[k + 2 * rr][l + 2 * pp] = blk[3 * l + k];
        }
    }
  double d[4], V[4][4];
  jacobi_sym<4, 8>(M, d, V);
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = d[i] < tol ? tol : d[i];
  double Mp[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) s += V[i][c] * d[c] * V[j][c];
      Mp[i][j] = s;
    }
  const double tp = (tol > 0.0 ? tol : 0.0) / 3.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int l = 0; l < 2; ++l)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          double s = (l == k) ? tp : 0.0;
#pragma unroll
          for (int pp = 0; pp < 2; ++pp)
#pragma unroll
            for (int rr = 0; rr < 2; ++rr) s += Q[a][pp] * Q[b][rr] * Mp[l + 2 * pp][k + 2 * rr];
          K[(2 * a + l) + 6 * (2 * b + k)] = s;
        }
}

// what the per-element kernel leaves: 0 energy (1 value), 1 forces (6), 2 Hessian (36), 3 mass (36)
template <int MODEL>
__device__ void tri2_element(int what, int nT, const int32_t* E, const double* q, const double* dX,
                             const double* vol, const double* params, int e, double* out) {
  double phi[3][3], F[9];
  tri2_load(nT, E, q, dX, e, phi, F);
  const double p0 = params[e], p1 = params[e + (int64_t)nT];
  if (what == 0)
    out[0] = tri2_energy<MODEL>(F, p0, p1) * vol[e];
  else if (what == 1)
    tri2_forces<MODEL>(F, phi, vol[e], p0, p1, out);
  else
    tri2_hessian<MODEL>(F, phi, vol[e], p0, p1, out);
}

// per element: what (see tri2_element) into out + stride * e; simple_psd_fix(tol) when psd
__global__ void __launch_bounds__(64)
    k_tri2_element(int what, int model, int nT, const int32_t* __restrict__ E, const double* __restrict__ q,
                   const double* __restrict__ dX, const double* __restrict__ vol, const double* __restrict__ params,
                   int psd, double tol, double* __restrict__ out) {
  __shared__ double stage[64 * 37];  // matrices (what 2, 3): the workgroup's 64 element matrices, pitch 37
  const int t = (int)threadIdx.x;
  const int e0 = blockIdx.x * blockDim.x;
  const int e = e0 + t;
  if (what < 2) {  // (what is the same for the whole grid)
    if (e >= nT) return;
    double* o = out + (int64_t)(what == 0 ? 1 : 6) * e;
    switch (model) {
      case M_NH: tri2_element<M_NH>(what, nT, E, q, dX, vol, params, e, o); break;
      case M_STVK: tri2_element<M_STVK>(what, nT, E, q, dX, vol, params, e, o); break;
      case M_COROT: tri2_element<M_COROT>(what, nT, E, q, dX, vol, params, e, o); break;
      default: tri2_element<M_ARAP>(what, nT, E, q, dX, vol, params, e, o); break;
    }
    return;
  }
  if (e < nT) {
    double K[36];
    if (what == 3) {  // consistent mass: (rho vol / 6) I2 on the diagonal blocks, (rho vol / 12) I2 off them
      const double mass = params[e] * vol[e];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c)
          K[r + 6 * c] = (r % 2 == c % 2) ? ((r / 2 == c / 2) ? mass / 6.0 : mass / 12.0) : 0.0;
    } else {
      if (psd) {
        double phi[3][3], F[9];
        tri2_load(nT, E, q, dX, e, phi, F);
        const double p0 = params[e], p1 = params[e + (int64_t)nT], v = vol[e];
        switch (model) {
          case M_NH: tri2_hessian_psd<M_NH>(F, phi, v, p0, p1, tol, K); break;
          case M_STVK: tri2_hessian_psd<M_STVK>(F, phi, v, p0, p1, tol, K); break;
          case M_COROT: tri2_hessian_psd<M_COROT>(F, phi, v, p0, p1, tol, K); break;
          default: tri2_hessian_psd<M_ARAP>(F, phi, v, p0, p1, tol, K); break;
        }
      } else {
        switch (model) {
          case M_NH: tri2_element<M_NH>(what, nT, E, q, dX, vol, params, e, K); break;
          case M_STVK: tri2_element<M_STVK>(what, nT, E, q, dX, vol, params, e, K); break;
          case M_COROT: tri2_element<M_COROT>(what, nT, E, q, dX, vol, params, e, K); break;
          default: tri2_element<M_ARAP>(what, nT, E, q, dX, vol, params, e, K); break;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 36; ++i) stage[37 * t + i] = K[i];
  }
  __syncthreads();
  // the block's matrices are contiguous in out: element e0 + u's entry i at 36 (e0 + u) + i
  const int cnt = 36 * min(64, nT - e0);
  double* out0 = out + 36ll * e0;
  for (int idx = t; idx < cnt; idx += 64) {
    const int u = idx / 36, i = idx - 36 * u;
    out0[idx] = stage[37 * u + i];
  }
}

// The shell clamp through a 6x6 reduced matrix.  With consistent inputs (shape gradients and normal
// derivatives that sum to zero over the three vertices) the 9x9 K has the three translations
// (1, 1, 1) / sqrt 3 x e_c in its null space, so with U (3 x 2) a fixed orthonormal basis of the
// complement of (1, 1, 1), K = (U x I3) M (U x I3)^T, M = (U x I3)^T K (U x I3) (6x6), and
// simple_psd_fix(K, tol) = (U x I3) M' (U x I3)^T + max(tol, 0) / 3 (1 1^T x I3), M' = M with its
// eigenvalues raised to tol (cyclic Jacobi).  When K does not annihilate the translations (to 1e-12 of
// its largest entry) the full 9x9 clamp runs instead.
__device__ __forceinline__ void tri3_psd(double* K, double tol) {
  double kmax = 0.0, tres = 0.0;
#pragma unroll
  for (int i = 0; i < 81; ++i) kmax = fmax(kmax, fabs(K[i]));
#pragma unroll
  for (int r = 0; r < 9; ++r)
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) tres = fmax(tres, fabs(K[r + 9 * dd] + K[r + 9 * (3 + dd)] + K[r + 9 * (6 + dd)]));
  if (!(tres <= 1e-12 * kmax)) {
    tri_psd<9>(K, tol);
    return;
  }
  constexpr double s2 = 0.70710678118654752440, s6 = 0.40824829046386301637;  // 1/sqrt(2), 1/sqrt(6)
  const double U[3][2] = {{s2, s6}, {-s2, s6}, {0.0, -2.0 * s6}};
  double M[6][6];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) s += U[a][i] * U[b][j] * K[(3 * a + c) + 9 * (3 * b + dd)];
          M[3 * i + c][3 * j + dd] = s;
        }
#pragma unroll
  for (int r = 0; r < 6; ++r)  // exactly symmetric
#pragma unroll
    for (int c = r + 1; c < 6; ++c) {
      const double x = 0.5 * (M[r][c] + M[c][r]);
      M[r][c] = x;
      M[c][r] = x;
    }
  double d[6], V[6][6];
  jacobi_sym<6, 12>(M, d, V);
#pragma unroll
  for (int i = 0; i < 6; ++i) d[i] = d[i] < tol ? tol : d[i];
  double Mp[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) s += V[i][c] * d[c] * V[j][c];
      Mp[i][j] = s;
      Mp[j][i] = s;
    }
  const double tp = (tol > 0.0 ? tol : 0.0) / 3.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
          double s = (c == dd) ? tp : 0.0;
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) s += U[a][i] * U[b][j] * Mp[3 * i + c][3 * j + dd];
          K[(3 * a + c) + 9 * (3 * b + dd)] = s;
          K[(3 * b + dd) + 9 * (3 * a + c)] = s;
        }
}

// tri3d StVK shell element Hessian (9 x 9) of element e.  Column 3 a + c of B is vec(e_c phi_a^T +
// D_col N^T) (D_col = dndq's column 3 a + c), so by bilinearity K(3 a1 + c1, 3 a2 + c2) = vol (H(phi_a1,
// phi_a2)[c1][c2] + D2 . H(phi_a1, N)[c1][.] + D1 . H(phi_a2, N)[c2][.] + D1^T H(N, N) D2), with
// H(x, y)[l][k] = d2psi(e_l x^T, e_k y^T) the closed-form block (hess_block) -- the reference's
// B^T d2psi B vol without forming B.  The workgroup's 64 matrices leave through LDS as one contiguous block.
__global__ void __launch_bounds__(64)
    k_tri3_stvk_hessian(int nT, const int32_t* __restrict__ E, const double* __restrict__ q,
                        const double* __restrict__ dX, const double* __restrict__ N, const double* __restrict__ n,
                        const double* __restrict__ dndq, const double* __restrict__ vol,
                        const double* __restrict__ params, int psd, double tol, double* __restrict__ out) {
  __shared__ double stage[64 * 82];
  const int t = (int)threadIdx.x;
  const int e0 = blockIdx.x * blockDim.x;
  const int e = e0 + t;
  if (e < nT) {
    double qe[9], Ne[3], ne[3], phi[3][3], D[27];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ne[k] = N[e + (int64_t)k * nT];
      ne[k] = n[e + (int64_t)k * nT];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
#pragma unroll
      for (int c = 0; c < 3; ++c) qe[3 * a + c] = q[3ll * v + c];
#pragma unroll
      for (int j = 0; j < 3; ++j) phi[a][j] = dX[e + (int64_t)(a + 3 * j) * nT];
    }
#pragma unroll
    for (int k = 0; k < 27; ++k) D[k] = dndq[e + (int64_t)k * nT];  // D[i + 3 col]
    // F3 = B q_e + vec(n N^T): F(i, j) = n_i N_j + sum_a phi_a[j] q_a[i] + N_j (D q_e)_i
    Pre<double> P;
    double Dq[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
#pragma unroll
      for (int col = 0; col < 9; ++col) s += D[i + 3 * col] * qe[col];
      Dq[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        P.F[i + 3 * j] = ne[i] * Ne[j] + phi[0][j] * qe[i] + phi[1][j] * qe[3 + i] + phi[2][j] * qe[6 + i] +
                         Ne[j] * Dq[i];
    P.vol = 1.0;
    pre_hessian<M_STVK>(P, params[e], params[e + (int64_t)nT], true);
    double HPN[3][9], HNN[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) hess_block<M_STVK>(P, phi[a], Ne, HPN[a]);
    hess_block<M_STVK>(P, Ne, Ne, HNN);
    const double v = vol[e];
    double K[81];
#pragma unroll
    for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
      for (int a2 = a1; a2 < 3; ++a2) {
        double HPP[9];
        hess_block<M_STVK>(P, phi[a1], phi[a2], HPP);
#pragma unroll
        for (int c1 = 0; c1 < 3; ++c1)
#pragma unroll
          for (int c2 = 0; c2 < 3; ++c2) {
            const double* D1 = D + 3 * (3 * a1 + c1);
            const double* D2 = D + 3 * (3 * a2 + c2);
            double x = HPP[3 * c1 + c2];
#pragma unroll
            for (int k = 0; k < 3; ++k) x += D2[k] * HPN[a1][3 * c1 + k] + D1[k] * HPN[a2][3 * c2 + k];
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
              for (int k = 0; k < 3; ++k) x += D1[l] * HNN[3 * l + k] * D2[k];
            const int r = 3 * a1 + c1, c = 3 * a2 + c2;
            K[r + 9 * c] = x * v;
            K[c + 9 * r] = x * v;
          }
      }
    if (psd) tri3_psd(K, tol);
#pragma unroll
    for (int i = 0; i < 81; ++i) stage[82 * t + i] = K[i];
  }
  __syncthreads();
  const int cnt = 81 * min(64, nT - e0);
  double* out0 = out + 81ll * e0;
  for (int idx = t; idx < cnt; idx += 64) {
    const int u = idx / 81, i = idx - 81 * u;
    out0[idx] = stage[82 * u + i];
  }
}

// energy: fixed-partition partial sums of the per-element energies (block b sums its contiguous range
// in index order), then one workgroup sums the partials in block order
__global__ void __launch_bounds__(256) k_tri_sum(int64_t n, const double* __restrict__ x, int nparts,
                                                 double* __restrict__ part) {
  __shared__ double sh[256];
  const int64_t per = (n + nparts - 1) / nparts;
  const int64_t i0 = (int64_t)blockIdx.x * per, i1 = min(n, i0 + per);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// gradient: g[dim v + c] = sum over v's incident elements, ascending, of the element's corner force
__global__ void __launch_bounds__(256)
    k_tri_gather_vec(int nV, int dim, const int32_t* __restrict__ iptr, const int32_t* __restrict__ inc,
                     const double* __restrict__ f, double* __restrict__ g) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const int ne = 3 * dim;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = iptr[v]; k < iptr[v + 1]; ++k) {
    const int code = inc[k], e = code >> 2, a = code & 3;
    for (int c = 0; c < dim; ++c) acc[c] += f[(int64_t)ne * e + dim * a + c];
  }
  for (int c = 0; c < dim; ++c) g[(int64_t)dim * v + c] = acc[c];
}

// Hessian / mass: the rows of vertex v, folded from the element matrices (ne x ne, column-major) of its
// incident elements in ascending element order; each lane owns its vertex's rows
__global__ void __launch_bounds__(256)
    k_tri_gather_mat(int nV, int dim, const int32_t* __restrict__ bptr, const int32_t* __restrict__ iptr,
                     const int32_t* __restrict__ inc, const uint32_t* __restrict__ irank,
                     const double* __restrict__ K, double* __restrict__ vals) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const int ne = 3 * dim, nb = bptr[v + 1] - bptr[v];
  double* row = vals + (int64_t)dim * dim * bptr[v];
  for (int i = 0; i < dim * dim * nb; ++i) row[i] = 0.0;
  for (int k = iptr[v]; k < iptr[v + 1]; ++k) {
    const int code = inc[k], e = code >> 2, a = code & 3;
    const uint32_t rk = irank[k];
    const double* Ke = K + (int64_t)ne * ne * e;
    for (int mm = 0; mm < 3; ++mm) {
      double* dst = row + dim * ((rk >> (8 * mm)) & 0xffu);
      for (int l = 0; l < dim; ++l)
        for (int c = 0; c < dim; ++c) dst[dim * l * nb + c] += Ke[(dim * a + l) + ne * (dim * mm + c)];
    }
  }
}

// The same fold with 64 vertices per workgroup and their rows accumulated in LDS: four lanes per vertex,
// lane mm of a vertex adding column block mm of each incident element's matrix (the element's three
// vertices are distinct, so the lanes of one step touch distinct slots; the steps run in ascending
// element order, the lanes of a vertex sharing a wave), then the rows leave as one contiguous run.  The
// sums are the same as k_tri_gather_mat's, bit for bit.
template <int DIM>
__global__ void __launch_bounds__(256)
    k_tri_gather_lds(int nV, const int32_t* __restrict__ bptr, const int32_t* __restrict__ iptr,
                     const int32_t* __restrict__ inc, const uint32_t* __restrict__ irank,
                     const double* __restrict__ K, double* __restrict__ vals) {
  extern __shared__ double acc[];
  constexpr int NE = 3 * DIM;
  const int v0 = blockIdx.x * 64;
  const int vend = min(v0 + 64, nV);
  const int b0 = bptr[v0];
  const int seglen = DIM * DIM * (bptr[vend] - b0);
  for (int i = threadIdx.x; i < seglen; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int v = v0 + ((int)threadIdx.x >> 2), mm = (int)threadIdx.x & 3;
  if (v < nV && mm < 3) {
    const int nb = bptr[v + 1] - bptr[v];
    double* row = acc + DIM * DIM * (bptr[v] - b0);
    for (int k = iptr[v]; k < iptr[v + 1]; ++k) {
      const int code = inc[k], e = code >> 2, a = code & 3;
      const double* Ke = K + (int64_t)NE * NE * e + DIM * a + NE * DIM * mm;
      double* dst = row + DIM * ((irank[k] >> (8 * mm)) & 0xffu);
#pragma unroll
      for (int l = 0; l < DIM; ++l)
#pragma unroll
        for (int c = 0; c < DIM; ++c) dst[DIM * l * nb + c] += Ke[l + NE * c];
    }
  }
  __syncthreads();
  double* out = vals + (int64_t)DIM * DIM * b0;
  for (int i = threadIdx.x; i < seglen; i += blockDim.x) out[i] = acc[i];
}

// the fold of the element matrices K (ne x ne per element) into vals: the LDS gather when the rows of 64
// vertices fit, else one lane per vertex in global memory
int tri_fold(const bh_trimesh* m, const double* K, double* vals, hipStream_t s) {
  const size_t lds = sizeof(double) * (size_t)m->dim * m->dim * m->maxSeg64;
  const unsigned g64 = (unsigned)((m->nV + 63) / 64);
  if (lds <= (size_t)max_lds_bytes()) {
    if (m->dim == 2) {
      gather_lds_attr(k_tri_gather_lds<2>);
      hipLaunchKernelGGL(k_tri_gather_lds<2>, dim3(g64), dim3(256), lds, s, m->nV, m->bptr, m->iptr, m->inc,
                         m->irank, K, vals);
    } else {
      gather_lds_attr(k_tri_gather_lds<3>);
      hipLaunchKernelGGL(k_tri_gather_lds<3>, dim3(g64), dim3(256), lds, s, m->nV, m->bptr, m->iptr, m->inc,
                         m->irank, K, vals);
    }
  } else {
    hipLaunchKernelGGL(k_tri_gather_mat, dim3((m->nV + 255) / 256), dim3(256), 0, s, m->nV, m->dim, m->bptr, m->iptr,
                       m->inc, m->irank, K, vals);
  }
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

bool tri2_model_ok(int model) { return model >= 0 && model < kTriModels; }

// the element matrices of kind what (2 Hessian, 3 mass) folded into vals; inputs on the device
int tri2_matrix(bh_trimesh* m, int what, int model, const double* q, const double* dX, const double* vol,
                const double* params, const bh_psd* psd, double* vals, hipStream_t s) {
  if (m->nV == 0) return BH_OK;
  double* K = nullptr;
  int st = tri_work(m, 36 * (size_t)m->nT, &K);
  if (st) return st;
  if (m->nT) {
    hipLaunchKernelGGL(k_tri2_element, dim3((m->nT + 63) / 64), dim3(64), 0, s, what, model, m->nT, m->E, q, dX, vol,
                       params, (psd && psd->enabled) ? 1 : 0, psd ? psd->tol : 0.0, K);
    BH_CHECK(hipGetLastError());
  }
  if ((st = tri_fold(m, K, vals, s))) return st;
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

// host arrays staged to the device: q (dim nV), dX (nT x cols), vol (nT), params (nT x 2)
struct TriStage {
  DevArr<double> q, dX, vol, params;
  int up(const bh_trimesh* m, const double* hq, const double* hdX, int cols, const double* hvol, const double* hp,
         hipStream_t s) {
    int st;
    if (hq && (st = q.upload(hq, (size_t)m->dim * m->nV, s))) return st;
    if (hdX && (st = dX.upload(hdX, (size_t)cols * m->nT, s))) return st;
    if (hvol && (st = vol.upload(hvol, (size_t)m->nT, s))) return st;
    if (hp && (st = params.upload(hp, 2 * (size_t)m->nT, s))) return st;
    return BH_OK;
  }
};

}  // namespace

extern "C" {

int bh_linear_tri2dmesh_dphi_dX(int nV, const double* V, int nT, const int* E, double* dphidX) {
  if (nV < 0 || nT < 0 || (nV && !V) || (nT && (!E || !dphidX))) return BH_BAD_ARG;
  int status = BH_OK;
  for (int e = 0; e < nT; ++e) {
    double x[3][2];
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
      if (v < 0 || v >= nV) return BH_BAD_ARG;
      for (int c = 0; c < 2; ++c) x[a][c] = V[v + (int64_t)c * nV];
    }
    const double a00 = x[1][0] - x[0][0], a10 = x[1][1] - x[0][1], a01 = x[2][0] - x[0][0], a11 = x[2][1] - x[0][1];
    const double det = a00 * a11 - a01 * a10;
    const double inv[2][2] = {{a11 / det, -a01 / det}, {-a10 / det, a00 / det}};
    for (int j = 0; j < 2; ++j) {
      dphidX[e + (int64_t)(1 + 3 * j) * nT] = inv[0][j];
      dphidX[e + (int64_t)(2 + 3 * j) * nT] = inv[1][j];
      dphidX[e + (int64_t)(0 + 3 * j) * nT] = -(inv[0][j] + inv[1][j]);
      if (!std::isfinite(inv[0][j]) || !std::isfinite(inv[1][j])) status = BH_NONFINITE;
    }
  }
  return status;
}

int bh_linear_tri3dmesh_dphi_dX(int nV, const double* V, int nT, const int* E, double* dphidX) {
  if (nV < 0 || nT < 0 || (nV && !V) || (nT && (!E || !dphidX))) return BH_BAD_ARG;
  int status = BH_OK;
  for (int e = 0; e < nT; ++e) {
    double x[3][3];
    for (int a = 0; a < 3; ++a) {
      const int v = E[e + (int64_t)a * nT];
      if (v < 0 || v >= nV) return BH_BAD_ARG;
      for (int c = 0; c < 3; ++c) x[a][c] = V[v + (int64_t)c * nV];
    }
    double T[3][2], G[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int c = 0; c < 3; ++c) {
      T[c][0] = x[1][c] - x[0][c];
      T[c][1] = x[2][c] - x[0][c];
    }
    for (int r = 0; r < 2; ++r)
      for (int t = 0; t < 2; ++t)
        for (int c = 0; c < 3; ++c) G[r][t] += T[c][r] * T[c][t];
    const double det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    const double Gi[2][2] = {{G[1][1] / det, -G[0][1] / det}, {-G[1][0] / det, G[0][0] / det}};
    for (int j = 0; j < 3; ++j) {
      const double r0 = Gi[0][0] * T[j][0] + Gi[0][1] * T[j][1], r1 = Gi[1][0] * T[j][0] + Gi[1][1] * T[j][1];
      dphidX[e + (int64_t)(1 + 3 * j) * nT] = r0;
      dphidX[e + (int64_t)(2 + 3 * j) * nT] = r1;
      dphidX[e + (int64_t)(0 + 3 * j) * nT] = -(r0 + r1);
      if (!std::isfinite(r0) || !std::isfinite(r1)) status = BH_NONFINITE;
    }
  }
  return status;
}

int bh_trimesh_create(int dim, int nV, int nT, const int* E, bh_trimesh** out) {
  if (!out || (dim != 2 && dim != 3) || nV < 0 || nT < 0 || (nT && !E)) return BH_BAD_ARG;
  *out = nullptr;
  if ((int64_t)nT >= (1ll << 29)) return BH_UNSUPPORTED;
  for (int64_t i = 0; i < 3ll * nT; ++i)
    if (E[i] < 0 || E[i] >= nV) return BH_BAD_ARG;
  bh_trimesh* m = new bh_trimesh();
  m->dim = dim;
  m->nV = nV;
  m->nT = nT;
  const int st = build_trimesh(m, E);
  if (st != BH_OK) {
    free_trimesh(m);
    delete m;
    return st;
  }
  *out = m;
  return BH_OK;
}

int bh_trimesh_destroy(bh_trimesh* m) {
  if (!m) return BH_BAD_ARG;
  (void)hipDeviceSynchronize();
  free_trimesh(m);
  delete m;
  return BH_OK;
}

int bh_trimesh_nnz(const bh_trimesh* m, int64_t* nnz) {
  if (!m || !nnz) return BH_BAD_ARG;
  *nnz = m->nnz;
  return BH_OK;
}

int bh_trimesh_pattern(const bh_trimesh* m, int32_t* rowptr, int32_t* colind) {
  if (!m || !rowptr || (m->nnz && !colind)) return BH_BAD_ARG;
  const int d = m->dim;
  int64_t k = 0;
  rowptr[0] = 0;
  for (int v = 0; v < m->nV; ++v)
    for (int l = 0; l < d; ++l) {
      for (int b = m->h_bptr[v]; b < m->h_bptr[v + 1]; ++b)
        for (int c = 0; c < d; ++c) colind[k++] = d * m->h_bcol[b] + c;
      rowptr[d * v + l + 1] = (int32_t)k;
    }
  return BH_OK;
}

int bh_linear_tri2dmesh_q_device(bh_trimesh* m, int model, const double* q, const double* dphidX,
                                 const double* volume, const double* params, double* d_energy, bh_stream_t stream) {
  if (!m || m->dim != 2 || !tri2_model_ok(model) || !d_energy || (m->nT && (!q || !dphidX || !volume || !params)))
    return BH_BAD_ARG;
  hipStream_t s = as_stream(stream);
  constexpr int kParts = 256;
  double* ee = nullptr;
  int st = tri_work(m, (size_t)m->nT + kParts, &ee);
  if (st) return st;
  double* part = ee + m->nT;
  if (m->nT) {
    hipLaunchKernelGGL(k_tri2_element, dim3((m->nT + 63) / 64), dim3(64), 0, s, 0, model, m->nT, m->E, q, dphidX,
                       volume, params, 0, 0.0, ee);
    BH_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_tri_sum, dim3(kParts), dim3(256), 0, s, (int64_t)m->nT, (const double*)ee, kParts, part);
  hipLaunchKernelGGL(k_tri_sum, dim3(1), dim3(256), 0, s, (int64_t)kParts, (const double*)part, 1, d_energy);
  BH_CHECK(hipGetLastError());
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

int bh_linear_tri2dmesh_dq_device(bh_trimesh* m, int model, const double* q, const double* dphidX,
                                  const double* volume, const double* params, double* d_g, bh_stream_t stream) {
  if (!m || m->dim != 2 || !tri2_model_ok(model) || !d_g || (m->nT && (!q || !dphidX || !volume || !params)))
    return BH_BAD_ARG;
  hipStream_t s = as_stream(stream);
  double* f = nullptr;
  int st = tri_work(m, 6 * (size_t)m->nT, &f);
  if (st) return st;
  if (m->nT) {
    hipLaunchKernelGGL(k_tri2_element, dim3((m->nT + 63) / 64), dim3(64), 0, s, 1, model, m->nT, m->E, q, dphidX,
                       volume, params, 0, 0.0, f);
    BH_CHECK(hipGetLastError());
  }
  if (m->nV) {
    hipLaunchKernelGGL(k_tri_gather_vec, dim3((m->nV + 255) / 256), dim3(256), 0, s, m->nV, 2, m->iptr, m->inc,
                       (const double*)f, d_g);
    BH_CHECK(hipGetLastError());
  }
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

int bh_linear_tri2dmesh_dq2_device(bh_trimesh* m, int model, const double* q, const double* dphidX,
                                   const double* volume, const double* params, const bh_psd* psd, double* d_vals,
                                   bh_stream_t stream) {
  if (!m || m->dim != 2 || !tri2_model_ok(model) || !d_vals || (m->nT && (!q || !dphidX || !volume || !params)))
    return BH_BAD_ARG;
  return tri2_matrix(m, 2, model, q, dphidX, volume, params, psd, d_vals, as_stream(stream));
}

int bh_linear_tri2dmesh_q(bh_trimesh* m, int model, const double* q, const double* dphidX, const double* volume,
                          const double* params, double* energy) {
  if (!m || !energy) return BH_BAD_ARG;
  TriStage S;
  DevArr<double> e;
  int st;
  if ((st = S.up(m, q, dphidX, 6, volume, params, nullptr)) || (st = e.alloc(1))) return st;
  if ((st = bh_linear_tri2dmesh_q_device(m, model, S.q.p, S.dX.p, S.vol.p, S.params.p, e.p, nullptr))) return st;
  BH_CHECK(hipMemcpy(energy, e.p, sizeof(double), hipMemcpyDeviceToHost));
  return BH_OK;
}

int bh_linear_tri2dmesh_dq(bh_trimesh* m, int model, const double* q, const double* dphidX, const double* volume,
                           const double* params, double* g) {
  if (!m || !g) return BH_BAD_ARG;
  TriStage S;
  DevArr<double> d;
  int st;
  if ((st = S.up(m, q, dphidX, 6, volume, params, nullptr)) || (st = d.alloc(2 * (size_t)m->nV))) return st;
  if ((st = bh_linear_tri2dmesh_dq_device(m, model, S.q.p, S.dX.p, S.vol.p, S.params.p, d.p, nullptr))) return st;
  if (m->nV) BH_CHECK(hipMemcpy(g, d.p, sizeof(double) * 2 * (size_t)m->nV, hipMemcpyDeviceToHost));
  return BH_OK;
}

int bh_linear_tri2dmesh_dq2(bh_trimesh* m, int model, const double* q, const double* dphidX, const double* volume,
                            const double* params, const bh_psd* psd, double* vals) {
  if (!m || !vals) return BH_BAD_ARG;
  TriStage S;
  DevArr<double> d;
  int st;
  if ((st = S.up(m, q, dphidX, 6, volume, params, nullptr)) || (st = d.alloc((size_t)m->nnz))) return st;
  if ((st = bh_linear_tri2dmesh_dq2_device(m, model, S.q.p, S.dX.p, S.vol.p, S.params.p, psd, d.p, nullptr))) return st;
  if (m->nnz) BH_CHECK(hipMemcpy(vals, d.p, sizeof(double) * (size_t)m->nnz, hipMemcpyDeviceToHost));
  return BH_OK;
}

int bh_linear_tri2dmesh_mass_matrix(bh_trimesh* m, const double* density, const double* volume, double* vals) {
  if (!m || m->dim != 2 || !vals || (m->nT && (!density || !volume))) return BH_BAD_ARG;
  hipStream_t s = nullptr;
  DevArr<double> rho, vol, d;
  int st;
  if ((st = rho.upload(density, (size_t)m->nT, s)) || (st = vol.upload(volume, (size_t)m->nT, s)) ||
      (st = d.alloc((size_t)m->nnz)))
    return st;
  // the mass kind reads params[e] as the density
  if ((st = tri2_matrix(m, 3, M_NH, nullptr, nullptr, vol.p, rho.p, nullptr, d.p, s))) return st;
  if (m->nnz) BH_CHECK(hipMemcpy(vals, d.p, sizeof(double) * (size_t)m->nnz, hipMemcpyDeviceToHost));
  return BH_OK;
}

int bh_linear_tri3dmesh_stvk_dq2_device(bh_trimesh* m, const double* q, const double* dphidX, const double* N,
                                        const double* n, const double* dndq, const double* volume,
                                        const double* params, const bh_psd* psd, double* d_vals, bh_stream_t stream) {
  if (!m || m->dim != 3 || !d_vals || (m->nT && (!q || !dphidX || !N || !n || !dndq || !volume || !params)))
    return BH_BAD_ARG;
  if (m->nV == 0) return BH_OK;
  hipStream_t s = as_stream(stream);
  double* K = nullptr;
  int st = tri_work(m, 81 * (size_t)m->nT, &K);
  if (st) return st;
  if (m->nT) {
    hipLaunchKernelGGL(k_tri3_stvk_hessian, dim3((m->nT + 63) / 64), dim3(64), 0, s, m->nT, m->E, q, dphidX, N, n,
                       dndq, volume, params, (psd && psd->enabled) ? 1 : 0, psd ? psd->tol : 0.0, K);
    BH_CHECK(hipGetLastError());
  }
  if ((st = tri_fold(m, K, d_vals, s))) return st;
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

int bh_linear_tri3dmesh_stvk_dq2(bh_trimesh* m, const double* q, const double* dphidX, const double* N,
                                 const double* n, const double* dndq, const double* volume, const double* params,
                                 const bh_psd* psd, double* vals) {
  if (!m || m->dim != 3 || !vals || (m->nT && (!q || !dphidX || !N || !n || !dndq || !volume || !params)))
    return BH_BAD_ARG;
  if (m->nV == 0) return BH_OK;
  hipStream_t s = nullptr;
  TriStage S;
  DevArr<double> dN, dn, dD, d;
  int st;
  if ((st = S.up(m, q, dphidX, 9, volume, params, s)) || (st = dN.upload(N, 3 * (size_t)m->nT, s)) ||
      (st = dn.upload(n, 3 * (size_t)m->nT, s)) || (st = dD.upload(dndq, 27 * (size_t)m->nT, s)) ||
      (st = d.alloc((size_t)m->nnz)))
    return st;
  if ((st = bh_linear_tri3dmesh_stvk_dq2_device(m, S.q.p, S.dX.p, dN.p, dn.p, dD.p, S.vol.p, S.params.p, psd, d.p,
                                                nullptr)))
    return st;
  if (m->nnz) BH_CHECK(hipMemcpy(vals, d.p, sizeof(double) * (size_t)m->nnz, hipMemcpyDeviceToHost));
  return BH_OK;
}

}  // extern "C"
