// Device Newton solver (included at the end of bartels_hip.hip; same translation unit).
//
// Replaces the reference's newtons_method_bisection (include/optimization_gradient_based.h:33-110)
// with search_direction_newton_sparse (src/search_direction_newton_sparse.cpp:5-26) and
// linesearch_backtracking_bisection (src/linesearch_backtracking_bisection.cpp:5-31):
//   * gradient / Hessian / energy come from the gather kernels above (FP64);
//   * the direct sparse factorisation becomes a 3x3-block-Jacobi preconditioned CG to a relative
//     residual opts->pcg_tol, run entirely on the device: the CG scalars live in device memory,
//     every reduction is a fixed-order two-level tree (deterministic), and the host only polls a
//     convergence flag every kPcgBatch iterations;
//   * fixed DOFs are handled as P H P^T + I on the fixed rows (equivalent to the reduced system).

namespace {

int red_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(kRedMaxBlocks, (n + 4095) / 4096)); }

// partial dot products over a fixed partition of [0, n)
__global__ void __launch_bounds__(kRedBlock) k_dot_partials(const double* __restrict__ a, const double* __restrict__ b,
                                                            int64_t n, double* __restrict__ partials) {
  __shared__ double sh[4];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kRedBlock) s += a[i] * b[i];
  s = block_sum_256(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_sum_partials(const double* __restrict__ partials, int np,
                                                            double* __restrict__ out) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedBlock) s += partials[i];
  s = block_sum_256(s, sh);
  if (threadIdx.x == 0) *out = s;
}

int device_dot(const double* a, const double* b, int64_t n, double* partials, double* out, hipStream_t s) {
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot_partials, dim3(nb), dim3(kRedBlock), 0, s, a, b, n, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedBlock), 0, s, partials, nb, out);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// g <- mask ? 0 : g - f
__global__ void __launch_bounds__(256) k_mask_gradient(int64_t n, const uint8_t* __restrict__ fixed,
                                                       const double* __restrict__ f, double* __restrict__ g) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = g[i] - (f ? f[i] : 0.0);
  if (fixed && fixed[i]) v = 0.0;
  g[i] = v;
}

// inverse of the 3x3 diagonal block of each vertex (fixed DOFs -> identity rows / columns)
template <typename V>
__global__ void __launch_bounds__(256) k_block_jacobi(int nV, const int32_t* __restrict__ bptr,
                                                      const int32_t* __restrict__ bcol, const V* __restrict__ vals,
                                                      const double* __restrict__ hdiag,
                                                      const uint8_t* __restrict__ fixed, int point,
                                                      double* __restrict__ Dinv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const int b0 = bptr[v], nb = bptr[v + 1] - b0;
  int r = 0;
  while (r < nb && bcol[b0 + r] != v) ++r;
  double A[9];  // row-major block; a vertex that belongs to no tet gets the identity
  if (r < nb && hdiag) {
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = hdiag[9ll * v + i];
  } else if (r < nb) {
    const V* base = vals + 9ll * b0 + 3ll * r;
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k) A[3 * l + k] = (double)base[3ll * l * nb + k];
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  if (point) {
    A[1] = A[2] = A[3] = A[5] = A[6] = A[7] = 0.0;
  }
  if (fixed) {
#pragma unroll
    for (int l = 0; l < 3; ++l)
      if (fixed[3ll * v + l]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          A[3 * l + k] = (l == k) ? 1.0 : 0.0;
          A[3 * k + l] = (l == k) ? 1.0 : 0.0;
        }
      }
  }
  const double c0 = A[4] * A[8] - A[5] * A[7], c1 = A[5] * A[6] - A[3] * A[8], c2 = A[3] * A[7] - A[4] * A[6];
  const double det = A[0] * c0 + A[1] * c1 + A[2] * c2;
  double* o = Dinv + 9ll * v;
  if (!(fabs(det) > 0.0) || !isfinite(det)) {
#pragma unroll
    for (int i = 0; i < 9; ++i) o[i] = (i % 4 == 0) ? (A[i] != 0.0 ? 1.0 / A[i] : 1.0) : 0.0;
    return;
  }
  const double id = 1.0 / det;
  o[0] = c0 * id;
  o[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  o[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  o[3] = c1 * id;
  o[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  o[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  o[6] = c2 * id;
  o[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  o[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

__global__ void __launch_bounds__(256) k_block_inv(int nV, const double* __restrict__ hd, int point,
                                                   double* __restrict__ Dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  double A[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) A[k] = hd[9ll * i + k];
  inv3_block(A, point, Dinv + 9ll * i);
}

// ---- PCG in the Chronopoulos-Gear form -----------------------------------------------------------
// (Chronopoulos & Gear 1989): with u = M^{-1} r, w = H u and s = H p carried by recurrence, the three
// inner products an iteration needs (r.u, r.r, w.u) are all available after the product w = H u, so
// one reduction per iteration suffices:
//   beta = gamma' / gamma,  alpha = gamma' / (delta' - beta gamma' / alpha)
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,  u = M^{-1} r,  w = H u.
// The vector recurrences are fused into the kernel that applies M^{-1} (block Jacobi: one pass; a
// factor: the forward colour sweep), and the partial sums go to three fixed regions of the
// workspace, reduced by one small kernel that also runs the convergence test.
// partial-sum regions of the workspace (r.u, r.r, w.u), each large enough for one partial per block
// of any kernel of the solve, and the second-stage partials after them
inline int64_t part_region(int nV) { return spmv_blocks(nV) + 2048; }
constexpr int kRed1Blocks = 64;

// block-Jacobi step, one thread per DOF row (coalesced), 64 vertices per block (their three rows
// meet in LDS for the 3x3 inverse): (start) x = 0, r = b, p = s = 0; (loop) the CG recurrences;
// then u = Dinv r, partials r.u and r.r
constexpr int kJacVerts = 64;
__global__ void __launch_bounds__(3 * kJacVerts)
    k_cg_vec_jacobi(int nV, int start, const double* __restrict__ b, const double* __restrict__ Dinv,
                    double* __restrict__ x, double* __restrict__ r, double* __restrict__ u, double* __restrict__ p,
                    double* __restrict__ s, const double* __restrict__ w, double* __restrict__ part_ru,
                    double* __restrict__ part_rr, const PcgState* __restrict__ st) {
  __shared__ double rs[3 * kJacVerts];
  __shared__ double sh[4];
  if (!start && cg_skip(st)) return;
  const int64_t i = blockIdx.x * (int64_t)(3 * kJacVerts) + threadIdx.x;
  const int64_t n = 3ll * nV;
  double ri = 0.0;
  if (i < n) {
    if (start) {
      x[i] = 0.0;
      p[i] = 0.0;
      s[i] = 0.0;
      ri = b[i];
    } else {
      const double alpha = st->alpha, beta = st->beta;
      const double pi = u[i] + beta * p[i];
      const double si = w[i] + beta * s[i];
      p[i] = pi;
      s[i] = si;
      x[i] += alpha * pi;
      ri = r[i] - alpha * si;
    }
    r[i] = ri;
  }
  rs[threadIdx.x] = ri;
  __syncthreads();
  double ru = 0.0, rr = 0.0;
  if (i < n) {
    const int l = (int)(i % 3);
    const int base = threadIdx.x - l;
    const double* D = Dinv + 3 * i;  // row l of the vertex's 3x3 inverse (9 v + 3 l = 3 i)
    const double ui = D[0] * rs[base] + D[1] * rs[base + 1] + D[2] * rs[base + 2];
    u[i] = ui;
    ru = ri * ui;
    rr = ri * ri;
  }
  ru = block_sum<3 * kJacVerts>(ru, sh);
  rr = block_sum<3 * kJacVerts>(rr, sh);
  if (threadIdx.x == 0) {
    part_ru[blockIdx.x] = ru;
    part_rr[blockIdx.x] = rr;
  }
}
inline int jac_blocks(int nV) { return std::max(1, (nV + kJacVerts - 1) / kJacVerts); }
inline int vec_blocks(int nV) { return std::max(1, std::min((nV + kRedBlock - 1) / kRedBlock, 2048)); }

// second stage of the reduction (kRed1Blocks partials per quantity) and the CG scalar step, run by
// one workgroup of 1024 threads
__device__ void cg_scalar_step(const double* __restrict__ part, int start, double tol, int maxit,
                               PcgState* __restrict__ st, double* sh) {
  if (!start && st->done) {
    // a non-positive-curvature exit in the first iteration applied its alpha = 1 update: count it once
    if (st->done == 3 && threadIdx.x == 0 && st->alpha != 0.0) {
      st->alpha = 0.0;
      st->iters += 1;
    }
    return;
  }
  double a = 0.0, b = 0.0, c = 0.0;
  if (threadIdx.x < kRed1Blocks) {
    a = part[threadIdx.x];
    b = part[kRed1Blocks + threadIdx.x];
    c = part[2 * kRed1Blocks + threadIdx.x];
  }
  a = block_sum<1024>(a, sh);
  b = block_sum<1024>(b, sh);
  c = block_sum<1024>(c, sh);
  if (threadIdx.x != 0) return;
  const double g = a, rr = b, d = c;
  if (start) {
    st->bb = rr;
    st->rr = rr;
    st->tol2 = tol * tol;
    st->iters = 0;
    st->maxit = maxit;
    st->beta = 0.0;
    st->gamma = g;
    st->delta = d;
    if (rr == 0.0) {
      st->alpha = 0.0;
      st->done = 1;
    } else if (!(g > 0.0)) {  // r.u <= 0 (or not finite): the preconditioner is not positive definite
      st->alpha = 0.0;
      st->done = 3;
    } else if (!(d > 0.0)) {  // non-positive curvature: take the preconditioned gradient
      st->alpha = 1.0;
      st->done = 3;
    } else {
      st->alpha = g / d;
      st->done = 0;
    }
    return;
  }
  st->iters += 1;
  st->rr = rr;
  if (rr <= st->tol2 * st->bb) {
    st->done = 1;
  } else if (st->iters >= st->maxit) {
    st->done = 2;
  } else if (!(g > 0.0)) {  // r.u <= 0 or not finite: indefinite preconditioner, keep the current iterate
    st->alpha = 0.0;
    st->done = 3;
  } else {
    const double beta = g / st->gamma;
    const double den = d - beta * g / st->alpha;  // = p'.H p' for the next direction
    st->beta = beta;
    st->gamma = g;
    st->delta = d;
    if (!(den > 0.0)) {  // non-positive curvature: keep the current iterate
      st->alpha = 0.0;
      st->done = 3;
    } else {
      st->alpha = g / den;
    }
  }
}

// first stage of the CG reduction: kRed1Blocks blocks sum fixed slices of the three regions
__global__ void __launch_bounds__(1024)
    k_cg_reduce1(const double* __restrict__ pru, const double* __restrict__ prr, const double* __restrict__ pwu,
                 int nru, int nrr, int nwu, double* __restrict__ out, int* __restrict__ arrived, int start,
                 double tol, int maxit, PcgState* __restrict__ st) {
  __shared__ double sh[16];
  __shared__ int last;
  const int cnt[3] = {nru, nrr, nwu};
  const double* srcs[3] = {pru, prr, pwu};
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const double* src = srcs[g];
    const int per = (cnt[g] + kRed1Blocks - 1) / kRed1Blocks;
    const int lo = blockIdx.x * per, hi = min(cnt[g], lo + per);
    double a = 0.0;
    for (int k = lo + threadIdx.x; k < hi; k += 1024) a += src[k];
    a = block_sum<1024>(a, sh);
    if (threadIdx.x == 0) out[g * kRed1Blocks + blockIdx.x] = a;
  }
  // the workgroup that arrives last runs the second stage and the scalar step (its partials are
  // summed in workgroup order, so the result does not depend on the arrival order)
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(arrived, 1) == kRed1Blocks - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  cg_scalar_step(out, start, tol, maxit, st, sh);
  if (threadIdx.x == 0) *arrived = 0;
}

// start of a factor-preconditioned solve: x = 0, r = b, p = s = 0, partial r.r
__global__ void __launch_bounds__(kRedBlock)
    k_cg_init(int64_t n, const double* __restrict__ b, double* __restrict__ x, double* __restrict__ r,
              double* __restrict__ p, double* __restrict__ s, double* __restrict__ part_rr) {
  __shared__ double sh[4];
  double rr = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedBlock) {
    const double bi = b[i];
    x[i] = 0.0;
    p[i] = 0.0;
    s[i] = 0.0;
    r[i] = bi;
    rr += bi * bi;
  }
  rr = block_sum_256(rr, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = rr;
}

// Eisenstat-form restart: p~ = s~ = 0
__global__ void __launch_bounds__(kRedBlock) k_eis_restart(int64_t n, double* __restrict__ pt,
                                                          double* __restrict__ st) {
  for (int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedBlock) {
    pt[i] = 0.0;
    st[i] = 0.0;
  }
}

// Eisenstat-form start: x~ = 0, p~ = s~ = 0
__global__ void __launch_bounds__(kRedBlock) k_eis_init(int64_t n, double* __restrict__ xt, double* __restrict__ pt,
                                                       double* __restrict__ st) {
  for (int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedBlock) {
    xt[i] = 0.0;
    pt[i] = 0.0;
    st[i] = 0.0;
  }
}

// y = a + s * b
__global__ void __launch_bounds__(256) k_axpy_to(int64_t n, const double* __restrict__ a, double s,
                                                 const double* __restrict__ b, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] + s * b[i];
}

__global__ void k_negate(int64_t n, const double* __restrict__ a, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = -a[i];
}

// backward-Euler inertia pieces: d = q - q_pred; y += a x; vals += a M
__global__ void __launch_bounds__(256) k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] - b[i];
}
__global__ void __launch_bounds__(256) k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
template <typename V>
__global__ void __launch_bounds__(256) k_add_mass(int64_t n, double a, const double* __restrict__ mass,
                                                  V* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) vals[i] = (V)((double)vals[i] + a * mass[i]);
}

// The diagonal blocks of the precision form's Hessian, rebuilt in FP64 from its off-diagonal blocks.
// Every element Hessian is translation invariant (it depends on q only through F, and the shape
// gradients of a tet sum to zero) and the PSD clamp adds exactly tol/4 (1 1^T (x) I) to every tet, so
// block row v of H sums to tol * (tets at v) * I, plus inertia * (row sum of M) with the
// backward-Euler term.  H_vv = that row sum - sum_{n != v} H_vn keeps the FP32 operator exactly
// translation invariant: its soft (smooth, low-stiffness) modes are then as accurate as the FP32
// off-diagonal entries themselves, where a rounded FP32 diagonal perturbs them by ~1e-7 of the
// largest stiffness (on the 6.07M-tet beam that alone cost 1.8x the CG iterations).  The sum of
// the off-diagonal blocks is symmetric in exact arithmetic (it is the row sum minus the symmetric
// H_vv) but not after FP32 rounding; its symmetric part is kept, so the assembled operator is exactly
// symmetric wherever the off-diagonal blocks are exact transposes of each other (CG on the
// unsymmetrised diagonal ended twice as far from the FP64 Newton iterate).
template <typename V>
__global__ void __launch_bounds__(256)
    k_implied_diag(int nV, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                   const V* __restrict__ vals, const int32_t* __restrict__ vdeg, double psd_tol,
                   const double* __restrict__ mass, double inertia, double* __restrict__ hdiag) {
  // eight lanes per vertex, lane j takes blocks j, j + 8, ... of the block row (neighbouring 3-value
  // pieces of each row), then a fixed xor butterfly over the eight lanes
  constexpr int L = 8;
  const int64_t gid = blockIdx.x * 256ll + threadIdx.x;
  const int v = (int)(gid / L), j = (int)(gid % L);
  double S[9], Ms[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = Ms[i] = 0.0;
  if (v < nV) {
    const int b0 = bptr[v], nb = bptr[v + 1] - b0;
    for (int r = j; r < nb; r += L) {
      const bool off = bcol[b0 + r] != v;
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int64_t at = 9ll * b0 + 3ll * l * nb + 3 * r + k;
          if (off) S[3 * l + k] += (double)vals[at];
          if (mass) Ms[3 * l + k] += mass[at];
        }
    }
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      S[i] += __shfl_xor(S[i], off, 64);
      Ms[i] += __shfl_xor(Ms[i], off, 64);
    }
  if (v >= nV || j != 0) return;
  const double g = psd_tol * (double)vdeg[v];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int t = 3 * (i % 3) + i / 3;  // transposed position
    hdiag[9ll * v + i] = ((i % 4 == 0) ? g : 0.0) + inertia * 0.5 * (Ms[i] + Ms[t]) - 0.5 * (S[i] + S[t]);
  }
}

struct NewtonWs {
  double *vals, *g, *b, *x, *r, *z, *p, *s, *Ap, *qt, *Dinv, *hdiag, *part, *scal;
  float* vals32;    // precision form: FP32 Hessian values
  float* params32;  // precision form: FP32 parameters (up to 6 nT: param_cols of the model)
  PcgState* st;
  int* arrived;     // workgroups of the reduction done so far (zero between reductions)
};

int newton_workspace(bh_mesh* m, NewtonWs& w) {
  const int64_t n = 3ll * m->nV;
  const int64_t nParts = 3 * part_region(m->nV) + 3 * kRed1Blocks + 16;
  const size_t need = sizeof(double) * ((size_t)m->nnz + 9 * (size_t)n + 18 * (size_t)m->nV + nParts + 16) +
                      sizeof(float) * ((size_t)m->nnz + 6 * (size_t)m->nT) + sizeof(PcgState) + 512;
  bool fresh = false;
  if (m->wsBytes < need) {
    if (m->ws) BH_CHECK(hipFree(m->ws));
    m->ws = nullptr;
    BH_CHECK(hipMalloc(&m->ws, need));
    m->wsBytes = need;
    fresh = true;
  }
  double* q = m->ws;
  w.vals = q; q += m->nnz;
  w.g = q; q += n;
  w.b = q; q += n;
  w.x = q; q += n;
  w.r = q; q += n;
  w.z = q; q += n;
  w.p = q; q += n;
  w.s = q; q += n;
  w.Ap = q; q += n;
  w.qt = q; q += n;
  w.Dinv = q; q += 9ll * m->nV;
  w.hdiag = q; q += 9ll * m->nV;
  w.part = q; q += nParts;
  w.scal = q; q += 16;
  float* f = reinterpret_cast<float*>(q);
  w.vals32 = f; f += m->nnz;
  w.params32 = f; f += 6ll * m->nT;
  w.st = reinterpret_cast<PcgState*>(reinterpret_cast<uintptr_t>(f + 64) & ~uintptr_t(63));
  w.arrived = reinterpret_cast<int*>(w.st + 1);
  if (fresh) BH_CHECK(hipMemset(w.arrived, 0, sizeof(int)));
  return BH_OK;
}

// objective(q) = Psi(q) - f.q  -> *out (device)
int objective(bh_mesh* m, const bh_newton_opts* o, const double* params, const double* f, const double* q, double* out,
              NewtonWs& w, hipStream_t s) {
  int st = (o->element_precision == 2)
               ? energy_dispatch(m, o->model, o->d_model_ids, q, (const float*)w.params32, out, s)
               : energy_dispatch(m, o->model, o->d_model_ids, q, params, out, s);
  if (st) return st;
  if (f) {
    st = device_dot(f, q, 3ll * m->nV, w.part, w.scal + 8, s);
    if (st) return st;
    hipLaunchKernelGGL(k_axpy_to, dim3(1), dim3(1), 0, s, (int64_t)1, out, -1.0, w.scal + 8, out);
    BH_CHECK(hipGetLastError());
  }
  if (o->d_mass_vals) {
    // + 1/2 inertia d^T M d, d = q - q_pred (w.z, w.Ap are free outside the PCG solve)
    const int64_t n = 3ll * m->nV;
    hipLaunchKernelGGL(k_sub, dim3((n + 255) / 256), dim3(256), 0, s, n, q, o->d_q_pred, w.z);
    hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks_for<double>(m->nV)), dim3(kRedBlock), 0, s, m->nV,
                       m->bptr, m->bcol, o->d_mass_vals, (const double*)nullptr, (const uint8_t*)nullptr, w.z, w.Ap,
                       (double*)nullptr, (const PcgState*)nullptr);
    st = device_dot(w.z, w.Ap, n, w.part, w.scal + 9, s);
    if (st) return st;
    hipLaunchKernelGGL(k_axpy_to, dim3(1), dim3(1), 0, s, (int64_t)1, out, 0.5 * o->inertia, w.scal + 9, out);
    BH_CHECK(hipGetLastError());
  }
  return BH_OK;
}

int read_scalar(const double* d, double* h, hipStream_t s) {
  BH_CHECK(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

// the two reduction stages and the scalar step (partial counts of the three regions)
inline void cg_reduce_regions(const NewtonWs& w, int nV, const double* pru, int nru, const double* prr, int nrr,
                              const double* pwu, int nwu, int start, double tol, int maxit, hipStream_t s) {
  const int64_t R = part_region(nV);
  double* stage2 = w.part + 3 * R;
  hipLaunchKernelGGL(k_cg_reduce1, dim3(kRed1Blocks), dim3(1024), 0, s, pru, prr, pwu, nru, nrr, nwu, stage2,
                     w.arrived, start, tol, maxit, w.st);
}
inline void cg_reduce(const NewtonWs& w, int nV, int nru, int nrr, int nwu, int start, const bh_newton_opts* o,
                      hipStream_t s) {
  const int64_t R = part_region(nV);
  cg_reduce_regions(w, nV, w.part, nru, w.part + R, nrr, w.part + 2 * R, nwu, start, o->pcg_tol,
                    o->pcg_max_iterations, s);
}

int cg_wait(const NewtonWs& w, PcgState* hs, hipStream_t s) {
  BH_CHECK(hipGetLastError());
  BH_CHECK(hipMemcpyAsync(hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  return BH_OK;
}

// PCG solve of H x = b on the masked system with block (or point) Jacobi; the iteration count goes to
// *iters, the exit state to *done
template <typename V>
int pcg_solve(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag, const uint8_t* fixed,
              NewtonWs& w, int* iters, int* done, hipStream_t s) {
  const int nV = m->nV;
  const int gj = jac_blocks(nV);
  const int gs = spmv_blocks_for<V>(nV);
  const int64_t R = part_region(nV);
  double* pru = w.part;
  double* prr = w.part + R;
  double* pwu = w.part + 2 * R;
  hipLaunchKernelGGL(k_block_jacobi<V>, dim3((nV + 255) / 256), dim3(256), 0, s, nV, m->bptr, m->bcol, vals, hdiag,
                     fixed, o->precond == BH_PRECOND_POINT_JACOBI ? 1 : 0, w.Dinv);
  hipLaunchKernelGGL(k_cg_vec_jacobi, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, 1, w.b, w.Dinv, w.x, w.r, w.z, w.p,
                     w.s, w.Ap, pru, prr, w.st);
  hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(gs), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals, hdiag, fixed, w.z,
                     w.Ap, pwu, (const PcgState*)nullptr);
  cg_reduce(w, nV, gj, gj, gs, 1, o, s);
  PcgState hs{};
  int st;
  for (;;) {
    for (int k = 0; k < kPcgBatch; ++k) {
      hipLaunchKernelGGL(k_cg_vec_jacobi, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, 0, w.b, w.Dinv, w.x, w.r, w.z,
                         w.p, w.s, w.Ap, pru, prr, w.st);
      hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(gs), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals, hdiag, fixed,
                         w.z, w.Ap, pwu, w.st);
      cg_reduce(w, nV, gj, gj, gs, 0, o, s);
    }
    if ((st = cg_wait(w, &hs, s))) return st;
    if (hs.done) break;
  }
  *iters = hs.iters;
  *done = hs.done;
  return BH_OK;
}

// ---- two-level block Jacobi (BH_PRECOND_JACOBI_2L) -------------------------------------------
// M^-1 = D^-1 + Z E^-1 Z^T: block Jacobi plus an additive coarse correction on aggregates of
// consecutive vertices (a multiple of the vector kernel's 64 vertices each, so every workgroup of
// that kernel lies in one aggregate).  Z holds the quadratic modes of each aggregate I: column
// (l, i) is e_l m_i(d), d = s_I (q_v - c_I) the position about the aggregate's centroid scaled by its
// inverse RMS radius, m = (1, d0, d1, d2, d0^2, d1^2, d2^2, d0 d1, d1 d2, d0 d2): thirty columns,
// which contain the translations, rotations and linear modes and the bending of a slab.  Rows of
// fixed DOFs are zero.  E = Z^T A Z (30 nagg square, dense) is formed once per system and inverted
// in place (Gauss-Jordan on the SPD matrix, no pivoting); per CG iteration the correction costs a
// reduction of Z^T r per aggregate, one dense product with E^-1 and a vertex pass u += Z y.  On a
// 96x24x24 beam: block Jacobi 2173 CG iterations, 16 aggregates with the 12 affine modes 271, 40
// affine aggregates 266, 8 quadratic aggregates 198, 16 quadratic 187.  E is accumulated with
// atomic adds, so its rounding, and with it the iterates, can differ between runs in the last bits.
constexpr int kMono = 10;
constexpr int kModes = 3 * kMono;  // column (l, i) = kMono l + i
struct Coarse {
  int nagg = 0, per = 0, k = 0;
  double *cent, *E, *g, *y, *pzr;  // cent: c_I and s_I, four per aggregate; pzr mode-major
  float* dinv;                     // the block-Jacobi inverse in FP32 (exactly symmetric blocks)
  int32_t* pos;                    // vertex -> position of a colour order (the factor solvers)
};

__device__ __forceinline__ void monomials(const double d[3], double m[kMono]) {
  m[0] = 1.0;
  m[1] = d[0];
  m[2] = d[1];
  m[3] = d[2];
  m[4] = d[0] * d[0];
  m[5] = d[1] * d[1];
  m[6] = d[2] * d[2];
  m[7] = d[0] * d[1];
  m[8] = d[1] * d[2];
  m[9] = d[0] * d[2];
}

// the scaled position of vertex v about aggregate I's centroid
__device__ __forceinline__ void agg_coords(const double* __restrict__ cent, int I, double q0, double q1, double q2,
                                           double d[3]) {
  const double* c = cent + 4 * I;
  d[0] = (q0 - c[0]) * c[3];
  d[1] = (q1 - c[1]) * c[3];
  d[2] = (q2 - c[2]) * c[3];
}

// aggregate layout for nV vertices and a requested count (0: one per 4096 vertices, 4 to 8; the
// dense inverse of E is formed by one workgroup and grows as k^3: ~20 ms per Newton step at
// k = 384, ~170 ms at k = 768, hence the cap at k = 480)
constexpr int kMaxAggregates = 16;
inline void coarse_layout(int nV, int want, int& nagg, int& per) {
  if (want <= 0) want = std::min(8, std::max(4, nV / 4096));
  want = std::min(want, kMaxAggregates);
  per = (nV + want - 1) / want;
  per = ((per + kJacVerts - 1) / kJacVerts) * kJacVerts;
  nagg = std::max(1, (nV + per - 1) / per);
}

int coarse_workspace(bh_mesh* m, int want, int gj, Coarse& c) {
  coarse_layout(m->nV, want, c.nagg, c.per);
  c.k = kModes * c.nagg;
  const size_t nd = 4 * (size_t)c.nagg + (size_t)c.k * c.k + 2 * (size_t)c.k + kModes * (size_t)gj + 64;
  const size_t need = sizeof(double) * nd + sizeof(float) * 9 * (size_t)m->nV + sizeof(int32_t) * (size_t)m->nV;
  if (m->cwsBytes < need) {
    if (m->cws) BH_CHECK(hipFree(m->cws));
    m->cws = nullptr;
    BH_CHECK(hipMalloc(&m->cws, need));
    m->cwsBytes = need;
  }
  double* q = m->cws;
  c.cent = q; q += 4 * c.nagg;
  c.E = q; q += (size_t)c.k * c.k;
  c.g = q; q += c.k;
  c.y = q; q += c.k;
  c.pzr = q;
  c.dinv = (float*)(m->cws + nd);
  c.pos = (int32_t*)(c.dinv + 9 * (size_t)m->nV);
  return BH_OK;
}

// centroid c_I and inverse RMS radius s_I of each aggregate (one workgroup per aggregate, fixed-order
// sums)
__global__ void __launch_bounds__(256) k_agg_centroid(int nV, int per, const double* __restrict__ q,
                                                      double* __restrict__ cent) {
  __shared__ double sh[4];
  __shared__ double cs[3];
  const int v0 = blockIdx.x * per, v1 = min(nV, v0 + per);
  const double cnt = (double)max(1, v1 - v0);
  double a[3] = {0.0, 0.0, 0.0};
  for (int v = v0 + threadIdx.x; v < v1; v += 256)
#pragma unroll
    for (int l = 0; l < 3; ++l) a[l] += q[3ll * v + l];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const double t = block_sum_256(a[l], sh);
    if (threadIdx.x == 0) cs[l] = t / cnt;
  }
  __syncthreads();
  double r2 = 0.0;
  for (int v = v0 + threadIdx.x; v < v1; v += 256)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double d = q[3ll * v + l] - cs[l];
      r2 += d * d;
    }
  r2 = block_sum_256(r2, sh);
  if (threadIdx.x == 0) {
    const double rms = sqrt(r2 / cnt);
#pragma unroll
    for (int l = 0; l < 3; ++l) cent[4 * blockIdx.x + l] = cs[l];
    cent[4 * blockIdx.x + 3] = rms > 0.0 ? 1.0 / rms : 1.0;
  }
}

__global__ void __launch_bounds__(256) k_fill_zero(int64_t n, double* __restrict__ a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

__global__ void __launch_bounds__(256) k_narrow(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// the masked block A_vn of the CG operator: fixed rows and columns zero (Z's rows of fixed DOFs are
// zero), the diagonal block from hdiag when given.  Masked entries are selected at the load (a
// conditional store into the array after the loads left the fixed-column mask without effect in the
// compiled code).
template <typename V>
__device__ __forceinline__ void masked_row_block(const V* __restrict__ vals, const double* __restrict__ hdiag,
                                                 const uint8_t* __restrict__ fixed, int v, int n, int b0, int nb,
                                                 int r, const bool fv[3], double A[9]) {
  bool fn[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) fn[c] = fixed && fixed[3ll * n + c];
  const double* src = (n == v && hdiag) ? hdiag + 9ll * v : nullptr;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double a = src ? src[3 * l + c] : (double)vals[9ll * b0 + 3ll * l * nb + 3 * r + c];
      A[3 * l + c] = (fv[l] || fn[c]) ? 0.0 : a;
    }
}

// B += A_vn Z_n (3 x 30, row-major): column (l, i) is A e_l m_i(d_n)
__device__ __forceinline__ void quad_right(const double A[9], const double mn[kMono], double B[3 * kModes]) {
#pragma unroll
  for (int rr = 0; rr < 3; ++rr)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double a = A[3 * rr + l];
#pragma unroll
      for (int i = 0; i < kMono; ++i) B[kModes * rr + kMono * l + i] += a * mn[i];
    }
}

// E += Z^T A Z.  A workgroup takes 64 rows v of one aggregate I; for each neighbouring aggregate
// J = I - 1, I, I + 1 every lane sums B_v = sum_{n in J} A_vn Z_n in registers, the 64 B_v and the
// rows' monomials meet in LDS, and the block E(I, J) += sum_v Z_v^T B_v (entry ((l, i), col) =
// sum_v m_i(v) B_v[l][col]) is contracted there, one atomic add per entry and workgroup.  Blocks of
// aggregates farther away (none on slab meshes) are added per block.
template <typename V>
__global__ void __launch_bounds__(64)
    k_coarse_E(int nV, int per, int nagg, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
               const V* __restrict__ vals, const double* __restrict__ hdiag, const uint8_t* __restrict__ fixed,
               const double* __restrict__ q, const double* __restrict__ cent, double* __restrict__ E) {
  __shared__ double ms[64][kMono];
  __shared__ double bs[64][3 * kModes + 1];
  const int t = threadIdx.x;
  const int v = blockIdx.x * 64 + t;
  const int I = (blockIdx.x * 64) / per;
  const int k = kModes * nagg;
  const bool in = v < nV;
  int b0 = 0, nb = 0;
  bool fv[3] = {false, false, false};
  double mv[kMono];
#pragma unroll
  for (int i = 0; i < kMono; ++i) mv[i] = 0.0;
  if (in) {
    b0 = bptr[v];
    nb = bptr[v + 1] - b0;
#pragma unroll
    for (int l = 0; l < 3; ++l) fv[l] = fixed && fixed[3ll * v + l];
    double d[3];
    agg_coords(cent, I, q[3ll * v], q[3ll * v + 1], q[3ll * v + 2], d);
    monomials(d, mv);
  }
#pragma unroll
  for (int i = 0; i < kMono; ++i) ms[t][i] = mv[i];
  for (int J = max(0, I - 1); J <= min(nagg - 1, I + 1); ++J) {
    double B[3 * kModes];
#pragma unroll
    for (int e = 0; e < 3 * kModes; ++e) B[e] = 0.0;
    for (int r = 0; r < nb; ++r) {
      const int n = bcol[b0 + r];
      if (n / per != J) continue;
      double A[9];
      masked_row_block(vals, hdiag, fixed, v, n, b0, nb, r, fv, A);
      double d[3], mn[kMono];
      agg_coords(cent, J, q[3ll * n], q[3ll * n + 1], q[3ll * n + 2], d);
      monomials(d, mn);
      quad_right(A, mn, B);
    }
    __syncthreads();  // the previous contraction has read bs
#pragma unroll
    for (int e = 0; e < 3 * kModes; ++e) bs[t][e] = B[e];
    __syncthreads();
    for (int o = t; o < kModes * kModes; o += 64) {
      const int row = o / kModes, col = o % kModes;
      const int l = row / kMono, i = row % kMono;
      double a = 0.0;
      for (int s = 0; s < 64; ++s) a += ms[s][i] * bs[s][kModes * l + col];
      atomicAdd(&E[(size_t)(kModes * I + row) * k + kModes * J + col], a);
    }
  }
  // aggregates beyond the neighbours (general vertex orders): per block, straight to E
  for (int r = 0; r < nb; ++r) {
    const int n = bcol[b0 + r];
    const int J = n / per;
    if (J >= I - 1 && J <= I + 1) continue;
    double A[9];
    masked_row_block(vals, hdiag, fixed, v, n, b0, nb, r, fv, A);
    double d[3], mn[kMono];
    agg_coords(cent, J, q[3ll * n], q[3ll * n + 1], q[3ll * n + 2], d);
    monomials(d, mn);
    double B[3 * kModes];
#pragma unroll
    for (int e = 0; e < 3 * kModes; ++e) B[e] = 0.0;
    quad_right(A, mn, B);
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < kMono; ++i)
        for (int col = 0; col < kModes; ++col)
          atomicAdd(&E[(size_t)(kModes * I + kMono * l + i) * k + kModes * J + col], mv[i] * B[kModes * l + col]);
  }
}

// E <- E^-1 in place (Gauss-Jordan without pivoting, one workgroup).  A pivot at or below 1e-12
// of its original diagonal entry belongs to a mode with no free DOF or one that the earlier modes
// already span (an aggregate whose free vertices share a coordinate plane makes a linear mode a
// multiple of a translation): it gets a zero row and column, which drops the mode from Z and keeps
// the correction positive semi-definite.
__global__ void __launch_bounds__(1024) k_coarse_inv(int k, double* __restrict__ E) {
  extern __shared__ double col[];  // 3k: column p, row p, the original diagonal
  double* row = col + k;
  double* dg = col + 2 * k;
  for (int i = threadIdx.x; i < k; i += 1024) dg[i] = E[(size_t)i * k + i];
  for (int p = 0; p < k; ++p) {
    for (int i = threadIdx.x; i < k; i += 1024) {
      col[i] = E[(size_t)i * k + p];
      row[i] = E[(size_t)p * k + i];
    }
    __syncthreads();
    const double d = col[p];
    const double piv = (d > 1e-12 * dg[p] && d > 0.0) ? 1.0 / d : 0.0;
    // wave wv takes rows wv, wv + 16, ...; its lanes the columns.  A thread loads a 4 x 4 tile of its
    // entries before it updates them (16 loads in flight instead of one)
    constexpr int R = 4, C = 4;
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    for (int i0 = wv; i0 < k; i0 += 16 * R)
      for (int j0 = lane; j0 < k; j0 += 64 * C) {
        double a[R][C];
#pragma unroll
        for (int x = 0; x < R; ++x)
#pragma unroll
          for (int y = 0; y < C; ++y) {
            const int i = i0 + 16 * x, j = j0 + 64 * y;
            a[x][y] = (i < k && j < k) ? E[(size_t)i * k + j] : 0.0;
          }
#pragma unroll
        for (int x = 0; x < R; ++x)
#pragma unroll
          for (int y = 0; y < C; ++y) {
            const int i = i0 + 16 * x, j = j0 + 64 * y;
            if (i >= k || j >= k) continue;
            double e;
            if (i == p && j == p)
              e = piv;
            else if (i == p)
              e = row[j] * piv;
            else if (j == p)
              e = -col[i] * piv;
            else
              e = a[x][y] - col[i] * row[j] * piv;
            E[(size_t)i * k + j] = e;
          }
      }
    __syncthreads();
  }
}

// the coarse right-hand side g = Z^T r: the partials of each aggregate's vector-kernel workgroups,
// summed in workgroup order (one workgroup per aggregate and mode)
__global__ void __launch_bounds__(256) k_coarse_g(int per, int nblocks, const double* __restrict__ pzr,
                                                  double* __restrict__ g, const PcgState* __restrict__ st) {
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int I = blockIdx.x / kModes, c = blockIdx.x % kModes;
  const int w0 = I * (per / kJacVerts), w1 = min(nblocks, w0 + per / kJacVerts);
  double a = 0.0;
  for (int w = w0 + threadIdx.x; w < w1; w += 256) a += pzr[(int64_t)c * nblocks + w];
  a = block_sum_256(a, sh);
  if (threadIdx.x == 0) g[blockIdx.x] = a;
}

// y = E^-1 g (one wave per row, its lanes along the row: coalesced) and the partials of g.y
// (= r . Z y, the coarse part of r.u), one per workgroup of four rows
constexpr int kCoarseRowsPerBlock = 4;
__global__ void __launch_bounds__(64 * kCoarseRowsPerBlock)
    k_coarse_y(int k, const double* __restrict__ Einv, const double* __restrict__ g, double* __restrict__ y,
               double* __restrict__ part, const PcgState* __restrict__ st) {
  __shared__ double sh[kCoarseRowsPerBlock];
  if (st && cg_skip(st)) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = blockIdx.x * kCoarseRowsPerBlock + wv;
  double a = 0.0;
  if (i < k)
    for (int j = lane; j < k; j += 64) a += Einv[(size_t)i * k + j] * g[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
  if (lane == 0) {
    if (i < k) y[i] = a;
    sh[wv] = (i < k) ? a * g[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kCoarseRowsPerBlock; ++w) t += sh[w];
    part[blockIdx.x] = t;
  }
}

// u += Z y (rows of fixed DOFs untouched): one thread per DOF row in vertex order, 64 vertices per
// workgroup (the vertex's positions meet in LDS, each read once).  u is in vertex order, or in the
// solve order of a colour-order solver when pos (vertex -> position) is given.
__global__ void __launch_bounds__(3 * kJacVerts)
    k_coarse_add(int nV, int per, const int32_t* __restrict__ pos, const double* __restrict__ q,
                 const double* __restrict__ cent, const double* __restrict__ y, const uint8_t* __restrict__ fixed,
                 double* __restrict__ u, const PcgState* __restrict__ st) {
  __shared__ double qs[3 * kJacVerts];
  if (st && cg_skip(st)) return;
  const int64_t i = blockIdx.x * (int64_t)(3 * kJacVerts) + threadIdx.x;
  const bool in = i < 3ll * nV;
  qs[threadIdx.x] = in ? q[i] : 0.0;
  __syncthreads();
  if (!in || (fixed && fixed[i])) return;
  const int v = (int)(i / 3), l = (int)(i % 3);
  const int I = v / per;
  const int base = threadIdx.x - l;
  double d[3], m[kMono];
  agg_coords(cent, I, qs[base], qs[base + 1], qs[base + 2], d);
  monomials(d, m);
  const double* a = y + kModes * I + kMono * l;
  double z = 0.0;
#pragma unroll
  for (int c = 0; c < kMono; ++c) z += a[c] * m[c];
  u[pos ? 3ll * pos[v] + l : i] += z;
}

// the vector kernel of block-Jacobi PCG plus the partials of Z^T r per workgroup (r of fixed DOFs
// is zero: b is masked and the fixed rows of the operator are the identity).  Z^T r of the
// workgroup is the 3 x 10 sum of r_v (x) m(d_v): r and q of the 64 vertices meet in LDS, six lanes
// per entry sum it in strides (forming the monomial themselves), and one lane per entry adds the six
// partials
__global__ void __launch_bounds__(3 * kJacVerts)
    k_cg_vec_jacobi_2l(int nV, int start, int per, const double* __restrict__ b, const float* __restrict__ Dinv,
                       double* __restrict__ x, double* __restrict__ r, double* __restrict__ u, double* __restrict__ p,
                       double* __restrict__ s, const double* __restrict__ w, const double* __restrict__ q,
                       const double* __restrict__ cent, double* __restrict__ part_ru, double* __restrict__ part_rr,
                       double* __restrict__ pzr, const PcgState* __restrict__ st) {
  constexpr int kSub = 6;
  static_assert(kModes * kSub <= 3 * kJacVerts, "six lanes per coarse entry");
  __shared__ double rs[3 * kJacVerts];
  __shared__ double qs[3 * kJacVerts];
  __shared__ double ms[kJacVerts][kMono + 1];  // the monomials of each vertex, formed once
  __shared__ double red[kModes][kSub];
  __shared__ double sh[3][2];
  if (!start && cg_skip(st)) return;
  const int I = (blockIdx.x * kJacVerts) / per;  // the workgroup's aggregate
  const int64_t i = blockIdx.x * (int64_t)(3 * kJacVerts) + threadIdx.x;
  const int64_t n = 3ll * nV;
  double ri = 0.0;
  if (i < n) {
    if (start) {
      x[i] = 0.0;
      p[i] = 0.0;
      s[i] = 0.0;
      ri = b[i];
    } else {
      const double alpha = st->alpha, beta = st->beta;
      const double pi = u[i] + beta * p[i];
      const double si = w[i] + beta * s[i];
      p[i] = pi;
      s[i] = si;
      x[i] += alpha * pi;
      ri = r[i] - alpha * si;
    }
    r[i] = ri;
  }
  rs[threadIdx.x] = ri;
  qs[threadIdx.x] = (i < n) ? q[i] : 0.0;
  __syncthreads();
  if (threadIdx.x < kJacVerts) {
    double d[3], mm[kMono];
    agg_coords(cent, I, qs[3 * threadIdx.x], qs[3 * threadIdx.x + 1], qs[3 * threadIdx.x + 2], d);
    monomials(d, mm);
#pragma unroll
    for (int c = 0; c < kMono; ++c) ms[threadIdx.x][c] = mm[c];
  }
  double ru = 0.0, rr = 0.0;
  if (i < n) {
    const int l = (int)(i % 3);
    const int base = threadIdx.x - l;
    const float* D = Dinv + 3 * i;
    const double ui = (double)D[0] * rs[base] + (double)D[1] * rs[base + 1] + (double)D[2] * rs[base + 2];
    u[i] = ui;
    ru = ri * ui;
    rr = ri * ri;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ru += __shfl_xor(ru, off, 64);
    rr += __shfl_xor(rr, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6][0] = ru;
    sh[threadIdx.x >> 6][1] = rr;
  }
  __syncthreads();
  // entry Q = (l, c) of the 3 x 10 sum: lanes kSub Q .. kSub Q + kSub - 1, lane j takes vertices
  // j, j + kSub, ...
  if (threadIdx.x < kModes * kSub) {
    const int Q = threadIdx.x / kSub, j = threadIdx.x % kSub;
    const int l = Q / kMono, c = Q % kMono;
    double a = 0.0;
    for (int vv = j; vv < kJacVerts; vv += kSub) a += rs[3 * vv + l] * ms[vv][c];  // r is zero beyond the mesh
    red[Q][j] = a;
  }
  __syncthreads();
  if (threadIdx.x < kModes) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kSub; ++j) a += red[threadIdx.x][j];
    pzr[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = a;  // mode-major: k_coarse_g reads runs
  }
  if (threadIdx.x == 0) {
    part_ru[blockIdx.x] = sh[0][0] + sh[1][0] + sh[2][0];
    part_rr[blockIdx.x] = sh[0][1] + sh[1][1] + sh[2][1];
  }
}

// ---- the coarse correction for the colour-order (factor) solvers -----------------------------
// Their vectors live in the solve order (position i holds vertex colVerts[i]); these kernels walk
// the vertices in vertex order, 64 per workgroup as in the block-Jacobi form, and reach a vertex's
// rows through pos = colVerts^-1.

__global__ void __launch_bounds__(256) k_inverse_order(int nV, const int32_t* __restrict__ colVerts,
                                                       int32_t* __restrict__ pos) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nV) pos[colVerts[i]] = i;
}

// the partials of Z^T r per workgroup of 64 vertices (r in the solve order), mode-major
__global__ void __launch_bounds__(3 * kJacVerts)
    k_coarse_zr_order(int nV, int per, const int32_t* __restrict__ pos, const double* __restrict__ r,
                      const double* __restrict__ q, const double* __restrict__ cent, double* __restrict__ pzr,
                      const PcgState* __restrict__ st) {
  constexpr int kSub = 6;
  __shared__ double rs[3 * kJacVerts];
  __shared__ double qs[3 * kJacVerts];
  __shared__ double ms[kJacVerts][kMono + 1];  // the monomials of each vertex, formed once
  __shared__ double red[kModes][kSub];
  if (st && cg_skip(st)) return;
  const int I = (blockIdx.x * kJacVerts) / per;
  const int64_t i = blockIdx.x * (int64_t)(3 * kJacVerts) + threadIdx.x;
  const bool in = i < 3ll * nV;
  rs[threadIdx.x] = in ? r[3ll * pos[i / 3] + i % 3] : 0.0;
  qs[threadIdx.x] = in ? q[i] : 0.0;
  __syncthreads();
  if (threadIdx.x < kJacVerts) {
    double d[3], mm[kMono];
    agg_coords(cent, I, qs[3 * threadIdx.x], qs[3 * threadIdx.x + 1], qs[3 * threadIdx.x + 2], d);
    monomials(d, mm);
#pragma unroll
    for (int c = 0; c < kMono; ++c) ms[threadIdx.x][c] = mm[c];
  }
  __syncthreads();
  if (threadIdx.x < kModes * kSub) {
    const int Q = threadIdx.x / kSub, j = threadIdx.x % kSub;
    const int l = Q / kMono, c = Q % kMono;
    double a = 0.0;
    for (int vv = j; vv < kJacVerts; vv += kSub) a += rs[3 * vv + l] * ms[vv][c];
    red[Q][j] = a;
  }
  __syncthreads();
  if (threadIdx.x < kModes) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < kSub; ++j) a += red[threadIdx.x][j];
    pzr[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
  }
}

// E = Z^T A Z and its inverse for the current system (vertex order, CSR operator)
template <typename V>
int coarse_setup(bh_mesh* m, const V* vals, const double* hdiag, const uint8_t* fixed, const double* q,
                 const Coarse& c, hipStream_t s) {
  const int nV = m->nV;
  hipLaunchKernelGGL(k_agg_centroid, dim3(c.nagg), dim3(256), 0, s, nV, c.per, q, c.cent);
  const int64_t kk = (int64_t)c.k * c.k;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0, s, kk, c.E);
  hipLaunchKernelGGL((k_coarse_E<V>), dim3((nV + 63) / 64), dim3(64), 0, s, nV, c.per, c.nagg, m->bptr, m->bcol, vals,
                     hdiag, fixed, q, (const double*)c.cent, c.E);
  hipLaunchKernelGGL(k_coarse_inv, dim3(1), dim3(1024), sizeof(double) * 3 * c.k, s, c.k, c.E);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// Two-level block-Jacobi PCG in vertex order on the CSR operator (FP64 or FP32 values, optional
// FP64 diagonal).  The FP32 CSR operator is exactly symmetric when the Hessian came through the
// PSD records; the element-math gathers leave it symmetric to FP32 rounding.
template <typename V>
int pcg_solve_2l(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag, const uint8_t* fixed,
                 const double* q, NewtonWs& w, int* iters, int* done, hipStream_t s) {
  const int nV = m->nV;
  const int gj = jac_blocks(nV);
  const int gs = spmv_blocks_for<V>(nV);
  const int64_t R = part_region(nV);
  Coarse c;
  int st = coarse_workspace(m, o->coarse_aggregates, gj, c);
  if (st) return st;
  const int gy = (c.k + kCoarseRowsPerBlock - 1) / kCoarseRowsPerBlock;
  if (gj + gy > R) return BH_UNSUPPORTED;
  double* pru = w.part;
  double* prr = w.part + R;
  double* pwu = w.part + 2 * R;
  hipLaunchKernelGGL(k_block_jacobi<V>, dim3((nV + 255) / 256), dim3(256), 0, s, nV, m->bptr, m->bcol, vals, hdiag,
                     fixed, 0, w.Dinv);
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((9ll * nV + 255) / 256)), dim3(256), 0, s, 9ll * nV,
                     (const double*)w.Dinv, c.dinv);
  hipLaunchKernelGGL(k_agg_centroid, dim3(c.nagg), dim3(256), 0, s, nV, c.per, q, c.cent);
  const int64_t kk = (int64_t)c.k * c.k;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((kk + 255) / 256)), dim3(256), 0, s, kk, c.E);
  hipLaunchKernelGGL((k_coarse_E<V>), dim3((nV + 63) / 64), dim3(64), 0, s, nV, c.per, c.nagg, m->bptr, m->bcol, vals,
                     hdiag, fixed, q, (const double*)c.cent, c.E);
  hipLaunchKernelGGL(k_coarse_inv, dim3(1), dim3(1024), sizeof(double) * 3 * c.k, s, c.k, c.E);
  BH_CHECK(hipGetLastError());
  auto precond_tail = [&](int start) {
    const PcgState* sk = start ? nullptr : w.st;  // kernels skip once the solve is done
    hipLaunchKernelGGL(k_cg_vec_jacobi_2l, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, start, c.per, w.b,
                       (const float*)c.dinv, w.x, w.r, w.z, w.p, w.s, w.Ap, q, (const double*)c.cent, pru, prr,
                       c.pzr, w.st);
    hipLaunchKernelGGL(k_coarse_g, dim3(c.k), dim3(256), 0, s, c.per, gj, (const double*)c.pzr, c.g, sk);
    hipLaunchKernelGGL(k_coarse_y, dim3(gy), dim3(64 * kCoarseRowsPerBlock), 0, s, c.k, (const double*)c.E,
                       (const double*)c.g, c.y, pru + gj, sk);
    hipLaunchKernelGGL(k_coarse_add, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, c.per, (const int32_t*)nullptr, q,
                       (const double*)c.cent, (const double*)c.y, fixed, w.z, sk);
    hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(gs), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals, hdiag, fixed, w.z,
                       w.Ap, pwu, start ? (const PcgState*)nullptr : (const PcgState*)w.st);
    cg_reduce(w, nV, gj + gy, gj, gs, start, o, s);
  };
  precond_tail(1);
  PcgState hs{};
  for (;;) {
    for (int k = 0; k < kPcgBatch; ++k) precond_tail(0);
    if ((st = cg_wait(w, &hs, s))) return st;
    if (hs.done) break;
  }
  *iters = hs.iters;
  *done = hs.done;
  return BH_OK;
}

// Block-Jacobi PCG on the split operator in the solve order of P (the precision form: the split
// operator takes every upper block as the transpose of the lower one, so it is exactly symmetric,
// where the FP32 CSR pattern holds H_vn and H_nv computed separately; CG on that slightly
// non-symmetric matrix needed up to 1.8x the iterations).  split: build the split Hessian first.
template <typename V>
int pcg_solve_split_jacobi(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag,
                           const uint8_t* fixed, const bh_precond* P, int split, NewtonWs& w, int* iters, int* done,
                           hipStream_t s) {
  const int nV = m->nV;
  const int gj = jac_blocks(nV);
  const int gsp = sweep_grid(nV);
  const int gr = (int)((3ll * nV + 255) / 256);
  const int64_t R = part_region(nV);
  double* pru = w.part;
  double* prr = w.part + R;
  double* pwu = w.part + 2 * R;
  int st;
  if (split && (st = precond_split_hessian<V>(P, vals, hdiag, fixed, s))) return st;
  const SplitH<V> h = split_view<V>(P);
  double *b = svec(P, SV_B), *x = svec(P, SV_X), *r = svec(P, SV_R), *p = svec(P, SV_P);
  double *sv = svec(P, SV_S), *u = svec(P, SV_U), *wv = svec(P, SV_W);
  hipLaunchKernelGGL(k_block_inv, dim3((nV + 255) / 256), dim3(256), 0, s, nV, (const double*)P->hdiag,
                     o->precond == BH_PRECOND_POINT_JACOBI ? 1 : 0, w.Dinv);
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 1, (const double*)w.b, b);
  hipLaunchKernelGGL(k_cg_vec_jacobi, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, 1, b, w.Dinv, x, r, u, p, sv, wv, pru,
                     prr, w.st);
  hipLaunchKernelGGL(k_split_spmv<V>, dim3(gsp), dim3(256), 0, s, nV, P->nlptr, P->nlcol, (const V*)h.hl, P->nuptr,
                     P->nucol, (const V*)h.hu, h.ld, (const double*)P->hdiag, (const double*)u, wv, pwu,
                     (const PcgState*)nullptr);
  cg_reduce(w, nV, gj, gj, gsp, 1, o, s);
  PcgState hs{};
  for (;;) {
    for (int k = 0; k < kPcgBatch; ++k) {
      hipLaunchKernelGGL(k_cg_vec_jacobi, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, 0, b, w.Dinv, x, r, u, p, sv, wv,
                         pru, prr, w.st);
      hipLaunchKernelGGL(k_split_spmv<V>, dim3(gsp), dim3(256), 0, s, nV, P->nlptr, P->nlcol, (const V*)h.hl,
                         P->nuptr, P->nucol, (const V*)h.hu, h.ld, (const double*)P->hdiag, (const double*)u, wv, pwu,
                         w.st);
      cg_reduce(w, nV, gj, gj, gsp, 0, o, s);
    }
    if ((st = cg_wait(w, &hs, s))) return st;
    if (hs.done) break;
  }
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 0, (const double*)x, w.x);
  BH_CHECK(hipGetLastError());
  *iters = hs.iters;
  *done = hs.done;
  return BH_OK;
}

// PCG with the Eisenstat-form preconditioner M = (Dt + L) Dt^{-1} (Dt + L)^T in the solve order of P (Dt = D:
// block symmetric Gauss-Seidel; or the network's Dt):
// CG on A~ = P^{-1} A P^{-T} with Dt as its preconditioner (one backward and one forward sweep per
// iteration, Chronopoulos-Gear recurrences, one reduction).  The transformed residual is an
// M^{-1}-norm estimate of the true one; once it is within kEisCheck * tol the residual of A x = b
// itself (x = P^{-T} x~) is computed at the end of a batch and decides convergence.  The first check
// comes once the transformed residual has fallen kEisCheck^2 below its start; every check then measures
// rho = (|r|^2 / |b|^2) / (gamma / gamma0), and the next one waits until rho * gamma / gamma0 predicts
// the tolerance (within a factor 4).  When the true residual stagnates between two checks (it fell by
// less than 4x), the recursive residual has drifted from it (most at tight tolerances): it is
// replaced by the one just computed (r~ = P^{-1} (b - A x)) and the recurrences restart from the
// current x~.  (Restarting at every failed check instead cost ~50 % more iterations at 1e-6.)
constexpr double kEisCheck = 100.0;

template <typename V>
int pcg_solve_eisenstat(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag,
                        const uint8_t* fixed, bh_precond* P, int learned, NewtonWs& w, int* iters, int* done,
                        hipStream_t s) {
  const int nV = m->nV;
  const int64_t n = 3ll * nV;
  const int gr = (int)((n + 255) / 256);
  const int gi = vec_blocks(nV);
  const int gsp = sweep_grid(nV);
  const int64_t R = part_region(nV);
  if (precond_sweep_blocks(P) > R) return BH_UNSUPPORTED;  // more colours than the regions hold
  double* pru = w.part;
  double* prr = w.part + R;
  double* pwu = w.part + 2 * R;
  int st = precond_split_hessian<V>(P, vals, hdiag, fixed, s);
  if (st) return st;
  if (learned) {
    if (!o->mgn) return BH_BAD_ARG;
    st = mgn_sgs<V>(o->mgn, P, vals, fixed, s);
  } else {
    const double w = (o->relaxation > 0.0 && o->relaxation < 2.0) ? o->relaxation : 1.0;
    st = precond_sgs(P, 1.0 / w, s);
  }
  if (st) return st;
  const SplitH<V> h = split_view<V>(P);
  double* b = svec(P, SV_B);
  double* xs = svec(P, SV_Y);  // x = P^{-T} x~ (FP64)
  EisVecs v{svec(P, SV_R), svec(P, SV_X), svec(P, SV_P), svec(P, SV_S),
            svec(P, SV_U), svec(P, SV_V), svec(P, SV_T), svec(P, SV_W)};
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 1, (const double*)w.b, b);
  double bb = 0.0;
  if ((st = device_dot(b, b, n, w.part, w.scal + 10, s)) || (st = read_scalar(w.scal + 10, &bb, s))) return st;
  const double tol = o->pcg_tol;
  hipLaunchKernelGGL(k_eis_init, dim3(gi), dim3(kRedBlock), 0, s, n, v.xt, v.pt, v.st);
  int ng = 0, nd = 0;
  eis_fwd<V, double>(P, 0, v, b, v.rt, nullptr, nullptr, s);           // r~ = P^{-1} b
  eis_bwd<V, double>(P, 1, v, nullptr, v.t, pru, nullptr, s, &ng, pwu);  // u = Dt r~, t = P^{-T} u
  eis_fwd<V, double>(P, 1, v, nullptr, v.w, pwu, nullptr, s, &nd, true);  // w~ = A~ u
  cg_reduce_regions(w, nV, pru, ng, pru, ng, pwu, nd, 1, 0.0, o->pcg_max_iterations, s);
  BH_CHECK(hipGetLastError());
  PcgState hs{};
  bool converged = bb == 0.0, fresh = false;
  double rho = 1.0 / (kEisCheck * kEisCheck);  // predicted (|r|^2/|b|^2) per unit of gamma/gamma0
  int itBase = 0;                              // iterations before the last restart
  double prevRr = -1.0;                        // true residual at the previous check
  while (!converged) {
    for (int k = 0; k < kPcgBatch; ++k) {
      eis_bwd<V, double>(P, 2, v, nullptr, v.t, pru, w.st, s, nullptr, pwu);
      eis_fwd<V, double>(P, 1, v, nullptr, v.w, pwu, w.st, s, nullptr, true);
      cg_reduce_regions(w, nV, pru, ng, pru, ng, pwu, nd, 0, 0.0, o->pcg_max_iterations - itBase, s);
    }
    if ((st = cg_wait(w, &hs, s))) return st;
    fresh = false;
    const double g = hs.bb > 0.0 ? hs.rr / hs.bb : 0.0;
    if (rho * g <= 4.0 * tol * tol || hs.done) {
      // x = P^{-T} x~ and the residual of A x = b (FP64)
      eis_bwd<V, double>(P, 0, v, v.xt, xs, nullptr, nullptr, s);
      double* r = svec(P, SV_Z);
      hipLaunchKernelGGL(k_split_spmv<V>, dim3(gsp), dim3(256), 0, s, nV, P->nlptr, P->nlcol, (const V*)h.hl,
                         P->nuptr, P->nucol, (const V*)h.hu, h.ld, (const double*)P->hdiag, (const double*)xs, r,
                         (double*)nullptr, (const PcgState*)nullptr);
      hipLaunchKernelGGL(k_resid, dim3(gr), dim3(256), 0, s, n, (const double*)b, r);
      double rr = 0.0;
      if ((st = device_dot(r, r, n, w.part, w.scal + 11, s)) || (st = read_scalar(w.scal + 11, &rr, s))) return st;
      fresh = true;
      const bool stagnant = prevRr > 0.0 && rr > 0.25 * prevRr;
      prevRr = rr;
      if (rr <= tol * tol * bb) {
        converged = true;
      } else if (!stagnant) {
        if (g > 0.0) rho = (rr / bb) / g;
      } else if (!hs.done && itBase + hs.iters < o->pcg_max_iterations) {
        // residual replacement and restart from x~ (r~ = P^{-1} r, p~ = s~ = 0)
        itBase += hs.iters;
        eis_fwd<V, double>(P, 0, v, r, v.rt, nullptr, nullptr, s);
        hipLaunchKernelGGL(k_eis_restart, dim3(gi), dim3(kRedBlock), 0, s, n, v.pt, v.st);
        eis_bwd<V, double>(P, 1, v, nullptr, v.t, pru, nullptr, s, nullptr, pwu);
        eis_fwd<V, double>(P, 1, v, nullptr, v.w, pwu, nullptr, s, nullptr, true);
        cg_reduce_regions(w, nV, pru, ng, pru, ng, pwu, nd, 1, 0.0, o->pcg_max_iterations - itBase, s);
        BH_CHECK(hipGetLastError());
        rho = rr / bb;  // at the restart gamma / gamma0 = 1
        prevRr = -1.0;
        fresh = false;
        continue;
      }
    }
    if (hs.done) break;
  }
  if (!fresh) eis_bwd<V, double>(P, 0, v, v.xt, xs, nullptr, nullptr, s);
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 0, (const double*)xs, w.x);
  BH_CHECK(hipGetLastError());
  *iters = itBase + hs.iters;
  *done = converged ? 1 : (hs.done ? hs.done : 2);
  return BH_OK;
}

// PCG with a colour-sweep block-triangular preconditioner, run in the preconditioner's solve order
// (colours contiguous): the masked Hessian is split into lower / transposed-lower / diagonal blocks in
// that order, b is permuted in and x out.  Same recurrences as above; the vector update runs inside
// the forward sweep, r.z comes from the backward sweep.
template <typename V>
int pcg_solve_factor(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag, const uint8_t* fixed,
                     const bh_precond* P, const double* q, NewtonWs& w, int* iters, int* done, hipStream_t s) {
  const int nV = m->nV;
  const int64_t n = 3ll * nV;
  const int gsp = sweep_grid(nV);
  const int gi = vec_blocks(nV);
  const int gr = (int)((n + 255) / 256);
  const int64_t R = part_region(nV);
  if (precond_sweep_blocks(P) > R || gsp > R) return BH_UNSUPPORTED;  // more colours than the regions hold
  double* pru = w.part;
  double* prr = w.part + R;
  double* pwu = w.part + 2 * R;
  int st = precond_split_hessian<V>(P, vals, hdiag, fixed, s);
  if (st) return st;
  const SplitH<V> h = split_view<V>(P);
  double *b = svec(P, SV_B), *x = svec(P, SV_X), *r = svec(P, SV_R), *p = svec(P, SV_P);
  double *sv = svec(P, SV_S), *u = svec(P, SV_U), *wv = svec(P, SV_W);
  int np[2] = {0, 0};
  // optional coarse correction u += Z E^-1 Z^T r after the factor's solve (the factor's r.u partials
  // come first in their region, the g.y partials after them)
  Coarse c;
  int gy = 0;
  const int gj = jac_blocks(nV);
  if (o->coarse_correction) {
    if ((st = coarse_workspace(m, o->coarse_aggregates, gj, c))) return st;
    gy = (c.k + kCoarseRowsPerBlock - 1) / kCoarseRowsPerBlock;
    if (precond_sweep_blocks(P) + gy > R) return BH_UNSUPPORTED;
    hipLaunchKernelGGL(k_inverse_order, dim3((nV + 255) / 256), dim3(256), 0, s, nV, P->colVerts, c.pos);
    if ((st = coarse_setup<V>(m, vals, hdiag, fixed, q, c, s))) return st;
  }
  auto coarse_tail = [&](const PcgState* sk) {
    if (!o->coarse_correction) return;
    hipLaunchKernelGGL(k_coarse_zr_order, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, c.per, (const int32_t*)c.pos,
                       (const double*)r, q, (const double*)c.cent, c.pzr, sk);
    hipLaunchKernelGGL(k_coarse_g, dim3(c.k), dim3(256), 0, s, c.per, gj, (const double*)c.pzr, c.g, sk);
    hipLaunchKernelGGL(k_coarse_y, dim3(gy), dim3(64 * kCoarseRowsPerBlock), 0, s, c.k, (const double*)c.E,
                       (const double*)c.g, c.y, pru + np[1], sk);
    hipLaunchKernelGGL(k_coarse_add, dim3(gj), dim3(3 * kJacVerts), 0, s, nV, c.per, (const int32_t*)c.pos, q,
                       (const double*)c.cent, (const double*)c.y, fixed, u, sk);
  };
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 1, (const double*)w.b, b);
  hipLaunchKernelGGL(k_cg_init, dim3(gi), dim3(kRedBlock), 0, s, n, b, x, r, p, sv, prr);
  CgVecs start{x, p, sv, u, wv, nullptr, nullptr};
  if ((st = precond_apply_solve(P, r, u, s, &start, 0, pru, np))) return st;
  coarse_tail(nullptr);
  hipLaunchKernelGGL(k_split_spmv<V>, dim3(gsp), dim3(256), 0, s, nV, P->nlptr, P->nlcol, (const V*)h.hl, P->nuptr,
                     P->nucol, (const V*)h.hu, h.ld, (const double*)P->hdiag, (const double*)u, wv, pwu,
                     (const PcgState*)nullptr);
  cg_reduce(w, nV, np[1] + gy, gi, gsp, 1, o, s);
  CgVecs loop{x, p, sv, u, wv, prr, w.st};
  PcgState hs{};
  for (;;) {
    for (int k = 0; k < kPcgBatch; ++k) {
      if ((st = precond_apply_solve(P, r, u, s, &loop, 1, pru, np))) return st;
      coarse_tail(w.st);
      hipLaunchKernelGGL(k_split_spmv<V>, dim3(gsp), dim3(256), 0, s, nV, P->nlptr, P->nlcol, (const V*)h.hl,
                         P->nuptr, P->nucol, (const V*)h.hu, h.ld, (const double*)P->hdiag, (const double*)u, wv, pwu,
                         w.st);
      cg_reduce(w, nV, np[1] + gy, np[0], gsp, 0, o, s);
    }
    if ((st = cg_wait(w, &hs, s))) return st;
    if (hs.done) break;
  }
  hipLaunchKernelGGL(k_permute_rows, dim3(gr), dim3(256), 0, s, nV, P->colVerts, 0, (const double*)x, w.x);
  BH_CHECK(hipGetLastError());
  *iters = hs.iters;
  *done = hs.done;
  return BH_OK;
}

}  // namespace

namespace {
__global__ void __launch_bounds__(256) k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// The Jacobi fallback of the factor and Gauss-Seidel solvers: a solve that hit its iteration cap
// (done 2) or broke down (done 3: r.u or p.Ap not positive or not finite, which an indefinite factor
// or diagonal produces) is redone from zero with block-Jacobi PCG, whose preconditioner is the
// (PSD-clamped) block diagonal of H itself.  split = 0 reuses the split Hessian the failed solve made
// (precision form), 1 makes it.
template <typename V>
int jacobi_fallback_solve(bh_mesh* m, const bh_newton_opts* o, const V* vals, const double* hdiag,
                          const uint8_t* fixed, int split, NewtonWs& w, int* pits, int* pdone, int* fallback,
                          hipStream_t s) {
  int p2 = 0, st;
  if (hdiag) {
    if (!o->precond_handle) return BH_BAD_ARG;
    st = pcg_solve_split_jacobi(m, o, vals, hdiag, fixed, o->precond_handle, split, w, &p2, pdone, s);
  } else {
    st = pcg_solve(m, o, vals, hdiag, fixed, w, &p2, pdone, s);
  }
  if (st) return st;
  *pits += p2;
  *fallback = 1;
  return BH_OK;
}

// one Newton direction: H(q) (PSD-fixed) and PCG on H d = -g, with the chosen preconditioner
template <typename V>
int newton_direction(bh_mesh* m, const bh_newton_opts* o, const V* params, V* vals, const uint8_t* fixed,
                     NewtonWs& w, int* pits, int* pdone, int* fallback, hipStream_t s, const double* q) {
  const bh_psd* psd = o->psd.enabled ? &o->psd : nullptr;
  int st = hessian_dispatch(m, o->model, o->d_model_ids, q, params, psd, vals, s);
  if (st) return st;
  if (o->d_mass_vals) {
    hipLaunchKernelGGL(k_add_mass<V>, dim3((m->nnz + 255) / 256), dim3(256), 0, s, m->nnz, o->inertia, o->d_mass_vals,
                       vals);
    BH_CHECK(hipGetLastError());
  }
  // precision form: FP64 diagonal blocks rebuilt from the FP32 off-diagonal blocks (exact row sums)
  const double* hdiag = nullptr;
  if (sizeof(V) == 4 && m->nV) {
    hipLaunchKernelGGL(k_implied_diag<V>, dim3((unsigned)((8ll * m->nV + 255) / 256)), dim3(256), 0, s, m->nV,
                       m->bptr, m->bcol, (const V*)vals, m->vdeg, psd ? std::max(psd->tol, 0.0) : 0.0, o->d_mass_vals,
                       o->d_mass_vals ? o->inertia : 0.0, w.hdiag);
    BH_CHECK(hipGetLastError());
    hdiag = w.hdiag;
  }
  *fallback = 0;
  if (o->precond == BH_PRECOND_SGS || o->precond == BH_PRECOND_LEARNED_SGS) {
    if (!o->precond_handle) return BH_BAD_ARG;
    if ((st = pcg_solve_eisenstat(m, o, vals, hdiag, fixed, o->precond_handle,
                                  o->precond == BH_PRECOND_LEARNED_SGS ? 1 : 0, w, pits, pdone, s)))
      return st;
    if (*pdone != 1 && o->jacobi_fallback) return jacobi_fallback_solve(m, o, vals, hdiag, fixed, 0, w, pits, pdone,
                                                                        fallback, s);
    return BH_OK;
  }
  if (o->precond == BH_PRECOND_BIC0 || o->precond == BH_PRECOND_FACTOR || o->precond == BH_PRECOND_LEARNED) {
    if (!o->precond_handle) return BH_BAD_ARG;
    if (o->precond == BH_PRECOND_BIC0) {
      // IC(0) may break down (non-positive pivots) on strongly heterogeneous problems: refactor with a
      // growing diagonal shift until every pivot is positive
      double shift = o->precond_shift;
      int fails = 0;
      for (int attempt = 0;; ++attempt) {
        fails = 0;
        if ((st = precond_factor_bic0(o->precond_handle, vals, hdiag, fixed, shift, &fails, s))) return st;
        if (fails == 0 || attempt >= 6) break;
        shift = (shift > 0.0) ? 4.0 * shift : 1e-3;
      }
      // pivots still not positive after the last shift: the factor holds infinite inverse blocks, so it
      // is never applied -- block Jacobi solves this system, or the solve fails
      if (fails) {
        *pits = 0;
        if (!o->jacobi_fallback) return BH_SOLVE_FAILED;
        return jacobi_fallback_solve(m, o, vals, hdiag, fixed, 1, w, pits, pdone, fallback, s);
      }
    }
    if (o->precond == BH_PRECOND_LEARNED) {
      if (!o->mgn) return BH_BAD_ARG;
      if ((st = mgn_factor<V>(o->mgn, o->precond_handle, vals, fixed, s))) return st;
    }
    if ((st = pcg_solve_factor(m, o, vals, hdiag, fixed, o->precond_handle, q, w, pits, pdone, s))) return st;
    if (*pdone != 1 && o->jacobi_fallback) return jacobi_fallback_solve(m, o, vals, hdiag, fixed, 0, w, pits, pdone,
                                                                        fallback, s);
    return BH_OK;
  }
  if (o->precond == BH_PRECOND_JACOBI_2L) return pcg_solve_2l(m, o, vals, hdiag, fixed, q, w, pits, pdone, s);
  if (hdiag) {  // precision form: block Jacobi on the symmetric split operator (colour lists made once)
    if (!m->solveOrder && (st = bh_precond_create(m, &m->solveOrder))) return st;
    return pcg_solve_split_jacobi(m, o, vals, hdiag, fixed, m->solveOrder, 1, w, pits, pdone, s);
  }
  return pcg_solve(m, o, vals, hdiag, fixed, w, pits, pdone, s);
}
}  // namespace

extern "C" {

int bh_spmv_device(bh_mesh* m, const double* d_vals, const double* d_x, double* d_y, bh_stream_t s) {
  if (!m || !d_vals || !d_x || !d_y) return BH_BAD_ARG;
  if (m->nV == 0) return BH_OK;
  hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks_for<double>(m->nV)), dim3(kRedBlock), 0, as_stream(s),
                     m->nV, m->bptr, m->bcol, d_vals, (const double*)nullptr, nullptr, d_x, d_y, nullptr, nullptr);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

void bh_newton_default_opts(bh_newton_opts* o) {
  if (!o) return;
  o->model = BH_NEOHOOKEAN;
  o->d_model_ids = nullptr;
  o->psd.enabled = 0;
  o->psd.tol = 1e-6;
  o->gradient_tol = 1e-3;
  o->relative_tol = 0;
  o->max_iterations = 100;
  o->max_iterations_ls = 100;
  o->alpha = 1.0;
  o->c = 1e-8;
  o->p = 0.5;
  o->armijo_with_alpha = 0;
  o->reject_nonfinite = 0;
  o->pcg_tol = 1e-10;
  o->pcg_max_iterations = 20000;
  o->precond = BH_PRECOND_JACOBI;
  o->precond_handle = nullptr;
  o->precond_shift = 0.0;
  o->element_precision = 0;
  o->jacobi_fallback = 1;
  o->mgn = nullptr;
  o->on_system = nullptr;
  o->user = nullptr;
  o->d_mass_vals = nullptr;
  o->d_q_pred = nullptr;
  o->inertia = 0.0;
  o->relaxation = 1.0;
  o->coarse_aggregates = 0;
  o->coarse_correction = 0;
}



int bh_newton_static_device(bh_mesh* m, const bh_newton_opts* o, const double* params, const uint8_t* fixed,
                            const double* f, double* q, bh_newton_report* rep, double* trace, bh_stream_t stream) {
  if (!m || !o || !params || !q) return BH_BAD_ARG;
  if (check_model_args(o->model, o->d_model_ids)) return BH_BAD_ARG;
  if (o->d_mass_vals && (!o->d_q_pred || !(o->inertia >= 0.0))) return BH_BAD_ARG;
  hipStream_t s = as_stream(stream);
  NewtonWs w;
  int st = newton_workspace(m, w);
  if (st) return st;
  const int64_t n = 3ll * m->nV;
  const int gn = (int)((n + 255) / 256);
  bh_newton_report R{};
  double g0 = -1.0;
  int it = 0;
  int status = BH_MAXIT;
  // precision form: 1 = FP32 element Hessian and CG operator (energy and gradient in FP64, so the
  // equilibrium is the FP64 one); 2 = energy, gradient and Hessian all from FP32 element math
  const bool f32 = o->element_precision == 1 || o->element_precision == 2;
  const bool f32g = o->element_precision == 2;
  if (o->element_precision < 0 || o->element_precision > 2) return BH_BAD_ARG;
  if (f32) {
    const int64_t np = (int64_t)(o->d_model_ids ? 2 : param_cols(o->model)) * m->nT;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, np, params, w.params32);
    BH_CHECK(hipGetLastError());
  }
  do {
    // gradient of the objective, masked
    st = f32g ? gradient_dispatch(m, o->model, o->d_model_ids, q, (const float*)w.params32, w.g, s)
             : gradient_dispatch(m, o->model, o->d_model_ids, q, params, w.g, s);
    if (st) return st;
    if (o->d_mass_vals) {  // + inertia M (q - q_pred)
      hipLaunchKernelGGL(k_sub, dim3(gn), dim3(256), 0, s, n, q, o->d_q_pred, w.z);
      hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks_for<double>(m->nV)), dim3(kRedBlock), 0, s, m->nV,
                         m->bptr, m->bcol, o->d_mass_vals, (const double*)nullptr, (const uint8_t*)nullptr, w.z,
                         w.Ap, (double*)nullptr, (const PcgState*)nullptr);
      hipLaunchKernelGGL(k_axpy, dim3(gn), dim3(256), 0, s, n, o->inertia, w.Ap, w.g);
    }
    hipLaunchKernelGGL(k_mask_gradient, dim3(gn), dim3(256), 0, s, n, fixed, f, w.g);
    st = device_dot(w.g, w.g, n, w.part, w.scal, s);
    if (st) return st;
    double gg;
    if ((st = read_scalar(w.scal, &gg, s))) return st;
    const double gnorm = std::sqrt(gg);
    if (g0 < 0) g0 = gnorm;
    double fx;
    if ((st = objective(m, o, params, f, q, w.scal + 1, w, s))) return st;
    if ((st = read_scalar(w.scal + 1, &fx, s))) return st;
    if (trace) {
      trace[4 * it + 0] = fx;
      trace[4 * it + 1] = gnorm;
      trace[4 * it + 2] = 0.0;
      trace[4 * it + 3] = 0.0;
    }
    R.energy = fx;
    R.gradient_norm = gnorm;
    const double thr = o->relative_tol ? o->gradient_tol * g0 : o->gradient_tol;
    if (gg < thr * thr) {
      status = BH_OK;
      R.converged = 1;
      break;
    }
    // Hessian (with the PSD callback) and direction d = -H^{-1} g
    hipLaunchKernelGGL(k_negate, dim3(gn), dim3(256), 0, s, n, w.g, w.b);
    int pits = 0, pdone = 0, fb = 0;
    st = f32 ? newton_direction<float>(m, o, w.params32, w.vals32, fixed, w, &pits, &pdone, &fb, s, q)
             : newton_direction<double>(m, o, params, w.vals, fixed, w, &pits, &pdone, &fb, s, q);
    if (st) return st;
    R.pcg_iterations_total += pits;
    R.pcg_fallbacks += fb;
    if (pdone != 1) R.pcg_failures += 1;
    if (o->on_system) {
      BH_CHECK(hipStreamSynchronize(s));
      o->on_system(o->user, it, f32 ? w.vals32 : nullptr, f32 ? nullptr : w.vals, w.b, w.x);
    }
    // line search on the objective (reference semantics unless the options say otherwise)
    st = device_dot(w.g, w.x, n, w.part, w.scal + 2, s);
    if (st) return st;
    double gd;
    if ((st = read_scalar(w.scal + 2, &gd, s))) return st;
    double alpha = o->alpha;
    int ls = 0;
    for (;;) {
      hipLaunchKernelGGL(k_axpy_to, dim3(gn), dim3(256), 0, s, n, q, alpha, w.x, w.qt);
      double ft;
      if ((st = objective(m, o, params, f, w.qt, w.scal + 3, w, s))) return st;
      if ((st = read_scalar(w.scal + 3, &ft, s))) return st;
      const double rhs = fx + o->c * (o->armijo_with_alpha ? alpha : 1.0) * gd;
      bool reject = ft > rhs;
      if (o->reject_nonfinite && !std::isfinite(ft)) reject = true;
      if (!reject || ls >= o->max_iterations_ls) break;
      alpha *= o->p;
      ++ls;
    }
    // x += alpha d (applied even when the search hit its cap, as the reference does)
    hipLaunchKernelGGL(k_axpy_to, dim3(gn), dim3(256), 0, s, n, q, alpha, w.x, q);
    BH_CHECK(hipGetLastError());
    R.ls_iterations_total += ls;
    if (trace) {
      trace[4 * it + 2] = alpha;
      trace[4 * it + 3] = pits;
    }
    ++it;
  } while (it <= o->max_iterations);
  BH_CHECK(hipStreamSynchronize(s));
  R.iterations = it;
  R.gradient_norm0 = g0;
  if (rep) *rep = R;
  return status;
}

}  // extern "C"
