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
                                                      const uint8_t* __restrict__ fixed, int point,
                                                      double* __restrict__ Dinv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const int b0 = bptr[v], nb = bptr[v + 1] - b0;
  int r = 0;
  while (r < nb && bcol[b0 + r] != v) ++r;
  double A[9];  // row-major block; a vertex that belongs to no tet gets the identity
  if (r < nb) {
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

// x = 0, r = b, z = Dinv r, p = z, and partials of r.z and r.r
__global__ void __launch_bounds__(kRedBlock) k_pcg_init(int nV, const double* __restrict__ b,
                                                        const double* __restrict__ Dinv, double* __restrict__ x,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        double* __restrict__ p, double* __restrict__ part) {
  __shared__ double sh[4];
  const int v = blockIdx.x * kRedBlock + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (v < nV) {
    double rv[3], zv[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) rv[l] = b[3ll * v + l];
    const double* D = Dinv + 9ll * v;
#pragma unroll
    for (int l = 0; l < 3; ++l) zv[l] = D[3 * l] * rv[0] + D[3 * l + 1] * rv[1] + D[3 * l + 2] * rv[2];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      x[3ll * v + l] = 0.0;
      r[3ll * v + l] = rv[l];
      z[3ll * v + l] = zv[l];
      p[3ll * v + l] = zv[l];
      rz += rv[l] * zv[l];
      rr += rv[l] * rv[l];
    }
  }
  rz = block_sum_256(rz, sh);
  rr = block_sum_256(rr, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rz;
    part[2 * blockIdx.x + 1] = rr;
  }
}

__global__ void __launch_bounds__(kRedBlock) k_pcg_start(const double* __restrict__ part, int np, double tol,
                                                         int maxit, PcgState* st) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedBlock) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = block_sum_256(a, sh);
  b = block_sum_256(b, sh);
  if (threadIdx.x == 0) {
    st->rz = a;
    st->rr = b;
    st->bb = b;
    st->tol2 = tol * tol;
    st->iters = 0;
    st->maxit = maxit;
    st->done = (b == 0.0) ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kRedBlock) k_pcg_alpha(const double* __restrict__ part, int np, PcgState* st) {
  __shared__ double sh[4];
  if (st->done) return;
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedBlock) a += part[i];
  a = block_sum_256(a, sh);
  if (threadIdx.x == 0) {
    st->pAp = a;
    if (!(a > 0.0)) {  // non-positive curvature: keep the current iterate (or the preconditioned gradient)
      st->alpha = (st->iters == 0) ? 1.0 : 0.0;
      st->done = 3;
    } else {
      st->alpha = st->rz / a;
    }
  }
}

__global__ void __launch_bounds__(kRedBlock)
    k_pcg_update(int nV, const double* __restrict__ Dinv, double* __restrict__ x, double* __restrict__ r,
                 double* __restrict__ z, const double* __restrict__ p, const double* __restrict__ Ap,
                 double* __restrict__ part, PcgState* st) {
  __shared__ double sh[4];
  const int done = st->done;
  if (done && done != 3) return;
  const double alpha = st->alpha;
  const int v = blockIdx.x * kRedBlock + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (v < nV) {
    double rv[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      x[3ll * v + l] += alpha * p[3ll * v + l];
      rv[l] = r[3ll * v + l] - alpha * Ap[3ll * v + l];
      r[3ll * v + l] = rv[l];
    }
    const double* D = Dinv + 9ll * v;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double zl = D[3 * l] * rv[0] + D[3 * l + 1] * rv[1] + D[3 * l + 2] * rv[2];
      z[3ll * v + l] = zl;
      rz += rv[l] * zl;
      rr += rv[l] * rv[l];
    }
  }
  rz = block_sum_256(rz, sh);
  rr = block_sum_256(rr, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rz;
    part[2 * blockIdx.x + 1] = rr;
  }
}

__global__ void __launch_bounds__(kRedBlock) k_pcg_beta(const double* __restrict__ part, int np, PcgState* st) {
  __shared__ double sh[4];
  if (st->done) {
    // a non-positive-curvature exit applied its last (alpha) update in k_pcg_update: count it once
    if (st->done == 3 && threadIdx.x == 0 && st->alpha != 0.0) {
      st->alpha = 0.0;
      st->iters += 1;
    }
    return;
  }
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedBlock) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = block_sum_256(a, sh);
  b = block_sum_256(b, sh);
  if (threadIdx.x == 0) {
    st->beta = a / st->rz;
    st->rz = a;
    st->rr = b;
    st->iters += 1;
    if (b <= st->tol2 * st->bb)
      st->done = 1;
    else if (st->iters >= st->maxit)
      st->done = 2;
  }
}

__global__ void __launch_bounds__(256) k_pcg_p(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                               const PcgState* __restrict__ st) {
  if (st->done) return;
  const double beta = st->beta;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + beta * p[i];
}

// generic-preconditioner CG pieces: x = 0, r = b with partial r.r; x += a p, r -= a Ap with partial r.r
__global__ void __launch_bounds__(kRedBlock) k_pcg_init_r(int64_t n, const double* __restrict__ b,
                                                          double* __restrict__ x, double* __restrict__ r,
                                                          double* __restrict__ part) {
  __shared__ double sh[4];
  const int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    rr = bi * bi;
  }
  rr = block_sum_256(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

__global__ void __launch_bounds__(kRedBlock) k_pcg_update_xr(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ Ap, double* __restrict__ part,
                                                             const PcgState* __restrict__ st) {
  __shared__ double sh[4];
  const int done = st->done;
  if (done && done != 3) return;
  const double alpha = st->alpha;
  const int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x;
  double rr = 0.0;
  if (i < n) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    rr = ri * ri;
  }
  rr = block_sum_256(rr, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = rr;
}

// st from separate r.r and r.z partial arrays (start: also bb; loop: beta and the convergence test)
__global__ void __launch_bounds__(kRedBlock) k_pcg_scalars2(const double* __restrict__ prr, int nrr,
                                                            const double* __restrict__ prz, int nrz, int start,
                                                            double tol, int maxit, PcgState* st) {
  __shared__ double sh[4];
  if (!start && st->done) {
    if (st->done == 3 && threadIdx.x == 0 && st->alpha != 0.0) {
      st->alpha = 0.0;
      st->iters += 1;
    }
    return;
  }
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nrz; i += kRedBlock) a += prz[i];
  for (int i = threadIdx.x; i < nrr; i += kRedBlock) b += prr[i];
  a = block_sum_256(a, sh);
  b = block_sum_256(b, sh);
  if (threadIdx.x == 0) {
    if (start) {
      st->rz = a;
      st->rr = b;
      st->bb = b;
      st->tol2 = tol * tol;
      st->iters = 0;
      st->maxit = maxit;
      st->done = (b == 0.0) ? 1 : 0;
    } else {
      st->beta = a / st->rz;
      st->rz = a;
      st->rr = b;
      st->iters += 1;
      if (b <= st->tol2 * st->bb)
        st->done = 1;
      else if (st->iters >= st->maxit)
        st->done = 2;
    }
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

struct NewtonWs {
  double *vals, *g, *b, *x, *r, *z, *p, *Ap, *qt, *Dinv, *part, *scal;
  float* vals32;    // precision form: FP32 Hessian values
  float* params32;  // precision form: FP32 parameters (2 nT)
  PcgState* st;
};

int newton_workspace(bh_mesh* m, NewtonWs& w) {
  const int64_t n = 3ll * m->nV;
  const int64_t nParts = 2ll * std::max<int64_t>(kRedMaxBlocks, (3ll * m->nV + kRedBlock - 1) / kRedBlock + 1) +
                         kRedMaxBlocks + spmv_blocks(m->nV) + 16;
  const size_t need = sizeof(double) * ((size_t)m->nnz + 8 * (size_t)n + 9 * (size_t)m->nV + nParts + 16) +
                      sizeof(float) * ((size_t)m->nnz + 2 * (size_t)m->nT) + sizeof(PcgState) + 512;
  if (m->wsBytes < need) {
    if (m->ws) BH_CHECK(hipFree(m->ws));
    m->ws = nullptr;
    BH_CHECK(hipMalloc(&m->ws, need));
    m->wsBytes = need;
  }
  double* q = m->ws;
  w.vals = q; q += m->nnz;
  w.g = q; q += n;
  w.b = q; q += n;
  w.x = q; q += n;
  w.r = q; q += n;
  w.z = q; q += n;
  w.p = q; q += n;
  w.Ap = q; q += n;
  w.qt = q; q += n;
  w.Dinv = q; q += 9ll * m->nV;
  w.part = q; q += nParts;
  w.scal = q; q += 16;
  float* f = reinterpret_cast<float*>(q);
  w.vals32 = f; f += m->nnz;
  w.params32 = f; f += 2ll * m->nT;
  w.st = reinterpret_cast<PcgState*>(reinterpret_cast<uintptr_t>(f + 64) & ~uintptr_t(63));
  return BH_OK;
}

// objective(q) = Psi(q) - f.q  -> *out (device)
int objective(bh_mesh* m, const bh_newton_opts* o, const double* params, const double* f, const double* q, double* out,
              NewtonWs& w, hipStream_t s) {
  int st = (o->element_precision == 1)
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
    hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks(m->nV)), dim3(kRedBlock), 0, s, m->nV, m->bptr, m->bcol,
                       o->d_mass_vals, (const uint8_t*)nullptr, w.z, w.Ap, (double*)nullptr, (const PcgState*)nullptr);
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

// PCG solve of H x = b on the masked system; returns the iteration count in *iters, status in *done
template <typename V>
int pcg_solve(bh_mesh* m, const bh_newton_opts* o, const V* vals, const uint8_t* fixed, NewtonWs& w, int* iters,
              int* done, hipStream_t s) {
  const int nV = m->nV;
  const int64_t n = 3ll * nV;
  const int gv = (nV + kRedBlock - 1) / kRedBlock;
  const int gr = (int)((n + kRedBlock - 1) / kRedBlock);
  const int gs = spmv_blocks(nV);
  (void)gr;
  hipLaunchKernelGGL(k_block_jacobi<V>, dim3((nV + 255) / 256), dim3(256), 0, s, nV, m->bptr, m->bcol, vals, fixed,
                     o->precond == BH_PRECOND_POINT_JACOBI ? 1 : 0, w.Dinv);
  hipLaunchKernelGGL(k_pcg_init, dim3(gv), dim3(kRedBlock), 0, s, nV, w.b, w.Dinv, w.x, w.r, w.z, w.p, w.part);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kRedBlock), 0, s, w.part, gv, o->pcg_tol, o->pcg_max_iterations,
                     w.st);
  BH_CHECK(hipGetLastError());
  PcgState hs{};
  for (int it = 0;; it += kPcgBatch) {
    for (int k = 0; k < kPcgBatch; ++k) {
      hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(gs), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals, fixed, w.p, w.Ap,
                         w.part, w.st);
      hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(kRedBlock), 0, s, w.part, gs, w.st);
      hipLaunchKernelGGL(k_pcg_update, dim3(gv), dim3(kRedBlock), 0, s, nV, w.Dinv, w.x, w.r, w.z, w.p, w.Ap, w.part,
                         w.st);
      hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(kRedBlock), 0, s, w.part, gv, w.st);
      hipLaunchKernelGGL(k_pcg_p, dim3(gr), dim3(256), 0, s, n, w.z, w.p, w.st);
    }
    BH_CHECK(hipGetLastError());
    BH_CHECK(hipMemcpyAsync(&hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    if (hs.done) break;
  }
  *iters = hs.iters;
  *done = hs.done;
  return BH_OK;
}

// PCG with a colour-sweep block-triangular preconditioner (same device-resident scalars)
template <typename V>
int pcg_solve_factor(bh_mesh* m, const bh_newton_opts* o, const V* vals, const uint8_t* fixed, const bh_precond* P,
                     NewtonWs& w, int* iters, int* done, hipStream_t s) {
  const int nV = m->nV;
  const int64_t n = 3ll * nV;
  const int gr = (int)((n + kRedBlock - 1) / kRedBlock);
  const int gs = spmv_blocks(nV);
  const int nz = red_blocks(n);
  double* part2 = w.part + gr + 8;  // second partial array (r.z), after the r.r / p.Ap partials
  int* dflag = &w.st->done;
  hipLaunchKernelGGL(k_pcg_init_r, dim3(gr), dim3(kRedBlock), 0, s, n, w.b, w.x, w.r, w.part);
  int st = precond_apply(P, w.r, w.z, nullptr, s);
  if (st) return st;
  BH_CHECK(hipMemcpyAsync(w.p, w.z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_dot_partials, dim3(nz), dim3(kRedBlock), 0, s, w.r, w.z, n, part2);
  hipLaunchKernelGGL(k_pcg_scalars2, dim3(1), dim3(kRedBlock), 0, s, w.part, gr, part2, nz, 1, o->pcg_tol,
                     o->pcg_max_iterations, w.st);
  BH_CHECK(hipGetLastError());
  PcgState hs{};
  for (;;) {
    for (int k = 0; k < kPcgBatch; ++k) {
      hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(gs), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals, fixed, w.p, w.Ap,
                         w.part, w.st);
      hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(kRedBlock), 0, s, w.part, gs, w.st);
      hipLaunchKernelGGL(k_pcg_update_xr, dim3(gr), dim3(kRedBlock), 0, s, n, w.x, w.r, w.p, w.Ap, w.part, w.st);
      if ((st = precond_apply(P, w.r, w.z, dflag, s))) return st;
      hipLaunchKernelGGL(k_dot_partials, dim3(nz), dim3(kRedBlock), 0, s, w.r, w.z, n, part2);
      hipLaunchKernelGGL(k_pcg_scalars2, dim3(1), dim3(kRedBlock), 0, s, w.part, gr, part2, nz, 0, o->pcg_tol,
                         o->pcg_max_iterations, w.st);
      hipLaunchKernelGGL(k_pcg_p, dim3(gr), dim3(256), 0, s, n, w.z, w.p, w.st);
    }
    BH_CHECK(hipGetLastError());
    BH_CHECK(hipMemcpyAsync(&hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    if (hs.done) break;
  }
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
  *fallback = 0;
  if (o->precond == BH_PRECOND_BIC0 || o->precond == BH_PRECOND_FACTOR || o->precond == BH_PRECOND_LEARNED) {
    if (!o->precond_handle) return BH_BAD_ARG;
    if (o->precond == BH_PRECOND_BIC0) {
      // IC(0) may break down (non-positive pivots) on strongly heterogeneous problems: refactor with a
      // growing diagonal shift until every pivot is positive
      double shift = o->precond_shift;
      for (int attempt = 0;; ++attempt) {
        int fails = 0;
        if ((st = precond_factor_bic0(o->precond_handle, vals, fixed, shift, &fails, s))) return st;
        if (fails == 0 || attempt >= 6) break;
        shift = (shift > 0.0) ? 4.0 * shift : 1e-3;
      }
    }
    if (o->precond == BH_PRECOND_LEARNED) {
      if (!o->mgn) return BH_BAD_ARG;
      if ((st = mgn_factor<V>(o->mgn, o->precond_handle, vals, fixed, s))) return st;
    }
    if ((st = pcg_solve_factor(m, o, vals, fixed, o->precond_handle, w, pits, pdone, s))) return st;
    if (*pdone == 2 && o->jacobi_fallback) {
      int p2 = 0;
      if ((st = pcg_solve(m, o, vals, fixed, w, &p2, pdone, s))) return st;
      *pits += p2;
      *fallback = 1;
    }
    return BH_OK;
  }
  return pcg_solve(m, o, vals, fixed, w, pits, pdone, s);
}
}  // namespace

extern "C" {

int bh_spmv_device(bh_mesh* m, const double* d_vals, const double* d_x, double* d_y, bh_stream_t s) {
  if (!m || !d_vals || !d_x || !d_y) return BH_BAD_ARG;
  if (m->nV == 0) return BH_OK;
  hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks(m->nV)), dim3(kRedBlock), 0, as_stream(s), m->nV, m->bptr,
                     m->bcol, d_vals, nullptr, d_x, d_y, nullptr, nullptr);
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
  const bool f32 = o->element_precision == 1;
  if (f32) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3((2ll * m->nT + 255) / 256), dim3(256), 0, s, 2ll * m->nT, params,
                       w.params32);
    BH_CHECK(hipGetLastError());
  }
  do {
    // gradient of the objective, masked
    st = f32 ? gradient_dispatch(m, o->model, o->d_model_ids, q, (const float*)w.params32, w.g, s)
             : gradient_dispatch(m, o->model, o->d_model_ids, q, params, w.g, s);
    if (st) return st;
    if (o->d_mass_vals) {  // + inertia M (q - q_pred)
      hipLaunchKernelGGL(k_sub, dim3(gn), dim3(256), 0, s, n, q, o->d_q_pred, w.z);
      hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(spmv_blocks(m->nV)), dim3(kRedBlock), 0, s, m->nV, m->bptr, m->bcol,
                         o->d_mass_vals, (const uint8_t*)nullptr, w.z, w.Ap, (double*)nullptr,
                         (const PcgState*)nullptr);
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
