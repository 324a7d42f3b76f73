// Block lower-triangular preconditioners on the Hessian's own 3x3-block pattern (included at the end
// of bartels_hip.hip; same translation unit).
//
// Vertices are coloured greedily so that no two neighbours share a colour; in colour order the
// factor L has, for every vertex v, a lower-triangular 3x3 diagonal block L_vv and one 3x3 block
// L_vn per neighbour n of lower colour -- exactly the lower half of H's pattern.  M = L L^T is
// applied as two colour sweeps (forward L y = r, backward L^T z = y); inside one colour every
// vertex is independent, so each sweep is one parallel kernel per colour.
//   * block IC(0): L from H by an incomplete Cholesky restricted to the pattern (colour by colour);
//   * learned: L written by the mesh-graph network (bh_precond_set_factor).
// Storage (FP32): per vertex the inverse of L_vv (6 lower-triangular entries), the lower blocks of
// its row (lptr/lcol/lval) and, for the backward sweep, the transposed blocks of its column
// (uptr/ucol/ulink into lval), so each sweep streams contiguous per-vertex lists.

struct bh_precond {
  bh_mesh* mesh = nullptr;
  int ncolors = 0;
  std::vector<int32_t> h_colPtr;
  int32_t* colVerts = nullptr;  // vertices grouped by colour
  int32_t* lptr = nullptr;      // nV + 1
  int32_t* lcol = nullptr;      // lower neighbour (colour(n) < colour(v)), sorted by colour then id
  int32_t* lslot = nullptr;     // block index of (v, n) in the mesh pattern
  int32_t* uptr = nullptr;      // nV + 1
  int32_t* ucol = nullptr;      // upper neighbour n (colour(n) > colour(v))
  int32_t* ulink = nullptr;     // index of the lower entry (n, v) in n's list
  int32_t* dslot = nullptr;     // block index of (v, v)
  float* lval = nullptr;        // 9 per lower entry, row-major
  float* uval = nullptr;        // the same blocks in upper-list order (for the backward sweep)
  float* dinv = nullptr;        // 6 per vertex: inverse of L_vv (lower triangular, row-major packed)
  int32_t* fail = nullptr;      // factorisation breakdown counter
  int64_t nlower = 0;
  float* ytmp = nullptr;        // forward-sweep result (3 nV)
  float* ztmp = nullptr;        // FP32 shadow of the backward-sweep result (3 nV)
};

namespace {

// ---- factor application ------------------------------------------------------------------------
// packed lower-triangular 3x3: [l00, l10, l11, l20, l21, l22]
__device__ __forceinline__ void ltri_mv(const float* L, const double* x, double* y) {
  y[0] = (double)L[0] * x[0];
  y[1] = (double)L[1] * x[0] + (double)L[2] * x[1];
  y[2] = (double)L[3] * x[0] + (double)L[4] * x[1] + (double)L[5] * x[2];
}

// Sweeps, 8 lanes per vertex: lane j takes list entries j, j+8, ... (a vertex's list is contiguous,
// so the lanes read neighbouring 36-byte blocks), the three partial sums are reduced across the 8
// lanes with a fixed xor butterfly, and lane 0 applies the diagonal inverse.  The factor is FP32
// and so is the sweep arithmetic: the intermediate y and an FP32 shadow of z are what the gathers
// read (12 bytes per neighbour); r comes in and z goes out in FP64 for the CG recurrences.
constexpr int kSweepLanes = 8;

// forward sweep over one colour: y_v = L_vv^{-1} (r_v - sum_n L_vn y_n)
__global__ void __launch_bounds__(256)
    k_sweep_fwd(const int32_t* __restrict__ verts, int n, const int32_t* __restrict__ lptr,
                const int32_t* __restrict__ lcol, const float* __restrict__ lval, const float* __restrict__ dinv,
                const double* __restrict__ r, float* __restrict__ y, const int* __restrict__ done) {
  if (done && *done) return;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / kSweepLanes, j = gid % kSweepLanes;
  int v = 0;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  if (i < n) {
    v = verts[i];
    for (int k = lptr[v] + j; k < lptr[v + 1]; k += kSweepLanes) {
      const float* B = lval + 9ll * k;
      const float* yn = y + 3ll * lcol[k];
      const float y0 = yn[0], y1 = yn[1], y2 = yn[2];
      s0 += B[0] * y0 + B[1] * y1 + B[2] * y2;
      s1 += B[3] * y0 + B[4] * y1 + B[5] * y2;
      s2 += B[6] * y0 + B[7] * y1 + B[8] * y2;
    }
  }
#pragma unroll
  for (int off = kSweepLanes / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (i < n && j == 0) {
    const float* L = dinv + 6ll * v;
    const float t0 = (float)r[3ll * v] - s0, t1 = (float)r[3ll * v + 1] - s1, t2 = (float)r[3ll * v + 2] - s2;
    y[3ll * v] = L[0] * t0;
    y[3ll * v + 1] = L[1] * t0 + L[2] * t1;
    y[3ll * v + 2] = L[3] * t0 + L[4] * t1 + L[5] * t2;
  }
}

// backward sweep over one colour: z_v = L_vv^{-T} (y_v - sum_{n: colour(n) > colour(v)} L_nv^T z_n), with
// L_nv read from uval (the lower blocks copied into the upper-list order)
__global__ void __launch_bounds__(256)
    k_sweep_bwd(const int32_t* __restrict__ verts, int n, const int32_t* __restrict__ uptr,
                const int32_t* __restrict__ ucol, const float* __restrict__ uval, const float* __restrict__ dinv,
                const float* __restrict__ y, float* __restrict__ z32, double* __restrict__ z,
                const int* __restrict__ done) {
  if (done && *done) return;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / kSweepLanes, j = gid % kSweepLanes;
  int v = 0;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  if (i < n) {
    v = verts[i];
    for (int k = uptr[v] + j; k < uptr[v + 1]; k += kSweepLanes) {
      const float* B = uval + 9ll * k;  // L_nv, use its transpose
      const float* zn = z32 + 3ll * ucol[k];
      const float z0 = zn[0], z1 = zn[1], z2 = zn[2];
      s0 += B[0] * z0 + B[3] * z1 + B[6] * z2;
      s1 += B[1] * z0 + B[4] * z1 + B[7] * z2;
      s2 += B[2] * z0 + B[5] * z1 + B[8] * z2;
    }
  }
#pragma unroll
  for (int off = kSweepLanes / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (i < n && j == 0) {
    const float* L = dinv + 6ll * v;
    const float t0 = y[3ll * v] - s0, t1 = y[3ll * v + 1] - s1, t2 = y[3ll * v + 2] - s2;
    // L^T t with L packed lower [l00, l10, l11, l20, l21, l22]
    const float o0 = L[0] * t0 + L[1] * t1 + L[3] * t2;
    const float o1 = L[2] * t1 + L[4] * t2;
    const float o2 = L[5] * t2;
    z32[3ll * v] = o0;
    z32[3ll * v + 1] = o1;
    z32[3ll * v + 2] = o2;
    z[3ll * v] = o0;
    z[3ll * v + 1] = o1;
    z[3ll * v + 2] = o2;
  }
}

// uval[t] = lval[ulink[t]]: the factor's lower blocks in the order of the upper lists
__global__ void __launch_bounds__(256) k_fill_uval(int64_t nu, const int32_t* __restrict__ ulink,
                                                   const float* __restrict__ lval, float* __restrict__ uval) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nu) return;
  const float* src = lval + 9ll * ulink[t];
#pragma unroll
  for (int i = 0; i < 9; ++i) uval[9 * t + i] = src[i];
}

// ---- block IC(0), one colour at a time ----------------------------------------------------------------
// H block (v, n) of the masked system (fixed DOFs -> identity rows / columns), row-major
template <typename V>
__device__ __forceinline__ void masked_block(const V* __restrict__ vals, const int32_t* __restrict__ bptr, int v,
                                             int slot, const uint8_t* __restrict__ fixed, int n, double* A) {
  const int b0 = bptr[v], nb = bptr[v + 1] - b0;
  const int r = slot - b0;
  const V* base = vals + 9ll * b0 + 3ll * r;
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) A[3 * l + k] = (double)base[3ll * l * nb + k];
  if (fixed) {
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const bool fr = fixed[3ll * v + l] != 0, fc = fixed[3ll * n + k] != 0;
        if (fr || fc) A[3 * l + k] = (v == n && l == k) ? 1.0 : 0.0;
      }
  }
}

template <typename V>
__global__ void __launch_bounds__(256)
    k_bic0_color(const int32_t* __restrict__ verts, int nverts, const int32_t* __restrict__ bptr,
                 const V* __restrict__ vals, const uint8_t* __restrict__ fixed, const int32_t* __restrict__ lptr,
                 const int32_t* __restrict__ lcol, const int32_t* __restrict__ lslot,
                 const int32_t* __restrict__ dslot, double shift, float* __restrict__ lval, float* __restrict__ dinv,
                 int32_t* __restrict__ fail) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nverts) return;
  const int v = verts[i];
  const int k0 = lptr[v], k1 = lptr[v + 1];
  double D[9];
  masked_block(vals, bptr, v, dslot[v], fixed, v, D);
  const double sd0 = D[0], sd4 = D[4], sd8 = D[8];
  D[0] += shift * fabs(sd0);
  D[4] += shift * fabs(sd4);
  D[8] += shift * fabs(sd8);
  for (int k = k0; k < k1; ++k) {
    const int n = lcol[k];
    double S[9];
    masked_block(vals, bptr, v, lslot[k], fixed, n, S);
    // S -= sum_j L_{v, n_j} L_{n, n_j}^T over common lower neighbours n_j (earlier in v's list)
    for (int kj = k0; kj < k; ++kj) {
      const int nj = lcol[kj];
      int kn = -1;
      for (int t = lptr[n]; t < lptr[n + 1]; ++t)
        if (lcol[t] == nj) {
          kn = t;
          break;
        }
      if (kn < 0) continue;
      const float* A = lval + 9ll * kj;
      const float* B = lval + 9ll * kn;
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          S[3 * l + c] -= (double)A[3 * l] * B[3 * c] + (double)A[3 * l + 1] * B[3 * c + 1] + (double)A[3 * l + 2] * B[3 * c + 2];
    }
    // L_vn = S L_nn^{-T}
    const float* Ni = dinv + 6ll * n;  // inverse of L_nn, lower triangular
    double Lvn[9];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double x[3] = {S[3 * l], S[3 * l + 1], S[3 * l + 2]};
      double o[3];
      ltri_mv(Ni, x, o);  // (S N^T)_l. = N S_l.^T  -- row l of S times N^T
      Lvn[3 * l] = o[0];
      Lvn[3 * l + 1] = o[1];
      Lvn[3 * l + 2] = o[2];
    }
    float* dst = lval + 9ll * k;
#pragma unroll
    for (int t = 0; t < 9; ++t) dst[t] = (float)Lvn[t];
    // D -= L_vn L_vn^T
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        D[3 * l + c] -= Lvn[3 * l] * Lvn[3 * c] + Lvn[3 * l + 1] * Lvn[3 * c + 1] + Lvn[3 * l + 2] * Lvn[3 * c + 2];
  }
  // Cholesky of D (3x3), store the inverse of the lower factor
  double l00 = D[0];
  bool ok = l00 > 0.0;
  l00 = sqrt(fmax(l00, 1e-300));
  const double l10 = D[3] / l00, l20 = D[6] / l00;
  double d11 = D[4] - l10 * l10;
  ok = ok && d11 > 0.0;
  const double l11 = sqrt(fmax(d11, 1e-300));
  const double l21 = (D[7] - l20 * l10) / l11;
  double d22 = D[8] - l20 * l20 - l21 * l21;
  ok = ok && d22 > 0.0;
  const double l22 = sqrt(fmax(d22, 1e-300));
  if (!ok) atomicAdd(fail, 1);
  // inverse of [[l00,0,0],[l10,l11,0],[l20,l21,l22]]
  const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
  const double i10 = -l10 * i00 * i11;
  const double i21 = -l21 * i11 * i22;
  const double i20 = -(l20 * i00 + l21 * i10) * i22;
  float* o = dinv + 6ll * v;
  o[0] = (float)i00;
  o[1] = (float)i10;
  o[2] = (float)i11;
  o[3] = (float)i20;
  o[4] = (float)i21;
  o[5] = (float)i22;
}

int precond_build(bh_mesh* m, bh_precond* P) {
  const int nV = m->nV;
  const std::vector<int32_t>& bp = m->h_bptr;
  const std::vector<int32_t>& bc = m->h_bcol;
  // greedy colouring in vertex order
  std::vector<int32_t> color(nV, -1);
  std::vector<char> used;
  int nc = 0;
  for (int v = 0; v < nV; ++v) {
    used.assign(nc + 1, 0);
    for (int k = bp[v]; k < bp[v + 1]; ++k) {
      const int c = color[bc[k]];
      if (c >= 0) used[c] = 1;
    }
    int c = 0;
    while (used[c]) ++c;
    color[v] = c;
    nc = std::max(nc, c + 1);
  }
  P->ncolors = nc;
  std::vector<int32_t> colPtr(nc + 1, 0), colVerts(nV);
  for (int v = 0; v < nV; ++v) colPtr[color[v] + 1]++;
  for (int c = 0; c < nc; ++c) colPtr[c + 1] += colPtr[c];
  {
    std::vector<int32_t> pos(colPtr.begin(), colPtr.end() - 1);
    for (int v = 0; v < nV; ++v) colVerts[pos[color[v]]++] = v;
  }
  std::vector<int32_t> lptr(nV + 1, 0), uptr(nV + 1, 0), lcol, lslot, ucol, dslot(nV, -1);
  for (int v = 0; v < nV; ++v) {
    std::vector<std::pair<int, int>> low;  // (colour, vertex) -> sorted
    for (int k = bp[v]; k < bp[v + 1]; ++k) {
      const int n = bc[k];
      if (n == v)
        dslot[v] = k;
      else if (color[n] < color[v])
        low.emplace_back(color[n], n);
    }
    std::sort(low.begin(), low.end());
    for (auto& pr : low) {
      lcol.push_back(pr.second);
      lslot.push_back((int32_t)(std::lower_bound(bc.begin() + bp[v], bc.begin() + bp[v + 1], pr.second) - bc.begin()));
    }
    lptr[v + 1] = (int32_t)lcol.size();
  }
  std::vector<int32_t> ulink;
  for (int v = 0; v < nV; ++v) {
    for (int k = bp[v]; k < bp[v + 1]; ++k) {
      const int n = bc[k];
      if (n != v && color[n] > color[v]) {
        int idx = -1;
        for (int t = lptr[n]; t < lptr[n + 1]; ++t)
          if (lcol[t] == v) {
            idx = t;
            break;
          }
        ucol.push_back(n);
        ulink.push_back(idx);
      }
    }
    uptr[v + 1] = (int32_t)ucol.size();
  }
  P->nlower = (int64_t)lcol.size();
  P->h_colPtr = colPtr;
  auto up = [](int32_t** d, const std::vector<int32_t>& h) -> int {
    BH_CHECK(hipMalloc(d, sizeof(int32_t) * std::max<size_t>(1, h.size())));
    if (!h.empty()) BH_CHECK(hipMemcpy(*d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
    return BH_OK;
  };
  int st;
  if ((st = up(&P->colVerts, colVerts)) || (st = up(&P->lptr, lptr)) || (st = up(&P->lcol, lcol)) ||
      (st = up(&P->lslot, lslot)) || (st = up(&P->uptr, uptr)) || (st = up(&P->ucol, ucol)) ||
      (st = up(&P->ulink, ulink)) || (st = up(&P->dslot, dslot)))
    return st;
  BH_CHECK(hipMalloc(&P->lval, sizeof(float) * 9 * std::max<int64_t>(1, P->nlower)));
  BH_CHECK(hipMalloc(&P->uval, sizeof(float) * 9 * std::max<int64_t>(1, P->nlower)));
  BH_CHECK(hipMalloc(&P->dinv, sizeof(float) * 6 * std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->fail, sizeof(int32_t)));
  BH_CHECK(hipMalloc(&P->ytmp, sizeof(float) * 3 * std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->ztmp, sizeof(float) * 3 * std::max(1, nV)));
  return BH_OK;
}

void precond_free(bh_precond* P) {
  void* ptrs[] = {P->colVerts, P->lptr, P->lcol, P->lslot, P->uptr, P->ucol, P->ulink, P->dslot,
                  P->lval, P->uval, P->dinv, P->fail, P->ytmp, P->ztmp};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// after every change of lval: the upper-list copy read by the backward sweep
int precond_finalize(bh_precond* P, hipStream_t s) {
  if (P->nlower == 0) return BH_OK;
  hipLaunchKernelGGL(k_fill_uval, dim3((P->nlower + 255) / 256), dim3(256), 0, s, P->nlower, P->ulink, P->lval, P->uval);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// z = (L L^T)^{-1} r
int precond_apply(const bh_precond* P, const double* r, double* z, const int* done, hipStream_t s) {
  const int nc = P->ncolors;
  for (int c = 0; c < nc; ++c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    hipLaunchKernelGGL(k_sweep_fwd, dim3((kSweepLanes * n + 255) / 256), dim3(256), 0, s, P->colVerts + P->h_colPtr[c],
                       n, P->lptr, P->lcol, P->lval, P->dinv, r, P->ytmp, done);
  }
  for (int c = nc - 1; c >= 0; --c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    hipLaunchKernelGGL(k_sweep_bwd, dim3((kSweepLanes * n + 255) / 256), dim3(256), 0, s, P->colVerts + P->h_colPtr[c],
                       n, P->uptr, P->ucol, P->uval, P->dinv, P->ytmp, P->ztmp, z, done);
  }
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

int precond_finalize(bh_precond* P, hipStream_t s);

template <typename V>
int precond_factor_bic0(bh_precond* P, const V* vals, const uint8_t* fixed, double shift, int* failures,
                        hipStream_t s) {
  bh_mesh* m = P->mesh;
  BH_CHECK(hipMemsetAsync(P->fail, 0, sizeof(int32_t), s));
  for (int c = 0; c < P->ncolors; ++c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    hipLaunchKernelGGL(k_bic0_color<V>, dim3((n + 255) / 256), dim3(256), 0, s, P->colVerts + P->h_colPtr[c], n,
                       m->bptr, vals, fixed, P->lptr, P->lcol, P->lslot, P->dslot, shift, P->lval, P->dinv, P->fail);
  }
  BH_CHECK(hipGetLastError());
  if (int st = precond_finalize(P, s)) return st;
  if (failures) {
    int32_t h = 0;
    BH_CHECK(hipMemcpyAsync(&h, P->fail, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    *failures = h;
  }
  return BH_OK;
}

}  // namespace

extern "C" {

int bh_precond_create(bh_mesh* m, bh_precond** out) {
  if (!m || !out) return BH_BAD_ARG;
  *out = nullptr;
  bh_precond* P = new bh_precond();
  P->mesh = m;
  const int st = precond_build(m, P);
  if (st) {
    precond_free(P);
    delete P;
    return st;
  }
  *out = P;
  return BH_OK;
}

int bh_precond_destroy(bh_precond* P) {
  if (!P) return BH_BAD_ARG;
  (void)hipDeviceSynchronize();
  precond_free(P);
  delete P;
  return BH_OK;
}

int bh_precond_info(const bh_precond* P, int* ncolors, int64_t* nlower) {
  if (!P) return BH_BAD_ARG;
  if (ncolors) *ncolors = P->ncolors;
  if (nlower) *nlower = P->nlower;
  return BH_OK;
}

int bh_precond_factor_bic0_device(bh_precond* P, const double* d_vals, const uint8_t* d_fixed_dof, double shift,
                                  int* failures, bh_stream_t s) {
  if (!P || !d_vals) return BH_BAD_ARG;
  return precond_factor_bic0(P, d_vals, d_fixed_dof, shift, failures, as_stream(s));
}

int bh_precond_set_factor_device(bh_precond* P, const float* d_lower, const float* d_diag_inv, bh_stream_t s) {
  if (!P || !d_lower || !d_diag_inv) return BH_BAD_ARG;
  hipStream_t hs = as_stream(s);
  if (P->nlower)
    BH_CHECK(hipMemcpyAsync(P->lval, d_lower, sizeof(float) * 9 * P->nlower, hipMemcpyDeviceToDevice, hs));
  BH_CHECK(hipMemcpyAsync(P->dinv, d_diag_inv, sizeof(float) * 6 * (size_t)P->mesh->nV, hipMemcpyDeviceToDevice, hs));
  return precond_finalize(P, hs);
}

int bh_precond_lists_device(const bh_precond* P, const int32_t** lptr, const int32_t** lcol, const int32_t** lslot,
                            const int32_t** dslot, const int32_t** colverts) {
  if (!P) return BH_BAD_ARG;
  if (lptr) *lptr = P->lptr;
  if (lcol) *lcol = P->lcol;
  if (lslot) *lslot = P->lslot;
  if (dslot) *dslot = P->dslot;
  if (colverts) *colverts = P->colVerts;
  return BH_OK;
}

int bh_precond_apply_device(bh_precond* P, const double* d_r, double* d_z, bh_stream_t s) {
  if (!P || !d_r || !d_z) return BH_BAD_ARG;
  return precond_apply(P, d_r, d_z, nullptr, as_stream(s));
}

}  // extern "C"
