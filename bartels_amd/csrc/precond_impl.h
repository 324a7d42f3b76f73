// Block lower-triangular preconditioners on the Hessian's own 3x3-block pattern (included at the end
// of bartels_hip.hip; same translation unit).
//
// Vertices are coloured (DSatur) so that no two neighbours share a colour; in colour order the
// factor L has, for every vertex v, a lower-triangular 3x3 diagonal block L_vv and one 3x3 block
// L_vn per neighbour n of lower colour -- exactly the lower half of H's pattern.  M = L L^T is
// applied as two colour sweeps (forward L y = r, backward L^T z = y); inside one colour every
// vertex is independent, so each sweep is one parallel kernel per colour.
//   * block IC(0): L from H by an incomplete Cholesky restricted to the pattern (colour by colour);
//   * learned: L written by the mesh-graph network (bh_precond_set_factor).
// Storage: the factor is written in vertex order (lptr/lcol/lval, FP32, and the inverse of each L_vv as
// 6 packed entries) by IC(0) or the network; the solve runs in the colour order -- position i holds
// vertex colVerts[i], so every colour is one contiguous range of positions, rows, factor blocks and
// Hessian blocks.  precond_finalize copies the factor into that order (lvs; uvs holds the same blocks
// in the order of the upper lists, for the backward sweep), and precond_split_hessian copies the
// masked Hessian into lower blocks, the transposed-lower copy for the upper part and FP64 diagonal
// blocks, which is what the solve's products read.

struct bh_precond {
  bh_mesh* mesh = nullptr;
  int ncolors = 0;
  std::vector<int32_t> h_colPtr;
  int32_t* colVerts = nullptr;  // vertices grouped by colour (position -> vertex)
  int32_t* lptr = nullptr;      // nV + 1
  int32_t* lcol = nullptr;      // lower neighbour (colour(n) < colour(v)), sorted by colour then id
  int32_t* lslot = nullptr;     // block index of (v, n) in the mesh pattern
  int32_t* uptr = nullptr;      // nV + 1
  int32_t* ucol = nullptr;      // upper neighbour n (colour(n) > colour(v))
  int32_t* ulink = nullptr;     // index of the lower entry (n, v) in n's list
  int32_t* dslot = nullptr;     // block index of (v, v)
  float* lval = nullptr;        // 9 per lower entry, row-major
  float* dinv = nullptr;        // 6 per vertex: inverse of L_vv (lower triangular, row-major packed)
  int32_t* fail = nullptr;      // factorisation breakdown counter
  int64_t nlower = 0;
  // the solve order
  int32_t* nlptr = nullptr;     // nV + 1 per position
  int32_t* nlcol = nullptr;     // position of the lower neighbour
  int32_t* nlsrc = nullptr;     // index of the entry in the vertex-order lists
  int32_t* nuptr = nullptr;     // nV + 1 per position
  int32_t* nucol = nullptr;     // position of the upper neighbour
  int32_t* nulink = nullptr;    // solve-order index of the lower entry (n, v) in n's list
  float* lvs = nullptr;         // factor lower blocks (9 each, row-major)
  float* uvs = nullptr;         // the lower blocks (n, i) in i's upper-list order
  float* dinvs = nullptr;       // 6 per position
  double* hsplit = nullptr;     // masked Hessian: lower blocks and their transposed copy (FP32 or FP64)
  double* hdiag = nullptr;      // masked diagonal blocks (FP64, 9 per position)
  double* sv = nullptr;         // solve vectors, 3 nV each: b, x, r, p, s, u, w, y, t, v, z
  // Eisenstat form: the block diagonal Dt and its inverse (FP32, packed symmetric, 6 per position)
  // and K = 2 Dt' - D in FP64 (6 per position), with Dt' the exact inverse of the stored FP32
  // inverse (the identity A = P + P^T - K that the sweeps rely on then holds to FP64 rounding)
  float* dt = nullptr;
  float* dtinv = nullptr;
  double* kk = nullptr;
  int eisLanes = 4;             // lanes per position in the Eisenstat sweeps (4 or 8)
};

namespace {

// ---- factor application (solve order) ----------------------------------------------------------
// packed lower-triangular 3x3: [l00, l10, l11, l20, l21, l22]
__device__ __forceinline__ void ltri_mv(const float* L, const double* x, double* y) {
  y[0] = (double)L[0] * x[0];
  y[1] = (double)L[1] * x[0] + (double)L[2] * x[1];
  y[2] = (double)L[3] * x[0] + (double)L[4] * x[1] + (double)L[5] * x[2];
}

// Sweeps, 8 lanes per position: lane j takes list entries j, j+8, ... (a position's list is
// contiguous and so are the lists of a colour, so the lanes stream neighbouring 36-byte blocks), the
// three partial sums are reduced across the 8 lanes with a fixed xor butterfly, and lane 0 applies the
// diagonal inverse.  The factor is stored in FP32 and widened; the sweep arithmetic and the vectors y,
// z are FP64, so M^{-1} stays a fixed linear operator to FP64 rounding (an FP32 z perturbs CG like an
// inexact preconditioner and costs iterations at tight tolerances).  FP16 factor blocks were tried and
// dropped: their 5e-4 rounding swamps the near-cancellation that makes an incomplete factor accurate
// on the smooth, low-stiffness modes, and IC(0) then needed more iterations than block Jacobi.
constexpr int kSweepLanes = 8;
constexpr int kSweepVpb = 256 / kSweepLanes;  // positions per block

// CG vectors a forward sweep updates before it reads r (Chronopoulos-Gear recurrences, rows of the
// colour's positions only): p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s
struct CgVecs {
  double *x, *p, *s;
  const double *u, *w;
  double* part_rr;  // partial r.r per block (after the update)
  const PcgState* st;
};

// forward sweep over the colour's positions [c0, c0 + n): y_i = L_ii^{-1} (r_i - sum_n L_in y_n),
// optionally after the CG update
__global__ void __launch_bounds__(256)
    k_sweep_fwd(int c0, int n, const int32_t* __restrict__ nlptr, const int32_t* __restrict__ nlcol,
                const float* __restrict__ lv, const float* __restrict__ dinv, double* __restrict__ r,
                double* __restrict__ y, CgVecs cg, int update) {
  __shared__ double sh[4];
  if (cg.st && cg_skip(cg.st)) return;
  const int j = threadIdx.x % kSweepLanes;
  const int li = blockIdx.x * kSweepVpb + threadIdx.x / kSweepLanes;
  const int i = c0 + li;
  double rj = 0.0, rr = 0.0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (li < n) {
    if (j < 3) {
      const int64_t row = 3ll * i + j;
      rj = r[row];
      if (update) {
        const double alpha = cg.st->alpha, beta = cg.st->beta;
        const double pj = cg.u[row] + beta * cg.p[row];
        const double sj = cg.w[row] + beta * cg.s[row];
        cg.p[row] = pj;
        cg.s[row] = sj;
        cg.x[row] += alpha * pj;
        rj -= alpha * sj;
        r[row] = rj;
        rr = rj * rj;
      }
    }
    for (int k = nlptr[i] + j; k < nlptr[i + 1]; k += kSweepLanes) {
      const float* B = lv + 9ll * k;
      const double* yn = y + 3ll * nlcol[k];
      const double y0 = yn[0], y1 = yn[1], y2 = yn[2];
      s0 += (double)B[0] * y0 + (double)B[1] * y1 + (double)B[2] * y2;
      s1 += (double)B[3] * y0 + (double)B[4] * y1 + (double)B[5] * y2;
      s2 += (double)B[6] * y0 + (double)B[7] * y1 + (double)B[8] * y2;
    }
  }
#pragma unroll
  for (int off = kSweepLanes / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  const int lb = (int)(threadIdx.x & 63) - j;  // group base lane within the wavefront
  const double r0 = __shfl(rj, lb, 64), r1 = __shfl(rj, lb + 1, 64), r2 = __shfl(rj, lb + 2, 64);
  if (li < n && j == 0) {
    const float* L = dinv + 6ll * i;
    const double t0 = r0 - s0, t1 = r1 - s1, t2 = r2 - s2;
    y[3ll * i] = (double)L[0] * t0;
    y[3ll * i + 1] = (double)L[1] * t0 + (double)L[2] * t1;
    y[3ll * i + 2] = (double)L[3] * t0 + (double)L[4] * t1 + (double)L[5] * t2;
  }
  if (cg.part_rr) {
    const double d = block_sum_256(rr, sh);
    if (threadIdx.x == 0) cg.part_rr[blockIdx.x] = d;
  }
}

// backward sweep over the colour's positions: z_i = L_ii^{-T} (y_i - sum_{n upper} L_ni^T z_n), with L_ni
// read from uvs (the lower blocks in upper-list order); optional partial r.z
__global__ void __launch_bounds__(256)
    k_sweep_bwd(int c0, int n, const int32_t* __restrict__ nuptr, const int32_t* __restrict__ nucol,
                const float* __restrict__ uv, const float* __restrict__ dinv, const double* __restrict__ y,
                double* __restrict__ z, const double* __restrict__ r, double* __restrict__ part_rz,
                const PcgState* __restrict__ st) {
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int j = threadIdx.x % kSweepLanes;
  const int li = blockIdx.x * kSweepVpb + threadIdx.x / kSweepLanes;
  const int i = c0 + li;
  double rz = 0.0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (li < n) {
    for (int k = nuptr[i] + j; k < nuptr[i + 1]; k += kSweepLanes) {
      const float* B = uv + 9ll * k;  // L_ni, use its transpose
      const double* zn = z + 3ll * nucol[k];
      const double z0 = zn[0], z1 = zn[1], z2 = zn[2];
      s0 += (double)B[0] * z0 + (double)B[3] * z1 + (double)B[6] * z2;
      s1 += (double)B[1] * z0 + (double)B[4] * z1 + (double)B[7] * z2;
      s2 += (double)B[2] * z0 + (double)B[5] * z1 + (double)B[8] * z2;
    }
  }
#pragma unroll
  for (int off = kSweepLanes / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  if (li < n && j == 0) {
    const float* L = dinv + 6ll * i;
    const double t0 = y[3ll * i] - s0, t1 = y[3ll * i + 1] - s1, t2 = y[3ll * i + 2] - s2;
    // L^T t with L packed lower [l00, l10, l11, l20, l21, l22]
    const double o0 = (double)L[0] * t0 + (double)L[1] * t1 + (double)L[3] * t2;
    const double o1 = (double)L[2] * t1 + (double)L[4] * t2;
    const double o2 = (double)L[5] * t2;
    z[3ll * i] = o0;
    z[3ll * i + 1] = o1;
    z[3ll * i + 2] = o2;
    if (r) rz = r[3ll * i] * o0 + r[3ll * i + 1] * o1 + r[3ll * i + 2] * o2;
  }
  if (part_rz) {
    const double d = block_sum_256(rz, sh);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = d;
  }
}

// the split Hessian's blocks are stored component-major: component q of block k at h[q * ld + k]
// (ld = number of lower blocks), so the lanes of a wave, which walk consecutive blocks, read each
// component as one contiguous run
template <typename V>
__device__ __forceinline__ void load_block(const V* __restrict__ h, int64_t ld, int64_t k, double B[9]) {
#pragma unroll
  for (int q = 0; q < 9; ++q) B[q] = (double)h[q * ld + k];
}

// Ap = H p in the solve order from the split Hessian (lower blocks, the transposed copy of the lower
// blocks for the upper part, diagonal blocks), partial p.Ap.  8 lanes per position over the joined
// lower + upper list; lane 0 adds the diagonal block.  The masked system is symmetric by construction.
template <typename V>
__global__ void __launch_bounds__(256)
    k_split_spmv(int nV, const int32_t* __restrict__ nlptr, const int32_t* __restrict__ nlcol,
                 const V* __restrict__ hl, const int32_t* __restrict__ nuptr, const int32_t* __restrict__ nucol,
                 const V* __restrict__ hu, int64_t ld, const double* __restrict__ hd, const double* __restrict__ p,
                 double* __restrict__ Ap, double* __restrict__ part, const PcgState* __restrict__ st) {
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int j = threadIdx.x % kSweepLanes;
  const int i = blockIdx.x * kSweepVpb + threadIdx.x / kSweepLanes;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (i < nV) {
    const int l0 = nlptr[i], nl = nlptr[i + 1] - l0;
    const int u0 = nuptr[i], nt = nl + nuptr[i + 1] - u0;
    for (int t = j; t < nt; t += kSweepLanes) {
      if (t < nl) {
        double B[9];
        load_block(hl, ld, l0 + t, B);
        const double* xp = p + 3ll * nlcol[l0 + t];
        const double x0 = xp[0], x1 = xp[1], x2 = xp[2];
        s0 += (double)B[0] * x0 + (double)B[1] * x1 + (double)B[2] * x2;
        s1 += (double)B[3] * x0 + (double)B[4] * x1 + (double)B[5] * x2;
        s2 += (double)B[6] * x0 + (double)B[7] * x1 + (double)B[8] * x2;
      } else {
        const int k = u0 + t - nl;
        double B[9];
        load_block(hu, ld, k, B);  // H_ni: use its transpose
        const double* xp = p + 3ll * nucol[k];
        const double x0 = xp[0], x1 = xp[1], x2 = xp[2];
        s0 += (double)B[0] * x0 + (double)B[3] * x1 + (double)B[6] * x2;
        s1 += (double)B[1] * x0 + (double)B[4] * x1 + (double)B[7] * x2;
        s2 += (double)B[2] * x0 + (double)B[5] * x1 + (double)B[8] * x2;
      }
    }
    if (j == 0) {
      const double* B = hd + 9ll * i;
      const double x0 = p[3ll * i], x1 = p[3ll * i + 1], x2 = p[3ll * i + 2];
      s0 += (double)B[0] * x0 + (double)B[1] * x1 + (double)B[2] * x2;
      s1 += (double)B[3] * x0 + (double)B[4] * x1 + (double)B[5] * x2;
      s2 += (double)B[6] * x0 + (double)B[7] * x1 + (double)B[8] * x2;
    }
  }
#pragma unroll
  for (int off = kSweepLanes / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  double pap = 0.0;
  if (i < nV && j < 3) {
    const int64_t row = 3ll * i + j;
    const double sv = (j == 0) ? s0 : (j == 1) ? s1 : s2;
    Ap[row] = sv;
    pap = p[row] * sv;
  }
  if (part) {
    const double d = block_sum_256(pap, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
  }
}

// ---- Eisenstat form ------------------------------------------------------------------------------
// M = P Dt^{-1} P^T with P = Dt + L, L the strictly lower blocks of the masked H in colour order and Dt
// a block diagonal (D: block symmetric Gauss-Seidel, or written by the network).  With A = L + D + L^T = P + P^T - K, K = 2 Dt - D,
//   P^{-1} A P^{-T} v = t + P^{-1} (v - K t),   t = P^{-T} v        (Eisenstat 1981),
// one backward sweep over the upper blocks and one forward sweep over the lower blocks: a single read
// of H per CG iteration and no separate product with H.  CG runs on P^{-1} A P^{-T} with Dt as its
// block-diagonal preconditioner (in exact arithmetic PCG with M on A); the residual of A x = b is
// checked directly (x = P^{-T} x~, r = b - A x) whenever the transformed residual says it may be met.

// packed upper triangle of a symmetric 3x3 (s00, s01, s02, s11, s12, s22) times x
template <typename T>
__device__ __forceinline__ void mv3s(const T* S, double x0, double x1, double x2, double* y) {
  const double a = S[0], b = S[1], c = S[2], d = S[3], e = S[4], f = S[5];
  y[0] = a * x0 + b * x1 + c * x2;
  y[1] = b * x0 + d * x1 + e * x2;
  y[2] = c * x0 + e * x1 + f * x2;
}

// Store an Eisenstat diagonal Dt (FP64 in, symmetric, one position), each as a packed upper triangle:
// FP32 Dt and its inverse, FP64 K = 2 Dt' - D (Dt' = the inverse of the stored FP32 inverse).  The
// symmetric packing reads 24 B per FP32 block and 48 B for K instead of 36 and 72.
__device__ __forceinline__ void eis_store_diag(int64_t i, const double* Dt, const double* __restrict__ hd,
                                               float* __restrict__ dt, float* __restrict__ dtinv,
                                               double* __restrict__ kk) {
  constexpr int kUp[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  double A[9], Ai[9], Ae[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) A[e] = Dt[e];
  inv3_block(A, 0, Ai);
  double As[9];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int r = kUp[k][0], c = kUp[k][1];
    dt[6 * i + k] = (float)(0.5 * (Dt[3 * r + c] + Dt[3 * c + r]));
    const float fi = (float)(0.5 * (Ai[3 * r + c] + Ai[3 * c + r]));
    dtinv[6 * i + k] = fi;
    As[3 * r + c] = As[3 * c + r] = (double)fi;
  }
  inv3_block(As, 0, Ae);  // Dt' = (stored inverse)^{-1}
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int r = kUp[k][0], c = kUp[k][1];
    kk[6 * i + k] = (Ae[3 * r + c] + Ae[3 * c + r]) - hd[9 * i + 3 * r + c];
  }
}

// block symmetric Gauss-Seidel: Dt = D, the masked diagonal blocks.  (A D-ILU diagonal,
// Dt_i = D_i - sum H_in Dt_n^{-1} H_in^T, was tried first: on these elasticity Hessians, which are not
// M-matrices, it turns indefinite and CG stalls.)
__global__ void __launch_bounds__(256) k_sgs_diag(int nV, const double* __restrict__ hd, double scale,
                                                  float* __restrict__ dt, float* __restrict__ dtinv,
                                                  double* __restrict__ kk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  double D[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = scale * hd[9ll * i + e];
  eis_store_diag(i, D, hd, dt, dtinv, kk);
}

// transformed-space CG vectors and the sweep outputs (solve order), all FP64: with the search and
// product vectors in FP32 the Chronopoulos-Gear recurrences lost conjugacy and CG needed 7x the
// iterations on the 6.07M-tet beam.
struct EisVecs {
  double *rt, *xt;                 // r~, x~
  double *pt, *st, *ut, *wt;       // p~, s~ (= A~ p~), u = Dt r~, w~ = A~ u
  double *t, *w;                   // backward / forward sweep results of the product (gather sources)
};


// backward sweep over one colour (colours high to low), LANES lanes per position, writing tb (type TT,
// also the gather source):
//   mode 0: tb = P^{-T} src
//   mode 1: u = Dt r~, tb = P^{-T} u, partial r~.u                       (tb = v.t)
//   mode 2: first the CG update of the position's rows (p~ = u + beta p~, s~ = w~ + beta s~,
//           x~ += alpha p~, r~ -= alpha s~), then mode 1
template <typename V, int LANES, typename TT>
__global__ void __launch_bounds__(256)
    k_eis_bwd(int c0, int n, const int32_t* __restrict__ nuptr, const int32_t* __restrict__ nucol,
              const V* __restrict__ hu, int64_t ld, const float* __restrict__ dt, const float* __restrict__ dtinv, int mode,
              EisVecs v, const double* __restrict__ src, TT* __restrict__ tb, double* __restrict__ part,
              const PcgState* __restrict__ st, const double* __restrict__ kk, double* __restrict__ fpart) {
  static_assert(LANES >= 3, "lanes 0..2 carry the three rows of a position");
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int j = threadIdx.x % LANES;
  const int li = blockIdx.x * (256 / LANES) + threadIdx.x / LANES;
  const int i = c0 + li;
  double aj = 0.0;  // lane j < 3: row j of the right-hand side
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (li < n) {
    if (j < 3) {
      const int64_t row = 3ll * i + j;
      if (mode == 0) {
        aj = src[row];
      } else {
        double rj = v.rt[row];
        if (mode == 2) {
          const double alpha = st->alpha, beta = st->beta;
          const double pj = v.ut[row] + beta * v.pt[row];
          const double sj = v.wt[row] + beta * v.st[row];
          v.pt[row] = pj;
          v.st[row] = sj;
          v.xt[row] += alpha * pj;
          rj -= alpha * sj;
          v.rt[row] = rj;
        }
        aj = rj;
      }
    }
    for (int k = nuptr[i] + j; k < nuptr[i + 1]; k += LANES) {
      double B[9];
      load_block(hu, ld, k, B);  // H_ni; the product needs H_in = H_ni^T
      const TT* tn = tb + 3ll * nucol[k];
      const double t0 = tn[0], t1 = tn[1], t2 = tn[2];
      s0 += (double)B[0] * t0 + (double)B[3] * t1 + (double)B[6] * t2;
      s1 += (double)B[1] * t0 + (double)B[4] * t1 + (double)B[7] * t2;
      s2 += (double)B[2] * t0 + (double)B[5] * t1 + (double)B[8] * t2;
    }
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  const int lb = (int)(threadIdx.x & 63) - j;
  const double a0 = __shfl(aj, lb, 64), a1 = __shfl(aj, lb + 1, 64), a2 = __shfl(aj, lb + 2, 64);
  double g = 0.0, fd = 0.0;
  if (li < n && j == 0) {
    double rhs[3] = {a0, a1, a2};
    if (mode >= 1) {
      double u[3];
      mv3s(dt + 6ll * i, a0, a1, a2, u);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        v.ut[3ll * i + l] = u[l];
        rhs[l] = u[l];
      }
      g = a0 * rhs[0] + a1 * rhs[1] + a2 * rhs[2];
    }
    double tt[3];
    mv3s(dtinv + 6ll * i, rhs[0] - s0, rhs[1] - s1, rhs[2] - s2, tt);
    tb[3ll * i] = (TT)tt[0];
    tb[3ll * i + 1] = (TT)tt[1];
    tb[3ll * i + 2] = (TT)tt[2];
    if (kk) {
      // the forward sweep's step on this colour when it is the first of the solve order (no lower
      // blocks): w = D~^{-1} (u - K t), w~ = t + w, partial w~.u -- what k_eis_fwd mode 1 computes
      double kt[3], ww[3];
      mv3s(kk + 6ll * i, (double)(TT)tt[0], (double)(TT)tt[1], (double)(TT)tt[2], kt);
      mv3s(dtinv + 6ll * i, rhs[0] - kt[0], rhs[1] - kt[1], rhs[2] - kt[2], ww);
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        v.w[3ll * i + l] = ww[l];
        const double wt = (double)(TT)tt[l] + ww[l];
        v.wt[3ll * i + l] = wt;
        fd += wt * rhs[l];
      }
    }
  }
  if (part) {
    const double d = block_sum_256(g, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
  }
  if (fpart) {
    const double d = block_sum_256(fd, sh);
    if (threadIdx.x == 0) fpart[blockIdx.x] = d;
  }
}

// forward sweep over one colour (colours low to high), LANES lanes per position, writing db (type TT,
// also the gather source):
//   mode 0: db = P^{-1} src
//   mode 1: db = w = P^{-1} (u - K t), w~ = t + w, partial w~.u           (db = v.w)
template <typename V, int LANES, typename TT>
__global__ void __launch_bounds__(256)
    k_eis_fwd(int c0, int n, const int32_t* __restrict__ nlptr, const int32_t* __restrict__ nlcol,
              const V* __restrict__ hl, int64_t ld, const float* __restrict__ dtinv, const double* __restrict__ kk, int mode,
              EisVecs v, const double* __restrict__ src, TT* __restrict__ db, double* __restrict__ part,
              const PcgState* __restrict__ st) {
  static_assert(LANES >= 3, "lanes 0..2 carry the three rows of a position");
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int j = threadIdx.x % LANES;
  const int li = blockIdx.x * (256 / LANES) + threadIdx.x / LANES;
  const int i = c0 + li;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (li < n) {
    for (int k = nlptr[i] + j; k < nlptr[i + 1]; k += LANES) {
      double B[9];
      load_block(hl, ld, k, B);  // H_in
      const TT* wn = db + 3ll * nlcol[k];
      const double w0 = wn[0], w1 = wn[1], w2 = wn[2];
      s0 += (double)B[0] * w0 + (double)B[1] * w1 + (double)B[2] * w2;
      s1 += (double)B[3] * w0 + (double)B[4] * w1 + (double)B[5] * w2;
      s2 += (double)B[6] * w0 + (double)B[7] * w1 + (double)B[8] * w2;
    }
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  double d = 0.0;
  if (li < n && j == 0) {
    double rhs[3];
    double t[3] = {0.0, 0.0, 0.0}, u[3] = {0.0, 0.0, 0.0};
    if (mode == 0) {
      rhs[0] = src[3ll * i];
      rhs[1] = src[3ll * i + 1];
      rhs[2] = src[3ll * i + 2];
    } else {
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        t[l] = v.t[3ll * i + l];
        u[l] = v.ut[3ll * i + l];
      }
      double kt[3];
      mv3s(kk + 6ll * i, t[0], t[1], t[2], kt);
#pragma unroll
      for (int l = 0; l < 3; ++l) rhs[l] = u[l] - kt[l];
    }
    double ww[3];
    mv3s(dtinv + 6ll * i, rhs[0] - s0, rhs[1] - s1, rhs[2] - s2, ww);
    db[3ll * i] = (TT)ww[0];
    db[3ll * i + 1] = (TT)ww[1];
    db[3ll * i + 2] = (TT)ww[2];
    if (mode == 1) {
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const double wt = t[l] + ww[l];
        v.wt[3ll * i + l] = wt;
        d += wt * u[l];
      }
    }
  }
  if (part) {
    const double dd = block_sum_256(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = dd;
  }
}

// r = b - y in place of y
__global__ void __launch_bounds__(256) k_resid(int64_t n, const double* __restrict__ b, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = b[i] - y[i];
}

// dst block t = src block link[t] (9 values each)
// the upper list's blocks: hu[., t] = hl[., link[t]] (the transposes are taken by the readers)
template <typename V>
__global__ void __launch_bounds__(256) k_copy_blocks(int64_t nb, int64_t ld, const int32_t* __restrict__ link,
                                                     const V* __restrict__ src, V* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nb) return;
  const int64_t k = link[t];
#pragma unroll
  for (int q = 0; q < 9; ++q) dst[q * ld + t] = src[q * ld + k];
}

// the factor in the solve order, one thread per stored value so that the writes are contiguous: the
// 6 packed inverse entries of L_ii per position i (vertex colVerts[i]), then the 9 values of every
// lower block in the solve order, then those of the upper lists (the lower blocks (n, i) of the
// neighbours, in each position's upper-list order)
__global__ void __launch_bounds__(256)
    k_factor_solve_order(int nV, int64_t nl, const int32_t* __restrict__ colVerts, const float* __restrict__ dinv,
                         const float* __restrict__ lval, const int32_t* __restrict__ nlsrc,
                         const int32_t* __restrict__ nulink, float* __restrict__ dinvs, float* __restrict__ lv,
                         float* __restrict__ uv) {
  int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < 6ll * nV) {
    const int i = (int)(g / 6), q = (int)(g - 6ll * i);
    dinvs[g] = dinv[6ll * colVerts[i] + q];
    return;
  }
  g -= 6ll * nV;
  if (g < 9 * nl) {
    const int64_t t = g / 9;
    lv[g] = lval[9ll * nlsrc[t] + (g - 9 * t)];
    return;
  }
  g -= 9 * nl;
  if (g < 9 * nl) {
    const int64_t t = g / 9;
    uv[g] = lval[9ll * nlsrc[nulink[t]] + (g - 9 * t)];
  }
}

// rows of a vector between vertex order and the solve order (to = 1: vertex -> position)
__global__ void __launch_bounds__(256) k_permute_rows(int nV, const int32_t* __restrict__ colVerts, int to,
                                                      const double* __restrict__ a, double* __restrict__ b) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3ll * nV) return;
  const int i = (int)(t / 3), l = (int)(t % 3);
  const int64_t v = 3ll * colVerts[i] + l;
  if (to)
    b[t] = a[v];
  else
    b[v] = a[t];
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

// the masked diagonal block of v: from hdiag (FP64, 9 per vertex) when given, else from the pattern
template <typename V>
__device__ __forceinline__ void masked_diag(const V* __restrict__ vals, const int32_t* __restrict__ bptr, int v,
                                            int slot, const double* __restrict__ hdiag,
                                            const uint8_t* __restrict__ fixed, double* A) {
  if (!hdiag) {
    masked_block(vals, bptr, v, slot, fixed, v, A);
    return;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = hdiag[9ll * v + i];
  if (fixed) {
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (fixed[3ll * v + l] || fixed[3ll * v + k]) A[3 * l + k] = (l == k) ? 1.0 : 0.0;
  }
}

// the masked Hessian in the solve order: for position i (vertex v = colVerts[i]) its diagonal block
// (symmetrised) and its lower blocks, one thread per position
template <typename V>
__global__ void __launch_bounds__(256)
    k_split_hessian(int nV, const int32_t* __restrict__ colVerts, const int32_t* __restrict__ bptr,
                    const V* __restrict__ vals, const uint8_t* __restrict__ fixed, const int32_t* __restrict__ dslot,
                    const int32_t* __restrict__ nlptr, const int32_t* __restrict__ nlsrc,
                    const int32_t* __restrict__ lslot, const int32_t* __restrict__ lcol,
                    const double* __restrict__ hdiag, int64_t ld, V* __restrict__ hl, double* __restrict__ hd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const int v = colVerts[i];
  double A[9];
  masked_diag(vals, bptr, v, dslot[v], hdiag, fixed, A);
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) hd[9ll * i + 3 * l + k] = 0.5 * (A[3 * l + k] + A[3 * k + l]);
  for (int t = nlptr[i]; t < nlptr[i + 1]; ++t) {
    const int k = nlsrc[t];
    masked_block(vals, bptr, v, lslot[k], fixed, lcol[k], A);
#pragma unroll
    for (int e = 0; e < 9; ++e) hl[e * ld + t] = (V)A[e];
  }
}

template <typename V>
__global__ void __launch_bounds__(256)
    k_bic0_color(const int32_t* __restrict__ verts, int nverts, const int32_t* __restrict__ bptr,
                 const V* __restrict__ vals, const uint8_t* __restrict__ fixed, const int32_t* __restrict__ lptr,
                 const int32_t* __restrict__ lcol, const int32_t* __restrict__ lslot,
                 const int32_t* __restrict__ dslot, const double* __restrict__ hdiag, double shift,
                 float* __restrict__ lval, float* __restrict__ dinv, int32_t* __restrict__ fail) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nverts) return;
  const int v = verts[i];
  const int k0 = lptr[v], k1 = lptr[v + 1];
  double D[9];
  masked_diag(vals, bptr, v, dslot[v], hdiag, fixed, D);
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
  // DSatur colouring (Brélaz 1979): repeatedly colour the uncoloured vertex whose neighbours already
  // use the most distinct colours (ties: lowest index) with the smallest free colour.  On the Kuhn
  // beams this finds 4 balanced colours where greedy colouring in vertex order needs 8; every colour is
  // one kernel launch per sweep.  (Greedy colouring in vertex order is kept as an option.)
  std::vector<int32_t> color(nV, -1);
  int nc = 0;
  if (m->coloring == 2) {  // greedy in vertex order
    std::vector<char> used;
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
  } else {
    std::vector<uint64_t> sat(nV, 0);  // bit c: some neighbour has colour c (colours >= 64 share bit 63)
    std::priority_queue<std::pair<int, int>> heap;  // (saturation, -vertex), stale entries skipped
    for (int v = 0; v < nV; ++v) heap.emplace(0, -v);
    std::vector<char> used;
    while (!heap.empty()) {
      const auto top = heap.top();
      heap.pop();
      const int v = -top.second;
      if (color[v] >= 0 || top.first != __builtin_popcountll(sat[v])) continue;
      used.assign(nc + 1, 0);
      for (int k = bp[v]; k < bp[v + 1]; ++k) {
        const int c = color[bc[k]];
        if (c >= 0) used[c] = 1;
      }
      int c = 0;
      while (used[c]) ++c;
      color[v] = c;
      nc = std::max(nc, c + 1);
      const uint64_t bit = uint64_t(1) << std::min(c, 63);
      for (int k = bp[v]; k < bp[v + 1]; ++k) {
        const int n = bc[k];
        if (color[n] < 0 && !(sat[n] & bit)) {
          sat[n] |= bit;
          heap.emplace(__builtin_popcountll(sat[n]), -n);
        }
      }
    }
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
  P->eisLanes = m->sweepLanes;
  // the solve order: position i <- vertex colVerts[i]; lower / upper lists re-indexed by position
  std::vector<int32_t> pos(nV), nlptr(nV + 1, 0), nlcol, nlsrc, nuptr(nV + 1, 0), nucol, nulink;
  for (int i = 0; i < nV; ++i) pos[colVerts[i]] = i;
  std::vector<int32_t> newLower(P->nlower);  // vertex-order lower index -> solve-order index
  nlcol.reserve(P->nlower);
  nlsrc.reserve(P->nlower);
  for (int i = 0; i < nV; ++i) {
    const int v = colVerts[i];
    for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
      newLower[k] = (int32_t)nlcol.size();
      nlcol.push_back(pos[lcol[k]]);
      nlsrc.push_back(k);
    }
    nlptr[i + 1] = (int32_t)nlcol.size();
  }
  for (int i = 0; i < nV; ++i) {
    const int v = colVerts[i];
    for (int k = uptr[v]; k < uptr[v + 1]; ++k) {
      if (ulink[k] < 0) return BH_BAD_ARG;  // the pattern is not symmetric
      nucol.push_back(pos[ucol[k]]);
      nulink.push_back(newLower[ulink[k]]);
    }
    nuptr[i + 1] = (int32_t)nucol.size();
  }
  auto up = [](int32_t** d, const std::vector<int32_t>& h) -> int {
    BH_CHECK(hipMalloc(d, sizeof(int32_t) * std::max<size_t>(1, h.size())));
    if (!h.empty()) BH_CHECK(hipMemcpy(*d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
    return BH_OK;
  };
  int st;
  if ((st = up(&P->colVerts, colVerts)) || (st = up(&P->lptr, lptr)) || (st = up(&P->lcol, lcol)) ||
      (st = up(&P->lslot, lslot)) || (st = up(&P->uptr, uptr)) || (st = up(&P->ucol, ucol)) ||
      (st = up(&P->ulink, ulink)) || (st = up(&P->dslot, dslot)) || (st = up(&P->nlptr, nlptr)) ||
      (st = up(&P->nlcol, nlcol)) || (st = up(&P->nlsrc, nlsrc)) || (st = up(&P->nuptr, nuptr)) ||
      (st = up(&P->nucol, nucol)) || (st = up(&P->nulink, nulink)))
    return st;
  const int64_t nl1 = std::max<int64_t>(1, P->nlower);
  BH_CHECK(hipMalloc(&P->lval, sizeof(float) * 9 * nl1));
  BH_CHECK(hipMalloc(&P->dinv, sizeof(float) * 6 * std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->fail, sizeof(int32_t)));
  BH_CHECK(hipMalloc(&P->lvs, sizeof(float) * 9 * nl1));
  BH_CHECK(hipMalloc(&P->uvs, sizeof(float) * 9 * nl1));
  BH_CHECK(hipMalloc(&P->dinvs, sizeof(float) * 6 * std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->hsplit, sizeof(double) * 18 * nl1));
  BH_CHECK(hipMalloc(&P->hdiag, sizeof(double) * 9 * (int64_t)std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->sv, sizeof(double) * 11 * 3 * (int64_t)std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->dt, sizeof(float) * 6 * (int64_t)std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->dtinv, sizeof(float) * 6 * (int64_t)std::max(1, nV)));
  BH_CHECK(hipMalloc(&P->kk, sizeof(double) * 6 * (int64_t)std::max(1, nV)));
  return BH_OK;
}

void precond_free(bh_precond* P) {
  void* ptrs[] = {P->colVerts, P->lptr, P->lcol, P->lslot, P->uptr, P->ucol, P->ulink, P->dslot, P->lval, P->dinv,
                  P->fail, P->nlptr, P->nlcol, P->nlsrc, P->nuptr, P->nucol, P->nulink, P->lvs, P->uvs,
                  P->dinvs, P->hsplit, P->hdiag, P->sv, P->dt, P->dtinv, P->kk};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// solve vectors (3 nV doubles each)
enum { SV_B, SV_X, SV_R, SV_P, SV_S, SV_U, SV_W, SV_Y, SV_T, SV_V, SV_Z };
inline double* svec(const bh_precond* P, int k) { return P->sv + (int64_t)k * 3 * P->mesh->nV; }

// the split Hessian (type V): lower blocks, transposed copy, diagonal blocks
template <typename V>
struct SplitH {
  V *hl, *hu;
  int64_t ld;  // component stride: component q of block k at h[q * ld + k]
};
template <typename V>
SplitH<V> split_view(const bh_precond* P) {
  V* base = reinterpret_cast<V*>(P->hsplit);
  const int64_t ld = std::max<int64_t>(1, P->nlower);
  return SplitH<V>{base, base + 9 * ld, ld};
}

// after every change of lval / dinv: the factor in the solve order (both list orders)
int precond_finalize(bh_precond* P, hipStream_t s) {
  const int nV = P->mesh->nV;
  const int64_t n = 6ll * nV + 18 * P->nlower;
  if (n)
    hipLaunchKernelGGL(k_factor_solve_order, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nV, P->nlower,
                       P->colVerts, P->dinv, P->lval, P->nlsrc, P->nulink, P->dinvs, P->lvs, P->uvs);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// the masked Hessian in the solve order (once per Newton system)
template <typename V>
int precond_split_hessian(const bh_precond* P, const V* vals, const double* hdiag, const uint8_t* fixed,
                          hipStream_t s) {
  const bh_mesh* m = P->mesh;
  const int nV = m->nV;
  SplitH<V> h = split_view<V>(P);
  if (nV)
    hipLaunchKernelGGL(k_split_hessian<V>, dim3((nV + 255) / 256), dim3(256), 0, s, nV, P->colVerts, m->bptr, vals,
                       fixed, P->dslot, P->nlptr, P->nlsrc, P->lslot, P->lcol, hdiag, h.ld, h.hl, P->hdiag);
  if (P->nlower)
    hipLaunchKernelGGL((k_copy_blocks<V>), dim3((P->nlower + 255) / 256), dim3(256), 0, s, P->nlower, h.ld,
                       P->nulink, (const V*)h.hl, h.hu);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// blocks of one colour's sweep kernel (one pass per block)
inline int sweep_grid(int n) { return std::max(1, (n + kSweepVpb - 1) / kSweepVpb); }

// blocks of one whole sweep (= dot-product partials it leaves)
int64_t precond_sweep_blocks(const bh_precond* P) {
  int64_t t = 0;
  for (int c = 0; c < P->ncolors; ++c)
    if (P->h_colPtr[c + 1] > P->h_colPtr[c]) t += sweep_grid(P->h_colPtr[c + 1] - P->h_colPtr[c]);
  return t;
}

// z = (L L^T)^{-1} r in the solve order.  With cg set, the forward sweep first applies the CG vector
// update of each colour's rows and the sweeps leave r.r (forward) and r.z (backward) partials, colour
// after colour, in cg->part_rr and part_rz; nparts receives how many each wrote.
int precond_apply_solve(const bh_precond* P, double* r, double* z, hipStream_t s, const CgVecs* cg = nullptr,
                        int update = 0, double* part_rz = nullptr, int* nparts = nullptr) {
  const int nc = P->ncolors;
  double* y = svec(P, SV_Y);
  CgVecs c0{};
  int off = 0;
  for (int c = 0; c < nc; ++c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    const int g = sweep_grid(n);
    CgVecs cc = cg ? *cg : c0;
    if (cg && cg->part_rr) cc.part_rr = cg->part_rr + off;
    hipLaunchKernelGGL(k_sweep_fwd, dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nlptr, P->nlcol, P->lvs,
                       P->dinvs, r, y, cc, update);
    off += g;
  }
  if (nparts) nparts[0] = off;
  off = 0;
  for (int c = nc - 1; c >= 0; --c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    const int g = sweep_grid(n);
    hipLaunchKernelGGL(k_sweep_bwd, dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nuptr, P->nucol, P->uvs,
                       P->dinvs, (const double*)y, z, part_rz ? (const double*)r : nullptr,
                       part_rz ? part_rz + off : nullptr, cg ? cg->st : nullptr);
    off += g;
  }
  if (nparts) nparts[1] = off;
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// the block SSOR diagonal Dt = scale * D of the split Hessian (scale = 1 / w; 1: symmetric Gauss-Seidel)
inline int precond_sgs(bh_precond* P, double scale, hipStream_t s) {
  const int nV = P->mesh->nV;
  if (nV)
    hipLaunchKernelGGL(k_sgs_diag, dim3((nV + 255) / 256), dim3(256), 0, s, nV, (const double*)P->hdiag, scale, P->dt,
                       P->dtinv, P->kk);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// Eisenstat sweeps over all colours (tb / db: output and gather buffer; double for mode 0, v.t / v.w
// otherwise); nparts (optional) receives the partial count
// fpart non-null (modes >= 1, tb = v.t): the last launch (the solve order's first colour) also takes
// the forward sweep's step on that colour, writing its dot-product partials to fpart; eis_fwd then
// starts at the second colour (first_done)
template <typename V, typename TT>
void eis_bwd(const bh_precond* P, int mode, const EisVecs& v, const double* src, TT* tb, double* part,
             const PcgState* st, hipStream_t s, int* nparts = nullptr, double* fpart = nullptr) {
  const SplitH<V> h = split_view<V>(P);
  const int lanes = P->eisLanes, vpb = 256 / lanes;
  int off = 0;
  for (int c = P->ncolors - 1; c >= 0; --c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    const int g = (n + vpb - 1) / vpb;
    const bool fuse = fpart && c == 0;
    const double* kk = fuse ? (const double*)P->kk : nullptr;
    double* fp = fuse ? fpart : nullptr;
    if (lanes == 4)
      hipLaunchKernelGGL((k_eis_bwd<V, 4, TT>), dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nuptr, P->nucol,
                         (const V*)h.hu, h.ld, (const float*)P->dt, (const float*)P->dtinv, mode, v, src, tb,
                         part ? part + off : nullptr, st, kk, fp);
    else
      hipLaunchKernelGGL((k_eis_bwd<V, 8, TT>), dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nuptr, P->nucol,
                         (const V*)h.hu, h.ld, (const float*)P->dt, (const float*)P->dtinv, mode, v, src, tb,
                         part ? part + off : nullptr, st, kk, fp);
    off += g;
  }
  if (nparts) *nparts = off;
}
template <typename V, typename TT>
void eis_fwd(const bh_precond* P, int mode, const EisVecs& v, const double* src, TT* db, double* part,
             const PcgState* st, hipStream_t s, int* nparts = nullptr, bool first_done = false) {
  const SplitH<V> h = split_view<V>(P);
  const int lanes = P->eisLanes, vpb = 256 / lanes;
  int off = 0;
  for (int c = 0; c < P->ncolors; ++c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    const int g = (n + vpb - 1) / vpb;
    if (c == 0 && first_done) {  // taken by the backward sweep's last launch
      off += g;
      continue;
    }
    if (lanes == 4)
      hipLaunchKernelGGL((k_eis_fwd<V, 4, TT>), dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nlptr, P->nlcol,
                         (const V*)h.hl, h.ld, (const float*)P->dtinv, (const double*)P->kk, mode, v, src, db,
                         part ? part + off : nullptr, st);
    else
      hipLaunchKernelGGL((k_eis_fwd<V, 8, TT>), dim3(g), dim3(256), 0, s, P->h_colPtr[c], n, P->nlptr, P->nlcol,
                         (const V*)h.hl, h.ld, (const float*)P->dtinv, (const double*)P->kk, mode, v, src, db,
                         part ? part + off : nullptr, st);
    off += g;
  }
  if (nparts) *nparts = off;
}

// z = (L L^T)^{-1} r with r, z in vertex order
int precond_apply(const bh_precond* P, const double* r, double* z, hipStream_t s) {
  const int nV = P->mesh->nV;
  if (!nV) return BH_OK;
  const int g = (int)((3ll * nV + 255) / 256);
  double* rs = svec(P, SV_R);
  double* zs = svec(P, SV_U);
  hipLaunchKernelGGL(k_permute_rows, dim3(g), dim3(256), 0, s, nV, P->colVerts, 1, r, rs);
  int st = precond_apply_solve(P, rs, zs, s);
  if (st) return st;
  hipLaunchKernelGGL(k_permute_rows, dim3(g), dim3(256), 0, s, nV, P->colVerts, 0, (const double*)zs, z);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

int precond_finalize(bh_precond* P, hipStream_t s);

template <typename V>
int precond_factor_bic0(bh_precond* P, const V* vals, const double* hdiag, const uint8_t* fixed, double shift,
                        int* failures, hipStream_t s) {
  bh_mesh* m = P->mesh;
  BH_CHECK(hipMemsetAsync(P->fail, 0, sizeof(int32_t), s));
  for (int c = 0; c < P->ncolors; ++c) {
    const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
    if (!n) continue;
    hipLaunchKernelGGL(k_bic0_color<V>, dim3((n + 255) / 256), dim3(256), 0, s, P->colVerts + P->h_colPtr[c], n,
                       m->bptr, vals, fixed, P->lptr, P->lcol, P->lslot, P->dslot, hdiag, shift, P->lval, P->dinv,
                       P->fail);
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
  return precond_factor_bic0(P, d_vals, (const double*)nullptr, d_fixed_dof, shift, failures, as_stream(s));
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
  return precond_apply(P, d_r, d_z, as_stream(s));
}

}  // extern "C"
