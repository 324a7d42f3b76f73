// Learned preconditioner: a mesh-graph network that writes the block lower-triangular factor of
// bh_precond (included at the end of bartels_hip.hip, after precond_impl.h; same translation unit).
// No counterpart in the reference (its Newton direction is a direct sparse solve).
//
// Graph: the vertices of the mesh and the lower pairs k = (v, n) of the colour order (colour(n) <
// colour(v)), i.e. exactly the off-diagonal blocks of the factor L.
// Inputs, from the masked Hessian H (fixed DOFs -> identity rows / columns):
//   s_v = sqrt(tr(H_vv) / 3),  Hn = S^{-1} H S^{-1}  (S = diag(s_v I3)),  C_v = chol(Hn_vv)
//   node features (8): Hn_vv[00, 11, 22, 10, 20, 21], fixed DOFs of v / 3, (pattern degree - 1) / 16
//   pair features (10): Hn_vn row-major (9), log(s_v / s_n)
// Network (width 32, every MLP = Linear -> ReLU -> Linear, weights row-major [in][out]):
//   h = MLP_node_enc(x_v),  e = MLP_pair_enc(f_k)
//   3 x { e_k += MLP_edge([e_k, h_v, h_n]);  m_v = sum of e_k over the pairs of v (lower list, then
//         upper list);  h_v += MLP_node([h_v, m_v]) }
//   d_v = h_v W_dv + b_dv (6),  O_k = e_k W_do + b_do (9)       (decoders start at zero)
// Factor (normalised): Lhat_vv = C_v T_v with T_v lower triangular, T_ii = exp(d), strict lower = d;
//   Lhat_vn = O_k.  At zero decoders Lhat Lhat^T = blockdiag(Hn_vv): block Jacobi.  The factor of H
//   is L = S Lhat (L_vv = s_v Lhat_vv, L_vn = s_v Lhat_vn).  The tied factor form (form 2) adds
//   Hn_vn Lhat_nn^{-T} to O_k: at zero decoders L L^T = (D + L_H) D^-1 (D + L_H)^T, block symmetric
//   Gauss-Seidel as a factor (L_H the strictly lower blocks of H in the colour order, D its diagonal).
// Loss of a system H x = b:  with xh = S x, bh = S^{-1} b,  |Lhat Lhat^T xh - bh|^2 / |bh|^2.
// Training uses probe systems xh = w, bh = Hn w (w the normalised right-hand side of the plain
// solver's Newton system and standard normal vectors): a Hutchinson estimate of the relative
// Frobenius error |Lhat Lhat^T - Hn|_F^2 / |Hn|_F^2, whose scale does not depend on the condition
// number (the loss on the solver's own solution x grows like cond(H)^2 and does not train).
// Everything is FP32; every reduction has a fixed order (results are deterministic).

struct bh_mgn {
  int64_t nparam = 0;
  float* w = nullptr;   // parameters
  float* g = nullptr;   // gradient accumulator
  float* gsave = nullptr;  // the accumulator before the latest system's backward (restored if that fails)
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  int64_t step = 0;
  // per-graph workspace
  int wsV = -1;
  int64_t wsP = -1;
  float* ws = nullptr;
  size_t wsBytes = 0;
  double* dscal = nullptr;  // [loss numerator, |bh|^2]
  float* part = nullptr;    // weight-gradient partials
  int32_t* pv = nullptr;    // pair -> higher vertex (owner row)
  int64_t pvSize = 0;
  double* probe = nullptr;  // training probes: x_j, b_j
  size_t probeBytes = 0;
  int form = 0;             // 0: general factor; 1: Eisenstat (symmetric Gauss-Seidel) form, O tied to H;
                            // 2: tied factor, O = Hn_vn Lhat_nn^{-T} + the pair decoder's output
  int probeSolution = 0;    // 1: the vector handed to a training step is a solution x of H x = b (probe 0)
  int energyLoss = 0;       // 1 (Eisenstat form): the energy-ratio loss instead of the residual loss
  int teacher = 0;          // 1 (factor forms): the IC(0) teacher loss (training steps factor the Hessian)
  int kaporin = 0;          // 1 (factor forms): the Kaporin condition number of L L^T against H
  int inferBf16 = 1;        // inference MLPs: 1 BF16 on the matrix cores, 0 the FP32 training forward
};

namespace {

constexpr int kW = 32;           // latent width
constexpr int kLayers = 3;       // message-passing layers
constexpr int kNodeIn = 8, kPairIn = 10;
constexpr int kOuterBlocks = 512;  // blocks of the weight-gradient reductions

// flat parameter layout
struct MlpOff {
  int64_t W1, b1, W2, b2;
};
struct MgnLayout {
  MlpOff nenc, penc, edge[kLayers], node[kLayers];
  int64_t Wdv, bdv, Wdo, bdo, total;
};
inline MlpOff mlp_off(int64_t& o, int in) {
  MlpOff m;
  m.W1 = o; o += (int64_t)in * kW;
  m.b1 = o; o += kW;
  m.W2 = o; o += kW * kW;
  m.b2 = o; o += kW;
  return m;
}
inline MgnLayout mgn_layout() {
  MgnLayout L;
  int64_t o = 0;
  L.nenc = mlp_off(o, kNodeIn);
  L.penc = mlp_off(o, kPairIn);
  for (int l = 0; l < kLayers; ++l) {
    L.edge[l] = mlp_off(o, 3 * kW);
    L.node[l] = mlp_off(o, 2 * kW);
  }
  L.Wdv = o; o += kW * 6;
  L.bdv = o; o += 6;
  L.Wdo = o; o += kW * 9;
  L.bdo = o; o += 9;
  L.total = o;
  return L;
}

// ---- workspace ------------------------------------------------------------------------------------
struct MgnWs {
  float *s, *C, *xn, *fp;             // scales, chol(Hn_vv), node / pair features
  float *h[kLayers + 1], *e[kLayers + 1];
  float *hidN[kLayers + 1], *hidE[kLayers + 1];  // MLP hidden activations (index 0: encoders)
  float* m[kLayers];                  // aggregated messages
  float *d, *O, *y, *r;               // decoder outputs, L^T xh, residual
  float *gh, *ge, *gm, *gin, *ghid, *gd, *gO, *gy, *gz, *gh2, *ge2;
};

// carves the workspace from base (or only sizes it when base is null); returns the float count
size_t mgn_ws_carve(float* base, int nV, int64_t nP, MgnWs& w) {
  size_t used = 0;
  auto take = [&](size_t n) {
    float* r = base ? base + used : nullptr;
    used += (n + 15) & ~size_t(15);
    return r;
  };
  const size_t V = nV, P = nP;
  w.s = take(V);
  w.C = take(6 * V);
  w.xn = take(kNodeIn * V);
  w.fp = take(kPairIn * P);
  for (int l = 0; l <= kLayers; ++l) {
    w.h[l] = take(kW * V);
    w.e[l] = take(kW * P);
    w.hidN[l] = take(kW * V);
    w.hidE[l] = take(kW * P);
  }
  for (int l = 0; l < kLayers; ++l) w.m[l] = take(kW * V);
  w.d = take(6 * V);
  w.O = take(9 * P);
  w.y = take(3 * V);
  w.r = take(3 * V);
  w.gh = take(kW * V);
  w.gh2 = take(kW * V);
  w.gm = take(kW * V);
  w.ge = take(kW * P);
  w.ge2 = take(kW * P);
  w.gin = take(kW * std::max(3 * P, 2 * V));
  w.ghid = take(kW * std::max(V, P));
  w.gd = take(6 * V);
  w.gO = take(9 * P);
  w.gy = take(3 * V);
  w.gz = take(3 * V);
  return used;
}

// ---- features -------------------------------------------------------------------------------------
// masked 3x3 block (v, n) of H, row-major, from a block index of the mesh pattern
template <typename V>
__device__ __forceinline__ void mgn_block(const V* __restrict__ vals, const int32_t* __restrict__ bptr, int v, int slot,
                                          const uint8_t* __restrict__ fixed, int n, float* A) {
  const int b0 = bptr[v], nb = bptr[v + 1] - b0;
  const V* base = vals + 9ll * b0 + 3ll * (slot - b0);
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float a = (float)base[3ll * l * nb + k];
      if (fixed && (fixed[3ll * v + l] || fixed[3ll * n + k])) a = (v == n && l == k) ? 1.f : 0.f;
      A[3 * l + k] = a;
    }
}

// per vertex: s_v, C_v, node features; per lower pair of v: pair features
template <typename V>
__global__ void __launch_bounds__(256)
    k_mgn_prep(int nV, const int32_t* __restrict__ bptr, const V* __restrict__ vals, const uint8_t* __restrict__ fixed,
               const int32_t* __restrict__ dslot, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
               const int32_t* __restrict__ lslot, float* __restrict__ s, float* __restrict__ C,
               float* __restrict__ xn) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  float D[9];
  mgn_block(vals, bptr, v, dslot[v], fixed, v, D);
  const float sv = sqrtf(fmaxf((D[0] + D[4] + D[8]) * (1.f / 3.f), 1e-30f));
  const float i2 = 1.f / (sv * sv);
#pragma unroll
  for (int i = 0; i < 9; ++i) D[i] *= i2;
  // Cholesky of the normalised diagonal block (pivots floored: a tiny positive value keeps L_vv
  // invertible if a block is only semidefinite)
  const float c00 = sqrtf(fmaxf(D[0], 1e-12f));
  const float c10 = D[3] / c00, c20 = D[6] / c00;
  const float c11 = sqrtf(fmaxf(D[4] - c10 * c10, 1e-12f));
  const float c21 = (D[7] - c20 * c10) / c11;
  const float c22 = sqrtf(fmaxf(D[8] - c20 * c20 - c21 * c21, 1e-12f));
  s[v] = sv;
  float* Cv = C + 6ll * v;
  Cv[0] = c00; Cv[1] = c10; Cv[2] = c11; Cv[3] = c20; Cv[4] = c21; Cv[5] = c22;
  int nf = 0;
  if (fixed) nf = (int)(fixed[3ll * v] != 0) + (int)(fixed[3ll * v + 1] != 0) + (int)(fixed[3ll * v + 2] != 0);
  float* x = xn + (int64_t)kNodeIn * v;
  x[0] = D[0]; x[1] = D[4]; x[2] = D[8]; x[3] = D[3]; x[4] = D[6]; x[5] = D[7];
  x[6] = nf * (1.f / 3.f);
  x[7] = (float)(bptr[v + 1] - bptr[v] - 1) * (1.f / 16.f);
}

// the rows [k0, k0 + cnt) of a row-major array of W floats per row, written from LDS (row t of the block
// at stage[W t]) as one contiguous run
template <int W>
__device__ __forceinline__ void store_rows(const float* __restrict__ stage, float* __restrict__ dst, int64_t k0,
                                           int cnt) {
  float* out = dst + W * k0;
  for (int i = threadIdx.x; i < W * cnt; i += blockDim.x) out[i] = stage[i];
}

// pair features, one thread per lower pair k = (v, n) (v = pv[k]); the 10 values of the workgroup's 256
// pairs leave through LDS as one contiguous run
template <typename V>
__global__ void __launch_bounds__(256)
    k_mgn_pairfeat(int64_t nP, const int32_t* __restrict__ pv, const int32_t* __restrict__ bptr,
                   const V* __restrict__ vals, const uint8_t* __restrict__ fixed, const int32_t* __restrict__ lcol,
                   const int32_t* __restrict__ lslot, const float* __restrict__ s, float* __restrict__ fp) {
  __shared__ float stage[256 * kPairIn];
  const int64_t k0 = blockIdx.x * (int64_t)blockDim.x;
  const int64_t k = k0 + threadIdx.x;
  if (k < nP) {
    const int v = pv[k], n = lcol[k];
    float A[9];
    mgn_block(vals, bptr, v, lslot[k], fixed, n, A);
    const float sv = s[v], sn = s[n];
    const float inv = 1.f / (sv * sn);
    float* f = stage + kPairIn * threadIdx.x;
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = A[i] * inv;
    f[9] = logf(sv / sn);
  }
  __syncthreads();
  store_rows<kPairIn>(stage, fp, k0, (int)min<int64_t>(blockDim.x, nP - k0));
}

__global__ void __launch_bounds__(256) k_mgn_pair_owner(int nV, const int32_t* __restrict__ lptr,
                                                        int32_t* __restrict__ pv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  for (int k = lptr[v]; k < lptr[v + 1]; ++k) pv[k] = v;
}

// ---- MLP rows -------------------------------------------------------------------------------------
// An MLP input is either one segment of IN <= 32 features or IN / 32 segments of 32: segment i of
// row r is p_i[(idx_i ? idx_i[r] : r) * width + j].  Widths are compile-time, so the gathered row
// stays in registers.
struct Seg3 {
  const float* p[3];
  const int32_t* idx[3];
};

template <int IN>
__device__ __forceinline__ void gather_in(const Seg3& S, int64_t r, float* x) {
  if constexpr (IN <= kW) {
    const int64_t row = S.idx[0] ? (int64_t)S.idx[0][r] : r;
    const float* src = S.p[0] + row * IN;
#pragma unroll
    for (int j = 0; j < IN; ++j) x[j] = src[j];
  } else {
#pragma unroll
    for (int sg = 0; sg < IN / kW; ++sg) {
      const int64_t row = S.idx[sg] ? (int64_t)S.idx[sg][r] : r;
      const float* src = S.p[sg] + row * kW;
#pragma unroll
      for (int j = 0; j < kW; ++j) x[sg * kW + j] = src[j];
    }
  }
}

// out = [resid +] relu(x W1 + b1) W2 + b2; hid = relu(...) kept for the backward.  One row per
// lane; the weights sit in LDS and are read as float4 rows (the same address across the wave, a
// broadcast); the input is streamed segment by segment from global memory.
template <int IN>
__global__ void __launch_bounds__(256)
    k_mlp_fwd(int64_t rows, Seg3 S, const float* __restrict__ W1, const float* __restrict__ b1,
              const float* __restrict__ W2, const float* __restrict__ b2, const float* __restrict__ resid,
              float* __restrict__ hid, float* __restrict__ out) {
  __shared__ __align__(16) float sW1[IN * kW];
  __shared__ __align__(16) float sW2[kW * kW];
  __shared__ float sb1[kW], sb2[kW];
  for (int i = threadIdx.x; i < IN * kW; i += blockDim.x) sW1[i] = W1[i];
  for (int i = threadIdx.x; i < kW * kW; i += blockDim.x) sW2[i] = W2[i];
  if (threadIdx.x < kW) {
    sb1[threadIdx.x] = b1[threadIdx.x];
    sb2[threadIdx.x] = b2[threadIdx.x];
  }
  __syncthreads();
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float hh[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) hh[j] = sb1[j];
  constexpr int NSEG = IN <= kW ? 1 : IN / kW;
  constexpr int SW = IN <= kW ? IN : kW;
#pragma unroll
  for (int sg = 0; sg < NSEG; ++sg) {
    const int64_t row = S.idx[sg] ? (int64_t)S.idx[sg][r] : r;
    const float* src = S.p[sg] + row * SW;
#pragma unroll 2
    for (int i = 0; i < SW; ++i) {
      const float xi = src[i];
      const float4* wr = reinterpret_cast<const float4*>(sW1 + (sg * SW + i) * kW);
#pragma unroll
      for (int q = 0; q < kW / 4; ++q) {
        const float4 wv = wr[q];
        hh[4 * q] += xi * wv.x;
        hh[4 * q + 1] += xi * wv.y;
        hh[4 * q + 2] += xi * wv.z;
        hh[4 * q + 3] += xi * wv.w;
      }
    }
  }
  float* hp = hid + r * kW;
#pragma unroll
  for (int j = 0; j < kW; ++j) {
    hh[j] = fmaxf(hh[j], 0.f);
    hp[j] = hh[j];
  }
  float o[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) o[j] = sb2[j] + (resid ? resid[r * kW + j] : 0.f);
#pragma unroll
  for (int i = 0; i < kW; ++i) {
    const float hi = hh[i];
    const float4* wr = reinterpret_cast<const float4*>(sW2 + i * kW);
#pragma unroll
    for (int q = 0; q < kW / 4; ++q) {
      const float4 wv = wr[q];
      o[4 * q] += hi * wv.x;
      o[4 * q + 1] += hi * wv.y;
      o[4 * q + 2] += hi * wv.z;
      o[4 * q + 3] += hi * wv.w;
    }
  }
  float* op = out + r * kW;
#pragma unroll
  for (int j = 0; j < kW; ++j) op[j] = o[j];
}

// ---- inference MLP on the matrix cores ---------------------------------------------------------------
// out = [resid +] relu(x W1 + b1) W2 + b2 with x, W1, W2 and the hidden activations rounded to BF16
// (round to nearest even) and FP32 accumulation; the residual stream stays FP32.  A wave computes
// 32-row tiles transposed, H^T = W1^T X^T by v_mfma_f32_32x32x16_bf16 (A = W1^T, 32 outputs x 16
// inputs per chunk; B = X^T, 16 inputs x 32 rows): lane l ends with row l % 32 of the tile and, in
// accumulator register r, hidden feature 8 (r / 4) + 4 (l / 32) + r % 4.  Registers 0-7 and 8-15 are
// then the B operands of O^T = W2^T H^T as they stand, with W2's rows taken in the same order (a
// reduction index may be permuted freely when A and B agree on it).  The output lands in the same
// layout: four float4 pieces per lane.  Weights and biases are held in registers for the whole
// grid-stride loop over tiles.
typedef __bf16 bh_bf16x8 __attribute__((ext_vector_type(8)));
typedef float bh_f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaBlocks = 2048;  // grid cap of the tile loop: 8 workgroups of 4 waves per CU

// inputs 16 c + 8 h .. 16 c + 8 h + 7 of row r (zero past IN and for rows past the end)
template <int IN>
__device__ __forceinline__ bh_bf16x8 mlp_in_chunk(const Seg3& S, int64_t r, bool live, int c, int h) {
  bh_bf16x8 b;
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = (__bf16)0.f;
  if (!live) return b;
  if constexpr (IN <= 16) {
    const int64_t row = S.idx[0] ? (int64_t)S.idx[0][r] : r;
    const float* src = S.p[0] + row * IN;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (8 * h + i < IN) b[i] = (__bf16)src[8 * h + i];
  } else {
    const int sg = c >> 1;
    const int64_t row = S.idx[sg] ? (int64_t)S.idx[sg][r] : r;
    const float4* src = reinterpret_cast<const float4*>(S.p[sg] + row * kW + 16 * (c & 1) + 8 * h);
    const float4 u = src[0], v = src[1];
    b[0] = (__bf16)u.x; b[1] = (__bf16)u.y; b[2] = (__bf16)u.z; b[3] = (__bf16)u.w;
    b[4] = (__bf16)v.x; b[5] = (__bf16)v.y; b[6] = (__bf16)v.z; b[7] = (__bf16)v.w;
  }
  return b;
}

template <int IN>
__global__ void __launch_bounds__(256)
    k_mlp_mfma(int64_t rows, Seg3 S, const float* __restrict__ W1, const float* __restrict__ b1,
               const float* __restrict__ W2, const float* __restrict__ b2, const float* __restrict__ resid,
               float* __restrict__ out) {
  static_assert(IN <= 16 || IN % kW == 0, "inputs: one segment of at most 16 or segments of 32");
  constexpr int NC = (IN + 15) / 16;
  const int lane = threadIdx.x & 63, m = lane & 31, h = lane >> 5;
  bh_bf16x8 A1[NC], A2[2];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = 16 * c + 8 * h + i;
      A1[c][i] = (__bf16)(k < IN ? W1[k * kW + m] : 0.f);
    }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int i = 0; i < 8; ++i) A2[c][i] = (__bf16)W2[(16 * c + 8 * (i >> 2) + 4 * h + (i & 3)) * kW + m];
  float c1[16], c2[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int f = 8 * (r >> 2) + 4 * h + (r & 3);
    c1[r] = b1[f];
    c2[r] = b2[f];
  }
  const int64_t ntile = (rows + 31) / 32;
  for (int64_t t = blockIdx.x * 4ll + (threadIdx.x >> 6); t < ntile; t += 4ll * gridDim.x) {
    const int64_t r = 32 * t + m;
    const bool live = r < rows;
    bh_bf16x8 X[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) X[c] = mlp_in_chunk<IN>(S, r, live, c, h);
    bh_f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = c1[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[c], X[c], acc, 0, 0, 0);
    bh_bf16x8 H0, H1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      H0[i] = (__bf16)fmaxf(acc[i], 0.f);
      H1[i] = (__bf16)fmaxf(acc[8 + i], 0.f);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = c2[i];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[0], H0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[1], H1, acc, 0, 0, 0);
    if (live) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t off = r * kW + 8 * g + 4 * h;
        float4 o = make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
        if (resid) {
          const float4 x = *reinterpret_cast<const float4*>(resid + off);
          o.x += x.x; o.y += x.y; o.z += x.z; o.w += x.w;
        }
        *reinterpret_cast<float4*>(out + off) = o;
      }
    }
  }
}

// given gout: ghid = (gout W2^T) * [hid > 0]; gin = ghid W1^T (IN values per row)
template <int IN>
__global__ void __launch_bounds__(256)
    k_mlp_bwd(int64_t rows, const float* __restrict__ W1, const float* __restrict__ W2, const float* __restrict__ hid,
              const float* __restrict__ gout, float* __restrict__ ghid, float* __restrict__ gin) {
  __shared__ float sW1[IN * kW], sW2[kW * kW];
  for (int i = threadIdx.x; i < IN * kW; i += blockDim.x) sW1[i] = W1[i];
  for (int i = threadIdx.x; i < kW * kW; i += blockDim.x) sW2[i] = W2[i];
  __syncthreads();
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float go[kW], gh[kW];
#pragma unroll
  for (int j = 0; j < kW; ++j) go[j] = gout[r * kW + j];
#pragma unroll
  for (int i = 0; i < kW; ++i) {
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < kW; ++j) a += go[j] * sW2[i * kW + j];
    gh[i] = (hid[r * kW + i] > 0.f) ? a : 0.f;
  }
#pragma unroll
  for (int i = 0; i < kW; ++i) ghid[r * kW + i] = gh[i];
  if (!gin) return;
  for (int i = 0; i < IN; ++i) {
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < kW; ++j) a += gh[j] * sW1[i * kW + j];
    gin[r * IN + i] = a;
  }
}

// Deterministic weight-gradient partials: block b sums, over its fixed share of the rows, the outer
// products x_r^T g_r (XW x GW) and g_r (GW) -> part[b][XW * GW + GW].  Rows are staged through LDS
// in chunks of 32; thread t owns entries t, t + 256, ...
template <int XW, int GW>
__global__ void __launch_bounds__(256)
    k_outer(int64_t rows, Seg3 S, const float* __restrict__ g, float* __restrict__ part) {
  constexpr int NE = XW * GW + GW;
  constexpr int PER = (NE + 255) / 256;
  constexpr int CH = 32;
  __shared__ float sx[CH * XW], sg[CH * GW];
  float acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = 0.f;
  const int64_t per = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = blockIdx.x * per, r1 = min(rows, r0 + per);
  for (int64_t c = r0; c < r1; c += CH) {
    const int nr = (int)min<int64_t>(CH, r1 - c);
    __syncthreads();
    for (int t = threadIdx.x; t < nr; t += blockDim.x) {
      float x[XW];
      gather_in<XW>(S, c + t, x);
#pragma unroll
      for (int i = 0; i < XW; ++i) sx[t * XW + i] = x[i];
    }
    for (int t = threadIdx.x; t < nr * GW; t += blockDim.x) sg[t] = g[(c + t / GW) * GW + t % GW];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int ent = threadIdx.x + 256 * q;
      if (ent < NE) {
        float a = acc[q];
        if (ent < XW * GW) {
          const int i = ent / GW, j = ent % GW;
          for (int t = 0; t < nr; ++t) a += sx[t * XW + i] * sg[t * GW + j];
        } else {
          const int j = ent - XW * GW;
          for (int t = 0; t < nr; ++t) a += sg[t * GW + j];
        }
        acc[q] = a;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int ent = threadIdx.x + 256 * q;
    if (ent < NE) part[(int64_t)blockIdx.x * NE + ent] = acc[q];
  }
}

// grad[W] += sum_b part[b][0 .. XW*GW), grad[bias] += sum_b part[b][XW*GW ..)
__global__ void __launch_bounds__(256) k_outer_reduce(int nb, int XW, int GW, const float* __restrict__ part,
                                                      float* __restrict__ gW, float* __restrict__ gb) {
  const int NE = XW * GW + GW;
  const int ent = blockIdx.x * blockDim.x + threadIdx.x;
  if (ent >= NE) return;
  float a = 0.f;
  for (int b = 0; b < nb; ++b) a += part[(int64_t)b * NE + ent];
  if (ent < XW * GW)
    gW[ent] += a;
  else
    gb[ent - XW * GW] += a;
}

// ---- message aggregation -------------------------------------------------------------------------
// m_v = sum over v's lower pairs (list order) then upper pairs (list order) of e_k; 32 lanes per vertex
__global__ void __launch_bounds__(256)
    k_mgn_aggregate(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ uptr,
                    const int32_t* __restrict__ ulink, const float* __restrict__ e, float* __restrict__ m) {
  // 8 lanes per vertex, 4 features each; every row is added in list order (lower list, then upper),
  // with up to four rows' loads in flight ahead of their sums
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int v = (int)(gid >> 3), j = 4 * (int)(gid & 7);
  if (v >= nV) return;
  const float4* E = reinterpret_cast<const float4*>(e + j);
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
  auto add = [&](const float4& x) {
    a.x += x.x; a.y += x.y; a.z += x.z; a.w += x.w;
  };
  const int l0 = lptr[v], l1 = lptr[v + 1], u0 = uptr[v], u1 = uptr[v + 1];
  const int n = (l1 - l0) + (u1 - u0);
  auto row = [&](int t) -> int64_t { return t < l1 - l0 ? (int64_t)(l0 + t) : (int64_t)ulink[u0 + t - (l1 - l0)]; };
  int t = 0;
  for (; t + 4 <= n; t += 4) {
    const float4 x0 = E[row(t) * (kW / 4)], x1 = E[row(t + 1) * (kW / 4)];
    const float4 x2 = E[row(t + 2) * (kW / 4)], x3 = E[row(t + 3) * (kW / 4)];
    add(x0); add(x1); add(x2); add(x3);
  }
  for (; t < n; ++t) add(E[row(t) * (kW / 4)]);
  *reinterpret_cast<float4*>(m + (int64_t)v * kW + j) = a;
}

// pair gradients from a per-vertex gradient: ge_k (+)= gm_v + gm_n
__global__ void __launch_bounds__(256)
    k_mgn_agg_bwd(int64_t nP, const int32_t* __restrict__ pv, const int32_t* __restrict__ lcol,
                  const float* __restrict__ gm, const float* __restrict__ gin_e, float* __restrict__ ge) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t k = gid / kW;
  const int j = (int)(gid % kW);
  if (k >= nP) return;
  ge[k * kW + j] = gin_e[k * kW + j] + gm[(int64_t)pv[k] * kW + j] + gm[(int64_t)lcol[k] * kW + j];
}

// vertex gradients from the edge-MLP input gradient gin (96 per pair: [e, h_v, h_n]) and the node
// path: gh_v = ghn_v + sum_{lower pairs of v} gin_k[32:64] + sum_{upper pairs} gin_k[64:96]
__global__ void __launch_bounds__(256)
    k_mgn_gather_gh(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ uptr,
                    const int32_t* __restrict__ ulink, const float* __restrict__ gin, const float* __restrict__ ghn,
                    float* __restrict__ gh) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int v = (int)(gid / kW), j = (int)(gid % kW);
  if (v >= nV) return;
  float a = ghn[(int64_t)v * kW + j];
  for (int k = lptr[v]; k < lptr[v + 1]; ++k) a += gin[(int64_t)k * 3 * kW + kW + j];
  for (int t = uptr[v]; t < uptr[v + 1]; ++t) a += gin[(int64_t)ulink[t] * 3 * kW + 2 * kW + j];
  gh[(int64_t)v * kW + j] = a;
}

// pair path of the edge MLP backward: ge_{l-1} = ge_l + gin[0:32]
__global__ void __launch_bounds__(256) k_mgn_ge_prev(int64_t nP, const float* __restrict__ ge,
                                                     const float* __restrict__ gin, float* __restrict__ out) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= nP * kW) return;
  const int64_t k = gid / kW;
  const int j = (int)(gid % kW);
  out[gid] = ge[gid] + gin[k * 3 * kW + j];
}

// node path of the node MLP backward: gh_prev = gh + gin[0:32], gm = gin[32:64]
__global__ void __launch_bounds__(256) k_mgn_gh_split(int nV, const float* __restrict__ gh, const float* __restrict__ gin,
                                                      float* __restrict__ ghp, float* __restrict__ gm) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gid >= (int64_t)nV * kW) return;
  const int64_t v = gid / kW;
  const int j = (int)(gid % kW);
  ghp[gid] = gh[gid] + gin[v * 2 * kW + j];
  gm[gid] = gin[v * 2 * kW + kW + j];
}

// ---- decoders and the factor -------------------------------------------------------------------------
template <int GW>
__global__ void __launch_bounds__(256) k_linear(int64_t rows, const float* __restrict__ x, const float* __restrict__ W,
                                                const float* __restrict__ b, float* __restrict__ y) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float o[GW];
#pragma unroll
  for (int j = 0; j < GW; ++j) o[j] = b[j];
  for (int i = 0; i < kW; ++i) {
    const float xi = x[r * kW + i];
#pragma unroll
    for (int j = 0; j < GW; ++j) o[j] += xi * W[i * GW + j];
  }
#pragma unroll
  for (int j = 0; j < GW; ++j) y[r * GW + j] = o[j];
}

// gx (+)= g W^T  (rows x 32)
template <int GW>
__global__ void __launch_bounds__(256) k_linear_bwd(int64_t rows, const float* __restrict__ g,
                                                    const float* __restrict__ W, float* __restrict__ gx) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float gg[GW];
#pragma unroll
  for (int j = 0; j < GW; ++j) gg[j] = g[r * GW + j];
  for (int i = 0; i < kW; ++i) {
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < GW; ++j) a += gg[j] * W[i * GW + j];
    gx[r * kW + i] = a;
  }
}

// diagonal of T: exp(d), with d capped at 20 (the derivative is zero past the cap)
__device__ __forceinline__ float texp(float d) { return expf(fminf(d, 20.f)); }
__device__ __forceinline__ float dtexp(float d) { return d < 20.f ? expf(d) : 0.f; }

// Lhat_vv = C_v T(d_v), row-major 3x3 (lower triangular)
__device__ __forceinline__ void mgn_lvv(const float* C, const float* d, float* L) {
  const float t00 = texp(d[0]), t10 = d[1], t11 = texp(d[2]), t20 = d[3], t21 = d[4], t22 = texp(d[5]);
  const float c00 = C[0], c10 = C[1], c11 = C[2], c20 = C[3], c21 = C[4], c22 = C[5];
  L[0] = c00 * t00;
  L[1] = 0.f;
  L[2] = 0.f;
  L[3] = c10 * t00 + c11 * t10;
  L[4] = c11 * t11;
  L[5] = 0.f;
  L[6] = c20 * t00 + c21 * t10 + c22 * t20;
  L[7] = c21 * t11 + c22 * t21;
  L[8] = c22 * t22;
}

// out_n = Lhat_nn^T a_n + sum_{pairs k = (v, n)} O_k^T a_v, with a = sx ? s .* u : u
__global__ void __launch_bounds__(256)
    k_mgn_lt(int nV, const int32_t* __restrict__ uptr, const int32_t* __restrict__ ucol,
             const int32_t* __restrict__ ulink, const float* __restrict__ C, const float* __restrict__ d,
             const float* __restrict__ O, const float* __restrict__ s, const double* __restrict__ u, int scale,
             const float* __restrict__ uf, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nV) return;
  auto val = [&](int v, int l) -> float {
    if (uf) return uf[3ll * v + l];
    const float a = (float)u[3ll * v + l];
    return scale ? a * s[v] : a;
  };
  float L[9];
  mgn_lvv(C + 6ll * n, d + 6ll * n, L);
  const float a0 = val(n, 0), a1 = val(n, 1), a2 = val(n, 2);
  float o0 = L[0] * a0 + L[3] * a1 + L[6] * a2;
  float o1 = L[1] * a0 + L[4] * a1 + L[7] * a2;
  float o2 = L[2] * a0 + L[5] * a1 + L[8] * a2;
  for (int t = uptr[n]; t < uptr[n + 1]; ++t) {
    const float* B = O + 9ll * ulink[t];
    const int v = ucol[t];
    const float b0 = val(v, 0), b1 = val(v, 1), b2 = val(v, 2);
    o0 += B[0] * b0 + B[3] * b1 + B[6] * b2;
    o1 += B[1] * b0 + B[4] * b1 + B[7] * b2;
    o2 += B[2] * b0 + B[5] * b1 + B[8] * b2;
  }
  out[3ll * n] = o0;
  out[3ll * n + 1] = o1;
  out[3ll * n + 2] = o2;
}

// r_v = Lhat_vv y_v + sum_{lower k of v} O_k y_n - b_v / s_v; per-block partials of |r|^2, |bh|^2
__global__ void __launch_bounds__(256)
    k_mgn_resid(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                const float* __restrict__ C, const float* __restrict__ d, const float* __restrict__ O,
                const float* __restrict__ s, const float* __restrict__ y, const double* __restrict__ b,
                float* __restrict__ r, double* __restrict__ part) {
  __shared__ double sh[8];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (v < nV) {
    float L[9];
    mgn_lvv(C + 6ll * v, d + 6ll * v, L);
    const float y0 = y[3ll * v], y1 = y[3ll * v + 1], y2 = y[3ll * v + 2];
    float z0 = L[0] * y0 + L[1] * y1 + L[2] * y2;
    float z1 = L[3] * y0 + L[4] * y1 + L[5] * y2;
    float z2 = L[6] * y0 + L[7] * y1 + L[8] * y2;
    for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
      const float* B = O + 9ll * k;
      const int n = lcol[k];
      const float w0 = y[3ll * n], w1 = y[3ll * n + 1], w2 = y[3ll * n + 2];
      z0 += B[0] * w0 + B[1] * w1 + B[2] * w2;
      z1 += B[3] * w0 + B[4] * w1 + B[5] * w2;
      z2 += B[6] * w0 + B[7] * w1 + B[8] * w2;
    }
    const float is = 1.f / s[v];
    const float bh0 = (float)b[3ll * v] * is, bh1 = (float)b[3ll * v + 1] * is, bh2 = (float)b[3ll * v + 2] * is;
    const float r0 = z0 - bh0, r1 = z1 - bh1, r2 = z2 - bh2;
    r[3ll * v] = r0;
    r[3ll * v + 1] = r1;
    r[3ll * v + 2] = r2;
    rr = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
    bb = (double)bh0 * bh0 + (double)bh1 * bh1 + (double)bh2 * bh2;
  }
  rr = block_reduce_sum(rr, sh);
  __syncthreads();
  bb = block_reduce_sum(bb, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rr;
    part[2 * blockIdx.x + 1] = bb;
  }
}

__global__ void __launch_bounds__(256) k_mgn_loss_sum(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[8];
  double a = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a += part[2 * i];
    c += part[2 * i + 1];
  }
  a = block_reduce_sum(a, sh);
  __syncthreads();
  c = block_reduce_sum(c, sh);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = c;
  }
}

// gz = 2 r / (|bh|^2 ns): the gradient of the mean over ns systems of |r|^2 / |bh|^2
__global__ void __launch_bounds__(256) k_mgn_gz(int64_t n, const float* __restrict__ r, const double* __restrict__ sc,
                                                int ns, float* __restrict__ gz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  gz[i] = (float)(2.0 / (sc[1] * ns)) * r[i];
}

// gO_k = gz_v y_n^T + xh_v gy_n^T ; gd_v from gLhat_vv = lower(gz_v y_v^T + xh_v gy_v^T), gT = C^T gLhat
__global__ void __launch_bounds__(256)
    k_mgn_gfactor(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                  const float* __restrict__ C, const float* __restrict__ d, const float* __restrict__ s,
                  const double* __restrict__ x, const float* __restrict__ y, const float* __restrict__ gz,
                  const float* __restrict__ gy, int accumulate, float* __restrict__ gO, float* __restrict__ gd) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const float sv = s[v];
  const float xh[3] = {(float)x[3ll * v] * sv, (float)x[3ll * v + 1] * sv, (float)x[3ll * v + 2] * sv};
  const float gzv[3] = {gz[3ll * v], gz[3ll * v + 1], gz[3ll * v + 2]};
  for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
    const int n = lcol[k];
    float* o = gO + 9ll * k;
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float a = gzv[l] * y[3ll * n + c] + xh[l] * gy[3ll * n + c];
        o[3 * l + c] = accumulate ? o[3 * l + c] + a : a;
      }
  }
  float G[9];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) G[3 * l + c] = (c <= l) ? gzv[l] * y[3ll * v + c] + xh[l] * gy[3ll * v + c] : 0.f;
  // Lhat = C T: gT = C^T G (lower part kept)
  const float* Cv = C + 6ll * v;
  const float c00 = Cv[0], c10 = Cv[1], c11 = Cv[2], c20 = Cv[3], c21 = Cv[4], c22 = Cv[5];
  const float Cm[9] = {c00, 0.f, 0.f, c10, c11, 0.f, c20, c21, c22};
  float gT[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gT[3 * i + j] = Cm[i] * G[j] + Cm[3 + i] * G[3 + j] + Cm[6 + i] * G[6 + j];
  const float* dv = d + 6ll * v;
  float* g = gd + 6ll * v;
  const float a0 = gT[0] * dtexp(dv[0]), a1 = gT[3], a2 = gT[4] * dtexp(dv[2]), a3 = gT[6], a4 = gT[7],
              a5 = gT[8] * dtexp(dv[5]);
  if (accumulate) {
    g[0] += a0; g[1] += a1; g[2] += a2; g[3] += a3; g[4] += a4; g[5] += a5;
  } else {
    g[0] = a0; g[1] = a1; g[2] = a2; g[3] = a3; g[4] = a4; g[5] = a5;
  }
}

// The Eisenstat (symmetric Gauss-Seidel) form as a factor: L = (Dt + L_H) Lhat^{-T} in the normalised
// space, i.e. the diagonal blocks Lhat_vv = C_v T(d_v) as in the general form and the off-diagonal
// blocks tied to the Hessian, O_k = Hn_vn Lhat_nn^{-T} for the pair k = (v, n), n of lower colour
// (Hn_vn: the masked, normalised block, the first 9 pair features).  Then L L^T = (Dt + L_H) Dt^{-1}
// (Dt + L_H)^T with Dt = Lhat Lhat^T, the preconditioner that bh_newton applies with Eisenstat's trick.
__device__ __forceinline__ void ltri_inv(const float* L, float* I) {  // row-major lower-triangular 3x3
  const float i00 = 1.f / L[0], i11 = 1.f / L[4], i22 = 1.f / L[8];
  const float i10 = -L[3] * i00 * i11;
  const float i21 = -L[7] * i11 * i22;
  const float i20 = -(L[6] * i00 + L[7] * i10) * i22;
  I[0] = i00; I[1] = 0.f; I[2] = 0.f;
  I[3] = i10; I[4] = i11; I[5] = 0.f;
  I[6] = i20; I[7] = i21; I[8] = i22;
}

// The IC(0) teacher loss of the factor forms: |Lhat - T|_F^2 / |T|_F^2 over the blocks of the factor's
// pattern, T the block IC(0) factor of the same Hessian (bh_precond's lval / dinv after
// precond_factor_bic0) in the normalised space, T_vv = inv(dinv_v) / s_v, T_k = lval_k / s_v; entries in
// the row or column of a fixed component are left out (both factors are the identity there).
__device__ __forceinline__ void teacher_blocks(int v, const float* __restrict__ dinv, float sv, float T[9]) {
  const float* a = dinv + 6ll * v;
  const float I[9] = {a[0], 0.f, 0.f, a[1], a[2], 0.f, a[3], a[4], a[5]};
  ltri_inv(I, T);
  const float is = 1.f / sv;
#pragma unroll
  for (int i = 0; i < 9; ++i) T[i] *= is;
}

__global__ void __launch_bounds__(256)
    k_mgn_teacher_fwd(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                      const uint8_t* __restrict__ fixed, const float* __restrict__ C, const float* __restrict__ d,
                      const float* __restrict__ O, const float* __restrict__ s, const float* __restrict__ lval,
                      const float* __restrict__ dinv, double* __restrict__ part) {
  __shared__ double sh[8];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double num = 0.0, den = 0.0;
  if (v < nV) {
    const float sv = s[v];
    bool fv[3] = {false, false, false};
    if (fixed)
#pragma unroll
      for (int l = 0; l < 3; ++l) fv[l] = fixed[3ll * v + l] != 0;
    float L[9], T[9];
    mgn_lvv(C + 6ll * v, d + 6ll * v, L);
    teacher_blocks(v, dinv, sv, T);
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c <= l; ++c)
        if (!fv[l] && !fv[c]) {
          const double e = (double)L[3 * l + c] - T[3 * l + c];
          num += e * e;
          den += (double)T[3 * l + c] * T[3 * l + c];
        }
    const float is = 1.f / sv;
    for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
      const int n = lcol[k];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        if (fixed && (fv[i / 3] || fixed[3ll * n + i % 3])) continue;
        const double t = (double)lval[9ll * k + i] * is;
        const double e = (double)O[9ll * k + i] - t;
        num += e * e;
        den += t * t;
      }
    }
  }
  num = block_reduce_sum(num, sh);
  __syncthreads();
  den = block_reduce_sum(den, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = num;
    part[2 * blockIdx.x + 1] = den;
  }
}

// gradient of the teacher loss: gO_k = 2 (O_k - T_k) / |T|^2 (masked), gd_v through Lhat_vv = C_v T(d_v)
__global__ void __launch_bounds__(256)
    k_mgn_teacher_bwd(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                      const uint8_t* __restrict__ fixed, const float* __restrict__ C, const float* __restrict__ d,
                      const float* __restrict__ O, const float* __restrict__ s, const float* __restrict__ lval,
                      const float* __restrict__ dinv, const double* __restrict__ sc, float* __restrict__ gO,
                      float* __restrict__ gd) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const float scale = sc[1] > 0.0 ? (float)(2.0 / sc[1]) : 0.f;
  const float sv = s[v];
  bool fv[3] = {false, false, false};
  if (fixed)
#pragma unroll
    for (int l = 0; l < 3; ++l) fv[l] = fixed[3ll * v + l] != 0;
  const float is = 1.f / sv;
  for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
    const int n = lcol[k];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const bool mk = fixed && (fv[i / 3] || fixed[3ll * n + i % 3]);
      gO[9ll * k + i] = mk ? 0.f : scale * (O[9ll * k + i] - lval[9ll * k + i] * is);
    }
  }
  float L[9], T[9], G[9];
  mgn_lvv(C + 6ll * v, d + 6ll * v, L);
  teacher_blocks(v, dinv, sv, T);
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) G[3 * l + c] = (c <= l && !fv[l] && !fv[c]) ? scale * (L[3 * l + c] - T[3 * l + c]) : 0.f;
  // Lhat = C T: gT = C^T G (lower part kept), then through T(d)
  const float* Cv = C + 6ll * v;
  const float Cm[9] = {Cv[0], 0.f, 0.f, Cv[1], Cv[2], 0.f, Cv[3], Cv[4], Cv[5]};
  float gT[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gT[3 * i + j] = Cm[i] * G[j] + Cm[3 + i] * G[3 + j] + Cm[6 + i] * G[6 + j];
  const float* dv = d + 6ll * v;
  float* g = gd + 6ll * v;
  g[0] = gT[0] * dtexp(dv[0]);
  g[1] = gT[3];
  g[2] = gT[4] * dtexp(dv[2]);
  g[3] = gT[6];
  g[4] = gT[7];
  g[5] = gT[8] * dtexp(dv[5]);
}


// the tied blocks O = B Xi^T (B = Hn_vn, Xi = Lhat_nn^{-1}); add = 1 adds them to the pair decoder's
// output already in O (the tied factor form), add = 0 replaces it (the Eisenstat form)
__global__ void __launch_bounds__(256)
    k_mgn_tied_O(int64_t nP, const int32_t* __restrict__ pv, const int32_t* __restrict__ lcol,
                 const float* __restrict__ fp, const float* __restrict__ C, const float* __restrict__ d,
                 float* __restrict__ O, int add) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nP) return;
  const int n = lcol[k];
  float X[9], Xi[9];
  mgn_lvv(C + 6ll * n, d + 6ll * n, X);
  ltri_inv(X, Xi);
  const float* B = fp + (int64_t)kPairIn * k;
  // O = B Xi^T: O[l][c] = sum_m B[l][m] Xi[c][m]
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float t = B[3 * l] * Xi[3 * c] + B[3 * l + 1] * Xi[3 * c + 1] + B[3 * l + 2] * Xi[3 * c + 2];
      O[9 * k + 3 * l + c] = add ? O[9 * k + 3 * l + c] + t : t;
    }
  (void)pv;
}

// the tied blocks' share of the gradient of Lhat_nn, gathered over the pairs in which n is the lower
// vertex (its upper list): with O = B X^{-T}, dL/dX = -X^{-T} G^T O for G = dL/dO; added (through
// X = C T(d)) to gd_n.  In the Eisenstat form gO is cleared afterwards (the pair decoder does not act
// there); in the tied factor form it also reaches the pair decoder, whose output is added to O.
__global__ void __launch_bounds__(256)
    k_mgn_tied_bwd(int nV, const int32_t* __restrict__ uptr, const int32_t* __restrict__ ulink,
                   const float* __restrict__ C, const float* __restrict__ d, const float* __restrict__ fp,
                   const float* __restrict__ gO, float* __restrict__ gd) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nV) return;
  float X[9], Xi[9];
  mgn_lvv(C + 6ll * n, d + 6ll * n, X);
  ltri_inv(X, Xi);
  float A[9];  // sum over pairs of G^T O (O the tied blocks B Xi^T, formed again from the pair features)
#pragma unroll
  for (int e = 0; e < 9; ++e) A[e] = 0.f;
  for (int t = uptr[n]; t < uptr[n + 1]; ++t) {
    const int64_t k = ulink[t];
    const float* G = gO + 9 * k;
    const float* B = fp + (int64_t)kPairIn * k;
    float Ok[9];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Ok[3 * l + c] = B[3 * l] * Xi[3 * c] + B[3 * l + 1] * Xi[3 * c + 1] + B[3 * l + 2] * Xi[3 * c + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) A[3 * i + j] += G[i] * Ok[j] + G[3 + i] * Ok[3 + j] + G[6 + i] * Ok[6 + j];
  }
  // GX = -Xi^T A (lower part)
  float GX[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      GX[3 * i + j] = (j <= i) ? -(Xi[i] * A[j] + Xi[3 + i] * A[3 + j] + Xi[6 + i] * A[6 + j]) : 0.f;
  const float* Cv = C + 6ll * n;
  const float Cm[9] = {Cv[0], 0.f, 0.f, Cv[1], Cv[2], 0.f, Cv[3], Cv[4], Cv[5]};
  float gT[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gT[3 * i + j] = Cm[i] * GX[j] + Cm[3 + i] * GX[3 + j] + Cm[6 + i] * GX[6 + j];
  const float* dv = d + 6ll * n;
  float* g = gd + 6ll * n;
  g[0] += gT[0] * dtexp(dv[0]);
  g[1] += gT[3];
  g[2] += gT[4] * dtexp(dv[2]);
  g[3] += gT[6];
  g[4] += gT[7];
  g[5] += gT[8] * dtexp(dv[5]);
}

// Energy-ratio loss of the Eisenstat form, per probe x (normalised: xh = s x):
//   q = x^T M x = sum_v [xh_v^T Dh_v xh_v + 2 xh_v^T z_v + z_v^T Dh_v^{-1} z_v],  z_v = sum_{n upper of v} Hn_vn xh_n,
//   a = x^T H x,   loss = mean over probes of (log q - log a)^2,
// with Dh_v = Lhat_vv Lhat_vv^T.  It compares the energies of M and H on the probes (both ends of the
// spectrum of M^{-1} H when the probes are smooth and rough vectors) and is local per vertex, unlike
// the residual loss, whose minimiser on the tied form made CG slower than symmetric Gauss-Seidel.
__global__ void __launch_bounds__(256)
    k_mgn_energy_fwd(int nV, const int32_t* __restrict__ uptr, const int32_t* __restrict__ ucol,
                     const int32_t* __restrict__ ulink, const float* __restrict__ fp, const float* __restrict__ C,
                     const float* __restrict__ d, const float* __restrict__ s, const double* __restrict__ x,
                     const double* __restrict__ b, float* __restrict__ z, double* __restrict__ part) {
  __shared__ double sh[8];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double q = 0.0, a = 0.0;
  if (v < nV) {
    const double sv = s[v];
    const double xh[3] = {sv * x[3ll * v], sv * x[3ll * v + 1], sv * x[3ll * v + 2]};
    double zz[3] = {0.0, 0.0, 0.0};
    for (int t = uptr[v]; t < uptr[v + 1]; ++t) {
      const float* B = fp + (int64_t)kPairIn * ulink[t];  // Hn_nv, n the higher vertex
      const int n = ucol[t];
      const double sn = s[n];
      const double x0 = sn * x[3ll * n], x1 = sn * x[3ll * n + 1], x2 = sn * x[3ll * n + 2];
#pragma unroll
      for (int l = 0; l < 3; ++l) zz[l] += (double)B[l] * x0 + (double)B[3 + l] * x1 + (double)B[6 + l] * x2;
    }
    float L[9], Li[9];
    mgn_lvv(C + 6ll * v, d + 6ll * v, L);
    ltri_inv(L, Li);
    double ltx[3], liz[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      ltx[l] = (double)L[l] * xh[0] + (double)L[3 + l] * xh[1] + (double)L[6 + l] * xh[2];   // L^T xh
      liz[l] = (double)Li[3 * l] * zz[0] + (double)Li[3 * l + 1] * zz[1] + (double)Li[3 * l + 2] * zz[2];  // L^-1 z
    }
    q = ltx[0] * ltx[0] + ltx[1] * ltx[1] + ltx[2] * ltx[2] + 2.0 * (xh[0] * zz[0] + xh[1] * zz[1] + xh[2] * zz[2]) +
        liz[0] * liz[0] + liz[1] * liz[1] + liz[2] * liz[2];
    a = x[3ll * v] * b[3ll * v] + x[3ll * v + 1] * b[3ll * v + 1] + x[3ll * v + 2] * b[3ll * v + 2];
#pragma unroll
    for (int l = 0; l < 3; ++l) z[3ll * v + l] = (float)zz[l];
  }
  q = block_reduce_sum(q, sh);
  __syncthreads();
  a = block_reduce_sum(a, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = q;
    part[2 * blockIdx.x + 1] = a;
  }
}

// its gradient: dloss/dDh_v = c (xh xh^T - Dh^-1 z z^T Dh^-1), c = 2 (log q - log a) / (q ns); through
// Dh = Lhat Lhat^T (dloss/dLhat = 2 G Lhat, lower part) and Lhat = C T(d) into gd
__global__ void __launch_bounds__(256)
    k_mgn_energy_bwd(int nV, const float* __restrict__ C, const float* __restrict__ d, const float* __restrict__ s,
                     const double* __restrict__ x, const float* __restrict__ z, const double* __restrict__ qa, int ns,
                     int accumulate, float* __restrict__ gd) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const double q = qa[0], a = qa[1];
  const double c = (q > 0.0 && a > 0.0) ? 2.0 * (log(q) - log(a)) / (q * ns) : 0.0;
  const double sv = s[v];
  const double xh[3] = {sv * x[3ll * v], sv * x[3ll * v + 1], sv * x[3ll * v + 2]};
  float L[9], Li[9];
  mgn_lvv(C + 6ll * v, d + 6ll * v, L);
  ltri_inv(L, Li);
  const double zz[3] = {z[3ll * v], z[3ll * v + 1], z[3ll * v + 2]};
  double liz[3], w[3];
#pragma unroll
  for (int l = 0; l < 3; ++l) liz[l] = (double)Li[3 * l] * zz[0] + (double)Li[3 * l + 1] * zz[1] + (double)Li[3 * l + 2] * zz[2];
#pragma unroll
  for (int l = 0; l < 3; ++l) w[l] = (double)Li[l] * liz[0] + (double)Li[3 + l] * liz[1] + (double)Li[6 + l] * liz[2];  // Dh^-1 z
  double G[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[3 * i + j] = c * (xh[i] * xh[j] - w[i] * w[j]);
  float GL[9];  // lower part of 2 G Lhat
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      GL[3 * i + j] = (j <= i) ? (float)(2.0 * (G[3 * i] * L[j] + G[3 * i + 1] * L[3 + j] + G[3 * i + 2] * L[6 + j])) : 0.f;
  const float* Cv = C + 6ll * v;
  const float Cm[9] = {Cv[0], 0.f, 0.f, Cv[1], Cv[2], 0.f, Cv[3], Cv[4], Cv[5]};
  float gT[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gT[3 * i + j] = Cm[i] * GL[j] + Cm[3 + i] * GL[3 + j] + Cm[6 + i] * GL[6 + j];
  const float* dv = d + 6ll * v;
  float* g = gd + 6ll * v;
  const float a0 = gT[0] * dtexp(dv[0]), a1 = gT[3], a2 = gT[4] * dtexp(dv[2]), a3 = gT[6], a4 = gT[7],
              a5 = gT[8] * dtexp(dv[5]);
  if (accumulate) {
    g[0] += a0; g[1] += a1; g[2] += a2; g[3] += a3; g[4] += a4; g[5] += a5;
  } else {
    g[0] = a0; g[1] = a1; g[2] = a2; g[3] = a3; g[4] = a4; g[5] = a5;
  }
}

// sum of squares of a float vector in FP64, one block (fixed order): the finiteness check of a gradient
__global__ void __launch_bounds__(256) k_sumsq_f(int64_t n, const float* __restrict__ a, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) acc += (double)a[i] * (double)a[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void __launch_bounds__(256) k_zero_f(int64_t n, float* __restrict__ a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f;
}

// Probe systems for the Frobenius-type loss: in the normalised space a probe w gives xh = w and
// bh = Hn w, i.e. x = S^-1 w and b = H x.  w is the normalised right-hand side S^-1 rhs (probe 0
// when a right-hand side is given) or a standard normal vector from a counter-based hash
// (splitmix64 + Box-Muller) of (seed, probe, DOF); fixed DOFs are zero.
__host__ __device__ __forceinline__ uint64_t mgn_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void __launch_bounds__(256)
    k_mgn_probe_x(int nV, const float* __restrict__ s, const uint8_t* __restrict__ fixed,
                  const double* __restrict__ rhs, int is_solution, uint64_t seed, double* __restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 3ll * nV) return;
  const int v = (int)(i / 3);
  double w;
  if (rhs && is_solution) {  // a solution of the system itself: the smooth modes CG finds hardest
    x[i] = (fixed && fixed[i]) ? 0.0 : rhs[i];
    return;
  }
  if (rhs) {
    w = rhs[i] / (double)s[v];
  } else {
    const uint64_t a = mgn_mix(seed ^ mgn_mix((uint64_t)(2 * i))), b = mgn_mix(seed ^ mgn_mix((uint64_t)(2 * i + 1)));
    const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740993.0), u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    w = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  x[i] = (fixed && fixed[i]) ? 0.0 : w / (double)s[v];
}

// Inference: the pair decoder, the tied blocks and the factor's lower blocks in one pass, one thread
// per lower pair k = (v, n): O_k = e_k W_do + b_do (+ Hn_vn Lhat_nn^{-T} in the tied factor form, added
// after the decoder as the training path adds it), written to O (the forward's pair output) and, scaled
// by s_v with the fixed components' rows and columns zeroed, to lval.  The same arithmetic in the same
// order as k_linear<9>, k_mgn_tied_O and the pair part of k_mgn_factor_out.
__global__ void __launch_bounds__(256)
    k_mgn_pair_factor(int64_t nP, const int32_t* __restrict__ pv, const int32_t* __restrict__ lcol,
                      const uint8_t* __restrict__ fixed, const float* __restrict__ e, const float* __restrict__ Wdo,
                      const float* __restrict__ bdo, const float* __restrict__ fp, const float* __restrict__ C,
                      const float* __restrict__ d, const float* __restrict__ s, int tied, float* __restrict__ O,
                      float* __restrict__ lval) {
  __shared__ float sW[kW * 9 + 9];
  __shared__ float stage[2][256 * 9];  // O and lval rows of the workgroup's pairs
  for (int i = threadIdx.x; i < kW * 9; i += blockDim.x) sW[i] = Wdo[i];
  if (threadIdx.x < 9) sW[kW * 9 + threadIdx.x] = bdo[threadIdx.x];
  __syncthreads();
  const int64_t k0 = blockIdx.x * (int64_t)blockDim.x;
  const int64_t k = k0 + threadIdx.x;
  if (k < nP) {
    float o[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) o[j] = sW[kW * 9 + j];
    const float4* x4 = reinterpret_cast<const float4*>(e + k * kW);
#pragma unroll
    for (int q = 0; q < kW / 4; ++q) {
      const float4 x = x4[q];
      const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < 9; ++j) o[j] += xs[u] * sW[(4 * q + u) * 9 + j];
    }
    const int v = pv[k], n = lcol[k];
    if (tied) {
      float X[9], Xi[9];
      mgn_lvv(C + 6ll * n, d + 6ll * n, X);
      ltri_inv(X, Xi);
      const float* B = fp + (int64_t)kPairIn * k;
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          o[3 * l + c] += B[3 * l] * Xi[3 * c] + B[3 * l + 1] * Xi[3 * c + 1] + B[3 * l + 2] * Xi[3 * c + 2];
    }
    const float sv = s[v];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      stage[0][9 * threadIdx.x + i] = o[i];
      const bool fz = fixed && (fixed[3ll * v + i / 3] || fixed[3ll * n + i % 3]);
      stage[1][9 * threadIdx.x + i] = fz ? 0.f : sv * o[i];
    }
  }
  __syncthreads();
  const int cnt = (int)min<int64_t>(blockDim.x, nP - k0);
  store_rows<9>(stage[0], O, k0, cnt);
  store_rows<9>(stage[1], lval, k0, cnt);
}

// The factor of H for bh_precond: L_vn = s_v O_k, inverse of L_vv = s_v Lhat_vv (packed lower).
// Fixed DOFs get an identity row and column (zero off-diagonal entries, unit diagonal), so the
// preconditioner maps vectors that vanish on the fixed DOFs to vectors that vanish there too, as
// the masked system (identity rows / columns) requires.
__global__ void __launch_bounds__(256)
    k_mgn_factor_out(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                     const uint8_t* __restrict__ fixed, const float* __restrict__ C, const float* __restrict__ d,
                     const float* __restrict__ O, const float* __restrict__ s, float* __restrict__ lval,
                     float* __restrict__ dinv) {
  // lval null: the lower blocks were written by k_mgn_pair_factor
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const float sv = s[v];
  bool fv[3] = {false, false, false};
  if (fixed)
#pragma unroll
    for (int l = 0; l < 3; ++l) fv[l] = fixed[3ll * v + l] != 0;
  if (lval)
  for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
    const int n = lcol[k];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const bool fz = fixed && (fv[i / 3] || fixed[3ll * n + i % 3]);
      lval[9ll * k + i] = fz ? 0.f : sv * O[9ll * k + i];
    }
  }
  float L[9];
  mgn_lvv(C + 6ll * v, d + 6ll * v, L);
  // identity rows / columns of the fixed components (C is already the identity there)
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (fv[l] || fv[c]) L[3 * l + c] = (l == c) ? 1.f / sv : 0.f;
  const double l00 = (double)sv * L[0], l10 = (double)sv * L[3], l11 = (double)sv * L[4];
  const double l20 = (double)sv * L[6], l21 = (double)sv * L[7], l22 = (double)sv * L[8];
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

__global__ void __launch_bounds__(256) k_mgn_adam(int64_t n, float* __restrict__ w, float* __restrict__ g,
                                                  float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  g[i] = 0.f;
}

// ---- host drivers ----------------------------------------------------------------------------------
inline dim3 g256(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

int mgn_prepare(bh_mgn* M, const bh_precond* P) {
  const int nV = P->mesh->nV;
  const int64_t nP = P->nlower;
  if (M->wsV != nV || M->wsP != nP) {
    MgnWs dry;
    const size_t need = sizeof(float) * mgn_ws_carve(nullptr, nV, nP, dry);
    if (need > M->wsBytes) {
      if (M->ws) BH_CHECK(hipFree(M->ws));
      M->ws = nullptr;
      BH_CHECK(hipMalloc(&M->ws, need));
      M->wsBytes = need;
    }
    if (M->pvSize < nP) {
      if (M->pv) BH_CHECK(hipFree(M->pv));
      BH_CHECK(hipMalloc(&M->pv, sizeof(int32_t) * std::max<int64_t>(1, nP)));
      M->pvSize = nP;
    }
    M->wsV = nV;
    M->wsP = nP;
  }
  return BH_OK;
}

// infer: the pass feeds no backward, so the MLPs may run in BF16 on the matrix cores (M->inferBf16)
// and keep no hidden activations
template <typename V>
int mgn_forward(bh_mgn* M, const bh_precond* P, const V* vals, const uint8_t* fixed, MgnWs& w, hipStream_t s,
                bool infer = false) {
  const bh_mesh* m = P->mesh;
  const int nV = m->nV;
  const int64_t nP = P->nlower;
  const MgnLayout L = mgn_layout();
  const float* W = M->w;
  hipLaunchKernelGGL((k_mgn_prep<V>), g256(nV), dim3(256), 0, s, nV, m->bptr, vals, fixed, P->dslot, P->lptr, P->lcol,
                     P->lslot, w.s, w.C, w.xn);
  hipLaunchKernelGGL(k_mgn_pair_owner, g256(nV), dim3(256), 0, s, nV, P->lptr, M->pv);
  if (nP)
    hipLaunchKernelGGL((k_mgn_pairfeat<V>), g256(nP), dim3(256), 0, s, nP, (const int32_t*)M->pv, m->bptr, vals, fixed,
                       P->lcol, P->lslot, w.s, w.fp);
  BH_CHECK(hipGetLastError());
  auto mlp = [&](auto inTag, int64_t rows, Seg3 S, const MlpOff& o, const float* resid, float* hid, float* out) {
    constexpr int IN = decltype(inTag)::value;
    if (rows == 0) return;
    if (infer && M->inferBf16) {
      const int64_t tiles = (rows + 31) / 32;
      const unsigned nb = (unsigned)std::min<int64_t>((tiles + 3) / 4, kMfmaBlocks);
      hipLaunchKernelGGL((k_mlp_mfma<IN>), dim3(nb), dim3(256), 0, s, rows, S, W + o.W1, W + o.b1, W + o.W2,
                         W + o.b2, resid, out);
    } else {
      hipLaunchKernelGGL((k_mlp_fwd<IN>), g256(rows), dim3(256), 0, s, rows, S, W + o.W1, W + o.b1, W + o.W2,
                         W + o.b2, resid, hid, out);
    }
  };
  using I8 = std::integral_constant<int, kNodeIn>;
  using I10 = std::integral_constant<int, kPairIn>;
  using I64 = std::integral_constant<int, 2 * kW>;
  using I96 = std::integral_constant<int, 3 * kW>;
  mlp(I8{}, nV, Seg3{{w.xn, nullptr, nullptr}, {nullptr, nullptr, nullptr}}, L.nenc, nullptr,
      w.hidN[0], w.h[0]);
  mlp(I10{}, nP, Seg3{{w.fp, nullptr, nullptr}, {nullptr, nullptr, nullptr}}, L.penc, nullptr,
      w.hidE[0], w.e[0]);
  for (int l = 0; l < kLayers; ++l) {
    mlp(I96{}, nP, Seg3{{w.e[l], w.h[l], w.h[l]}, {nullptr, M->pv, P->lcol}}, L.edge[l], w.e[l],
        w.hidE[l + 1], w.e[l + 1]);
    hipLaunchKernelGGL(k_mgn_aggregate, g256((int64_t)nV * 8), dim3(256), 0, s, nV, P->lptr, P->uptr, P->ulink,
                       w.e[l + 1], w.m[l]);
    mlp(I64{}, nV, Seg3{{w.h[l], w.m[l], nullptr}, {nullptr, nullptr, nullptr}}, L.node[l], w.h[l],
        w.hidN[l + 1], w.h[l + 1]);
  }
  hipLaunchKernelGGL(k_linear<6>, g256(nV), dim3(256), 0, s, (int64_t)nV, w.h[kLayers], W + L.Wdv, W + L.bdv, w.d);
  if (nP && !infer)  // (inference applies the pair decoder in k_mgn_pair_factor, or does not need it)
    hipLaunchKernelGGL(k_linear<9>, g256(nP), dim3(256), 0, s, nP, w.e[kLayers], W + L.Wdo, W + L.bdo, w.O);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// The network's backward pass from the factor-side gradients w.gd (node decoder outputs) and w.gO (pair
// decoder outputs): the decoders, the message-passing layers in reverse, the encoders; the weight
// gradients are accumulated into M->g.
int mgn_network_backward(bh_mgn* M, const bh_precond* P, MgnWs& w, hipStream_t s) {
  const int nV = P->mesh->nV;
  const int64_t nP = P->nlower;
  const MgnLayout L = mgn_layout();
    float* G = M->g;
    float* wp = M->part;  // weight-gradient partials (after the loss partials were consumed)
    auto outer = [&](auto xwTag, auto gwTag, int64_t rows, Seg3 S, const float* g, float* gW, float* gb) {
      constexpr int XW = decltype(xwTag)::value, GW = decltype(gwTag)::value;
      if (rows == 0) return;
      hipLaunchKernelGGL((k_outer<XW, GW>), dim3(kOuterBlocks), dim3(256), 0, s, rows, S, g, wp);
      hipLaunchKernelGGL(k_outer_reduce, g256(XW * GW + GW), dim3(256), 0, s, kOuterBlocks, XW, GW, wp, gW, gb);
    };
    using I6 = std::integral_constant<int, 6>;
    using I9 = std::integral_constant<int, 9>;
    using I8 = std::integral_constant<int, kNodeIn>;
    using I10 = std::integral_constant<int, kPairIn>;
    using I32 = std::integral_constant<int, kW>;
    using I64 = std::integral_constant<int, 2 * kW>;
    using I96 = std::integral_constant<int, 3 * kW>;
    auto plain = [](const float* p, int) { return Seg3{{p, nullptr, nullptr}, {nullptr, nullptr, nullptr}}; };
    const float* W = M->w;
    // decoders
    outer(I32{}, I6{}, nV, plain(w.h[kLayers], kW), w.gd, G + L.Wdv, G + L.bdv);
    outer(I32{}, I9{}, nP, plain(w.e[kLayers], kW), w.gO, G + L.Wdo, G + L.bdo);
    hipLaunchKernelGGL(k_linear_bwd<6>, g256(nV), dim3(256), 0, s, (int64_t)nV, w.gd, W + L.Wdv, w.gh);
    if (nP) hipLaunchKernelGGL(k_linear_bwd<9>, g256(nP), dim3(256), 0, s, nP, w.gO, W + L.Wdo, w.ge);
    for (int l = kLayers - 1; l >= 0; --l) {
      // node MLP: h_{l+1} = h_l + MLP([h_l, m_l])
      const MlpOff& on = L.node[l];
      hipLaunchKernelGGL(k_mlp_bwd<2 * kW>, g256(nV), dim3(256), 0, s, (int64_t)nV, W + on.W1, W + on.W2,
                         w.hidN[l + 1], w.gh, w.ghid, w.gin);
      outer(I32{}, I32{}, nV, plain(w.hidN[l + 1], kW), w.gh, G + on.W2, G + on.b2);
      outer(I64{}, I32{}, nV, Seg3{{w.h[l], w.m[l], nullptr}, {nullptr, nullptr, nullptr}}, w.ghid,
            G + on.W1, G + on.b1);
      hipLaunchKernelGGL(k_mgn_gh_split, g256((int64_t)nV * kW), dim3(256), 0, s, nV, w.gh, w.gin, w.gh2, w.gm);
      // aggregation: ge_{l+1} += gm_v + gm_n
      if (nP) {
        hipLaunchKernelGGL(k_mgn_agg_bwd, g256(nP * kW), dim3(256), 0, s, nP, M->pv, P->lcol, w.gm, w.ge, w.ge2);
        // edge MLP: e_{l+1} = e_l + MLP([e_l, h_l[v], h_l[n]])
        const MlpOff& oe = L.edge[l];
        hipLaunchKernelGGL(k_mlp_bwd<3 * kW>, g256(nP), dim3(256), 0, s, nP, W + oe.W1, W + oe.W2, w.hidE[l + 1],
                           w.ge2, w.ghid, w.gin);
        outer(I32{}, I32{}, nP, plain(w.hidE[l + 1], kW), w.ge2, G + oe.W2, G + oe.b2);
        outer(I96{}, I32{}, nP, Seg3{{w.e[l], w.h[l], w.h[l]}, {nullptr, M->pv, P->lcol}}, w.ghid,
              G + oe.W1, G + oe.b1);
        hipLaunchKernelGGL(k_mgn_ge_prev, g256(nP * kW), dim3(256), 0, s, nP, w.ge2, w.gin, w.ge);
      }
      hipLaunchKernelGGL(k_mgn_gather_gh, g256((int64_t)nV * kW), dim3(256), 0, s, nV, P->lptr, P->uptr, P->ulink,
                         w.gin, w.gh2, w.gh);
    }
    // encoders (weights only)
    hipLaunchKernelGGL(k_mlp_bwd<kNodeIn>, g256(nV), dim3(256), 0, s, (int64_t)nV, W + L.nenc.W1, W + L.nenc.W2,
                       w.hidN[0], w.gh, w.ghid, (float*)nullptr);
    outer(I32{}, I32{}, nV, plain(w.hidN[0], kW), w.gh, G + L.nenc.W2, G + L.nenc.b2);
    outer(I8{}, I32{}, nV, plain(w.xn, kNodeIn), w.ghid, G + L.nenc.W1, G + L.nenc.b1);
    if (nP) {
      hipLaunchKernelGGL(k_mlp_bwd<kPairIn>, g256(nP), dim3(256), 0, s, nP, W + L.penc.W1, W + L.penc.W2, w.hidE[0],
                         w.ge, w.ghid, (float*)nullptr);
      outer(I32{}, I32{}, nP, plain(w.hidE[0], kW), w.ge, G + L.penc.W2, G + L.penc.b2);
      outer(I10{}, I32{}, nP, plain(w.fp, kPairIn), w.ghid, G + L.penc.W1, G + L.penc.b1);
    }
    BH_CHECK(hipGetLastError());
    return BH_OK;
}

// Mean loss over ns systems (b_j, x_j) and, with backward, its gradient accumulated into M->g:
// one forward, the factor-side gradients of every system summed, one backward through the network.
constexpr int kMaxSystems = 16;

// The network's backward for one system with a guard: the accumulated gradient is saved first, and if
// the accumulation is not finite afterwards it is restored and BH_NONFINITE returned, so one diverged
// system drops only its own contribution to a minibatch.  dscal[2 kMaxSystems] holds the check's sum.
int mgn_backward_checked(bh_mgn* M, const bh_precond* P, MgnWs& w, hipStream_t s) {
  const size_t bytes = sizeof(float) * (size_t)M->nparam;
  BH_CHECK(hipMemcpyAsync(M->gsave, M->g, bytes, hipMemcpyDeviceToDevice, s));
  int st = mgn_network_backward(M, P, w, s);
  if (st) return st;
  hipLaunchKernelGGL(k_sumsq_f, dim3(1), dim3(256), 0, s, M->nparam, (const float*)M->g, M->dscal + 2 * kMaxSystems);
  BH_CHECK(hipGetLastError());
  double g2 = 0.0;
  BH_CHECK(hipMemcpyAsync(&g2, M->dscal + 2 * kMaxSystems, sizeof(double), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  if (std::isfinite(g2)) return BH_OK;
  BH_CHECK(hipMemcpyAsync(M->g, M->gsave, bytes, hipMemcpyDeviceToDevice, s));
  return BH_NONFINITE;
}

// the accumulated gradient as it was before the latest guarded backward
int mgn_restore_gradient(bh_mgn* M, hipStream_t s) {
  BH_CHECK(hipMemcpyAsync(M->g, M->gsave, sizeof(float) * (size_t)M->nparam, hipMemcpyDeviceToDevice, s));
  return BH_OK;
}

template <typename V>
int mgn_loss(bh_mgn* M, const bh_precond* P, const V* vals, const uint8_t* fixed, int ns, const double* const* bs,
             const double* const* xs, int backward, double* loss, hipStream_t s) {
  if (ns < 1 || ns > kMaxSystems) return BH_BAD_ARG;
  int st = mgn_prepare(M, P);
  if (st) return st;
  MgnWs w;
  mgn_ws_carve(M->ws, P->mesh->nV, P->nlower, w);
  if ((st = mgn_forward(M, P, vals, fixed, w, s))) return st;
  const int nV = P->mesh->nV;
  const int64_t nP = P->nlower;
  if ((nV + 255) / 256 > 65536) return BH_UNSUPPORTED;  // loss partials
  if ((M->form == 1 || M->form == 2) && nP)
    hipLaunchKernelGGL(k_mgn_tied_O, g256(nP), dim3(256), 0, s, nP, M->pv, P->lcol, w.fp, w.C, w.d, w.O,
                       M->form == 2 ? 1 : 0);
  const MgnLayout L = mgn_layout();
  const int nb = (int)((nV + 255) / 256);
  double* part = reinterpret_cast<double*>(M->part);
  if (M->energyLoss) {
    for (int j = 0; j < ns; ++j) {
      hipLaunchKernelGGL(k_mgn_energy_fwd, g256(nV), dim3(256), 0, s, nV, P->uptr, P->ucol, P->ulink, w.fp, w.C, w.d,
                         w.s, xs[j], bs[j], w.y, part);
      hipLaunchKernelGGL(k_mgn_loss_sum, dim3(1), dim3(256), 0, s, nb, part, M->dscal + 2 * j);
      if (backward)
        hipLaunchKernelGGL(k_mgn_energy_bwd, g256(nV), dim3(256), 0, s, nV, w.C, w.d, w.s, xs[j], w.y,
                           M->dscal + 2 * j, ns, j > 0 ? 1 : 0, w.gd);
    }
    if (backward && nP) hipLaunchKernelGGL(k_zero_f, g256(9 * nP), dim3(256), 0, s, 9 * nP, w.gO);
  }
  if (M->teacher) {
    hipLaunchKernelGGL(k_mgn_teacher_fwd, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, fixed, w.C, w.d, w.O, w.s,
                       P->lval, P->dinv, part);
    hipLaunchKernelGGL(k_mgn_loss_sum, dim3(1), dim3(256), 0, s, nb, part, M->dscal);
    if (backward)
      hipLaunchKernelGGL(k_mgn_teacher_bwd, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, fixed, w.C, w.d, w.O,
                         w.s, P->lval, P->dinv, M->dscal, w.gO, w.gd);
    ns = 1;
  }
  for (int j = 0; j < ns && !M->energyLoss && !M->teacher; ++j) {
    // y = Lhat^T xh ; r = Lhat y - bh
    hipLaunchKernelGGL(k_mgn_lt, g256(nV), dim3(256), 0, s, nV, P->uptr, P->ucol, P->ulink, w.C, w.d, w.O, w.s, xs[j],
                       1, (const float*)nullptr, w.y);
    hipLaunchKernelGGL(k_mgn_resid, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, w.C, w.d, w.O, w.s, w.y, bs[j],
                       w.r, part);
    hipLaunchKernelGGL(k_mgn_loss_sum, dim3(1), dim3(256), 0, s, nb, part, M->dscal + 2 * j);
    if (backward) {
      hipLaunchKernelGGL(k_mgn_gz, g256(3ll * nV), dim3(256), 0, s, 3ll * nV, w.r, M->dscal + 2 * j, ns, w.gz);
      hipLaunchKernelGGL(k_mgn_lt, g256(nV), dim3(256), 0, s, nV, P->uptr, P->ucol, P->ulink, w.C, w.d, w.O, w.s,
                         (const double*)nullptr, 0, w.gz, w.gy);
      hipLaunchKernelGGL(k_mgn_gfactor, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, w.C, w.d, w.s, xs[j], w.y,
                         w.gz, w.gy, j > 0 ? 1 : 0, w.gO, w.gd);
    }
  }
  BH_CHECK(hipGetLastError());
  if (backward && (M->form == 1 || M->form == 2) && !M->energyLoss) {
    hipLaunchKernelGGL(k_mgn_tied_bwd, g256(nV), dim3(256), 0, s, nV, P->uptr, P->ulink, w.C, w.d, w.fp, w.gO, w.gd);
    if (nP && M->form == 1) hipLaunchKernelGGL(k_zero_f, g256(9 * nP), dim3(256), 0, s, 9 * nP, w.gO);
    BH_CHECK(hipGetLastError());
  }
  if (backward && (st = mgn_backward_checked(M, P, w, s))) return st;
  if (loss) {
    double h[2 * kMaxSystems];
    BH_CHECK(hipMemcpyAsync(h, M->dscal, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, s));
    BH_CHECK(hipStreamSynchronize(s));
    double a = 0.0;
    for (int j = 0; j < ns; ++j) {
      if (M->energyLoss) {
        const double e = (h[2 * j] > 0 && h[2 * j + 1] > 0) ? std::log(h[2 * j] / h[2 * j + 1]) : 0.0;
        a += e * e;
      } else {
        a += h[2 * j + 1] > 0 ? h[2 * j] / h[2 * j + 1] : 0.0;
      }
    }
    *loss = a / ns;
  }
  return BH_OK;
}

// ---- the Kaporin loss of the factor forms -----------------------------------------------------------
// log kappa(M^-1 H) = log(tr(M^-1 H) / n) - log det(M^-1 H) / n for M = L L^T, Kaporin's condition number
// (I. E. Kaporin, "New convergence results and preconditioning strategies for the conjugate gradient
// method", Numer. Linear Algebra Appl. 1(2), 1994), which bounds the CG iteration count.  With
// log det(M^-1 H) = log det H - 2 sum log L_ii (log det H does not depend on the network) and the
// Hutchinson estimate tr(L^-1 H L^-T) = mean over probes z of f = u^T H u, u = L^-T z (z standard normal on
// the free DOFs), the loss is log(mean f / n) + (2 / n) sum log L_ii.  Gradient on the factor's pattern:
// d f / d L = -2 u y^T with y = L^-1 H u, and d log L_ii / d L_ii = 1 / L_ii.  Everything runs in the
// preconditioner's solve order with its sweeps, after the network's factor is installed; the gradient
// then maps to the normalised factor (L = S Lhat: dLhat = s_v dL on the rows of v).  n = 3 nV.

__global__ void __launch_bounds__(256)
    k_kap_probe(int nV, const int32_t* __restrict__ colVerts, const uint8_t* __restrict__ fixed, uint64_t seed,
                double* __restrict__ z) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3ll * nV) return;
  const int64_t g = 3ll * colVerts[t / 3] + t % 3;  // the DOF in vertex order: the draw does not depend on
                                                      // the colouring
  const uint64_t a = mgn_mix(seed ^ mgn_mix((uint64_t)(2 * g))), b = mgn_mix(seed ^ mgn_mix((uint64_t)(2 * g + 1)));
  const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740993.0), u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
  z[t] = (fixed && fixed[g]) ? 0.0 : sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// per-block partial sums of a . b (fixed partition: deterministic)
__global__ void __launch_bounds__(256) k_kap_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ part) {
  __shared__ double sh[8];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += a[i] * b[i];
  acc = block_reduce_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}
__global__ void __launch_bounds__(256) k_kap_sum(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[8];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
  acc = block_reduce_sum(acc, sh);
  if (threadIdx.x == 0) *out = acc;
}

// accumulates sum over probes of u_i y_n^T on the lower pairs (vertex-order entry nlsrc[t]) and u_i y_i^T
// on the diagonal blocks (9 per vertex), positions i in the solve order
__global__ void __launch_bounds__(256)
    k_kap_accum(int nV, const int32_t* __restrict__ colVerts, const int32_t* __restrict__ nlptr,
                const int32_t* __restrict__ nlcol, const int32_t* __restrict__ nlsrc, const double* __restrict__ u,
                const double* __restrict__ y, int first, float* __restrict__ gO, float* __restrict__ gD) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const double u0 = u[3ll * i], u1 = u[3ll * i + 1], u2 = u[3ll * i + 2];
  const double ui[3] = {u0, u1, u2};
  for (int t = nlptr[i]; t < nlptr[i + 1]; ++t) {
    const int n = nlcol[t];
    float* o = gO + 9ll * nlsrc[t];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float a = (float)(ui[l] * y[3ll * n + c]);
        o[3 * l + c] = first ? a : o[3 * l + c] + a;
      }
  }
  float* dd = gD + 9ll * colVerts[i];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float a = (float)(ui[l] * y[3ll * i + c]);
      dd[3 * l + c] = first ? a : dd[3 * l + c] + a;
    }
}

// the loss gradient on the normalised factor and through Lhat_vv = C_v T(d_v); partial sums of
// log Lhat_ll over the free components (for the reported loss)
__global__ void __launch_bounds__(256)
    k_kap_final(int nV, const int32_t* __restrict__ lptr, const int32_t* __restrict__ lcol,
                const uint8_t* __restrict__ fixed, const float* __restrict__ C, const float* __restrict__ d,
                const float* __restrict__ s, float gscale, float dscale, float* __restrict__ gO,
                const float* __restrict__ gD, float* __restrict__ gd, double* __restrict__ part) {
  __shared__ double sh[8];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  double ld = 0.0;
  if (v < nV) {
    const float sv = s[v];
    bool fv[3] = {false, false, false};
    if (fixed)
#pragma unroll
      for (int l = 0; l < 3; ++l) fv[l] = fixed[3ll * v + l] != 0;
    for (int k = lptr[v]; k < lptr[v + 1]; ++k) {
      const int n = lcol[k];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const bool mk = fixed && (fv[i / 3] || fixed[3ll * n + i % 3]);
        gO[9ll * k + i] = mk ? 0.f : gscale * sv * gO[9ll * k + i];
      }
    }
    float L[9], G[9];
    mgn_lvv(C + 6ll * v, d + 6ll * v, L);
    const float* A = gD + 9ll * v;
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool keep = c <= l && !fv[l] && !fv[c];
        G[3 * l + c] = keep ? gscale * sv * A[3 * l + c] + (l == c ? dscale / L[3 * l + c] : 0.f) : 0.f;
      }
#pragma unroll
    for (int l = 0; l < 3; ++l)
      if (!fv[l]) ld += log((double)L[4 * l]);
    const float* Cv = C + 6ll * v;
    const float Cm[9] = {Cv[0], 0.f, 0.f, Cv[1], Cv[2], 0.f, Cv[3], Cv[4], Cv[5]};
    float gT[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) gT[3 * i + j] = Cm[i] * G[j] + Cm[3 + i] * G[3 + j] + Cm[6 + i] * G[6 + j];
    const float* dv = d + 6ll * v;
    float* g = gd + 6ll * v;
    g[0] = gT[0] * dtexp(dv[0]);
    g[1] = gT[3];
    g[2] = gT[4] * dtexp(dv[2]);
    g[3] = gT[6];
    g[4] = gT[7];
    g[5] = gT[8] * dtexp(dv[5]);
  }
  ld = block_reduce_sum(ld, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = ld;
}

constexpr int kKapDotBlocks = 256;

// One training step of the Kaporin loss on a Hessian (the factor forms): forward, the network's factor
// into P, nprobe probes, the gradient, the network's backward, Adam.
template <typename V>
int mgn_kaporin_step(bh_mgn* M, bh_precond* P, const V* vals, const uint8_t* fixed, int nprobe, uint64_t seed,
                     double lr, double* loss, hipStream_t s) {
  int st = mgn_prepare(M, P);
  if (st) return st;
  const bh_mesh* m = P->mesh;
  const int nV = m->nV;
  const int64_t nP = P->nlower, n3 = 3ll * nV;
  MgnWs w;
  mgn_ws_carve(M->ws, nV, nP, w);
  if ((st = mgn_forward(M, P, vals, fixed, w, s))) return st;
  if (M->form == 2 && nP)
    hipLaunchKernelGGL(k_mgn_tied_O, g256(nP), dim3(256), 0, s, nP, M->pv, P->lcol, w.fp, w.C, w.d, w.O, 1);
  hipLaunchKernelGGL(k_mgn_factor_out, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, fixed, w.C, w.d, w.O, w.s,
                     P->lval, P->dinv);
  BH_CHECK(hipGetLastError());
  if ((st = precond_finalize(P, s))) return st;
  double* z = svec(P, SV_Z);
  double* u = svec(P, SV_U);
  double* uv = svec(P, SV_T);
  double* huv = svec(P, SV_V);
  double* hu = svec(P, SV_W);
  double* y = svec(P, SV_Y);
  double* part = reinterpret_cast<double*>(M->part);
  float* gD = w.gm;  // 9 per vertex (free until the network's backward)
  const int g3 = (int)((n3 + 255) / 256);
  const int nc = P->ncolors;
  for (int j = 0; j < nprobe; ++j) {
    hipLaunchKernelGGL(k_kap_probe, dim3(g3), dim3(256), 0, s, nV, P->colVerts, fixed,
                       mgn_mix(seed + 0x1000ull * (uint64_t)j), z);
    for (int c = nc - 1; c >= 0; --c) {  // u = L^-T z
      const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
      if (!n) continue;
      hipLaunchKernelGGL(k_sweep_bwd, dim3(sweep_grid(n)), dim3(256), 0, s, P->h_colPtr[c], n, P->nuptr, P->nucol,
                         P->uvs, P->dinvs, (const double*)z, u, (const double*)nullptr, (double*)nullptr,
                         (const PcgState*)nullptr);
    }
    hipLaunchKernelGGL(k_permute_rows, dim3(g3), dim3(256), 0, s, nV, P->colVerts, 0, (const double*)u, uv);
    hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(spmv_blocks_for<V>(nV)), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals,
                       (const double*)nullptr, fixed, (const double*)uv, huv, (double*)nullptr,
                       (const PcgState*)nullptr);
    hipLaunchKernelGGL(k_permute_rows, dim3(g3), dim3(256), 0, s, nV, P->colVerts, 1, (const double*)huv, hu);
    hipLaunchKernelGGL(k_kap_dot, dim3(kKapDotBlocks), dim3(256), 0, s, n3, (const double*)u, (const double*)hu,
                       part);
    hipLaunchKernelGGL(k_kap_sum, dim3(1), dim3(256), 0, s, kKapDotBlocks, (const double*)part, M->dscal + j);
    CgVecs none{};
    for (int c = 0; c < nc; ++c) {  // y = L^-1 H u
      const int n = P->h_colPtr[c + 1] - P->h_colPtr[c];
      if (!n) continue;
      hipLaunchKernelGGL(k_sweep_fwd, dim3(sweep_grid(n)), dim3(256), 0, s, P->h_colPtr[c], n, P->nlptr, P->nlcol,
                         P->lvs, P->dinvs, hu, y, none, 0);
    }
    hipLaunchKernelGGL(k_kap_accum, g256(nV), dim3(256), 0, s, nV, P->colVerts, P->nlptr, P->nlcol, P->nlsrc,
                       (const double*)u, (const double*)y, j == 0 ? 1 : 0, w.gO, gD);
    BH_CHECK(hipGetLastError());
  }
  double f[kMaxSystems];
  BH_CHECK(hipMemcpyAsync(f, M->dscal, sizeof(double) * nprobe, hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  double tr = 0.0;
  for (int j = 0; j < nprobe; ++j) tr += f[j];
  tr /= nprobe;
  if (!(tr > 0.0) || !std::isfinite(tr)) return BH_NONFINITE;  // an indefinite or broken factor: no step
  const double n = (double)n3;
  const int nb = (int)((nV + 255) / 256);
  hipLaunchKernelGGL(k_kap_final, dim3(nb), dim3(256), 0, s, nV, P->lptr, P->lcol, fixed, w.C, w.d, w.s,
                     (float)(-2.0 / (tr * nprobe)), (float)(2.0 / n), w.gO, (const float*)gD, w.gd, part);
  hipLaunchKernelGGL(k_kap_sum, dim3(1), dim3(256), 0, s, nb, (const double*)part, M->dscal + kMaxSystems);
  if (M->form == 2)
    hipLaunchKernelGGL(k_mgn_tied_bwd, g256(nV), dim3(256), 0, s, nV, P->uptr, P->ulink, w.C, w.d, w.fp, w.gO, w.gd);
  BH_CHECK(hipGetLastError());
  if ((st = mgn_backward_checked(M, P, w, s))) return st;
  double logdet = 0.0;
  BH_CHECK(hipMemcpyAsync(&logdet, M->dscal + kMaxSystems, sizeof(double), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  const double lv = std::log(tr / n) + 2.0 * logdet / n;
  if (loss) *loss = lv;
  if (!std::isfinite(lv)) {  // drop this system's gradient, keep the accumulation before it
    if ((st = mgn_restore_gradient(M, s))) return st;
    return BH_NONFINITE;
  }
  if (lr == 0.0) return BH_OK;  // the gradient stays accumulated in M->g
  M->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const double c1 = 1.0 - std::pow(b1, (double)M->step), c2 = 1.0 - std::pow(b2, (double)M->step);
  hipLaunchKernelGGL(k_mgn_adam, g256(M->nparam), dim3(256), 0, s, M->nparam, M->w, M->g, M->m1, M->m2, (float)lr,
                     (float)b1, (float)b2, 1e-8f, (float)c1, (float)c2);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// One training step on a Hessian: nprobe probe systems (the normalised right-hand side first when
// rhs is given, then standard normal vectors drawn from seed), loss, backward, Adam.
template <typename V>
int mgn_train_step(bh_mgn* M, const bh_precond* P, const V* vals, const uint8_t* fixed, const double* rhs, int nprobe,
                   uint64_t seed, double lr, double* loss, hipStream_t s) {
  if (nprobe < 1 || nprobe > kMaxSystems) return BH_BAD_ARG;
  if (M->kaporin) return mgn_kaporin_step<V>(M, const_cast<bh_precond*>(P), vals, fixed, nprobe, seed, lr, loss, s);
  int st = mgn_prepare(M, P);
  if (st) return st;
  const bh_mesh* m = P->mesh;
  const int nV = m->nV;
  const size_t need = sizeof(double) * 2 * (size_t)nprobe * 3 * (size_t)nV;
  if (need > M->probeBytes) {
    if (M->probe) BH_CHECK(hipFree(M->probe));
    M->probe = nullptr;
    BH_CHECK(hipMalloc(&M->probe, need));
    M->probeBytes = need;
  }
  // the scales come from the forward pass of mgn_loss; probes need them first, so run the
  // feature kernel once here (it is cheap and deterministic)
  MgnWs w;
  mgn_ws_carve(M->ws, nV, P->nlower, w);
  hipLaunchKernelGGL((k_mgn_prep<V>), g256(nV), dim3(256), 0, s, nV, m->bptr, vals, fixed, P->dslot, P->lptr, P->lcol,
                     P->lslot, w.s, w.C, w.xn);
  const double* bs[kMaxSystems];
  const double* xs[kMaxSystems];
  for (int j = 0; j < nprobe; ++j) {
    double* x = M->probe + (size_t)(2 * j) * 3 * nV;
    double* b = x + 3ll * nV;
    const bool useRhs = rhs && j == 0;
    hipLaunchKernelGGL(k_mgn_probe_x, g256(3ll * nV), dim3(256), 0, s, nV, w.s, fixed, useRhs ? rhs : nullptr,
                       M->probeSolution, mgn_mix(seed + 0x1000ull * (uint64_t)j), x);
    hipLaunchKernelGGL(k_pcg_spmv<V>, dim3(spmv_blocks_for<V>(nV)), dim3(kRedBlock), 0, s, nV, m->bptr, m->bcol, vals,
                       (const double*)nullptr, fixed,
                       x, b, nullptr, nullptr);
    bs[j] = b;
    xs[j] = x;
  }
  BH_CHECK(hipGetLastError());
  if (M->teacher) {
    // the teacher: block IC(0) of this Hessian (diagonal-shift retry as the solver does), into P
    double shift = 0.0;
    for (int attempt = 0;; ++attempt) {
      int fails = 0;
      if ((st = precond_factor_bic0(const_cast<bh_precond*>(P), vals, (const double*)nullptr, fixed, shift, &fails, s)))
        return st;
      if (fails == 0) break;
      if (attempt >= 6) return BH_SOLVE_FAILED;
      shift = (shift > 0.0) ? 4.0 * shift : 1e-3;
    }
  }
  double lv = 0.0;
  if ((st = mgn_loss<V>(M, P, vals, fixed, M->teacher ? 1 : nprobe, bs, xs, 1, &lv, s))) return st;
  if (loss) *loss = lv;
  if (!std::isfinite(lv)) {  // a diverged step: drop its gradient (the accumulation before it stays)
    if ((st = mgn_restore_gradient(M, s))) return st;
    return BH_NONFINITE;
  }
  if (lr == 0.0) return BH_OK;  // the gradient stays accumulated in M->g
  M->step += 1;
  const double b1 = 0.9, b2 = 0.999;
  const double c1 = 1.0 - std::pow(b1, (double)M->step), c2 = 1.0 - std::pow(b2, (double)M->step);
  hipLaunchKernelGGL(k_mgn_adam, g256(M->nparam), dim3(256), 0, s, M->nparam, M->w, M->g, M->m1, M->m2, (float)lr,
                     (float)b1, (float)b2, 1e-8f, (float)c1, (float)c2);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

template <typename V>
int mgn_factor(bh_mgn* M, bh_precond* P, const V* vals, const uint8_t* fixed, hipStream_t s) {
  int st = mgn_prepare(M, P);
  if (st) return st;
  MgnWs w;
  mgn_ws_carve(M->ws, P->mesh->nV, P->nlower, w);
  if ((st = mgn_forward(M, P, vals, fixed, w, s, true))) return st;
  const int nV = P->mesh->nV;
  const MgnLayout L = mgn_layout();
  if (P->nlower)  // tied factor form: L_vn = s_v (Hn_vn Lhat_nn^{-T} + O_k); general form: s_v O_k
    hipLaunchKernelGGL(k_mgn_pair_factor, g256(P->nlower), dim3(256), 0, s, P->nlower, M->pv, P->lcol, fixed,
                       w.e[kLayers], M->w + L.Wdo, M->w + L.bdo, w.fp, w.C, w.d, w.s, M->form == 2 ? 1 : 0, w.O,
                       P->lval);
  hipLaunchKernelGGL(k_mgn_factor_out, g256(nV), dim3(256), 0, s, nV, P->lptr, P->lcol, fixed, w.C, w.d, w.O, w.s,
                     (float*)nullptr, P->dinv);
  BH_CHECK(hipGetLastError());
  return precond_finalize(P, s);
}

// The Eisenstat-form diagonal from the network, one thread per solve position i (vertex colVerts[i]):
// Dt_i = s_v^2 Lhat_vv Lhat_vv^T with Lhat_vv = C_v T(d_v) (C_v the Cholesky factor of the scaled
// masked H_vv), so zero decoders give Dt = D, block symmetric Gauss-Seidel; fixed components get
// identity rows / columns.  Writes Dt, its inverse and K = 2 Dt - D for the sweeps.
__global__ void __launch_bounds__(256)
    k_mgn_sgs_out(int nV, const int32_t* __restrict__ colVerts, const uint8_t* __restrict__ fixed,
                  const float* __restrict__ C, const float* __restrict__ d, const float* __restrict__ s,
                  const double* __restrict__ hd, float* __restrict__ dt, float* __restrict__ dtinv,
                  double* __restrict__ kk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nV) return;
  const int v = colVerts[i];
  const double sv = s[v];
  float L[9];
  mgn_lvv(C + 6ll * v, d + 6ll * v, L);
  bool fv[3] = {false, false, false};
  if (fixed)
#pragma unroll
    for (int l = 0; l < 3; ++l) fv[l] = fixed[3ll * v + l] != 0;
  double D[9];
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x = sv * sv *
                       ((double)L[3 * l] * L[3 * c] + (double)L[3 * l + 1] * L[3 * c + 1] + (double)L[3 * l + 2] * L[3 * c + 2]);
      D[3 * l + c] = (fv[l] || fv[c]) ? ((l == c) ? 1.0 : 0.0) : x;
    }
  eis_store_diag(i, D, hd, dt, dtinv, kk);
}

// the network's Eisenstat-form diagonal for the split Hessian already in P
template <typename V>
int mgn_sgs(bh_mgn* M, bh_precond* P, const V* vals, const uint8_t* fixed, hipStream_t s) {
  int st = mgn_prepare(M, P);
  if (st) return st;
  MgnWs w;
  mgn_ws_carve(M->ws, P->mesh->nV, P->nlower, w);
  if ((st = mgn_forward(M, P, vals, fixed, w, s, true))) return st;
  const int nV = P->mesh->nV;
  hipLaunchKernelGGL(k_mgn_sgs_out, g256(nV), dim3(256), 0, s, nV, P->colVerts, fixed, w.C, w.d, w.s,
                     (const double*)P->hdiag, P->dt, P->dtinv, P->kk);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

// portable parameter initialisation (splitmix64 uniforms): He-uniform first layers, second layers
// scaled by 1/4 (residual branches start small), zero biases and zero decoders
void mgn_init_host(std::vector<float>& w, uint64_t seed) {
  const MgnLayout L = mgn_layout();
  w.assign((size_t)L.total, 0.f);
  uint64_t ctr = 0;
  auto uni = [&]() {
    uint64_t z = seed + (++ctr) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int64_t off, int in, int out, double scale) {
    const double bound = std::sqrt(6.0 / in) * scale;
    for (int64_t i = 0; i < (int64_t)in * out; ++i) w[off + i] = (float)((2.0 * uni() - 1.0) * bound);
  };
  auto mlp = [&](const MlpOff& o, int in) {
    fill(o.W1, in, kW, 1.0);
    fill(o.W2, kW, kW, 0.25);
  };
  mlp(L.nenc, kNodeIn);
  mlp(L.penc, kPairIn);
  for (int l = 0; l < kLayers; ++l) {
    mlp(L.edge[l], 3 * kW);
    mlp(L.node[l], 2 * kW);
  }
}

}  // namespace

extern "C" {

int bh_mgn_create(uint64_t seed, bh_mgn** out) {
  if (!out) return BH_BAD_ARG;
  *out = nullptr;
  auto* M = new bh_mgn();
  const MgnLayout L = mgn_layout();
  M->nparam = L.total;
  std::vector<float> h;
  mgn_init_host(h, seed);
  const size_t bytes = sizeof(float) * (size_t)M->nparam;
  const size_t partBytes = sizeof(float) * (size_t)kOuterBlocks * (3 * kW * kW + kW) + sizeof(double) * 2 * 65536;
  if (hipMalloc(&M->w, bytes) != hipSuccess || hipMalloc(&M->g, bytes) != hipSuccess || hipMalloc(&M->gsave, bytes) != hipSuccess ||
      hipMalloc(&M->m1, bytes) != hipSuccess || hipMalloc(&M->m2, bytes) != hipSuccess ||
      hipMalloc(&M->part, partBytes) != hipSuccess || hipMalloc(&M->dscal, sizeof(double) * (2 * kMaxSystems + 1)) != hipSuccess) {
    delete M;
    return BH_DEVICE_ERROR;
  }
  BH_CHECK(hipMemcpy(M->w, h.data(), bytes, hipMemcpyHostToDevice));
  BH_CHECK(hipMemset(M->g, 0, bytes));
  BH_CHECK(hipMemset(M->m1, 0, bytes));
  BH_CHECK(hipMemset(M->m2, 0, bytes));
  *out = M;
  return BH_OK;
}

int bh_mgn_destroy(bh_mgn* M) {
  if (!M) return BH_BAD_ARG;
  (void)hipDeviceSynchronize();
  void* ptrs[] = {M->w, M->g, M->gsave, M->m1, M->m2, M->part, M->dscal, M->ws, M->pv, M->probe};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete M;
  return BH_OK;
}

int bh_mgn_set_form(bh_mgn* M, int form, int probe_is_solution, int energy_loss) {
  // energy_loss: 0 the residual loss, 1 the energy-ratio loss (form 1), 2 the IC(0) teacher (forms 0, 2)
  if (!M || form < 0 || form > 2 || energy_loss < 0 || energy_loss > 3 || (energy_loss == 1 && form != 1) ||
      (energy_loss >= 2 && form == 1))
    return BH_BAD_ARG;
  M->form = form;
  M->teacher = energy_loss == 2 ? 1 : 0;
  M->kaporin = energy_loss == 3 ? 1 : 0;
  M->probeSolution = probe_is_solution ? 1 : 0;
  M->energyLoss = energy_loss == 1 ? 1 : 0;
  return BH_OK;
}

int bh_mgn_set_inference(bh_mgn* M, int precision) {
  if (!M || precision < 0 || precision > 1) return BH_BAD_ARG;
  M->inferBf16 = precision;
  return BH_OK;
}

int bh_mgn_num_params(const bh_mgn* M, int64_t* n) {
  if (!M || !n) return BH_BAD_ARG;
  *n = M->nparam;
  return BH_OK;
}

int bh_mgn_get_params(const bh_mgn* M, float* host_params, float* host_grad) {
  if (!M) return BH_BAD_ARG;
  BH_CHECK(hipDeviceSynchronize());
  if (host_params) BH_CHECK(hipMemcpy(host_params, M->w, sizeof(float) * M->nparam, hipMemcpyDeviceToHost));
  if (host_grad) BH_CHECK(hipMemcpy(host_grad, M->g, sizeof(float) * M->nparam, hipMemcpyDeviceToHost));
  return BH_OK;
}

int bh_mgn_set_params(bh_mgn* M, const float* host_params) {
  if (!M || !host_params) return BH_BAD_ARG;
  BH_CHECK(hipDeviceSynchronize());
  BH_CHECK(hipMemcpy(M->w, host_params, sizeof(float) * M->nparam, hipMemcpyHostToDevice));
  BH_CHECK(hipMemset(M->g, 0, sizeof(float) * M->nparam));
  BH_CHECK(hipMemset(M->m1, 0, sizeof(float) * M->nparam));
  BH_CHECK(hipMemset(M->m2, 0, sizeof(float) * M->nparam));
  M->step = 0;
  return BH_OK;
}

int bh_mgn_loss_device(bh_mgn* M, bh_precond* P, const float* d_vals, const uint8_t* d_fixed_dof, const double* d_b,
                       const double* d_x, int backward, double* loss, bh_stream_t stream) {
  if (!M || !P || !d_vals || !d_b || !d_x) return BH_BAD_ARG;
  if (P->mesh->nV == 0) return BH_OK;
  const double* bs[1] = {d_b};
  const double* xs[1] = {d_x};
  return mgn_loss<float>(M, P, d_vals, d_fixed_dof, 1, bs, xs, backward, loss, as_stream(stream));
}

int bh_mgn_train_step_device(bh_mgn* M, bh_precond* P, const float* d_vals, const uint8_t* d_fixed_dof,
                             const double* d_rhs, int nprobe, uint64_t seed, double lr, double* loss,
                             bh_stream_t stream) {
  if (!M || !P || !d_vals || !(lr >= 0)) return BH_BAD_ARG;
  if (P->mesh->nV == 0) return BH_OK;
  return mgn_train_step<float>(M, P, d_vals, d_fixed_dof, d_rhs, nprobe, seed, lr, loss, as_stream(stream));
}

int bh_mgn_adam_step(bh_mgn* M, double lr, double beta1, double beta2, double eps, bh_stream_t stream) {
  if (!M || !(lr > 0)) return BH_BAD_ARG;
  M->step += 1;
  const double c1 = 1.0 - std::pow(beta1, (double)M->step), c2 = 1.0 - std::pow(beta2, (double)M->step);
  hipLaunchKernelGGL(k_mgn_adam, g256(M->nparam), dim3(256), 0, as_stream(stream), M->nparam, M->w, M->g, M->m1,
                     M->m2, (float)lr, (float)beta1, (float)beta2, (float)eps, (float)c1, (float)c2);
  BH_CHECK(hipGetLastError());
  return BH_OK;
}

int bh_mgn_factor_device(bh_mgn* M, bh_precond* P, const float* d_vals, const uint8_t* d_fixed_dof,
                         bh_stream_t stream) {
  if (!M || !P || !d_vals) return BH_BAD_ARG;
  if (P->mesh->nV == 0) return BH_OK;
  return mgn_factor<float>(M, P, d_vals, d_fixed_dof, as_stream(stream));
}

int bh_mgn_debug_latents(bh_mgn* M, int layer, const float** h, const float** e, const float** m) {
  if (!M || !M->ws || layer < -1 || layer > kLayers) return BH_BAD_ARG;
  MgnWs w;
  mgn_ws_carve(M->ws, M->wsV, M->wsP, w);
  if (h) *h = layer < 0 ? w.xn : w.h[layer];
  if (e) *e = layer < 0 ? w.fp : w.e[layer];
  if (m) *m = (layer >= 0 && layer < kLayers) ? w.m[layer] : nullptr;
  return BH_OK;
}

int bh_mgn_debug_buffers(bh_mgn* M, const float** s, const float** C, const float** d, const float** O) {
  if (!M || !M->ws) return BH_BAD_ARG;
  MgnWs w;
  mgn_ws_carve(M->ws, M->wsV, M->wsP, w);
  if (s) *s = w.s;
  if (C) *C = w.C;
  if (d) *d = w.d;
  if (O) *O = w.O;
  return BH_OK;
}

}  // extern "C"