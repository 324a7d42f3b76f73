// Shared device linear-algebra pieces (included at the end of bartels_hip.hip, before the
// preconditioner, network and Newton parts; same translation unit): the CG scalar block, the
// fixed-shape block reduction and the 16-lane block-pattern SpMV.
namespace {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;
constexpr int kPcgBatch = 16;

// Device-resident CG scalars (Chronopoulos-Gear form: one reduction of r.u, r.r and w.u per
// iteration).  done: 0 running, 1 converged, 2 iteration cap, 3 non-positive curvature.
struct PcgState {
  double gamma, delta, alpha, beta, rr, bb, tol2;
  int done, iters, maxit;
};

// kernels of an iteration stop once the solve is over, except for the one update that applies
// alpha = 1 after non-positive curvature in the first iteration
__device__ __forceinline__ bool cg_skip(const PcgState* st) {
  const int d = st->done;
  return d != 0 && !(d == 3 && st->alpha != 0.0);
}

// sum over a block of NT threads (NT a multiple of 64), fixed order; the result is valid in thread 0
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}
__device__ __forceinline__ double block_sum_256(double v, double* sh) { return block_sum<256>(v, sh); }

// inverse of 3x3 blocks (row-major, 9 per block; point = 1 keeps only the diagonal); singular blocks
// fall back to the inverse of their diagonal
__device__ __forceinline__ void inv3_block(double* A, int point, double* o) {
  if (point) A[1] = A[2] = A[3] = A[5] = A[6] = A[7] = 0.0;
  const double c0 = A[4] * A[8] - A[5] * A[7], c1 = A[5] * A[6] - A[3] * A[8], c2 = A[3] * A[7] - A[4] * A[6];
  const double det = A[0] * c0 + A[1] * c1 + A[2] * c2;
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

// Ap = mask(H p), partial p.Ap per block.  16 lanes per vertex: lane j takes blocks j, j+16, ... of
// the vertex's block row (its three rows are one contiguous CSR segment, so the 16 lanes read
// neighbouring 3-value pieces), then the three row sums are reduced across the 16 lanes with a fixed
// xor butterfly.  FP32 values are widened; products are summed in FP64.  One pass per block (the
// loads of a row are a dependent chain; many resident blocks hide it).  hdiag (optional, FP64, 9 per
// vertex) replaces the stored diagonal blocks.
// lanes per vertex: 16 with FP64 blocks, 8 with FP32 blocks (two blocks per lane on the Kuhn
// meshes; on the 6.07M-tet beam the FP32 product got 5 % faster per Newton step with 8 lanes, the
// FP64 one 16 % slower, and unrolling the block loop or 4 lanes were slower in both)
template <typename V>
constexpr int spmv_lanes() {
  return sizeof(V) == 4 ? 8 : 16;
}
constexpr int kSpmvLanes = 16;  // the most lanes of any instance: sizes the partial-sum regions
inline int spmv_blocks(int nV) {
  return (int)std::max<int64_t>(1, (kSpmvLanes * (int64_t)nV + kRedBlock - 1) / kRedBlock);
}
template <typename V>
inline int spmv_blocks_for(int nV) {
  return (int)std::max<int64_t>(1, (spmv_lanes<V>() * (int64_t)nV + kRedBlock - 1) / kRedBlock);
}

template <typename V>
__global__ void __launch_bounds__(kRedBlock)
    k_pcg_spmv(int nV, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
               const V* __restrict__ vals, const double* __restrict__ hdiag, const uint8_t* __restrict__ fixed,
               const double* __restrict__ p,
               double* __restrict__ Ap, double* __restrict__ part, const PcgState* __restrict__ st) {
  constexpr int L = spmv_lanes<V>();
  __shared__ double sh[4];
  if (st && cg_skip(st)) return;
  const int64_t gid = blockIdx.x * (int64_t)kRedBlock + threadIdx.x;
  const int v = (int)(gid / L), j = (int)(gid % L);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (v < nV) {
    const int b0 = bptr[v], nb = bptr[v + 1] - b0;
    const V* row = vals + 9ll * b0;
    for (int r = j; r < nb; r += L) {
      const int c = bcol[b0 + r];
      const double* xp = p + 3ll * c;
      const double x0 = xp[0], x1 = xp[1], x2 = xp[2];
      if (hdiag && c == v) {
        const double* D = hdiag + 9ll * v;
        s0 += D[0] * x0 + D[1] * x1 + D[2] * x2;
        s1 += D[3] * x0 + D[4] * x1 + D[5] * x2;
        s2 += D[6] * x0 + D[7] * x1 + D[8] * x2;
        continue;
      }
      const V* r0 = row + 3 * r;
      const V* r1 = r0 + 3 * nb;
      const V* r2 = r0 + 6 * nb;
      s0 += (double)r0[0] * x0 + (double)r0[1] * x1 + (double)r0[2] * x2;
      s1 += (double)r1[0] * x0 + (double)r1[1] * x1 + (double)r1[2] * x2;
      s2 += (double)r2[0] * x0 + (double)r2[1] * x1 + (double)r2[2] * x2;
    }
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 64);
    s1 += __shfl_xor(s1, off, 64);
    s2 += __shfl_xor(s2, off, 64);
  }
  double pap = 0.0;
  if (v < nV && j < 3) {
    const int64_t rowi = 3ll * v + j;
    double sv = (j == 0) ? s0 : (j == 1) ? s1 : s2;
    const double pr = p[rowi];
    if (fixed && fixed[rowi]) sv = pr;
    Ap[rowi] = sv;
    pap = pr * sv;
  }
  if (part) {
    const double d = block_sum_256(pap, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = d;
  }
}

}  // namespace
