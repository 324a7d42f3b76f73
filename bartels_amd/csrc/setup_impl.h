// On-device set-up of a mesh (included at the end of bartels_hip.hip, same translation unit): the
// vertex -> incident-tet lists, the 3x3-block pattern and the block ranks that the host build_mesh makes,
// built from device arrays with radix sorts, scans and a sort-unique of the 16 nT vertex pairs, bit for bit
// the host's arrays.  Replaces the triplet pattern of sim::assemble (src/assemble.cpp:57-94) made once.
#include <hipcub/hipcub.hpp>

namespace {

__global__ void __launch_bounds__(256) k_setup_incidence_keys(int nV, int nT, const int32_t* __restrict__ E,
                                                              uint32_t* __restrict__ key, int32_t* __restrict__ code,
                                                              int32_t* __restrict__ cnt, int32_t* __restrict__ bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // i = 4 e + a: ascending e, then a
  if (i >= 4ll * nT) return;
  const int e = (int)(i >> 2), a = (int)(i & 3);
  const int v = E[e + (int64_t)a * nT];
  if (v < 0 || v >= nV) {
    atomicOr(bad, 1);
    key[i] = 0u;
    code[i] = 0;
    return;
  }
  key[i] = (uint32_t)v;
  code[i] = (e << 2) | a;
  atomicAdd(cnt + v, 1);
}

// the 16 vertex pairs (v, u) of every incidence of v, as 64-bit keys v << 32 | u
__global__ void __launch_bounds__(256) k_setup_pair_keys(int64_t ninc, int nT, const int32_t* __restrict__ E,
                                                         const uint32_t* __restrict__ incV,
                                                         const int32_t* __restrict__ inc, uint64_t* __restrict__ pk) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= ninc) return;
  const uint64_t v = incV[k];
  const int e = inc[k] >> 2;
#pragma unroll
  for (int b = 0; b < 4; ++b) pk[4 * k + b] = (v << 32) | (uint32_t)E[e + (int64_t)b * nT];
}

__global__ void __launch_bounds__(256) k_setup_split_pairs(int64_t n, const uint64_t* __restrict__ uk,
                                                           int32_t* __restrict__ bcol, int32_t* __restrict__ bcnt) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  bcol[j] = (int32_t)(uk[j] & 0xffffffffull);
  atomicAdd(bcnt + (int)(uk[j] >> 32), 1);
}

// per incidence: the ranks of the tet's four vertices in the neighbour list of the incidence's vertex
__global__ void __launch_bounds__(256) k_setup_ranks(int64_t ninc, int nT, const int32_t* __restrict__ E,
                                                     const uint32_t* __restrict__ incV, const int32_t* __restrict__ inc,
                                                     const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                                                     uint32_t* __restrict__ rank) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= ninc) return;
  const int v = (int)incV[k];
  const int e = inc[k] >> 2;
  const int lo0 = bptr[v], hi0 = bptr[v + 1];
  uint32_t packed = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int u = E[e + (int64_t)b * nT];
    int lo = lo0, hi = hi0;  // first position with bcol >= u
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (bcol[mid] < u)
        lo = mid + 1;
      else
        hi = mid;
    }
    packed |= (uint32_t)(lo - lo0) << (8 * b);
  }
  rank[k] = packed;
}

// per group of 64 vertices: the longest incidence list (gLen) and 64 gLen slots (for the scan);
// per group of vpw vertices: the largest CSR segment of the gathers (maxSeg), and the largest valence
__global__ void __launch_bounds__(256) k_setup_groups(int nV, const int32_t* __restrict__ incPtr,
                                                      const int32_t* __restrict__ bptr, int32_t* __restrict__ gLen,
                                                      int64_t* __restrict__ slots, int32_t* __restrict__ vdeg,
                                                      int32_t* __restrict__ stats /*seg64, seg32, seg16, maxnb*/) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int nG = (nV + 63) / 64;
  if (g >= nG) return;
  int L = 0, nb = 0;
  for (int v = 64 * g; v < min(nV, 64 * g + 64); ++v) {
    const int d = incPtr[v + 1] - incPtr[v];
    vdeg[v] = d;
    L = max(L, d);
    nb = max(nb, bptr[v + 1] - bptr[v]);
  }
  gLen[g] = L;
  slots[g] = 64ll * L;
  const int v0 = 64 * g;
  atomicMax(stats + 0, 9 * (bptr[min(v0 + 64, nV)] - bptr[v0]));
  for (int h = 0; h < 2; ++h) atomicMax(stats + 1, 9 * (bptr[min(v0 + 32 * h + 32, nV)] - bptr[min(v0 + 32 * h, nV)]));
  for (int h = 0; h < 4; ++h) atomicMax(stats + 2, 9 * (bptr[min(v0 + 16 * h + 16, nV)] - bptr[min(v0 + 16 * h, nV)]));
  atomicMax(stats + 3, nb);
}

__global__ void __launch_bounds__(256) k_setup_interleave(int nV, const int32_t* __restrict__ incPtr,
                                                          const int64_t* __restrict__ gBase,
                                                          const int32_t* __restrict__ inc,
                                                          const uint32_t* __restrict__ rank, int32_t* __restrict__ incE,
                                                          uint32_t* __restrict__ incR) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nV) return;
  const int64_t base = gBase[v >> 6] + (v & 63);
  for (int k = incPtr[v]; k < incPtr[v + 1]; ++k) {
    const int64_t idx = base + 64ll * (k - incPtr[v]);
    incE[idx] = inc[k];
    incR[idx] = rank[k];
  }
}

__global__ void __launch_bounds__(256) k_setup_fill(int64_t n, int32_t* __restrict__ incE, uint32_t* __restrict__ incR) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  incE[i] = -1;
  incR[i] = 0u;
}

__global__ void __launch_bounds__(256) k_setup_check_finite(int64_t n, const double* __restrict__ a,
                                                            int32_t* __restrict__ bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && !isfinite(a[i])) atomicOr(bad, 2);
}

// device buffers of the set-up, freed on every exit path
struct SetupBufs {
  std::vector<void*> p;
  template <typename T>
  int get(T** out, size_t n) {
    void* q = nullptr;
    BH_CHECK(hipMalloc(&q, sizeof(T) * std::max<size_t>(1, n)));
    p.push_back(q);
    *out = static_cast<T*>(q);
    return BH_OK;
  }
  ~SetupBufs() {
    for (void* q : p) (void)hipFree(q);
  }
};

int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b)) ++b;
  return b;
}

template <typename T>
int d2h(std::vector<T>& h, const T* d, size_t n) {
  h.resize(n);
  if (n) BH_CHECK(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
  return BH_OK;
}

// the arrays of build_mesh from device E (nT x 4 column-major): m->E is already on the device
int build_mesh_device(bh_mesh* m, hipStream_t s) {
  const int nV = m->nV, nT = m->nT;
  const int64_t ninc = 4ll * nT, npair = 16ll * nT;
  const int nG = (nV + 63) / 64;
  SetupBufs B;
  int st;
  uint32_t *keyA, *keyB;
  int32_t *codeA, *inc, *cnt, *incPtr, *flag, *bcnt, *bptr, *bcol, *gLen, *stats, *vdeg;
  uint64_t *pk, *pk2, *uk;
  int64_t *nuniq, *slots, *gBase;
  uint32_t* rank;
  if ((st = B.get(&keyA, ninc)) || (st = B.get(&keyB, ninc)) || (st = B.get(&codeA, ninc)) ||
      (st = B.get(&inc, ninc)) || (st = B.get(&cnt, nV + 1)) || (st = B.get(&incPtr, nV + 1)) ||
      (st = B.get(&flag, 1)) || (st = B.get(&stats, 4)) || (st = B.get(&nuniq, 1)))
    return st;
  BH_CHECK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nV + 1), s));
  BH_CHECK(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
  BH_CHECK(hipMemsetAsync(stats, 0, sizeof(int32_t) * 4, s));
  const unsigned gi = (unsigned)((ninc + 255) / 256);
  if (ninc) {
    hipLaunchKernelGGL(k_setup_incidence_keys, dim3(gi), dim3(256), 0, s, nV, nT, (const int32_t*)m->E, keyA, codeA,
                       cnt, flag);
    BH_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_setup_check_finite, dim3((unsigned)((12ll * nT + 255) / 256 + 1)), dim3(256), 0, s, 12ll * nT,
                     (const double*)m->dX, flag);
  hipLaunchKernelGGL(k_setup_check_finite, dim3((unsigned)((nT + 255) / 256 + 1)), dim3(256), 0, s, (int64_t)nT,
                     (const double*)m->vol, flag);
  int32_t hflag = 0;
  BH_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  if (hflag & 1) return BH_BAD_ARG;
  if (hflag & 2) return BH_NONFINITE;
  // incidence lists: a stable radix sort by vertex keeps ascending (e, a) within every vertex
  const int vbits = bits_for((uint64_t)std::max(1, nV));
  size_t tb = 0, tb2 = 0;
  void* temp = nullptr;
  BH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keyA, keyB, codeA, inc, (int)ninc, 0, vbits, s));
  BH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt, incPtr, nV + 1, s));
  size_t tbytes = std::max(tb, tb2);
  // pair sort and unique
  if ((st = B.get(&pk, npair)) || (st = B.get(&pk2, npair)) || (st = B.get(&uk, npair))) return st;
  const int pbits = 32 + vbits;
  BH_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, pk, pk2, (int)std::min<int64_t>(npair, INT32_MAX), 0, pbits,
                                             s));
  tbytes = std::max(tbytes, tb);
  BH_CHECK(hipcub::DeviceSelect::Unique(nullptr, tb, pk2, uk, nuniq, (int)std::min<int64_t>(npair, INT32_MAX), s));
  tbytes = std::max(tbytes, tb);
  if (npair > INT32_MAX) return BH_UNSUPPORTED;
  {
    int64_t* z = nullptr;
    BH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, z, z, nG + 1, s));
    tbytes = std::max(tbytes, tb);
  }
  char* tmp;
  if ((st = B.get(&tmp, tbytes))) return st;
  temp = tmp;
  if (ninc) BH_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, tbytes, keyA, keyB, codeA, inc, (int)ninc, 0, vbits, s));
  BH_CHECK(hipcub::DeviceScan::ExclusiveSum(temp, tbytes, cnt, incPtr, nV + 1, s));
  if (ninc) {
    hipLaunchKernelGGL(k_setup_pair_keys, dim3(gi), dim3(256), 0, s, ninc, nT, (const int32_t*)m->E,
                       (const uint32_t*)keyB, (const int32_t*)inc, pk);
    BH_CHECK(hipGetLastError());
    BH_CHECK(hipcub::DeviceRadixSort::SortKeys(temp, tbytes, pk, pk2, (int)npair, 0, pbits, s));
    BH_CHECK(hipcub::DeviceSelect::Unique(temp, tbytes, pk2, uk, nuniq, (int)npair, s));
  } else {
    BH_CHECK(hipMemsetAsync(nuniq, 0, sizeof(int64_t), s));
  }
  int64_t nb = 0;
  BH_CHECK(hipMemcpyAsync(&nb, nuniq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  if (9 * nb > (int64_t)INT32_MAX) return BH_UNSUPPORTED;  // int32 CSR indices
  if ((st = B.get(&bcnt, nV + 1))) return st;
  BH_CHECK(hipMemsetAsync(bcnt, 0, sizeof(int32_t) * (nV + 1), s));
  BH_CHECK(hipMalloc(&m->bcol, sizeof(int32_t) * std::max<int64_t>(1, nb)));
  BH_CHECK(hipMalloc(&m->bptr, sizeof(int32_t) * (nV + 1)));
  bcol = m->bcol;
  bptr = m->bptr;
  if (nb) {
    hipLaunchKernelGGL(k_setup_split_pairs, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, nb,
                       (const uint64_t*)uk, bcol, bcnt);
    BH_CHECK(hipGetLastError());
  }
  BH_CHECK(hipcub::DeviceScan::ExclusiveSum(temp, tbytes, bcnt, bptr, nV + 1, s));
  // ranks, groups, interleaved lists
  if ((st = B.get(&rank, ninc)) || (st = B.get(&gLen, nG)) || (st = B.get(&slots, nG + 1)) ||
      (st = B.get(&gBase, nG + 1)))
    return st;
  BH_CHECK(hipMalloc(&m->vdeg, sizeof(int32_t) * std::max(1, nV)));
  vdeg = m->vdeg;
  if (ninc) {
    hipLaunchKernelGGL(k_setup_ranks, dim3(gi), dim3(256), 0, s, ninc, nT, (const int32_t*)m->E, (const uint32_t*)keyB,
                       (const int32_t*)inc, (const int32_t*)bptr, (const int32_t*)bcol, rank);
    BH_CHECK(hipGetLastError());
  }
  BH_CHECK(hipMemsetAsync(slots, 0, sizeof(int64_t) * (nG + 1), s));
  if (nG) {
    hipLaunchKernelGGL(k_setup_groups, dim3((nG + 255) / 256), dim3(256), 0, s, nV, (const int32_t*)incPtr,
                       (const int32_t*)bptr, gLen, slots, vdeg, stats);
    BH_CHECK(hipGetLastError());
  }
  BH_CHECK(hipcub::DeviceScan::ExclusiveSum(temp, tbytes, slots, gBase, nG + 1, s));
  int64_t tot = 0;
  int32_t hstats[4] = {0, 0, 0, 0};
  BH_CHECK(hipMemcpyAsync(&tot, gBase + nG, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipMemcpyAsync(hstats, stats, sizeof(hstats), hipMemcpyDeviceToHost, s));
  BH_CHECK(hipStreamSynchronize(s));
  if (hstats[3] > 255) return BH_UNSUPPORTED;  // ranks are packed in 8 bits
  BH_CHECK(hipMalloc(&m->incE, sizeof(int32_t) * std::max<int64_t>(1, tot)));
  BH_CHECK(hipMalloc(&m->incR, sizeof(uint32_t) * std::max<int64_t>(1, tot)));
  hipLaunchKernelGGL(k_setup_fill, dim3((unsigned)((std::max<int64_t>(1, tot) + 255) / 256)), dim3(256), 0, s,
                     std::max<int64_t>(1, tot), m->incE, m->incR);
  if (nV) {
    hipLaunchKernelGGL(k_setup_interleave, dim3((nV + 255) / 256), dim3(256), 0, s, nV, (const int32_t*)incPtr,
                       (const int64_t*)gBase, (const int32_t*)inc, (const uint32_t*)rank, m->incE, m->incR);
  }
  BH_CHECK(hipGetLastError());
  // the per-group arrays the kernels read, and the host copies of the pattern
  BH_CHECK(hipMalloc(&m->gBase, sizeof(int64_t) * std::max(1, nG)));
  BH_CHECK(hipMalloc(&m->gLen, sizeof(int32_t) * std::max(1, nG)));
  if (nG) {
    BH_CHECK(hipMemcpyAsync(m->gBase, gBase, sizeof(int64_t) * nG, hipMemcpyDeviceToDevice, s));
    BH_CHECK(hipMemcpyAsync(m->gLen, gLen, sizeof(int32_t) * nG, hipMemcpyDeviceToDevice, s));
  }
  BH_CHECK(hipStreamSynchronize(s));
  if ((st = d2h(m->h_bptr, (const int32_t*)bptr, (size_t)nV + 1)) || (st = d2h(m->h_bcol, (const int32_t*)bcol, nb)))
    return st;
  m->maxSeg64 = hstats[0];
  m->maxSeg32 = hstats[1];
  m->maxSeg16 = hstats[2];
  for (int v0 = 0; v0 < nV; v0 += 8)
    m->maxSeg8 = std::max(m->maxSeg8, 9 * (m->h_bptr[std::min(v0 + 8, nV)] - m->h_bptr[v0]));
  m->nblocks = nb;
  m->nnz = 9 * nb;
  return BH_OK;
}

}  // namespace

extern "C" {

int bh_mesh_create_device(int nV, int nT, const int32_t* d_E, const double* d_V, const double* d_dphidX,
                          const double* d_volume, bh_mesh** out) {
  if (!out || nV < 0 || nT < 0 || (nT && !d_E)) return BH_BAD_ARG;
  if (nT && !(d_dphidX && d_volume) && !d_V) return BH_BAD_ARG;
  if ((d_dphidX == nullptr) != (d_volume == nullptr)) return BH_BAD_ARG;
  *out = nullptr;
  if ((int64_t)nT >= (1ll << 29)) return BH_UNSUPPORTED;  // element id packed with the corner in int32
  bh_mesh* m = new bh_mesh();
  m->nV = nV;
  m->nT = nT;
  hipStream_t s = nullptr;
  auto fail = [&](int st) {
    free_mesh(m);
    delete m;
    return st;
  };
  int st = BH_OK;
  if (hipMalloc(&m->E, sizeof(int32_t) * 4 * std::max(1, nT)) != hipSuccess ||
      hipMalloc(&m->dX, sizeof(double) * 12 * std::max(1, nT)) != hipSuccess ||
      hipMalloc(&m->vol, sizeof(double) * std::max(1, nT)) != hipSuccess)
    return fail(BH_DEVICE_ERROR);
  if (nT) {
    if (hipMemcpyAsync(m->E, d_E, sizeof(int32_t) * 4 * (size_t)nT, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail(BH_DEVICE_ERROR);
    if (d_dphidX) {
      if (hipMemcpyAsync(m->dX, d_dphidX, sizeof(double) * 12 * (size_t)nT, hipMemcpyDeviceToDevice, s) !=
              hipSuccess ||
          hipMemcpyAsync(m->vol, d_volume, sizeof(double) * (size_t)nT, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(BH_DEVICE_ERROR);
    }
  }
  // the pattern first: it checks E before dphidX reads through it
  if (d_dphidX || !nT) {
    if ((st = build_mesh_device(m, s))) return fail(st);
  } else {
    // dphidX and the signed volume from V on the device (k_dphidX), once E is known to be in range
    if (hipMemsetAsync(m->dX, 0, sizeof(double) * 12 * (size_t)nT, s) != hipSuccess ||
        hipMemsetAsync(m->vol, 0, sizeof(double) * (size_t)nT, s) != hipSuccess)
      return fail(BH_DEVICE_ERROR);
    if ((st = build_mesh_device(m, s))) return fail(st);
    hipLaunchKernelGGL(k_dphidX, dim3((nT + 255) / 256), dim3(256), 0, s, view(m), d_V, m->dX, m->vol);
    // a degenerate tet gives non-finite shape gradients: rejected as bh_mesh_create rejects them
    int32_t* flag = nullptr;
    int32_t hflag = 0;
    if (hipMalloc(&flag, sizeof(int32_t)) != hipSuccess) return fail(BH_DEVICE_ERROR);
    const bool ok = hipMemsetAsync(flag, 0, sizeof(int32_t), s) == hipSuccess;
    hipLaunchKernelGGL(k_setup_check_finite, dim3((unsigned)((12ll * nT + 255) / 256)), dim3(256), 0, s, 12ll * nT,
                       (const double*)m->dX, flag);
    hipLaunchKernelGGL(k_setup_check_finite, dim3((unsigned)((nT + 255) / 256)), dim3(256), 0, s, (int64_t)nT,
                       (const double*)m->vol, flag);
    const bool ok2 = hipGetLastError() == hipSuccess &&
                     hipMemcpyAsync(&hflag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s) == hipSuccess &&
                     hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(flag);
    if (!ok || !ok2) return fail(BH_DEVICE_ERROR);
    if (hflag) return fail(BH_NONFINITE);
  }
  if ((st = finish_mesh(m))) return fail(st);
  if (d_V && nV) {
    // the gathers' visiting order from the rest positions (host-side sort of the group centroids)
    std::vector<double> Vc(3 * (size_t)nV), Xi(3 * (size_t)nV);
    if (hipMemcpy(Vc.data(), d_V, sizeof(double) * Vc.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(BH_DEVICE_ERROR);
    for (int v = 0; v < nV; ++v)
      for (int c = 0; c < 3; ++c) Xi[3 * (size_t)v + c] = Vc[v + (size_t)c * nV];
    if ((st = build_gather_order(m, Xi.data()))) return fail(st);
  }
  *out = m;
  return BH_OK;
}

}  // extern "C"
