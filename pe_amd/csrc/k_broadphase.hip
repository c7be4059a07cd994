// Broadphase of the time step: device-wide scan, LDS-staged LSD radix sort of
// (cell key, body) pairs, cell ranges, and the pair enumeration that runs the
// narrowphase on every AABB-overlapping pair (HashGrids.h:1121-1358 semantics:
// every pair of bodies that are not both fixed and whose closed AABBs overlap is
// handed to the narrowphase exactly once).
#include "kernels.hpp"

namespace pe_amd {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;  // 2048 elements per block tile
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr double kSpanFill = 0.999999;

// Exclusive scan of one value per thread over a 256-thread block (4 wave64s).
// Returns the exclusive prefix; *total receives the block total.  Uses lds[0..3].
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* lds, uint32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t v = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (lane == 63) lds[wid] = v;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    const uint32_t s = lds[k];
    if (k < wid) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return v - x + pre;
}

// Fills (kernels rather than hipMemsetAsync so that a captured step graph holds no memset node).
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = (uint32_t)i;
}
void iota_u32(uint32_t* p, size_t n, hipStream_t s) {
  if (!n) return;
  size_t nb = (n + 255) / 256;
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_iota_u32, dim3((unsigned)nb), dim3(256), 0, s, p, n);
}
void fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
  if (!n) return;
  size_t nb = (n + 255) / 256;
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)nb), dim3(256), 0, s, p, value, n);
}
void fill_u64(unsigned long long* p, unsigned long long value, size_t n, hipStream_t s) {
  if (!n) return;
  size_t nb = (n + 255) / 256;
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)nb), dim3(256), 0, s, p, value, n);
}

// ---------------------------------------------------------------------------------------------
// Device-wide exclusive scan (reduce -> scan of block sums -> apply).
// ---------------------------------------------------------------------------------------------
__global__ void k_scan_reduce(const uint32_t* in, size_t n, uint32_t* sums) {
  __shared__ uint32_t lds[4];
  const size_t base = (size_t)blockIdx.x * kTile;
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const size_t idx = base + (size_t)j * kThreads + threadIdx.x;
    if (idx < n) acc += in[idx];
  }
  uint32_t tot;
  block_exclusive_scan(acc, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_scan_apply(const uint32_t* in, uint32_t* out, size_t n, const uint32_t* offs,
                             uint32_t* total) {
  __shared__ uint32_t lds[4];
  const size_t base = (size_t)blockIdx.x * kTile + (size_t)threadIdx.x * kItems;  // blocked
  uint32_t x[kItems];
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const size_t idx = base + j;
    x[j] = idx < n ? in[idx] : 0u;
    acc += x[j];
  }
  uint32_t tot;
  uint32_t pre = block_exclusive_scan(acc, lds, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const size_t idx = base + j;
    if (idx < n) out[idx] = pre;
    pre += x[j];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) *total = pre;
}

size_t scan_scratch_words(size_t n) {
  size_t words = 0;
  while (n > (size_t)kTile) {
    n = (n + kTile - 1) / kTile;
    words += 2 * n + 1;
  }
  return words + 1;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total,
                        uint32_t* scratch, hipStream_t s) {
  if (n == 0) {
    if (total) fill_u32(total, 0u, 1, s);
    return;
  }
  const size_t nb = (n + kTile - 1) / kTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kThreads), 0, s, in, out, n, nullptr, total);
    return;
  }
  uint32_t* sums = scratch;
  uint32_t* sums_scanned = scratch + nb;
  uint32_t* rest = scratch + 2 * nb + 1;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kThreads), 0, s, in, n, sums);
  exclusive_scan_u32(sums, sums_scanned, nb, nullptr, rest, s);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kThreads), 0, s, in, out, n, sums_scanned, total);
}

// ---------------------------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: per-tile digit histograms -> one scan over
// (digit, tile) -> stable scatter.  The scatter stages its 2048-element tile in LDS and
// orders it by digit with eight stable 1-bit splits, so the global writes of one digit
// run are contiguous.
// ---------------------------------------------------------------------------------------------
__global__ void k_rs_hist(const uint32_t* keys, size_t n, int shift, uint32_t* hist, uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const size_t idx = base + (size_t)j * kThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                             size_t n, int shift, const uint32_t* hist_scanned, uint32_t ntiles) {
  __shared__ uint32_t sk[kTile];
  __shared__ uint32_t sv[kTile];
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t lds[4];
  const size_t base = (size_t)blockIdx.x * kTile;
  const uint32_t valid = (uint32_t)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t li = j * kThreads + threadIdx.x;
    const size_t idx = base + li;
    sk[li] = idx < n ? kin[idx] : 0xFFFFFFFFu;
    sv[li] = idx < n ? vin[idx] : 0u;
  }
  cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int b = 0; b < 8; ++b) {
    const int bit = shift + b;
    uint32_t k[kItems], v[kItems];
    uint32_t zeros = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      k[j] = sk[threadIdx.x * kItems + j];
      v[j] = sv[threadIdx.x * kItems + j];
      zeros += ((k[j] >> bit) & 1u) ? 0u : 1u;
    }
    uint32_t total_zeros;
    const uint32_t zpre = block_exclusive_scan(zeros, lds, &total_zeros);
    uint32_t opre = threadIdx.x * kItems - zpre;
    uint32_t z = zpre;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const bool one = (k[j] >> bit) & 1u;
      const uint32_t pos = one ? (total_zeros + opre++) : (z++);
      sk[pos] = k[j];
      sv[pos] = v[j];
    }
    __syncthreads();
  }
  // tile-local digit starts
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t li = j * kThreads + threadIdx.x;
    if (li < valid) atomicAdd(&cnt[(sk[li] >> shift) & 255u], 1u);
  }
  __syncthreads();
  uint32_t tot;
  const uint32_t my = cnt[threadIdx.x];
  const uint32_t start = block_exclusive_scan(my, lds, &tot);
  __syncthreads();
  cnt[threadIdx.x] = start;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t li = j * kThreads + threadIdx.x;
    if (li < valid) {
      const uint32_t key = sk[li];
      const uint32_t d = (key >> shift) & 255u;
      const uint32_t g = hist_scanned[(size_t)d * ntiles + blockIdx.x] + (li - cnt[d]);
      kout[g] = key;
      vout[g] = sv[li];
    }
  }
}

size_t sort_scratch_words(size_t n) {
  const size_t ntiles = (n + kTile - 1) / kTile;
  const size_t h = 256 * (ntiles ? ntiles : 1);
  return 2 * h + scan_scratch_words(h);
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      size_t n, int bits, uint32_t* scratch, hipStream_t s) {
  if (n <= 1) return;
  const uint32_t ntiles = (uint32_t)((n + kTile - 1) / kTile);
  const size_t h = 256 * (size_t)ntiles;
  uint32_t* hist = scratch;
  uint32_t* hs = scratch + h;
  uint32_t* sc = scratch + 2 * h;
  uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
  const int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kThreads), 0, s, ki, n, shift, hist, ntiles);
    exclusive_scan_u32(hist, hs, h, nullptr, sc, s);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kThreads), 0, s, ki, vi, ko, vo, n, shift, hs, ntiles);
    uint32_t* t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
  }
  if (ki != keys) {
    hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
  }
}

void radix_sort_pairs_8bit(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                           size_t n, uint32_t* scratch, hipStream_t s) {
  if (n == 0) return;
  const uint32_t ntiles = (uint32_t)((n + kTile - 1) / kTile);
  const size_t h = 256 * (size_t)ntiles;
  uint32_t* hist = scratch;
  uint32_t* hs = scratch + h;
  uint32_t* sc = scratch + 2 * h;
  hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kThreads), 0, s, keys_in, n, 0, hist, ntiles);
  exclusive_scan_u32(hist, hs, h, nullptr, sc, s);
  hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kThreads), 0, s, keys_in, vals_in, keys_out, vals_out, n, 0, hs,
                     ntiles);
}

// ---------------------------------------------------------------------------------------------
// Hash-grid keys: a body goes to the finest level whose cell edge is at least its AABB
// size, and to the cell of its AABB's min corner (floor, periodic wrap into the table).
// Planes (infinite AABBs) are kept out of the grid (key = empty, sorted last).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bucket_of(const GridParams& G, int l, const double* a, long long* c) {
  c[0] = (long long)floor(a[0] * G.inv[l]);
  c[1] = (long long)floor(a[1] * G.inv[l]);
  c[2] = (long long)floor(a[2] * G.inv[l]);
  return ((uint32_t)(c[2] & (G.dimz - 1)) * G.dimy + (uint32_t)(c[1] & (G.dimy - 1))) * G.dimx +
         (uint32_t)(c[0] & (G.dimx - 1));
}

__global__ void k_cell_keys(Bodies B, GridParams G, uint32_t* keys, uint32_t* vals, uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n) return;
  vals[i] = i;
  if (B.shape[i] == kShapePlane) { keys[i] = kEmpty; return; }
  const double* a = B.aabb + 6 * (size_t)i;
  const double size = aabb_size(a);
  if (!(isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]) && isfinite(size))) {
    atomicOr(flags, ERR_NONFINITE);
    keys[i] = kEmpty;
    return;
  }
  // A body fills at most kSpanFill of its cell edge, so floor(x * (1/span)) of the min
  // corners of two overlapping bodies differ by at most one even after rounding.
  int l = 0;
  while (l < G.nlevels - 1 && size > G.span[l] * kSpanFill) ++l;
  if (size > G.span[l] * kSpanFill) atomicOr(flags, ERR_LEVEL_OVERFLOW);
  long long c[3];
  keys[i] = (uint32_t)l * G.table + bucket_of(G, l, a, c);
}

__global__ void k_cell_ranges(const uint32_t* keys, uint32_t ngrid, uint32_t* start, uint32_t* end) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ngrid) return;
  const uint32_t k = keys[s];
  if (k == kEmpty) return;  // a non-finite body (the step reports it and stops)
  if (s == 0 || keys[s - 1] != k) start[k] = s;
  if (s == ngrid - 1 || keys[s + 1] != k) end[k] = s + 1;
}

struct CountEmit {
  uint32_t n = 0;
  __device__ void operator()(uint32_t, uint32_t, d3, d3, double) { ++n; }
};

// Counts a body's contacts and stages the first kStage of them: the geometry ids, point, normal
// and distance (k_pairs_place adds the top-level bodies and the friction, as WriteEmit does).
struct StageEmit {
  const BroadphaseBuffers* P;
  uint32_t base;   // slot * kStage
  uint32_t n = 0;
  __device__ void operator()(uint32_t g1, uint32_t g2, d3 gpos, d3 nrm, double dist) {
    if (n < kStage) {
      const uint32_t at = base + n;
      P->st_b1[at] = g1;
      P->st_b2[at] = g2;
      st3(P->st_pos, at, gpos);
      st3(P->st_nrm, at, nrm);
      P->st_dist[at] = dist;
    }
    ++n;
  }
};

struct WriteEmit {
  const Bodies* B;
  const RawContacts* R;
  const double* csf;
  const double* cdf;
  uint32_t nmat;
  uint32_t at;
  // g1, g2: geometry ids (a body, or a union member); the contact joins their top-level bodies
  __device__ void operator()(uint32_t g1, uint32_t g2, d3 gpos, d3 n, double dist) {
    const double mu = contact_friction(*B, csf, cdf, nmat, g1, g2, gpos, n);
    R->b1[at] = geom_body(*B, g1);
    R->b2[at] = geom_body(*B, g2);
    st3(R->pos, at, gpos);
    st3(R->nrm, at, n);
    R->dist[at] = dist;
    R->mu[at] = mu;
    ++at;
  }
};

// Pair enumeration for the body in sorted slot s.  Same-level pairs are taken from the
// 27-cell neighbourhood and emitted by the body with the smaller index; a body also
// scans the 27-cell neighbourhood of its min-corner cell in every coarser level; then it
// is tested against every plane.  The filter is pe's: not both fixed, overlapping AABBs.
// Candidates are read from the sorted-order copies (P.saabb, P.sfixed): the bodies of one cell
// are consecutive there, so a cell's boxes are contiguous instead of gathered by body index.
// The narrowphase of a scene holding spheres and planes only: the calls narrowphase() makes for
// a grid sphere a against a sphere or a plane b, without the other shapes' routines, whose
// registers and stack would otherwise size the pair kernels.
template <class Emit>
__device__ __forceinline__ void narrowphase_spheres(const Bodies& B, uint32_t a, uint32_t b, Emit& emit) {
  if (B.shape[b] == kShapePlane) np_sphere_plane(B, body_sphere(B, a), b, emit);
  else np_sphere_sphere(body_sphere(B, a), body_sphere(B, b), emit);
}

template <bool SPHERES, class Emit>
__device__ void enumerate_pairs(const Bodies& B, const GridParams& G, const BroadphaseBuffers& P,
                                uint32_t s, Emit& emit, uint32_t* flags) {
  const uint32_t a = P.vals[s];
  const uint32_t la = P.keys[s] / G.table;
  const size_t W = P.stride;
  double abox[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) abox[k] = P.saabb[k * W + s];
  const bool fa = P.sfixed[s] != 0;
  for (int l = (int)la; l < G.nlevels; ++l) {
    long long c[3];
    bucket_of(G, l, abox, c);
    // The 27 cells' start slots are loaded together (independent loads, one round trip) into a
    // mask of the non-empty cells; the walk then takes those in the order dz, dy, dx.
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
      const uint32_t bucket = ((uint32_t)((c[2] + dz) & (G.dimz - 1)) * G.dimy +
                               (uint32_t)((c[1] + dy) & (G.dimy - 1))) * G.dimx +
                              (uint32_t)((c[0] + dx) & (G.dimx - 1));
      if (P.cell_start[(uint32_t)l * G.table + bucket] != kEmpty) mask |= 1u << k;
    }
    while (mask) {
      const int k = __builtin_ctz(mask);
      mask &= mask - 1u;
      const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
      const uint32_t bucket = ((uint32_t)((c[2] + dz) & (G.dimz - 1)) * G.dimy +
                               (uint32_t)((c[1] + dy) & (G.dimy - 1))) * G.dimx +
                              (uint32_t)((c[0] + dx) & (G.dimx - 1));
      const uint32_t key = (uint32_t)l * G.table + bucket;
      const uint32_t st = P.cell_start[key], en = P.cell_end[key];
      for (uint32_t t = st; t < en; ++t) {
        const uint32_t b = P.vals[t];
        if (l == (int)la && b <= a) continue;
        if (fa && P.sfixed[t]) continue;
        double bbox[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) bbox[k] = P.saabb[k * W + t];
        if (!aabb_overlap(abox, bbox)) continue;
        if constexpr (SPHERES) narrowphase_spheres(B, a, b, emit);
        else narrowphase(B, a, b, emit);
      }
    }
  }
  for (uint32_t k = 0; k < P.nplanes; ++k) {
    const uint32_t p = P.planes[k];
    if (fa) continue;  // planes are fixed
    if (!aabb_overlap(abox, B.aabb + 6 * (size_t)p)) continue;
    if constexpr (SPHERES) narrowphase_spheres(B, a, p, emit);
    else narrowphase(B, a, p, emit);
  }
}

template <bool SPHERES>
__global__ void k_pairs_count(Bodies B, GridParams G, BroadphaseBuffers P, uint32_t ngrid) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ngrid) return;
  StageEmit e{&P, s * kStage};
  enumerate_pairs<SPHERES>(B, G, P, s, e, P.flags);
  P.counts[s] = e.n;
  if (e.n > kStage) atomicOr(P.flags, (uint32_t)ERR_STAGE_FULL);
}

template <bool SPHERES>
__global__ void k_pairs_write(Bodies B, GridParams G, BroadphaseBuffers P, uint32_t ngrid, RawContacts R,
                              const double* csf, const double* cdf, uint32_t nmat) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  // the bodies whose contacts the count pass could not stage: the walk again, now writing (the
  // enumeration is deterministic, so it emits the same contacts in the same order).  (Measured
  // when this pass wrote every body's contacts: walking a compacted list of the bodies with
  // contacts took longer than skipping the others, 286 against 163 us on the 200k pile.)
  if (s >= ngrid || P.counts[s] <= kStage) return;
  WriteEmit e{&B, &R, csf, cdf, nmat, P.offsets[s]};
  enumerate_pairs<SPHERES>(B, G, P, s, e, P.flags);
}

__global__ void k_pairs_place(Bodies B, uint32_t ngrid, BroadphaseBuffers P, RawContacts R, const double* csf,
                              const double* cdf, uint32_t nmat) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ngrid) return;
  const uint32_t n = P.counts[s];
  if (n == 0 || n > kStage) return;
  const uint32_t at = P.offsets[s], from = s * kStage;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t g1 = P.st_b1[from + k], g2 = P.st_b2[from + k];
    const d3 gpos = mk3(P.st_pos[3 * (size_t)(from + k)], P.st_pos[3 * (size_t)(from + k) + 1],
                        P.st_pos[3 * (size_t)(from + k) + 2]);
    const d3 nrm = mk3(P.st_nrm[3 * (size_t)(from + k)], P.st_nrm[3 * (size_t)(from + k) + 1],
                       P.st_nrm[3 * (size_t)(from + k) + 2]);
    R.b1[at + k] = geom_body(B, g1);
    R.b2[at + k] = geom_body(B, g2);
    st3(R.pos, at + k, gpos);
    st3(R.nrm, at + k, nrm);
    R.dist[at + k] = P.st_dist[from + k];
    R.mu[at + k] = contact_friction(B, csf, cdf, nmat, g1, g2, gpos, nrm);
  }
}

__global__ void k_sorted_boxes(Bodies B, BroadphaseBuffers P, uint32_t ngrid) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ngrid) return;
  const uint32_t b = P.vals[s];
  const double* a = B.aabb + 6 * (size_t)b;
#pragma unroll
  for (int k = 0; k < 6; ++k) P.saabb[(size_t)k * P.stride + s] = a[k];
  P.sfixed[s] = B.fixed[b];
}

void launch_pairs_place(const Bodies& B, const BroadphaseBuffers& P, const RawContacts& R, const double* csf,
                        const double* cdf, uint32_t nmat, hipStream_t s) {
  const uint32_t ngrid = B.n - P.nplanes;
  if (ngrid)
    hipLaunchKernelGGL(k_pairs_place, dim3((ngrid + 255) / 256), dim3(256), 0, s, B, ngrid, P, R, csf, cdf, nmat);
}

void launch_cell_keys(const Bodies& B, const GridParams& G, BroadphaseBuffers& P, hipStream_t s) {
  if (B.n) hipLaunchKernelGGL(k_cell_keys, dim3((B.n + 255) / 256), dim3(256), 0, s, B, G, P.keys, P.vals, P.flags);
}

void launch_cell_ranges(uint32_t n, const GridParams& G, BroadphaseBuffers& P, hipStream_t s) {
  fill_u32(P.cell_start, kEmpty, (size_t)G.nlevels * G.table, s);
  const uint32_t ngrid = n - P.nplanes;
  if (ngrid) hipLaunchKernelGGL(k_cell_ranges, dim3((ngrid + 255) / 256), dim3(256), 0, s, P.keys, ngrid, P.cell_start, P.cell_end);
}

void launch_sorted_boxes(const Bodies& B, BroadphaseBuffers& P, hipStream_t s) {
  const uint32_t ngrid = B.n - P.nplanes;
  if (ngrid) hipLaunchKernelGGL(k_sorted_boxes, dim3((ngrid + 255) / 256), dim3(256), 0, s, B, P, ngrid);
}

void launch_pairs(const Bodies& B, const GridParams& G, BroadphaseBuffers& P, const RawContacts& R,
                  const double* csf, const double* cdf, uint32_t nmat, bool write, bool spheres_only, hipStream_t s) {
  const uint32_t ngrid = B.n - P.nplanes;
  if (!ngrid) return;
  const dim3 g((ngrid + 127) / 128), t(128);
  if (!write) {
    if (spheres_only) hipLaunchKernelGGL(k_pairs_count<true>, g, t, 0, s, B, G, P, ngrid);
    else hipLaunchKernelGGL(k_pairs_count<false>, g, t, 0, s, B, G, P, ngrid);
  } else {
    if (spheres_only) hipLaunchKernelGGL(k_pairs_write<true>, g, t, 0, s, B, G, P, ngrid, R, csf, cdf, nmat);
    else hipLaunchKernelGGL(k_pairs_write<false>, g, t, 0, s, B, G, P, ngrid, R, csf, cdf, nmat);
  }
}

}  // namespace pe_amd
