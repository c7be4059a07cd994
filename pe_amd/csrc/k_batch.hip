// Batched scenes: one workgroup advances one scene through whole time steps, and a
// persistent grid of workgroups pulls scenes from a queue (largest first).  Per scene
// and step, as CollisionSystem<HCSITS>::simulationStep (HCSITS.h:1472-1531):
//   detection    counting sort of the grid bodies into an LDS bucket table (one periodic
//                grid level whose cell edge bounds every AABB), 27-bucket neighbourhood
//                and plane tests, narrowphase -- every pair that is not both fixed and
//                whose closed AABBs overlap is handed to the narrowphase once
//                (HashGrids.h:1121-1358 semantics); count pass, scan, write pass, so the
//                contact order is deterministic;
//   body caching predicted velocities and solver body records;
//   colouring    Jones-Plassmann rounds with LDS claims (a contact whose priority is the
//                smallest among the uncoloured contacts of both its non-fixed bodies takes
//                the round's colour), then a stable colour-major ordering;
//   caching      contact records in sweep order (HCSITS.h:1741-1818);
//   sweeps       K sweeps, one workgroup barrier per colour (HCSITS.h:1891-1947);
//   integration  HCSITS.h:1954-2004, 4135-4184.
// Colours, sweep order and results do not depend on the workgroup size or grid.
#include "batch.hpp"
#include "step_ops.hpp"

namespace pe_amd {

namespace {

constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr double kBatchSpanFill = 0.999999;

// Inclusive/exclusive scan of one value per thread over the workgroup (blockDim.x a
// multiple of 64, at most 512).  red needs 8 words.
__device__ __forceinline__ uint32_t wg_exclusive_scan(uint32_t x, uint32_t* red, uint32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t v = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (lane == 63) red[wid] = v;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int k = 0; k < nw; ++k) {
    const uint32_t s = red[k];
    if (k < wid) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return v - x + pre;
}

__device__ __forceinline__ uint32_t wg_sum(uint32_t x, uint32_t* red) {
  uint32_t tot;
  wg_exclusive_scan(x, red, &tot);
  return tot;
}

// Workgroup max of non-negative doubles (wavefront shuffles, then one LDS word per wave).
__device__ __forceinline__ double wg_max(double x, unsigned long long* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(u, d, 64);
    u = o > u ? o : u;
  }
  if (lane == 0) red[wid] = u;
  __syncthreads();
  unsigned long long m = 0;
  for (int k = 0; k < nw; ++k) m = red[k] > m ? red[k] : m;
  __syncthreads();
  return __longlong_as_double((long long)m);
}

// Bijective mix of a contact's scene-local index: distinct, well-spread colouring priorities.
__device__ __forceinline__ uint32_t colour_priority(uint32_t c) {
  uint32_t h = c * 0x9E3779B1u;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  return h;
}

struct SceneCountEmit {
  uint32_t n = 0;
  __device__ void operator()(uint32_t, uint32_t, d3, d3, double) { ++n; }
};

struct SceneWriteEmit {
  const Bodies* B;
  const RawContacts* R;
  const BatchArgs* A;
  const SceneInfo* S;
  uint32_t at;
  __device__ void operator()(uint32_t b1, uint32_t b2, d3 gpos, d3 n, double dist) {
    double mu;
    if (A->scene_mu) mu = contact_is_dynamic(*B, b1, b2, gpos, n) ? S->mu_d : S->mu_s;
    else mu = contact_friction(*B, A->csf, A->cdf, A->nmat, b1, b2, gpos, n);
    R->b1[at] = b1;
    R->b2[at] = b2;
    st3(R->pos, at, gpos);
    st3(R->nrm, at, n);
    R->dist[at] = dist;
    R->mu[at] = mu;
    ++at;
  }
};

__device__ __forceinline__ uint32_t scene_bucket(const SceneInfo& S, const double* a) {
  const long long cx = (long long)floor(a[0] * S.inv);
  const long long cy = (long long)floor(a[1] * S.inv);
  const long long cz = (long long)floor(a[2] * S.inv);
  const uint32_t mx = (1u << S.lx) - 1, my = (1u << S.ly) - 1, mz = (1u << S.lz) - 1;
  return (((uint32_t)cz & mz) << (S.lx + S.ly)) | (((uint32_t)cy & my) << S.lx) | ((uint32_t)cx & mx);
}

// Pairs of grid body a (same rules as the single-world broadphase): bodies of the 27
// neighbouring buckets with a larger index, then the scene's planes.
template <class Emit>
__device__ void scene_pairs(const Bodies& B, const BatchArgs& A, const SceneInfo& S, const uint32_t* tab,
                            uint32_t a, Emit& emit) {
  const double* aa = B.aabb + 6 * (size_t)a;
  double abox[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) abox[k] = aa[k];
  const bool fa = B.fixed[a] != 0;
  const long long cx = (long long)floor(abox[0] * S.inv);
  const long long cy = (long long)floor(abox[1] * S.inv);
  const long long cz = (long long)floor(abox[2] * S.inv);
  const uint32_t mx = (1u << S.lx) - 1, my = (1u << S.ly) - 1, mz = (1u << S.lz) - 1;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const uint32_t key = (((uint32_t)(cz + dz) & mz) << (S.lx + S.ly)) |
                             (((uint32_t)(cy + dy) & my) << S.lx) | ((uint32_t)(cx + dx) & mx);
        const uint32_t st = tab[key], en = tab[key + 1];
        for (uint32_t t = st; t < en; ++t) {
          const uint32_t b = A.sorted[S.body0 + t];
          if (b <= a) continue;
          if (fa && B.fixed[b]) continue;
          if (!aabb_overlap(abox, B.aabb + 6 * (size_t)b)) continue;
          narrowphase(B, a, b, emit);
        }
      }
  if (fa) return;  // planes are fixed
  for (uint32_t k = 0; k < S.nplane; ++k) {
    const uint32_t p = A.plane_idx[S.plane0 + k];
    if (!aabb_overlap(abox, B.aabb + 6 * (size_t)p)) continue;
    narrowphase(B, a, p, emit);
  }
}

// One time step of scene sc.  Returns false (uniformly) if the scene hit an error; its
// flags are then set and it does not advance.
template <int MODEL, class Trig>
__device__ bool scene_step(const Bodies& B, const RawContacts& R, const SolverContacts& C, const BatchArgs& A,
                           const SceneInfo& S, uint32_t sc, uint32_t* tab, uint32_t* coff, uint32_t* red,
                           unsigned long long* red64, uint32_t* lflag) {
  const uint32_t T = blockDim.x, tid = threadIdx.x;
  const uint32_t b0 = S.body0, b1 = S.body0 + S.nbody;
  const uint32_t table = 1u << (S.lx + S.ly + S.lz);
  SceneOut& O = A.out[sc];

  // ---- detection: bucket counts
  for (uint32_t t = tid; t <= table; t += T) tab[t] = 0;
  if (tid == 0) {
    *lflag = 0;
    O.counters[0] = 0;
    O.counters[1] = 0;
  }
  __syncthreads();
  uint32_t ngrid_local = 0;
  for (uint32_t i = b0 + tid; i < b1; i += T) {
    if (B.shape[i] == kShapePlane) { A.bucket[i] = kEmptyKey; continue; }
    const double* a = B.aabb + 6 * (size_t)i;
    const double size = aabb_size(a);
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) finite = finite && isfinite(a[k]);
    if (!finite) { atomicOr(lflag, ERR_NONFINITE); A.bucket[i] = kEmptyKey; continue; }
    if (size > S.span * kBatchSpanFill) atomicOr(lflag, ERR_LEVEL_OVERFLOW);
    const uint32_t key = scene_bucket(S, a);
    A.bucket[i] = key;
    atomicAdd(&tab[key], 1u);
    ++ngrid_local;
  }
  const uint32_t ngrid = wg_sum(ngrid_local, red);  // (barriers inside)
  if (*lflag) {
    if (tid == 0) { O.flags |= *lflag; atomicOr(A.any_flags, *lflag); }
    return false;
  }
  // inclusive bucket ends, each thread a contiguous run of the table
  {
    const uint32_t per = (table + T - 1) / T;
    const uint32_t lo = tid * per, hi = min(lo + per, table);
    uint32_t sum = 0;
    for (uint32_t t = lo; t < hi; ++t) sum += tab[t];
    uint32_t tot;
    uint32_t run = wg_exclusive_scan(sum, red, &tot);
    for (uint32_t t = lo; t < hi; ++t) {
      run += tab[t];
      tab[t] = run;
    }
  }
  __syncthreads();
  for (uint32_t i = b0 + tid; i < b1; i += T) {
    const uint32_t key = A.bucket[i];
    if (key == kEmptyKey) continue;
    const uint32_t pos = atomicSub(&tab[key], 1u) - 1u;
    A.sorted[b0 + pos] = i;
  }
  if (tid == 0) tab[table] = ngrid;
  __syncthreads();
  // bucket lists in body order (the placement above is in arrival order)
  for (uint32_t t = tid; t < table; t += T) {
    const uint32_t st = tab[t], en = tab[t + 1];
    for (uint32_t x = st + 1; x < en; ++x) {
      const uint32_t v = A.sorted[b0 + x];
      uint32_t y = x;
      while (y > st && A.sorted[b0 + y - 1] > v) {
        A.sorted[b0 + y] = A.sorted[b0 + y - 1];
        --y;
      }
      A.sorted[b0 + y] = v;
    }
  }
  __syncthreads();

  // ---- detection: count, scan, write
  for (uint32_t i = b0 + tid; i < b1; i += T) {
    uint32_t n = 0;
    if (A.bucket[i] != kEmptyKey) {
      SceneCountEmit e;
      scene_pairs(B, A, S, tab, i, e);
      n = e.n;
    }
    A.count[i] = n;
  }
  __syncthreads();
  uint32_t nc = 0;
  for (uint32_t base = b0; base < b1; base += T) {
    const uint32_t i = base + tid;
    const uint32_t x = i < b1 ? A.count[i] : 0u;
    uint32_t tot;
    const uint32_t ex = wg_exclusive_scan(x, red, &tot);
    if (i < b1) A.count[i] = nc + ex;
    nc += tot;
  }
  if (nc > S.ccap) {
    if (tid == 0) { O.flags |= ERR_CAPACITY; atomicOr(A.any_flags, ERR_CAPACITY); }
    return false;
  }
  for (uint32_t i = b0 + tid; i < b1; i += T) {
    if (A.bucket[i] == kEmptyKey) continue;
    SceneWriteEmit e{&B, &R, &A, &S, S.c0 + A.count[i]};
    scene_pairs(B, A, S, tab, i, e);
  }
  // ---- body caching
  for (uint32_t i = b0 + tid; i < b1; i += T) predict_body(B, i, A.gravity, A.dt);
  __syncthreads();

  // ---- colouring (claims in LDS, indexed by the scene-local body)
  const uint32_t c0 = S.c0;
  for (uint32_t c = tid; c < nc; c += T) A.colour[c0 + c] = kUncoloured;
  uint32_t ncol = 0;
  for (;;) {
    for (uint32_t i = tid; i < S.nbody; i += T) tab[i] = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t c = tid; c < nc; c += T) {
      if (A.colour[c0 + c] != kUncoloured) continue;
      const uint32_t p = colour_priority(c);
      const uint32_t e1 = R.b1[c0 + c], e2 = R.b2[c0 + c];
      if (!B.fixed[e1]) atomicMin(&tab[e1 - b0], p);
      if (!B.fixed[e2]) atomicMin(&tab[e2 - b0], p);
    }
    __syncthreads();
    uint32_t left = 0;
    for (uint32_t c = tid; c < nc; c += T) {
      if (A.colour[c0 + c] != kUncoloured) continue;
      const uint32_t p = colour_priority(c);
      const uint32_t e1 = R.b1[c0 + c], e2 = R.b2[c0 + c];
      const bool win = (B.fixed[e1] || tab[e1 - b0] == p) && (B.fixed[e2] || tab[e2 - b0] == p);
      if (win) A.colour[c0 + c] = ncol;
      else ++left;
    }
    ++ncol;
    const uint32_t remaining = wg_sum(left, red);
    if (remaining == 0) break;
    if (ncol >= kMaxColours) {
      if (tid == 0) { O.flags |= ERR_COLOURS; atomicOr(A.any_flags, ERR_COLOURS); }
      return false;
    }
  }
  if (nc == 0) ncol = 0;
  // colour offsets
  for (uint32_t k = tid; k <= kMaxColours; k += T) coff[k] = 0;
  __syncthreads();
  for (uint32_t c = tid; c < nc; c += T) atomicAdd(&coff[A.colour[c0 + c] + 1], 1u);
  __syncthreads();
  if (tid == 0)
    for (uint32_t k = 1; k <= ncol; ++k) coff[k] += coff[k - 1];
  __syncthreads();
  // stable colour-major order: per chunk of T contacts, rank within the wave by colour
  // (7 ballots), per-wave counts in LDS, running cursors per colour
  {
    const uint32_t nw = T >> 6, lane = tid & 63, wid = tid >> 6;
    uint32_t* wc = tab;                        // nw x kMaxColours
    uint32_t* cursor = tab + nw * kMaxColours; // kMaxColours
    for (uint32_t k = tid; k < kMaxColours; k += T) cursor[k] = 0;
    for (uint32_t base = 0; base < nc; base += T) {
      for (uint32_t k = tid; k < nw * kMaxColours; k += T) wc[k] = 0;
      __syncthreads();
      const uint32_t c = base + tid;
      const bool valid = c < nc;
      const uint32_t col = valid ? A.colour[c0 + c] : 0u;
      unsigned long long mask = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 7; ++bit) {
        const unsigned long long bb = __ballot(((col >> bit) & 1u) != 0);
        mask &= ((col >> bit) & 1u) ? bb : ~bb;
      }
      const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
      if (valid && rank == 0) wc[wid * kMaxColours + col] = (uint32_t)__popcll(mask);
      __syncthreads();
      for (uint32_t k = tid; k < ncol; k += T) {
        uint32_t run = cursor[k];
        for (uint32_t w = 0; w < nw; ++w) {
          const uint32_t x = wc[w * kMaxColours + k];
          wc[w * kMaxColours + k] = run;
          run += x;
        }
        cursor[k] = run;
      }
      __syncthreads();
      if (valid) A.perm[c0 + coff[col] + wc[wid * kMaxColours + col] + rank] = c0 + c;
      __syncthreads();
    }
  }
  if (tid == 0)
    for (uint32_t k = 0; k <= ncol; ++k) A.colour_off[(size_t)sc * (kMaxColours + 1) + k] = coff[k];

  // ---- contact caching in sweep order (a slot whose contact index is out of range -- an
  // internal error -- stops the scene before any sweep reads it)
  for (uint32_t s = tid; s < nc; s += T) {
    const uint32_t c = A.perm[c0 + s];
    if (c - c0 >= nc) { atomicOr(lflag, ERR_INTERNAL); continue; }
    setup_contact(B, R, c, c0 + s, C, A.erp, A.dtinv, O.counters);
  }
  __syncthreads();
  if (*lflag) {
    if (tid == 0) { O.flags |= *lflag; atomicOr(A.any_flags, *lflag); }
    return false;
  }

  // ---- sweeps
  uint32_t iters = 0;
  double last = 0;
  if (nc) {
    for (int it = 0; it < A.iters; ++it) {
      double dmax = 0;
      for (uint32_t k = 0; k < ncol; ++k) {
        const uint32_t en = coff[k + 1];
        for (uint32_t s = coff[k] + tid; s < en; s += T)
          dmax = smax(dmax, relax_contact<MODEL, Trig>(C, c0 + s, B.sb, B.dvw, A.dtinv, A.max_sub));
        __syncthreads();
      }
      iters = it + 1;
      if (A.thr > 0 || it + 1 == A.iters) {
        last = wg_max(dmax, red64);
        if (A.thr > 0 && last < A.thr) break;
      }
    }
  }

  // ---- integration
  for (uint32_t i = b0 + tid; i < b1; i += T) integrate_body<Trig>(B, i, A.dt, A.damping);
  __syncthreads();
  if (tid == 0) {
    O.nc = nc;
    O.ncol = ncol;
    O.iters = iters;
    O.delta = last;
  }
  return true;
}

}  // namespace

template <int MODEL, class Trig>
__global__ void __launch_bounds__(kBatchMaxThreads) k_batch(Bodies B, RawContacts R, SolverContacts C, BatchArgs A) {
  __shared__ uint32_t tab[kBatchTable + 1];
  __shared__ uint32_t coff[kMaxColours + 1];
  __shared__ uint32_t red[16];
  __shared__ unsigned long long red64[8];
  __shared__ uint32_t lflag;
  __shared__ uint32_t lscene;
  const uint32_t target = A.queue[1];
  for (;;) {
    if (threadIdx.x == 0) lscene = atomicAdd(&A.queue[0], 1u);
    __syncthreads();
    const uint32_t q = lscene;
    __syncthreads();
    if (q >= A.nscenes) break;
    const uint32_t sc = A.order[q];
    const SceneInfo S = A.info[sc];
    uint32_t done = A.out[sc].steps;
    if (A.out[sc].flags & ~ERR_CAPACITY) continue;  // a failed scene stays where it failed
    __syncthreads();
    if (threadIdx.x == 0) A.out[sc].flags = 0;
    while (done < target) {
      if (!scene_step<MODEL, Trig>(B, R, C, A, S, sc, tab, coff, red, red64, &lflag)) break;
      ++done;
      if (threadIdx.x == 0) A.out[sc].steps = done;
    }
    __syncthreads();
  }
}

__global__ void k_batch_begin(uint32_t* queue, uint32_t add_steps) {
  if (threadIdx.x == 0) {
    queue[0] = 0;
    queue[1] += add_steps;
  }
}

void launch_batch_begin(uint32_t* queue, uint32_t add_steps, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, s, queue, add_steps);
}

template <class Trig>
static void batch_dispatch(int model, const Bodies& B, const RawContacts& R, const SolverContacts& C,
                           const BatchArgs& A, uint32_t grid, uint32_t threads, hipStream_t s) {
  const dim3 g(grid), b(threads);
  switch (model) {
    case 0: hipLaunchKernelGGL((k_batch<0, Trig>), g, b, 0, s, B, R, C, A); break;
    case 1: hipLaunchKernelGGL((k_batch<1, Trig>), g, b, 0, s, B, R, C, A); break;
    case 2: hipLaunchKernelGGL((k_batch<2, Trig>), g, b, 0, s, B, R, C, A); break;
    case 4: hipLaunchKernelGGL((k_batch<4, Trig>), g, b, 0, s, B, R, C, A); break;
    case 5: hipLaunchKernelGGL((k_batch<5, Trig>), g, b, 0, s, B, R, C, A); break;
    default: hipLaunchKernelGGL((k_batch<3, Trig>), g, b, 0, s, B, R, C, A); break;
  }
}

void launch_batch(int model, bool exact_trig, const Bodies& B, const RawContacts& R, const SolverContacts& C,
                  const BatchArgs& A, uint32_t grid, uint32_t threads, hipStream_t s) {
  if (exact_trig) batch_dispatch<TrigExact>(model, B, R, C, A, grid, threads, s);
  else batch_dispatch<TrigFast>(model, B, R, C, A, grid, threads, s);
}

}  // namespace pe_amd
