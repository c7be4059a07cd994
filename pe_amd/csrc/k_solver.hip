// Solver kernels: contact-graph colouring, contact caching and the NBGS sweeps.
#include <algorithm>

#include "step_ops.hpp"

namespace pe_amd {

// ---------------------------------------------------------------------------------------------
// Conflict graph.  Two contacts conflict when they share a non-fixed body (fixed bodies have
// zero inverse mass and inertia, so their velocity corrections never change).  The graph is
// held as body -> contacts adjacency lists.
// ---------------------------------------------------------------------------------------------
__global__ void k_adj_count(RawContacts R, const uint8_t* fixed, uint32_t nc, uint32_t* deg) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t b1 = R.b1[c], b2 = R.b2[c];
  if (!fixed[b1]) atomicAdd(&deg[b1], 1u);
  if (!fixed[b2]) atomicAdd(&deg[b2], 1u);
}

__global__ void k_adj_fill(RawContacts R, const uint8_t* fixed, uint32_t nc, const uint32_t* start,
                           uint32_t* cursor, uint32_t* adj) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t b1 = R.b1[c], b2 = R.b2[c];
  if (!fixed[b1]) adj[start[b1] + atomicAdd(&cursor[b1], 1u)] = c;
  if (!fixed[b2]) adj[start[b2] + atomicAdd(&cursor[b2], 1u)] = c;
}

void launch_adjacency(const Bodies& B, const RawContacts& R, uint32_t nc, ColourBuffers& C,
                      uint32_t* scan_scratch, hipStream_t s) {
  fill_u32(C.deg, 0u, B.n, s);
  fill_u32(C.adj_cursor, 0u, B.n, s);
  if (nc) hipLaunchKernelGGL(k_adj_count, dim3((nc + 255) / 256), dim3(256), 0, s, R, B.fixed, nc, C.deg);
  exclusive_scan_u32(C.deg, C.adj_start, B.n, C.adj_start + B.n, scan_scratch, s);
  if (nc) hipLaunchKernelGGL(k_adj_fill, dim3((nc + 255) / 256), dim3(256), 0, s, R, B.fixed, nc, C.adj_start,
                             C.adj_cursor, C.adj);
}

// Priority of a contact: conflict degree first (largest-degree-first keeps the colour count
// low), then a hash of the index; ties are impossible because the index is unique.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint64_t prio_key(uint32_t deg, uint32_t c) {
  return ((uint64_t)deg << 32) | (uint64_t)mix32(c);
}

__global__ void k_colour_init(RawContacts R, const uint8_t* fixed, uint32_t nc, const uint32_t* start,
                              uint32_t* prio, uint32_t* colour, uint32_t* cround, uint32_t* remaining) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *remaining = nc;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t b1 = R.b1[c], b2 = R.b2[c];
  uint32_t d = 0;
  if (!fixed[b1]) d += start[b1 + 1] - start[b1] - 1;
  if (!fixed[b2]) d += start[b2 + 1] - start[b2] - 1;
  prio[c] = d;
  colour[c] = kUncoloured;
  cround[c] = kUncoloured;
}

// One Jones-Plassmann round r: an uncoloured contact whose priority beats every uncoloured
// neighbour takes the smallest colour that no neighbour coloured in an earlier round has.
// A neighbour coloured in this same round counts as uncoloured (it was a local maximum, so
// it outranks this contact), hence two neighbours are never coloured in one round and the
// result depends neither on thread timing nor on the order of the adjacency lists.
__global__ void k_colour_round(RawContacts R, const uint8_t* fixed, uint32_t nc, const uint32_t* start,
                               const uint32_t* adj, const uint32_t* prio, uint32_t* colour, uint32_t* cround,
                               uint32_t round, uint32_t* remaining, uint32_t* flags) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  if (cround[c] < round) return;
  const uint64_t my = prio_key(prio[c], c);
  uint64_t used[kMaxColours / 64] = {0, 0};
  const uint32_t ends[2] = {R.b1[c], R.b2[c]};
  for (int e = 0; e < 2; ++e) {
    const uint32_t b = ends[e];
    if (fixed[b]) continue;
    for (uint32_t k = start[b]; k < start[b + 1]; ++k) {
      const uint32_t o = adj[k];
      if (o == c) continue;
      if (cround[o] < round) {
        const uint32_t oc = colour[o];
        used[oc >> 6] |= 1ull << (oc & 63);
      } else if (prio_key(prio[o], o) > my) {
        return;  // a neighbour goes first
      }
    }
  }
  uint32_t pick = kMaxColours;
  for (uint32_t w = 0; w < kMaxColours / 64; ++w) {
    if (~used[w]) { pick = w * 64 + (uint32_t)__builtin_ctzll(~used[w]); break; }
  }
  if (pick >= kMaxColours) { atomicOr(flags, ERR_COLOURS); pick = kMaxColours - 1; }
  colour[c] = pick;
  cround[c] = round;
  atomicSub(remaining, 1u);
}

// Colour histogram, and the last round that coloured a contact (max of cround over the coloured
// contacts; uncoloured ones hold kUncoloured and are skipped).
__global__ void k_colour_hist(const uint32_t* colour, const uint32_t* cround, uint32_t nc, uint32_t* hist,
                              uint32_t* last_round) {
  __shared__ uint32_t h[kMaxColours];
  __shared__ uint32_t lr;
  if (threadIdx.x < kMaxColours) h[threadIdx.x] = 0;
  if (threadIdx.x == 0) lr = 0;
  __syncthreads();
  uint32_t mr = 0;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    const uint32_t k = colour[c];
    if (k == kUncoloured) continue;
    atomicAdd(&h[k], 1u);
    mr = max(mr, cround[c]);
  }
  atomicMax(&lr, mr);
  __syncthreads();
  if (threadIdx.x < kMaxColours && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 0 && lr) atomicMax(last_round, lr);
}

void launch_colour_init(const RawContacts& R, const Bodies& B, uint32_t nc, ColourBuffers& C, hipStream_t s) {
  hipLaunchKernelGGL(k_colour_init, dim3((nc + 255) / 256 + 1), dim3(256), 0, s, R, B.fixed, nc, C.adj_start,
                     C.prio, C.colour, C.cround, C.remaining);
}

void launch_colour_round(const RawContacts& R, const Bodies& B, uint32_t nc, ColourBuffers& C, uint32_t round,
                         uint32_t* flags, hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_colour_round, dim3((nc + 255) / 256), dim3(256), 0, s, R, B.fixed, nc, C.adj_start,
                             C.adj, C.prio, C.colour, C.cround, round, C.remaining, flags);
}

void launch_colour_hist(uint32_t nc, ColourBuffers& C, hipStream_t s) {
  fill_u32(C.hist, 0u, kMaxColours, s);
  fill_u32(C.remaining + 1, 0u, 1, s);
  if (nc) {
    uint32_t nb = (nc + 255) / 256;
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_colour_hist, dim3(nb), dim3(256), 0, s, C.colour, C.cround, nc, C.hist, C.remaining + 1);
  }
}

// Contact caching (HCSITS.h:1741-1818) in solver order: slot s <- raw contact perm[s].
__global__ void k_setup(Bodies B, RawContacts R, const uint32_t* perm, uint32_t nc, SolverContacts C,
                        double erp, double dtinv, unsigned long long* counters) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nc) setup_contact(B, R, perm[s], s, C, erp, dtinv, counters);
}

void launch_setup(const Bodies& B, const RawContacts& R, const uint32_t* perm, uint32_t nc,
                  SolverContacts& C, double erp, double dtinv, unsigned long long* maxpen, hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_setup, dim3((nc + 255) / 256), dim3(256), 0, s, B, R, perm, nc, C, erp, dtinv, maxpen);
}

// ---------------------------------------------------------------------------------------------
// Sweeps.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void wave_max_atomic(double v, unsigned long long* out) {
  // non-negative doubles order like their bit patterns
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(u, d, 64);
    u = o > u ? o : u;
  }
  if ((threadIdx.x & 63) == 0 && u) atomicMax(out, u);
}

// One colour: every contact in [begin, end) is independent of the others.  A group of 2^SPEC
// lanes relaxes one contact (the golden-section search runs SPEC iterations per round on it).
template <int MODEL, class Trig, int SPEC>
__global__ void __launch_bounds__(256) k_sweep(SolverContacts C, uint32_t begin, uint32_t end, const double* sb,
                                               double* dvw, double dtinv, int maxSub, unsigned long long* delta,
                                               const uint32_t* stop) {
  if (stop != nullptr && *stop) return;   // an earlier sweep of this step met the abort threshold
  uint32_t i = begin + ((blockIdx.x * blockDim.x + threadIdx.x) >> SPEC);
  // A whole wavefront per contact: the index is wave-uniform, so the contact's and its bodies'
  // records load through the scalar path into SGPRs (reads only), easing VGPR pressure.
  if (SPEC == 6) i = __builtin_amdgcn_readfirstlane(i);
  double d = 0;
  if (i < end) d = relax_contact<MODEL, Trig, SPEC>(C, i, sb, dvw, dtinv, maxSub);
  wave_max_atomic(d, delta);
}

// All sweeps of a step in one workgroup: colour k's contacts are spread over the workgroup's
// groups of 2^SPEC lanes, a barrier separates colours (the contacts of one colour share no
// non-fixed body, so this is the colour launches' computation, bit for bit), and the abort test
// after each sweep needs no host.  For steps too small to fill the device.
template <int MODEL, class Trig, int SPEC, int THREADS>
__global__ void __launch_bounds__(THREADS) k_sweep_block(SolverContacts C, ColourOffsets O, uint32_t iters,
                                                      const double* sb, double* dvw, double dtinv, int maxSub,
                                                      double thr, unsigned long long* deltas) {
  __shared__ unsigned long long wmax[8];
  __shared__ unsigned long long smax;
  constexpr uint32_t G = 1u << SPEC;
  const uint32_t groups = blockDim.x / G, g = threadIdx.x / G;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  uint32_t it = 0;
  while (it < iters) {
    // non-negative doubles (and NaN above +inf) order like their bit patterns, as in the
    // colour launches' atomicMax
    unsigned long long u = 0;
    for (uint32_t k = 0; k < O.ncol; ++k) {
      for (uint32_t i = O.off[k] + g; i < O.off[k + 1]; i += groups) {
        const unsigned long long d =
            (unsigned long long)__double_as_longlong(relax_contact<MODEL, Trig, SPEC>(C, i, sb, dvw, dtinv, maxSub));
        u = d > u ? d : u;
      }
      __syncthreads();
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const unsigned long long o = __shfl_xor(u, d, 64);
      u = o > u ? o : u;
    }
    if (lane == 0) wmax[wave] = u;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = 0;
      for (uint32_t w = 0; w < nwaves; ++w) m = wmax[w] > m ? wmax[w] : m;
      deltas[it] = m;
      smax = m;
    }
    __syncthreads();
    ++it;
    if (thr > 0 && __longlong_as_double((long long)smax) < thr) break;
  }
  if (threadIdx.x == 0) deltas[iters] = it;
}

template <class Trig>
static void block_dispatch(int model, const SolverContacts& C, const ColourOffsets& O, uint32_t iters, const Bodies& B,
                           double dtinv, int maxSub, double thr, unsigned long long* deltas, hipStream_t s) {
  // 512 lanes (two wavefronts per SIMD) keep models 0 and 1 within their registers; model 3's
  // search needs a SIMD's whole register file per wavefront, so 256 lanes
  const dim3 g(1);
  switch (model) {
    case 0: hipLaunchKernelGGL((k_sweep_block<0, Trig, 0, 512>), g, dim3(512), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 1: hipLaunchKernelGGL((k_sweep_block<1, Trig, 0, 512>), g, dim3(512), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 2: hipLaunchKernelGGL((k_sweep_block<2, Trig, 0, 512>), g, dim3(512), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 4: hipLaunchKernelGGL((k_sweep_block<4, Trig, 0, 512>), g, dim3(512), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 5: hipLaunchKernelGGL((k_sweep_block<5, Trig, 4, 256>), g, dim3(256), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    default: hipLaunchKernelGGL((k_sweep_block<3, Trig, 4, 256>), g, dim3(256), 0, s, C, O, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
  }
}

void launch_sweep_block(int model, bool exact_trig, const SolverContacts& C, const ColourOffsets& O, uint32_t iters,
                        const Bodies& B, double dtinv, int maxSub, double thr, unsigned long long* deltas,
                        hipStream_t s) {
  if (exact_trig) block_dispatch<TrigExact>(model, C, O, iters, B, dtinv, maxSub, thr, deltas, s);
  else block_dispatch<TrigFast>(model, C, O, iters, B, dtinv, maxSub, thr, deltas, s);
}

// All sweeps in the given contact order, one contact at a time, by one group of 2^SPEC lanes.
template <int MODEL, class Trig, int SPEC>
__global__ void __launch_bounds__(64) k_sweep_serial(SolverContacts C, uint32_t nc, uint32_t iters, const double* sb, double* dvw,
                               double dtinv, int maxSub, double thr, unsigned long long* deltas) {
  if (threadIdx.x >= (1u << SPEC)) return;
  uint32_t it = 0;
  while (it < iters) {
    double dmax = 0;
    for (uint32_t i = 0; i < nc; ++i) dmax = smax(dmax, relax_contact<MODEL, Trig, SPEC>(C, i, sb, dvw, dtinv, maxSub));
    if (threadIdx.x == 0) deltas[it] = (unsigned long long)__double_as_longlong(dmax);
    ++it;
    if (thr > 0 && dmax < thr) break;
  }
  if (threadIdx.x == 0) deltas[iters] = it;
}

// ---------------------------------------------------------------------------------------------
// Dataflow sweeps.  The coloured sweep order (sweep-major, then colour-major slots) is one
// sequential Gauss-Seidel order; what it fixes is, for every body, the order of the updates that
// touch it.  Here each (sweep, slot) item waits only for the previous updates of its own bodies
// (per-body counters) instead of for a whole colour, so a sliding contact's long golden-section
// chain delays only the contacts that share its bodies.  Results equal the coloured launches bit
// for bit.  Items go round-robin to the wavefronts of a co-resident (cooperative) grid, which
// take them in increasing order: the smallest unfinished item always has its predecessors done
// and its wavefront free, so the sweep progresses; every wait also has a spin limit and a shared
// abort flag, so no wavefront can wait forever.
// ---------------------------------------------------------------------------------------------
__global__ void k_inv_perm(const uint32_t* perm, uint32_t nc, uint32_t* inv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nc) inv[perm[s]] = s;
}

// rank1[s] / rank2[s]: position of slot s among the slots of its body 1 / body 2 (non-fixed).
__global__ void k_body_ranks(const SolverContacts C, const uint8_t* fixed, uint32_t n, const uint32_t* start,
                             const uint32_t* adj, const uint32_t* inv, uint32_t* rank1, uint32_t* rank2) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n || fixed[b]) return;
  const uint32_t st = start[b], en = start[b + 1];
  for (uint32_t j = st; j < en; ++j) {
    const uint32_t sj = inv[adj[j]];
    uint32_t r = 0;
    for (uint32_t k = st; k < en; ++k) r += inv[adj[k]] < sj ? 1u : 0u;
    if ((C.b1[sj] & ~kFixedBit) == b) rank1[sj] = r;
    else rank2[sj] = r;
  }
}

// One look at a polled counter value: reached the target, or not yet (stop: the spin limit is
// exhausted or another wavefront raised the abort flag, checked every 64 polls).
__device__ __forceinline__ bool poll_step(uint32_t seen, uint32_t target, uint32_t& spins, uint32_t spin_limit,
                                          const uint32_t* abort_flag, bool& stop) {
  if (seen >= target) return true;
  stop = ++spins > spin_limit ||
         ((spins & 63u) == 0 &&
          __hip_atomic_load(const_cast<uint32_t*>(abort_flag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  return false;
}

// Polls with relaxed agent-scope loads.  No acquire: on gfx950 an agent-scope acquire
// invalidates the XCD's whole L2, which made every wavefront's reads miss (measured: 121 steps/s
// with an acquire per poll, 187 with one acquire per item, 215 with none).  None is needed for
// what an item reads after its wait: everything the sweeps write and other wavefronts read (the
// velocity corrections, the impulses) is accessed with agent-scope atomics, which are coherent
// across the XCDs per access (sc1 loads and stores); the item's loads of them are issued only
// after the wait loop has seen the counter (its exit depends on the loaded value); and the
// producer waits for its stores' completion before its counter update.  Everything else the
// relaxation reads was written by earlier kernels.
__device__ __forceinline__ bool wait_count(const uint32_t* ctr, uint32_t target, const uint32_t* abort_flag,
                                           uint32_t spin_limit) {
  uint32_t* c = const_cast<uint32_t*>(ctr);
  uint32_t a = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (a >= target) return true;
  uint32_t spins = 0;
  bool stop = false;
  // two polls in flight: the next one is issued before the previous one's value is looked at, so
  // a counter update is seen about half a round trip sooner (the loop alternates two registers:
  // a copy from one to the other would wait for the newer poll)
  for (;;) {
    __builtin_amdgcn_s_sleep(1);
    const uint32_t b = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (poll_step(a, target, spins, spin_limit, abort_flag, stop)) return true;
    if (stop) return false;
    __builtin_amdgcn_s_sleep(1);
    a = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (poll_step(b, target, spins, spin_limit, abort_flag, stop)) return true;
    if (stop) return false;
  }
}

// A contact's fields held in registers (the impulse excluded: it is loaded after the wait).
struct PrefetchedFields {
  double v[CF_COUNT];
  __device__ __forceinline__ double operator()(int k) const { return v[k]; }
};

// relax_contact<..., COHERENT = true> with the read-only inputs already loaded: the velocity
// corrections and the impulse by agent-scope loads, the new ones stored by lane 0 likewise.
template <int MODEL, class Trig, int SPEC>
__device__ __forceinline__ double relax_coherent(const SolverContacts& C, uint32_t i, const PrefetchedFields& cf,
                                                 const RelaxBody& B1, const RelaxBody& B2, uint32_t i1,
                                                 uint32_t i2, bool fx1, bool fx2, double* dvw, int maxSub) {
  const size_t cap = C.cap;
  double* f = C.f;
  double* q1 = dvw + (size_t)kDVW * i1;
  double* q2 = dvw + (size_t)kDVW * i2;
  d3 dv1 = mk3(0, 0, 0), dw1 = mk3(0, 0, 0), dv2 = mk3(0, 0, 0), dw2 = mk3(0, 0, 0);
  if (!fx1) {
    dv1 = mk3(sweep_ld<true>(q1), sweep_ld<true>(q1 + 1), sweep_ld<true>(q1 + 2));
    dw1 = mk3(sweep_ld<true>(q1 + 3), sweep_ld<true>(q1 + 4), sweep_ld<true>(q1 + 5));
  }
  if (!fx2) {
    dv2 = mk3(sweep_ld<true>(q2), sweep_ld<true>(q2 + 1), sweep_ld<true>(q2 + 2));
    dw2 = mk3(sweep_ld<true>(q2 + 3), sweep_ld<true>(q2 + 4), sweep_ld<true>(q2 + 5));
  }
  // the impulse's three components with one load instruction (lanes 0-2), then broadcast
  const uint32_t lane = threadIdx.x & 63u;
  double pl = 0.0;
  if (lane < 3) pl = sweep_ld<true>(f + (CF_P + lane) * cap + i);
  d3 p = mk3(wave_read_lane(pl, 0), wave_read_lane(pl, 1), wave_read_lane(pl, 2));
  const double delta = relax_update<MODEL, Trig, SPEC>(cf, B1, B2, dv1, dw1, dv2, dw2, p, maxSub);
  if (lane == 0) {
    sweep_st<true>(f + CF_P * cap + i, p.x);
    sweep_st<true>(f + (CF_P + 1) * cap + i, p.y);
    sweep_st<true>(f + (CF_P + 2) * cap + i, p.z);
    if (!fx1) {
      sweep_st<true>(q1, dv1.x); sweep_st<true>(q1 + 1, dv1.y); sweep_st<true>(q1 + 2, dv1.z);
      sweep_st<true>(q1 + 3, dw1.x); sweep_st<true>(q1 + 4, dw1.y); sweep_st<true>(q1 + 5, dw1.z);
    }
    if (!fx2) {
      sweep_st<true>(q2, dv2.x); sweep_st<true>(q2 + 1, dv2.y); sweep_st<true>(q2 + 2, dv2.z);
      sweep_st<true>(q2 + 3, dw2.x); sweep_st<true>(q2 + 4, dw2.y); sweep_st<true>(q2 + 5, dw2.z);
    }
  }
  return delta;
}

template <int MODEL, class Trig, int SPEC>
__global__ void __launch_bounds__(256) k_sweep_dataflow(SolverContacts C, uint32_t nc, uint32_t iters,
                                                        const double* sb, double* dvw, double dtinv, int maxSub,
                                                        const uint32_t* deg, const uint32_t* rank1,
                                                        const uint32_t* rank2, uint32_t* done,
                                                        unsigned long long* deltas, uint32_t* flags,
                                                        uint32_t spin_limit, const uint32_t* slots) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint64_t total = (uint64_t)iters * nc;
  // the largest impulse change of this wavefront's items of sweep qd, added to deltas[qd] once
  // per sweep (a wavefront's items are in increasing order, so its sweeps come one after the
  // other) instead of by every item on its release path
  double dmax = 0.0;
  uint32_t qd = (uint32_t)(wave / nc);
  for (uint64_t item = wave; item < total; item += nwaves) {
    const uint32_t q = (uint32_t)(item / nc);
    const uint32_t k = (uint32_t)(item - (uint64_t)q * nc);
    if (q != qd) {
      if (lane == 0 && dmax > 0) atomicMax(&deltas[qd], (unsigned long long)__double_as_longlong(dmax));
      dmax = 0.0;
      qd = q;
    }
    const uint32_t s = __builtin_amdgcn_readfirstlane(slots ? slots[k] : k);
    const uint32_t e1 = C.b1[s], e2 = C.b2[s];
    const uint32_t i1 = e1 & ~kFixedBit, i2 = e2 & ~kFixedBit;
    const bool f1 = (e1 & kFixedBit) != 0, f2 = (e2 & kFixedBit) != 0;
    // both bodies' targets before either wait, so that the second body's loads are not one more
    // round trip after the first wait
    // (the loads are unconditional, so they issue together: indices of fixed bodies are valid too)
    const uint32_t d1 = deg[i1], k1 = rank1[s], d2 = deg[i2], k2 = rank2[s];
    const uint32_t t1 = q * d1 + k1, t2 = q * d2 + k2;
    // The relaxation's read-only inputs -- the contact record and the bodies' records -- are
    // loaded before the wait, whose latency they then overlap; only the velocity corrections and
    // the impulse, which earlier items write, are loaded after it.
    PrefetchedFields cf;
#pragma unroll
    for (int k = 0; k < CF_P; ++k) cf.v[k] = C.f[(size_t)k * C.cap + s];
#pragma unroll
    for (int k = CF_P + 3; k < CF_COUNT; ++k) cf.v[k] = C.f[(size_t)k * C.cap + s];
    const RelaxBody B1 = relax_body(sb + (size_t)kSB * i1);
    const RelaxBody B2 = relax_body(sb + (size_t)kSB * i2);
    if ((!f1 && !wait_count(done + i1, t1, flags, spin_limit)) || (!f2 && !wait_count(done + i2, t2, flags, spin_limit))) {
      // give up and raise the flag, so that the wavefronts waiting for this one stop as well
      if (lane == 0) __hip_atomic_fetch_or(flags, ERR_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    // Compiler barrier: the relaxation's loads of the shared data are relaxed atomics of other
    // addresses, which the language would let the compiler hoist above the wait loop.  The fence
    // emits no instruction; what orders the loads at run time is the wait loop's exit, which
    // depends on the counter's loaded value (in-order issue, sc1 loads coherent per access).
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const double d = relax_coherent<MODEL, Trig, SPEC>(C, s, cf, B1, B2, i1, i2, f1, f2, dvw, maxSub);
    dmax = fmax(dmax, d);
    if (lane == 0) {
      // Release without an agent-scope release: that writes back the XCD's whole L2 (measured
      // 215 -> 248 steps/s without it).  The item's stores are all agent-scope atomics (device
      // scope bit set), which complete only once coherent at device scope, so waiting for them
      // (s_waitcnt 0; signal fences keep the compiler from moving memory accesses across) orders
      // them before the counter updates.
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (!f1) __hip_atomic_fetch_add(done + i1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (!f2) __hip_atomic_fetch_add(done + i2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (lane == 0 && dmax > 0) atomicMax(&deltas[qd], (unsigned long long)__double_as_longlong(dmax));
}

// Lanes per contact.  A wavefront issues one instruction every four cycles whatever its lanes
// do, so a lone contact's relaxation costs its instruction count; the speculated search with a
// whole wavefront per contact (SPEC 6) cuts the sequential sin/cos evaluations of a pass from 20
// to 4 rounds and measures 1.2x (device libm) to 2.8x (correctly rounded trig) faster per sliding
// contact than one lane per contact.  Colours are small in the latency-bound regime, so the
// largest group (64, 16, 4, 1) with which every wavefront of the colour is resident at once wins;
// big colours fall back to more contacts per wavefront (occupancy queried once).
template <int MODEL, class Trig, int SPEC>
static uint64_t resident_lanes(uint32_t ncu) {
  static int blocks = -1;
  if (blocks < 0) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void*)k_sweep<MODEL, Trig, SPEC>, 256, 0) != hipSuccess ||
        b < 1)
      b = 1;
    blocks = b;
  }
  return (uint64_t)blocks * 256u * ncu;
}

// A colour of a searching model (3, 5): the largest lane group per contact with which every
// wavefront of the colour is resident at once.
template <int MODEL, class Trig>
static void search_sweep(const SolverContacts& C, uint32_t begin, uint32_t end, const Bodies& B, double dtinv,
                         int maxSub, unsigned long long* delta, const uint32_t* stop, uint32_t ncu, hipStream_t s) {
  const uint64_t n = end - begin;
  const dim3 b(256);
  auto grid = [&](int spec) { return dim3((unsigned)(((n << spec) + 255) / 256)); };
  if ((n << 6) <= resident_lanes<MODEL, Trig, 6>(ncu))
    hipLaunchKernelGGL((k_sweep<MODEL, Trig, 6>), grid(6), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop);
  else if ((n << 4) <= resident_lanes<MODEL, Trig, 4>(ncu))
    hipLaunchKernelGGL((k_sweep<MODEL, Trig, 4>), grid(4), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop);
  else if ((n << 2) <= resident_lanes<MODEL, Trig, 2>(ncu))
    hipLaunchKernelGGL((k_sweep<MODEL, Trig, 2>), grid(2), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop);
  else
    hipLaunchKernelGGL((k_sweep<MODEL, Trig, 0>), grid(0), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop);
}

template <class Trig>
static void sweep_dispatch(int model, const SolverContacts& C, uint32_t begin, uint32_t end, const Bodies& B,
                           double dtinv, int maxSub, unsigned long long* delta, const uint32_t* stop, uint32_t ncu,
                           hipStream_t s) {
  const uint64_t n = end - begin;
  const dim3 b(256);
  auto grid = [&](int spec) { return dim3((unsigned)(((n << spec) + 255) / 256)); };
  switch (model) {
    case 0: hipLaunchKernelGGL((k_sweep<0, Trig, 0>), grid(0), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop); break;
    case 1: hipLaunchKernelGGL((k_sweep<1, Trig, 0>), grid(0), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop); break;
    case 2: hipLaunchKernelGGL((k_sweep<2, Trig, 0>), grid(0), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop); break;
    case 4: hipLaunchKernelGGL((k_sweep<4, Trig, 0>), grid(0), b, 0, s, C, begin, end, B.sb, B.dvw, dtinv, maxSub, delta, stop); break;
    case 5: search_sweep<5, Trig>(C, begin, end, B, dtinv, maxSub, delta, stop, ncu, s); break;
    default: search_sweep<3, Trig>(C, begin, end, B, dtinv, maxSub, delta, stop, ncu, s); break;
  }
}

template <class Trig, int SPEC>
static pehip_sweep_status dataflow_dispatch(int model, const SolverContacts& C, uint32_t nc, uint32_t iters,
                                           const Bodies& B, double dtinv, int maxSub, const DataflowBuffers& D,
                                           uint32_t ncu, hipStream_t s, DataflowLaunchInfo* info,
                                           const uint32_t* slots) {
  void* args_fn = nullptr;
  switch (model) {
    case 0: args_fn = (void*)k_sweep_dataflow<0, Trig, 0>; break;
    case 1: args_fn = (void*)k_sweep_dataflow<1, Trig, 0>; break;
    case 2: args_fn = (void*)k_sweep_dataflow<2, Trig, 0>; break;
    case 4: args_fn = (void*)k_sweep_dataflow<4, Trig, 0>; break;
    case 5: args_fn = (void*)k_sweep_dataflow<5, Trig, SPEC>; break;
    default: args_fn = (void*)k_sweep_dataflow<3, Trig, SPEC>; break;
  }
  int per_cu = 0;
  info->occupancy_err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)args_fn, 256, 0);
  info->per_cu = per_cu;
  if (info->occupancy_err != hipSuccess || per_cu < 1) return SWEEP_LAUNCH_FAILED;
  const uint64_t want = ((uint64_t)iters * nc * 64 + 255) / 256;  // one wavefront per item at most
  const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)per_cu * ncu, std::max<uint64_t>(want, 1));
  info->grid = grid;
  SolverContacts Cc = C;
  const double* sb = B.sb;
  double* dvw = B.dvw;
  uint32_t spin = D.spin_limit;
  void* args[] = {&Cc, &nc, &iters, &sb, &dvw, &dtinv, &maxSub, (void*)&D.deg, (void*)&D.rank1, (void*)&D.rank2,
                  (void*)&D.done, (void*)&D.deltas, (void*)&D.flags, &spin, (void*)&slots};
  info->launch_err = hipLaunchCooperativeKernel(args_fn, dim3(grid), dim3(256), args, 0, s);
  return info->launch_err == hipSuccess ? SWEEP_OK : SWEEP_LAUNCH_FAILED;
}

void launch_dataflow_prepare(const SolverContacts& C, uint32_t nc, const Bodies& B, const DataflowBuffers& D,
                             hipStream_t s) {
  if (!nc) return;
  hipLaunchKernelGGL(k_inv_perm, dim3((nc + 255) / 256), dim3(256), 0, s, D.perm, nc, D.inv);
  hipLaunchKernelGGL(k_body_ranks, dim3((B.n + 255) / 256), dim3(256), 0, s, C, B.fixed, B.n, D.adj_start, D.adj,
                     D.inv, D.rank1, D.rank2);
  fill_u32(D.done, 0u, B.n, s);
}

pehip_sweep_status launch_sweep_dataflow(int model, bool exact_trig, const SolverContacts& C, uint32_t nc,
                                         uint32_t iters, const Bodies& B, double dtinv, int maxSub,
                                         const DataflowBuffers& D, uint32_t ncu, hipStream_t s,
                                         DataflowLaunchInfo* info, const uint32_t* slots) {
  if (!nc || !iters) return SWEEP_OK;
  // a whole wavefront owns each item, so the speculated search costs no parallelism here
  if (exact_trig) return dataflow_dispatch<TrigExact, 6>(model, C, nc, iters, B, dtinv, maxSub, D, ncu, s, info, slots);
  return dataflow_dispatch<TrigFast, 6>(model, C, nc, iters, B, dtinv, maxSub, D, ncu, s, info, slots);
}

void launch_sweep(int model, bool exact_trig, const SolverContacts& C, uint32_t begin, uint32_t end,
                  const Bodies& B, double dtinv, int maxSub, unsigned long long* delta, const uint32_t* stop,
                  uint32_t ncu, hipStream_t s) {
  if (end <= begin) return;
  if (exact_trig) sweep_dispatch<TrigExact>(model, C, begin, end, B, dtinv, maxSub, delta, stop, ncu, s);
  else sweep_dispatch<TrigFast>(model, C, begin, end, B, dtinv, maxSub, delta, stop, ncu, s);
}

// After sweep `it`: the first sweep whose L-inf impulse change is below thr raises stop[0] and
// records the number of sweeps run in stop[1] (a NaN change never stops, as on the host).
__global__ void k_abort_check(const unsigned long long* deltas, uint32_t it, double thr, uint32_t* stop) {
  if (threadIdx.x != 0 || stop[0]) return;
  if (__longlong_as_double((long long)deltas[it]) < thr) {
    stop[1] = it + 1;
    stop[0] = 1;
  }
}

void launch_abort_check(const unsigned long long* deltas, uint32_t it, double thr, uint32_t* stop, hipStream_t s) {
  hipLaunchKernelGGL(k_abort_check, dim3(1), dim3(64), 0, s, deltas, it, thr, stop);
}

template <class Trig>
static void serial_dispatch(int model, const SolverContacts& C, uint32_t nc, uint32_t iters, const Bodies& B,
                            double dtinv, int maxSub, double thr, unsigned long long* deltas, hipStream_t s) {
  const dim3 g(1), b(64);
  switch (model) {
    case 0: hipLaunchKernelGGL((k_sweep_serial<0, Trig, 0>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 1: hipLaunchKernelGGL((k_sweep_serial<1, Trig, 0>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 2: hipLaunchKernelGGL((k_sweep_serial<2, Trig, 0>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 4: hipLaunchKernelGGL((k_sweep_serial<4, Trig, 0>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    case 5: hipLaunchKernelGGL((k_sweep_serial<5, Trig, 6>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
    default: hipLaunchKernelGGL((k_sweep_serial<3, Trig, 6>), g, b, 0, s, C, nc, iters, B.sb, B.dvw, dtinv, maxSub, thr, deltas); break;
  }
}

void launch_sweep_serial(int model, bool exact_trig, const SolverContacts& C, uint32_t nc, uint32_t iters,
                         const Bodies& B, double dtinv, int maxSub, double thr, unsigned long long* deltas,
                         hipStream_t s) {
  if (exact_trig) serial_dispatch<TrigExact>(model, C, nc, iters, B, dtinv, maxSub, thr, deltas, s);
  else serial_dispatch<TrigFast>(model, C, nc, iters, B, dtinv, maxSub, thr, deltas, s);
}

}  // namespace pe_amd
