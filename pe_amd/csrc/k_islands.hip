// Island sweeps: all K NBGS sweeps of each island of the contact graph in one wavefront or one
// workgroup, with the island's body records and velocity corrections in LDS.
//
// The coloured order relaxes, sweep after sweep, colour after colour, contacts that share no
// non-fixed body; what it fixes is, for every body, the order of the updates that touch it.  The
// non-fixed bodies linked by contacts fall apart into islands (connected components; fixed bodies
// link nothing, their velocity corrections never change), and no contact of one island touches a
// body of another.  So each island can run its K sweeps on its own, colour by colour within the
// island: every body then sees its updates in the coloured order, and the results equal the colour
// launches bit for bit.  pe's own notion of independent groups of bodies is its union-find
// (pe/core/batches/UnionFind.h:58).
//
// Small islands (up to kIslandSmall contacts) are packed into wavefronts by size class, one lane
// per contact: a wavefront of class k holds 64 >> k islands of 2^(k-1) < size <= 2^k contacts, so
// at most 64 + (64 >> k) <= 128 bodies.  Per sweep a wavefront runs as many rounds as its islands
// have colours; in round r each island's lanes of its r-th colour relax.  Islands of up to
// kIslandLarge contacts take a workgroup of kIslandLarge threads in the same way, a barrier between
// rounds.  The rest go to the dataflow sweeps.
#include <algorithm>

#include "kernels.hpp"

namespace pe_amd {

// Root of x's set.  Links only ever attach a root below a smaller index, so a parent pointer only
// decreases and any ancestor is a valid parent: path halving by atomicMin keeps the trees flat.
// Loads are agent-scope atomics: other XCDs link concurrently during the union pass.
__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  uint32_t p = __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != x) {
    const uint32_t g = __hip_atomic_load(parent + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g != p) __hip_atomic_fetch_min(parent + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = g;
  }
  return x;
}

__device__ __forceinline__ uint32_t slot_colour(const ColourOffsets& O, uint32_t s) {
  uint32_t lo = 0, hi = O.ncol;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s >= O.off[mid]) lo = mid;
    else hi = mid;
  }
  return lo;
}

// rank of colour c among the colours of mask m (4 words), and the number of colours in m
__device__ __forceinline__ uint32_t colour_rank(const uint32_t* m, uint32_t c) {
  uint32_t r = 0;
  for (uint32_t w = 0; w < (c >> 5); ++w) r += __popc(m[w]);
  return r + __popc(m[c >> 5] & ((1u << (c & 31)) - 1u));
}

__global__ void k_isl_init(uint32_t n, IslandBuffers I) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < ISL_HDR_WORDS) I.hdr[b] = 0;
  if (b >= n) return;
  I.parent[b] = b;
  I.ncont[b] = 0;
  I.cursor[b] = 0;
  I.cmask[4 * b] = I.cmask[4 * b + 1] = I.cmask[4 * b + 2] = I.cmask[4 * b + 3] = 0;
}

// Links the two non-fixed bodies of every contact.
__global__ void k_isl_union(SolverContacts C, uint32_t nc, IslandBuffers I) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  const uint32_t e1 = C.b1[s], e2 = C.b2[s];
  if ((e1 | e2) & kFixedBit) return;
  uint32_t a = e1, b = e2;
  for (;;) {
    a = uf_find(I.parent, a);
    b = uf_find(I.parent, b);
    if (a == b) return;
    if (a < b) { const uint32_t t = a; a = b; b = t; }
    // attach the larger root a below b, if a is still a root
    uint32_t expect = a;
    if (__hip_atomic_compare_exchange_strong(I.parent + a, &expect, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

// Per slot: its island's root; per root: contacts and colours.
__global__ void k_isl_count(SolverContacts C, uint32_t nc, ColourOffsets O, IslandBuffers I) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  const uint32_t e1 = C.b1[s], e2 = C.b2[s];
  const uint32_t b = (e1 & kFixedBit) ? (e2 & ~kFixedBit) : e1;
  const uint32_t r = uf_find(I.parent, b);
  I.sroot[s] = r;
  atomicAdd(&I.ncont[r], 1u);
  const uint32_t c = slot_colour(O, s);
  atomicOr(&I.cmask[4 * r + (c >> 5)], 1u << (c & 31));
}

// Per root: small (size class, position in the class), large (unit) or dataflow.
__global__ void k_isl_classify(uint32_t n, IslandBuffers I) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const uint32_t m = I.ncont[b];
  if (m == 0 || I.parent[b] != b) return;
  if (m <= kIslandSmall) {
    const uint32_t k = m <= 1 ? 0u : 32u - (uint32_t)__clz(m - 1);   // ceil(log2 m)
    const uint32_t pos = atomicAdd(&I.hdr[ISL_CLASS_COUNT + k], 1u);
    I.place[b] = (k << 24) | pos;
  } else if (m <= kIslandLarge) {
    I.place[b] = (1u << 30) | atomicAdd(&I.hdr[ISL_LARGE_UNITS], 1u);
  } else {
    I.place[b] = 2u << 30;
    atomicAdd(&I.hdr[ISL_DATAFLOW_CONTACTS], m);
  }
}

__global__ void k_isl_plan(IslandBuffers I) {
  if (threadIdx.x != 0) return;
  uint32_t base = 0;
  for (uint32_t k = 0; k < kIslandClasses; ++k) {
    I.hdr[ISL_CLASS_BASE + k] = base;
    const uint32_t per = kIslandSmall >> k;
    base += (I.hdr[ISL_CLASS_COUNT + k] + per - 1) / per;
  }
  I.hdr[ISL_SMALL_UNITS] = base;
}

// Slots into their units' lanes with their rounds; dataflow slots flagged.
__global__ void k_isl_assign_slots(uint32_t nc, ColourOffsets O, IslandBuffers I) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  const uint32_t r = I.sroot[s];
  const uint32_t pl = I.place[r];
  const uint32_t kind = pl >> 30;
  I.dflag[s] = kind == 2 ? 1u : 0u;
  if (kind == 2) return;
  const uint32_t* m = I.cmask + 4 * r;
  const uint32_t round = colour_rank(m, slot_colour(O, s));
  const uint32_t rounds = __popc(m[0]) + __popc(m[1]) + __popc(m[2]) + __popc(m[3]);
  const uint32_t j = atomicAdd(&I.cursor[r], 1u);
  if (kind == 0) {
    const uint32_t k = (pl >> 24) & 63u, pos = pl & 0xFFFFFFu, per = kIslandSmall >> k;
    const uint32_t unit = I.hdr[ISL_CLASS_BASE + k] + pos / per;
    const uint32_t lane = ((pos % per) << k) + j;
    I.sslot[unit * kIslandSmall + lane] = s;
    I.sround[unit * kIslandSmall + lane] = round;
    atomicMax(&I.srounds[unit], rounds);
  } else {
    const uint32_t unit = pl & 0x3FFFFFFFu;
    I.lslot[unit * kIslandLarge + j] = s;
    I.lround[unit * kIslandLarge + j] = round;
    if (j == 0) I.lrounds[unit] = rounds;
  }
}

// Non-fixed bodies with contacts into their units' body tables.
__global__ void k_isl_assign_bodies(uint32_t n, const uint8_t* fixed, const uint32_t* deg, IslandBuffers I) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n || fixed[b] || deg[b] == 0) return;
  const uint32_t r = uf_find(I.parent, b);
  const uint32_t pl = I.place[r];
  const uint32_t kind = pl >> 30;
  if (kind == 2) return;
  if (kind == 0) {
    const uint32_t k = (pl >> 24) & 63u, pos = pl & 0xFFFFFFu, per = kIslandSmall >> k;
    const uint32_t unit = I.hdr[ISL_CLASS_BASE + k] + pos / per;
    const uint32_t j = atomicAdd(&I.snb[unit], 1u);
    I.blocal[b] = j;
    I.sbody[unit * kIslandSmallBodies + j] = b;
  } else {
    const uint32_t unit = pl & 0x3FFFFFFFu;
    const uint32_t j = atomicAdd(&I.lnb[unit], 1u);
    I.blocal[b] = j;
    I.lbody[unit * kIslandLargeBodies + j] = b;
  }
}

__global__ void k_isl_compact(uint32_t nc, IslandBuffers I) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nc && I.dflag[s]) I.dslots[I.dpos[s]] = s;
}

void launch_island_prepare(const SolverContacts& C, uint32_t nc, const ColourOffsets& O, const Bodies& B,
                           const uint32_t* deg, IslandBuffers& I, hipStream_t s) {
  (void)deg;
  const uint32_t n = std::max<uint32_t>(B.n, ISL_HDR_WORDS);
  hipLaunchKernelGGL(k_isl_init, dim3((n + 255) / 256), dim3(256), 0, s, B.n, I);
  if (!nc) return;
  const dim3 gc((nc + 255) / 256), gb((B.n + 255) / 256), t(256);
  hipLaunchKernelGGL(k_isl_union, gc, t, 0, s, C, nc, I);
  hipLaunchKernelGGL(k_isl_count, gc, t, 0, s, C, nc, O, I);
  hipLaunchKernelGGL(k_isl_classify, gb, t, 0, s, B.n, I);
  hipLaunchKernelGGL(k_isl_plan, dim3(1), dim3(64), 0, s, I);
}

void launch_island_assign(const SolverContacts& C, uint32_t nc, const ColourOffsets& O, const Bodies& B,
                          const uint32_t* deg, IslandBuffers& I, uint32_t small_units, uint32_t large_units,
                          uint32_t dataflow_contacts, uint32_t* scan_scratch, hipStream_t s) {
  (void)C;
  if (!nc) return;
  fill_u32(I.sslot, kIslandNone, (size_t)small_units * kIslandSmall, s);
  fill_u32(I.snb, 0u, small_units, s);
  fill_u32(I.srounds, 0u, small_units, s);
  fill_u32(I.lslot, kIslandNone, (size_t)large_units * kIslandLarge, s);
  fill_u32(I.lnb, 0u, large_units, s);
  const dim3 gc((nc + 255) / 256), gb((B.n + 255) / 256), t(256);
  hipLaunchKernelGGL(k_isl_assign_slots, gc, t, 0, s, nc, O, I);
  hipLaunchKernelGGL(k_isl_assign_bodies, gb, t, 0, s, B.n, B.fixed, deg, I);
  if (dataflow_contacts) {
    exclusive_scan_u32(I.dflag, I.dpos, nc, I.hdr + ISL_HDR_WORDS - 1, scan_scratch, s);
    hipLaunchKernelGGL(k_isl_compact, gc, t, 0, s, nc, I);
  }
}

// ---------------------------------------------------------------------------------------------
// The sweep kernel.  Blocks [0, large_units) run one large island each with all their lanes;
// the following blocks run four small units each, one per wavefront, independently.
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kIslandThreads = 256;
static_assert(kIslandThreads == kIslandLarge, "a large unit has one lane per contact");
static_assert(kIslandThreads / 64 * kIslandSmallBodies == kIslandLargeBodies, "the LDS body table is shared");
constexpr int kLdsBody = kSB + 6;   // a body's record (kSB doubles) and its velocity corrections

// A contact's fields held in registers for the whole launch.
struct RegFields {
  double v[CF_COUNT];
  __device__ __forceinline__ double operator()(int k) const { return v[k]; }
};

// One unit: `lanes` lanes from `lane0` of the block, bodies in lds[0 .. nb), slot table `slot` /
// `round` (lanes entries), `rounds` rounds per sweep.  WAVE: the unit is one wavefront (ordering
// between rounds within the wavefront), else the whole block (barriers).
template <int MODEL, class Trig, bool WAVE>
__device__ __forceinline__ void island_unit(const SolverContacts& C, const uint32_t* slot, const uint32_t* round,
                                            uint32_t rounds, const uint32_t* body, uint32_t nb, uint32_t lane,
                                            uint32_t lanes, double* lds, const uint32_t* blocal, const double* sb,
                                            double* dvw, uint32_t iters, int maxSub, unsigned long long* deltas,
                                            unsigned long long* wmax) {
  for (uint32_t j = lane; j < nb; j += lanes) {
    const uint32_t b = body[j];
    double* L = lds + (size_t)kLdsBody * j;
#pragma unroll
    for (int k = 0; k < kSB; ++k) L[k] = sb[(size_t)kSB * b + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[kSB + k] = dvw[(size_t)kDVW * b + k];
  }
  const uint32_t s = slot[lane];
  const bool act = s != kIslandNone;
  uint32_t rnd = 0, i1 = 0, i2 = 0, l1 = 0, l2 = 0;
  bool fx1 = true, fx2 = true;
  RegFields cf;
  d3 p = mk3(0, 0, 0);
  if (act) {
    rnd = round[lane];
    const uint32_t e1 = C.b1[s], e2 = C.b2[s];
    i1 = e1 & ~kFixedBit;
    i2 = e2 & ~kFixedBit;
    fx1 = (e1 & kFixedBit) != 0;
    fx2 = (e2 & kFixedBit) != 0;
    if (!fx1) l1 = blocal[i1];
    if (!fx2) l2 = blocal[i2];
#pragma unroll
    for (int k = 0; k < CF_COUNT; ++k) cf.v[k] = C.f[(size_t)k * C.cap + s];
    p = mk3(cf.v[CF_P], cf.v[CF_P + 1], cf.v[CF_P + 2]);
  }
  if constexpr (WAVE) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  else __syncthreads();
  for (uint32_t q = 0; q < iters; ++q) {
    unsigned long long u = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
      if (act && rnd == r) {
        double* L1 = lds + (size_t)kLdsBody * l1;
        double* L2 = lds + (size_t)kLdsBody * l2;
        const RelaxBody B1 = relax_body(fx1 ? sb + (size_t)kSB * i1 : L1);
        const RelaxBody B2 = relax_body(fx2 ? sb + (size_t)kSB * i2 : L2);
        d3 dv1 = mk3(0, 0, 0), dw1 = mk3(0, 0, 0), dv2 = mk3(0, 0, 0), dw2 = mk3(0, 0, 0);
        if (!fx1) {
          dv1 = mk3(L1[kSB], L1[kSB + 1], L1[kSB + 2]);
          dw1 = mk3(L1[kSB + 3], L1[kSB + 4], L1[kSB + 5]);
        }
        if (!fx2) {
          dv2 = mk3(L2[kSB], L2[kSB + 1], L2[kSB + 2]);
          dw2 = mk3(L2[kSB + 3], L2[kSB + 4], L2[kSB + 5]);
        }
        const double d = relax_update<MODEL, Trig, 0>(cf, B1, B2, dv1, dw1, dv2, dw2, p, maxSub);
        if (!fx1) {
          L1[kSB] = dv1.x; L1[kSB + 1] = dv1.y; L1[kSB + 2] = dv1.z;
          L1[kSB + 3] = dw1.x; L1[kSB + 4] = dw1.y; L1[kSB + 5] = dw1.z;
        }
        if (!fx2) {
          L2[kSB] = dv2.x; L2[kSB + 1] = dv2.y; L2[kSB + 2] = dv2.z;
          L2[kSB + 3] = dw2.x; L2[kSB + 4] = dw2.y; L2[kSB + 5] = dw2.z;
        }
        // non-negative doubles (and NaN above +inf) order like their bit patterns, as in the
        // colour launches' atomicMax
        const unsigned long long du = (unsigned long long)__double_as_longlong(d);
        u = du > u ? du : u;
      }
      if constexpr (WAVE) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      else __syncthreads();
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const unsigned long long o = __shfl_xor(u, d, 64);
      u = o > u ? o : u;
    }
    if constexpr (WAVE) {
      if ((threadIdx.x & 63u) == 0 && u) atomicMax(&deltas[q], u);
    } else {
      // one atomic per block and sweep
      if ((threadIdx.x & 63u) == 0) wmax[threadIdx.x >> 6] = u;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long m = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
        if (m) atomicMax(&deltas[q], m);
      }
      __syncthreads();
    }
  }
  if (act) {
    C.f[(size_t)CF_P * C.cap + s] = p.x;
    C.f[(size_t)(CF_P + 1) * C.cap + s] = p.y;
    C.f[(size_t)(CF_P + 2) * C.cap + s] = p.z;
  }
  if constexpr (WAVE) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  else __syncthreads();
  for (uint32_t j = lane; j < nb; j += lanes) {
    const uint32_t b = body[j];
    const double* L = lds + (size_t)kLdsBody * j;
#pragma unroll
    for (int k = 0; k < 6; ++k) dvw[(size_t)kDVW * b + k] = L[kSB + k];
  }
}

template <int MODEL, class Trig>
__global__ void __launch_bounds__(kIslandThreads) k_sweep_islands(SolverContacts C, IslandBuffers I,
                                                                  uint32_t small_units, uint32_t large_units,
                                                                  const double* sb, double* dvw, uint32_t iters,
                                                                  int maxSub, unsigned long long* deltas) {
  __shared__ double lds[kIslandLargeBodies * kLdsBody];
  __shared__ unsigned long long wmax[kIslandThreads / 64];
  if (blockIdx.x < large_units) {
    const uint32_t u = blockIdx.x;
    island_unit<MODEL, Trig, false>(C, I.lslot + (size_t)u * kIslandLarge, I.lround + (size_t)u * kIslandLarge,
                                    I.lrounds[u], I.lbody + (size_t)u * kIslandLargeBodies, I.lnb[u], threadIdx.x,
                                    kIslandThreads, lds, I.blocal, sb, dvw, iters, maxSub, deltas, wmax);
  } else {
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t u = (blockIdx.x - large_units) * (kIslandThreads / 64) + w;
    if (u >= small_units) return;
    island_unit<MODEL, Trig, true>(C, I.sslot + (size_t)u * kIslandSmall, I.sround + (size_t)u * kIslandSmall,
                                   I.srounds[u], I.sbody + (size_t)u * kIslandSmallBodies, I.snb[u],
                                   threadIdx.x & 63u, 64, lds + (size_t)w * kIslandSmallBodies * kLdsBody, I.blocal,
                                   sb, dvw, iters, maxSub, deltas, wmax);
  }
}

template <class Trig>
static void islands_dispatch(int model, const SolverContacts& C, const IslandBuffers& I, uint32_t small_units,
                             uint32_t large_units, uint32_t iters, const Bodies& B, int maxSub,
                             unsigned long long* deltas, hipStream_t s) {
  const uint32_t blocks = large_units + (small_units + kIslandThreads / 64 - 1) / (kIslandThreads / 64);
  if (!blocks) return;
  const dim3 g(blocks), t(kIslandThreads);
  switch (model) {
    case 0: hipLaunchKernelGGL((k_sweep_islands<0, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
    case 1: hipLaunchKernelGGL((k_sweep_islands<1, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
    case 2: hipLaunchKernelGGL((k_sweep_islands<2, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
    case 4: hipLaunchKernelGGL((k_sweep_islands<4, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
    case 5: hipLaunchKernelGGL((k_sweep_islands<5, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
    default: hipLaunchKernelGGL((k_sweep_islands<3, Trig>), g, t, 0, s, C, I, small_units, large_units, B.sb, B.dvw, iters, maxSub, deltas); break;
  }
}

void launch_sweep_islands(int model, bool exact_trig, const SolverContacts& C, const IslandBuffers& I,
                          uint32_t small_units, uint32_t large_units, uint32_t iters, const Bodies& B, int maxSub,
                          unsigned long long* deltas, hipStream_t s) {
  if (!iters) return;
  if (exact_trig) islands_dispatch<TrigExact>(model, C, I, small_units, large_units, iters, B, maxSub, deltas, s);
  else islands_dispatch<TrigFast>(model, C, I, small_units, large_units, iters, B, maxSub, deltas, s);
}

}  // namespace pe_amd
