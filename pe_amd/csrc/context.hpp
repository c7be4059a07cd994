// The single-world context behind the C-ABI (internal to the library): device buffers of one
// pe world and the host-side step driver shared by context.hip and adjoint.hip.
#pragma once

#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pe_hip.h"
#include "batch.hpp"

namespace pe_amd {

template <class T>
inline hipError_t dalloc(T*& p, size_t n) {
  p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)&p, n * sizeof(T));
}
template <class T>
inline void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace pe_amd

using namespace pe_amd;  // internal header: the context code is written against pe_amd


struct pehip_ctx {
  pehip_params prm;
  int device = 0;
  uint32_t ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  bool profiling = false;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> launch_ev;  // profiling: one event before and after every sweep launch
  size_t n_launch_events = 0;

  // bodies
  uint32_t n = 0;
  uint8_t *shape = nullptr, *fixed = nullptr;
  uint32_t* mat = nullptr;
  uint64_t* sid = nullptr;
  double *param = nullptr, *pos = nullptr, *q = nullptr, *R = nullptr, *v = nullptr, *w = nullptr;
  double *pnormal = nullptr, *pd = nullptr, *aabb = nullptr, *inv_mass = nullptr, *Ib = nullptr, *Ibinv = nullptr;
  double *force = nullptr, *torque = nullptr, *sb = nullptr, *dvw = nullptr;
  std::vector<uint8_t> h_fixed, h_shape;
  bool spheres_only = false;  // every body a sphere or a plane (the pair kernels' narrowphase)
  // union members (spheres), grouped by union; mstart has n + 1 entries when nm > 0
  uint32_t nm = 0;
  uint32_t *mstart = nullptr, *mowner = nullptr, *mmat = nullptr;
  uint64_t* msid = nullptr;
  double *mrad = nullptr, *mrpos = nullptr, *mpos = nullptr;
  bool members_missing = false;  // unions uploaded, their members not yet

  // materials
  uint32_t nmat = 0;
  double *csf = nullptr, *cdf = nullptr;

  // broadphase
  GridParams grid{};
  int key_bits = 0;
  BroadphaseBuffers bp{};
  uint32_t *planes = nullptr, *scan_scratch = nullptr, *sort_scratch = nullptr;
  // One device block for the words the host reads back during a step, so that each read-back is
  // one copy: kStFlags (error flags; the abort flags at +4), kStTotal (contact count),
  // kStRemaining (colouring: uncoloured contacts, last round), kStHist (colour histogram).
  uint32_t* status = nullptr;
  uint32_t *flags = nullptr, *total = nullptr;   // into status

  // contacts: the per-contact arrays below are carved out of one device block, the context's own
  // (cblock, allocated with capacity own_ccap) or, during pehip_step_noalloc, the library's static
  // scratch on this device (scratch; capacity kScratchContacts)
  size_t ccap = 0;
  void* cblock = nullptr;
  size_t own_ccap = 0;
  bool contacts_in_scratch = false;
  bool no_alloc = false;        // inside pehip_step_noalloc: outgrowing the scratch is an error
  void* scratch = nullptr;      // device address of the static scratch on this context's device
  uint32_t nc = 0;
  RawContacts raw{};
  SolverContacts sol{};
  ColourBuffers col{};
  uint32_t *perm = nullptr, *sort_keys = nullptr, *sort_keys_tmp = nullptr, *sort_vals_tmp = nullptr;
  uint32_t *df_inv = nullptr, *df_rank1 = nullptr, *df_rank2 = nullptr, *df_done = nullptr;  // dataflow sweeps
  uint64_t dataflow_steps = 0;  // steps whose sweeps ran as one dataflow launch
  bool dataflow_refused = false;  // the device refused a cooperative launch: AUTO uses colour launches
  std::string dataflow_refusal;   // what the refused launch saw
  uint32_t *cscan_scratch = nullptr, *csort_scratch = nullptr;
  IslandBuffers isl{};           // island sweeps: per-body arrays allocated with the bodies, the
                                 // per-contact and per-unit ones carved with the contacts
  std::vector<uint32_t> colour_off;
  uint32_t colour_rounds = 4;   // rounds the last colouring needed (the next one's first batch)
  unsigned long long *deltas = nullptr, *maxpen = nullptr;
  double* ke_partial = nullptr;
  uint32_t* hstage = nullptr;  // pinned

  // reference-order request
  std::vector<uint32_t> order_b1, order_b2;
  bool order_pending = false;

  pehip_stats stats{};

  // differentiable runs (checkpoints of the last pehip_adjoint_forward)
  struct AdjointState* adj = nullptr;

  Bodies bodies() const {
    Bodies B;
    B.n = n; B.shape = shape; B.fixed = fixed; B.mat = mat; B.sid = sid; B.param = param;
    B.pos = pos; B.q = q; B.R = R; B.v = v; B.w = w; B.pnormal = pnormal; B.pd = pd; B.aabb = aabb;
    B.inv_mass = inv_mass; B.Ib = Ib; B.Ibinv = Ibinv; B.force = force; B.torque = torque;
    B.sb = sb; B.dvw = dvw;
    B.nm = nm; B.mstart = mstart; B.mowner = mowner; B.msid = msid; B.mmat = mmat; B.mrad = mrad;
    B.mrpos = mrpos; B.mpos = mpos;
    return B;
  }

  void free_members() {
    dfree(mstart); dfree(mowner); dfree(mmat); dfree(msid); dfree(mrad); dfree(mrpos); dfree(mpos);
    nm = 0;
    members_missing = false;
  }

  void free_bodies() {
    free_members();
    dfree(shape); dfree(fixed); dfree(mat); dfree(sid); dfree(param); dfree(pos); dfree(q); dfree(R);
    dfree(v); dfree(w); dfree(pnormal); dfree(pd); dfree(aabb); dfree(inv_mass); dfree(Ib); dfree(Ibinv);
    dfree(force); dfree(torque); dfree(sb); dfree(dvw); dfree(planes);
    dfree(bp.keys); dfree(bp.vals); dfree(bp.keys_tmp); dfree(bp.vals_tmp); dfree(bp.cell_start);
    dfree(bp.cell_end); dfree(bp.counts); dfree(bp.offsets); dfree(scan_scratch); dfree(sort_scratch);
    dfree(bp.saabb); dfree(bp.sfixed);
    dfree(bp.st_b1); dfree(bp.st_b2); dfree(bp.st_pos); dfree(bp.st_nrm); dfree(bp.st_dist);
    dfree(col.deg); dfree(col.adj_start); dfree(col.adj_cursor); dfree(ke_partial); dfree(df_done);
    dfree(isl.parent); dfree(isl.ncont); dfree(isl.cmask); dfree(isl.cursor); dfree(isl.place); dfree(isl.blocal);
    n = 0;
  }
  // Points the per-contact arrays into `base` laid out for `cap` contacts and returns the bytes
  // the layout needs (every array starts on a 256-byte boundary).  base = nullptr only measures;
  // clear = true sets every array pointer to nullptr instead.
  size_t carve_contacts(char* base, size_t cap, bool clear = false) {
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
      using T = std::remove_reference_t<decltype(*p)>;
      off = (off + 255) & ~(size_t)255;
      if (clear) p = nullptr;
      else if (base) p = reinterpret_cast<T*>(base + off);
      off += (count ? count : 1) * sizeof(T);
    };
    take(raw.b1, cap); take(raw.b2, cap); take(raw.pos, 3 * cap); take(raw.nrm, 3 * cap);
    take(raw.dist, cap); take(raw.mu, cap);
    take(sol.b1, cap); take(sol.b2, cap); take(sol.f, (size_t)CF_COUNT * cap);
    take(col.adj, 2 * cap); take(col.colour, cap); take(col.prio, cap); take(col.cround, cap);
    take(perm, cap); take(sort_keys, cap); take(sort_keys_tmp, cap); take(sort_vals_tmp, cap);
    take(df_inv, cap); take(df_rank1, cap); take(df_rank2, cap);
    take(cscan_scratch, scan_scratch_words(cap)); take(csort_scratch, sort_scratch_words(cap));
    const size_t su = island_small_units_bound(cap), lu = island_large_units_bound(cap);
    take(isl.sroot, cap); take(isl.dflag, cap); take(isl.dpos, cap); take(isl.dslots, cap);
    take(isl.sslot, su * kIslandSmall); take(isl.sround, su * kIslandSmall); take(isl.sbody, su * kIslandSmallBodies);
    take(isl.snb, su); take(isl.srounds, su);
    take(isl.lslot, lu * kIslandLarge); take(isl.lround, lu * kIslandLarge); take(isl.lbody, lu * kIslandLargeBodies);
    take(isl.lnb, lu); take(isl.lrounds, lu);
    take(isl.hdr, ISL_HDR_WORDS);
    if (clear) raw.cap = sol.cap = 0;
    else if (base) raw.cap = sol.cap = cap;
    return (off + 255) & ~(size_t)255;
  }
  void unbind_contacts() {
    carve_contacts(nullptr, 0, true);
    ccap = 0;
  }
  void free_contacts() {
    unbind_contacts();
    if (cblock) (void)hipFree(cblock);
    cblock = nullptr;
    own_ccap = 0;
    contacts_in_scratch = false;
  }
};

#define CK(call)                                                                 \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);              \
      return e_ == hipErrorOutOfMemory ? PEHIP_OUT_OF_MEMORY : PEHIP_HIP_ERROR;  \
    }                                                                            \
  } while (0)

#define CKS(st)                             \
  do {                                      \
    pehip_status s_ = (st);                 \
    if (s_ != PEHIP_OK) return s_;          \
  } while (0)

inline pehip_status fail(pehip_ctx* ctx, pehip_status st, const std::string& msg) {
  ctx->err = msg;
  return st;
}

inline pehip_status sync_check(pehip_ctx* ctx) {
  CK(hipGetLastError());
  CK(hipStreamSynchronize(ctx->stream));
  return PEHIP_OK;
}


// One time step of CollisionSystem<HCSITS>::simulationStep (HCSITS.h:1472-1531).
pehip_status pe_one_step(pehip_ctx* ctx, double dt);
// Frees the checkpoints of a differentiable run.
void pe_adjoint_free(pehip_ctx* ctx);
