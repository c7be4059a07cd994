// The single-world context behind the C-ABI (internal to the library): device buffers of one
// pe world and the host-side step driver shared by context.hip and adjoint.hip.
#pragma once

#include <string>
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
  uint32_t *flags = nullptr, *total = nullptr;

  // contacts
  size_t ccap = 0;
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
  std::vector<uint32_t> colour_off;
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
    dfree(col.deg); dfree(col.adj_start); dfree(col.adj_cursor); dfree(ke_partial); dfree(df_done);
    n = 0;
  }
  void free_contacts() {
    dfree(raw.b1); dfree(raw.b2); dfree(raw.pos); dfree(raw.nrm); dfree(raw.dist); dfree(raw.mu);
    dfree(sol.b1); dfree(sol.b2); dfree(sol.f);
    dfree(col.adj); dfree(col.colour); dfree(col.prio); dfree(col.cround);
    dfree(perm); dfree(sort_keys); dfree(sort_keys_tmp); dfree(sort_vals_tmp);
    dfree(df_inv); dfree(df_rank1); dfree(df_rank2);
    dfree(cscan_scratch); dfree(csort_scratch);
    ccap = 0;
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
