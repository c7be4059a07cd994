// Host side of the C-ABI: one context = one pe world in HBM.  Orchestrates the time
// step of CollisionSystem<HCSITS>::simulationStep (HCSITS.h:1472-1531) on the GPU:
// detection -> body caching -> colouring -> contact caching -> sweeps -> integration.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "context.hpp"

using namespace pe_amd;

namespace {

const char* kVersion = "pe_hip 0.1.0 (gfx950)";
// Automatic schedule: dataflow sweeps up to this many contacts (above it a colour holds enough
// contacts to fill the device and the colour launches are throughput-bound, not latency-bound).
constexpr uint32_t kDataflowMaxContacts = 65536;
// Polls (each with a short sleep) before a waiting wavefront gives up: seconds, far above any
// legitimate wait.
constexpr uint32_t kDataflowSpinLimit = 1u << 22;
// With an abort threshold the host reads the device's stop flag once per this many sweeps.
constexpr uint32_t kAbortPoll = 32;
// Automatic schedule with an abort threshold: one workgroup runs all sweeps up to this many
// contacts (relaxation models 0 and 1: one pass of its 512 lanes per colour), where a colour
// launch is mostly launch latency.  Beyond it the workgroup needs several passes per colour and
// the colour launches are faster.
constexpr uint32_t kSingleBlockMaxContacts = 512;
std::string g_create_error;

uint32_t next_pow2(uint64_t x) {
  uint32_t p = 1;
  while (p < x && p < (1u << 30)) p <<= 1;
  return p;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Static scratch of pehip_step_noalloc: the per-contact arrays of a step for at most
// kScratchContacts contacts, in a file-scope device array (one copy per device, held until the
// process ends).  Host code never passes `scratch` itself to a kernel or a copy (in host code the
// name is a stand-in, not a device address): pehip_create looks up the array's address on the
// context's device once with hipGetSymbolAddress and the context uses that pointer.
// ---------------------------------------------------------------------------------------------
constexpr size_t kScratchContacts = 16384;
constexpr size_t kScratchBytes = 12u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;
}  // namespace

// Device address of the static scratch on `device` (the current device), looked up once.
static pehip_status scratch_address(int device, void** out, std::string& why) {
  if (device < 0 || device >= kMaxDevices) { why = "device ordinal beyond the scratch table"; return PEHIP_INVALID_ARG; }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return PEHIP_HIP_ERROR;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return PEHIP_OK;
}

// The context's own contact arrays bound again after steps on the static scratch.
static pehip_status use_own_contacts(pehip_ctx* ctx) {
  if (!ctx->contacts_in_scratch) return PEHIP_OK;
  CK(hipStreamSynchronize(ctx->stream));
  ctx->unbind_contacts();
  ctx->contacts_in_scratch = false;
  if (ctx->cblock) {
    ctx->carve_contacts(static_cast<char*>(ctx->cblock), ctx->own_ccap);
    ctx->ccap = ctx->own_ccap;
  }
  return PEHIP_OK;
}

static pehip_status ensure_contact_capacity(pehip_ctx* ctx, size_t need) {
  if (ctx->no_alloc) {
    if (need <= ctx->ccap) return PEHIP_OK;
    return fail(ctx, PEHIP_INVALID_ARG, "pehip_step_noalloc: " + std::to_string(need) +
                                            " contacts exceed the static scratch's maximum of " +
                                            std::to_string(kScratchContacts));
  }
  CKS(use_own_contacts(ctx));
  if (need <= ctx->ccap) return PEHIP_OK;
  CK(hipStreamSynchronize(ctx->stream));
  ctx->free_contacts();
  ctx->nc = 0;  // the last step's contacts went with the arrays
  const size_t cap = std::max<size_t>(need + need / 2, 4096);
  const size_t bytes = ctx->carve_contacts(nullptr, cap);
  CK(hipMalloc(&ctx->cblock, bytes));
  ctx->carve_contacts(static_cast<char*>(ctx->cblock), cap);
  ctx->ccap = ctx->own_ccap = cap;
  return PEHIP_OK;
}

// Static AABB-size bounds and the grid hierarchy / bucket table of the uploaded scene.
static void plan_grid(pehip_ctx* ctx, const uint8_t* shape, const double* param, const double* pos) {
  double mx = 0, mn = 1e300;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  uint32_t ngrid = 0;
  for (uint32_t i = 0; i < ctx->n; ++i) {
    if (shape[i] == PEHIP_PLANE) continue;
    ++ngrid;
    const double* p = param + 3 * (size_t)i;
    double smax_b, smin_b;
    if (shape[i] == PEHIP_SPHERE) {
      smax_b = 2.0 * (p[0] + kContactThreshold);
      smin_b = smax_b;
    } else if (shape[i] == PEHIP_CAPSULE) {
      // AABB size between its diameter (lower bound) and axis-aligned length + diameter
      smax_b = p[1] + 2.0 * (p[0] + kContactThreshold);
      smin_b = 2.0 * p[0];
    } else if (shape[i] == PEHIP_UNION) {
      // inside the bounding sphere about the centre of mass; at least one member's diameter
      smax_b = 2.0 * (p[0] + kContactThreshold);
      smin_b = 2.0 * p[1];
    } else {
      smax_b = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) + 2.0 * kContactThreshold;
      smin_b = std::max(p[0], std::max(p[1], p[2]));
    }
    mx = std::max(mx, smax_b);
    mn = std::min(mn, smin_b);
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], pos[3 * (size_t)i + k]);
      hi[k] = std::max(hi[k], pos[3 * (size_t)i + k]);
    }
  }
  GridParams& G = ctx->grid;
  if (ngrid == 0) { mx = mn = 1.0; lo[0] = lo[1] = lo[2] = 0; hi[0] = hi[1] = hi[2] = 1; }
  const double top = mx * 1.001;
  int nl = 1 + (int)std::floor(std::log2(mx / mn) - 1e-9);
  nl = std::max(1, std::min(kMaxLevels, nl));
  G.nlevels = nl;
  for (int l = 0; l < kMaxLevels; ++l) {
    const int e = std::max(0, nl - 1 - l);
    G.span[l] = top / std::ldexp(1.0, e);
    G.inv[l] = 1.0 / G.span[l];
  }
  uint32_t d[3];
  for (int k = 0; k < 3; ++k) {
    const double cells = std::ceil((hi[k] - lo[k] + mx) / top) + 2.0;
    d[k] = std::max<uint32_t>(4, std::min<uint32_t>(512, next_pow2((uint64_t)cells)));
  }
  const uint64_t cap = std::max<uint64_t>(4096, 8ull * ngrid);
  while ((uint64_t)d[0] * d[1] * d[2] > cap) {
    int k = 0;
    if (d[1] > d[k]) k = 1;
    if (d[2] > d[k]) k = 2;
    if (d[k] <= 4) break;
    d[k] >>= 1;
  }
  G.dimx = d[0]; G.dimy = d[1]; G.dimz = d[2];
  G.table = d[0] * d[1] * d[2];
  const uint64_t keys = (uint64_t)nl * G.table + 1;
  int bits = 1;
  while ((1ull << bits) < keys) ++bits;
  ctx->key_bits = bits;
}

// The members of every union whose pose a state upload changes follow it as pe's setters move
// them: a new orientation places each member at pos + R rpos (Union::setOrientation ->
// Sphere::update(dq), Union.cpp:306-322, Sphere.cpp:1059); a new position alone translates each
// member by the union's displacement (Union::setPosition -> Sphere::update(dp), Union.cpp:251-272,
// Sphere.cpp:1031).  Unions whose pose is unchanged keep their members' positions bit for bit.
static pehip_status move_members(pehip_ctx* ctx, const double* pos, const double* q, const std::vector<double>& hR) {
  const uint32_t n = ctx->n, nm = ctx->nm;
  const size_t D = sizeof(double);
  std::vector<double> opos(3 * (size_t)n), oq(4 * (size_t)n), mp(3 * (size_t)nm), mr(3 * (size_t)nm);
  std::vector<uint32_t> start((size_t)n + 1);
  CK(hipMemcpy(opos.data(), ctx->pos, 3 * D * n, hipMemcpyDeviceToHost));
  CK(hipMemcpy(oq.data(), ctx->q, 4 * D * n, hipMemcpyDeviceToHost));
  CK(hipMemcpy(mp.data(), ctx->mpos, 3 * D * nm, hipMemcpyDeviceToHost));
  CK(hipMemcpy(mr.data(), ctx->mrpos, 3 * D * nm, hipMemcpyDeviceToHost));
  CK(hipMemcpy(start.data(), ctx->mstart, 4 * start.size(), hipMemcpyDeviceToHost));
  bool moved = false;
  for (uint32_t i = 0; i < n; ++i) {
    if (ctx->h_shape[i] != PEHIP_UNION) continue;
    const bool rot = std::memcmp(&oq[4 * (size_t)i], &q[4 * (size_t)i], 4 * D) != 0;
    const bool tra = std::memcmp(&opos[3 * (size_t)i], &pos[3 * (size_t)i], 3 * D) != 0;
    if (!rot && !tra) continue;
    moved = true;
    m33 R;
    for (int k = 0; k < 9; ++k) R.a[k] = hR[9 * (size_t)i + k];
    const d3 gp = mk3(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]);
    const d3 dp = gp - mk3(opos[3 * (size_t)i], opos[3 * (size_t)i + 1], opos[3 * (size_t)i + 2]);
    for (uint32_t m = start[i]; m < start[i + 1]; ++m) {
      double* x = &mp[3 * (size_t)m];
      const d3 y = rot ? gp + mat_vec(R, mk3(mr[3 * (size_t)m], mr[3 * (size_t)m + 1], mr[3 * (size_t)m + 2]))
                       : mk3(x[0], x[1], x[2]) + dp;
      x[0] = y.x; x[1] = y.y; x[2] = y.z;
    }
  }
  if (moved) CK(hipMemcpy(ctx->mpos, mp.data(), 3 * D * nm, hipMemcpyHostToDevice));
  return PEHIP_OK;
}

// The arrays of pehip_upload_bodies' validated bodies on the device, the broadphase's and
// colouring's per-body buffers, and the grid plan (the context holds no bodies on entry).
static pehip_status upload_body_arrays(pehip_ctx* ctx, uint32_t n, const std::vector<uint8_t>& fx,
                                       const std::vector<uint32_t>& hplanes, const uint8_t* shape,
                                       const uint64_t* sid, const uint32_t* material, const double* param,
                                       const double* pos, const double* q, const double* v, const double* w,
                                       const double* inv_mass, const double* inertia_body,
                                       const double* inv_inertia_body, const double* plane_normal,
                                       const double* plane_d) {
  ctx->n = n;
  ctx->h_fixed = fx;
  ctx->h_shape.assign(shape, shape + n);
  // rotation matrices from the quaternions (Quaternion::toRotationMatrix)
  std::vector<double> hR(9 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i) {
    qd qq;
    for (int k = 0; k < 4; ++k) qq.v[k] = q[4 * (size_t)i + k];
    const m33 m = quat_to_R(qq);
    for (int k = 0; k < 9; ++k) hR[9 * (size_t)i + k] = m.a[k];
  }
  CK(dalloc(ctx->shape, n)); CK(dalloc(ctx->fixed, n)); CK(dalloc(ctx->mat, n)); CK(dalloc(ctx->sid, n));
  CK(dalloc(ctx->param, 3 * (size_t)n)); CK(dalloc(ctx->pos, 3 * (size_t)n)); CK(dalloc(ctx->q, 4 * (size_t)n));
  CK(dalloc(ctx->R, 9 * (size_t)n)); CK(dalloc(ctx->v, 3 * (size_t)n)); CK(dalloc(ctx->w, 3 * (size_t)n));
  CK(dalloc(ctx->pnormal, 3 * (size_t)n)); CK(dalloc(ctx->pd, n)); CK(dalloc(ctx->aabb, 6 * (size_t)n));
  CK(dalloc(ctx->inv_mass, n)); CK(dalloc(ctx->Ib, 9 * (size_t)n)); CK(dalloc(ctx->Ibinv, 9 * (size_t)n));
  CK(dalloc(ctx->force, 3 * (size_t)n)); CK(dalloc(ctx->torque, 3 * (size_t)n));
  CK(dalloc(ctx->sb, (size_t)kSB * n)); CK(dalloc(ctx->dvw, (size_t)kDVW * n));
  CK(dalloc(ctx->planes, hplanes.size()));
  const size_t D = sizeof(double);
  CK(hipMemcpy(ctx->shape, shape, n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->fixed, fx.data(), n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->mat, material, 4 * (size_t)n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->sid, sid, 8 * (size_t)n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->param, param, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->pos, pos, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->q, q, 4 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->R, hR.data(), 9 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->v, v, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->w, w, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->pnormal, plane_normal, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->pd, plane_d, D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->inv_mass, inv_mass, D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->Ib, inertia_body, 9 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->Ibinv, inv_inertia_body, 9 * D * n, hipMemcpyHostToDevice));
  if (!hplanes.empty()) CK(hipMemcpy(ctx->planes, hplanes.data(), 4 * hplanes.size(), hipMemcpyHostToDevice));
  CK(hipMemset(ctx->force, 0, 3 * D * n));
  CK(hipMemset(ctx->torque, 0, 3 * D * n));
  CK(hipMemset(ctx->dvw, 0, kDVW * D * n));
  CK(hipMemset(ctx->sb, 0, kSB * D * n));

  plan_grid(ctx, shape, param, pos);
  const size_t cells = (size_t)ctx->grid.nlevels * ctx->grid.table;
  BroadphaseBuffers& P = ctx->bp;
  CK(dalloc(P.keys, n)); CK(dalloc(P.vals, n)); CK(dalloc(P.keys_tmp, n)); CK(dalloc(P.vals_tmp, n));
  CK(dalloc(P.cell_start, cells)); CK(dalloc(P.cell_end, cells));
  CK(dalloc(P.counts, (size_t)n + 1)); CK(dalloc(P.offsets, (size_t)n + 1));
  CK(dalloc(P.saabb, 6 * (size_t)n)); CK(dalloc(P.sfixed, n));
  CK(dalloc(P.st_b1, (size_t)kStage * n)); CK(dalloc(P.st_b2, (size_t)kStage * n));
  CK(dalloc(P.st_pos, 3 * (size_t)kStage * n)); CK(dalloc(P.st_nrm, 3 * (size_t)kStage * n));
  CK(dalloc(P.st_dist, (size_t)kStage * n));
  P.stride = n;
  ctx->spheres_only = std::all_of(shape, shape + n, [](uint8_t t) { return t == PEHIP_SPHERE || t == PEHIP_PLANE; });
  CK(dalloc(ctx->scan_scratch, scan_scratch_words((size_t)n + 1)));
  CK(dalloc(ctx->sort_scratch, sort_scratch_words(n)));
  CK(dalloc(ctx->col.deg, n)); CK(dalloc(ctx->col.adj_start, (size_t)n + 1)); CK(dalloc(ctx->col.adj_cursor, n));
  CK(dalloc(ctx->ke_partial, (n + 255) / 256));
  CK(dalloc(ctx->df_done, n));
  CK(dalloc(ctx->isl.parent, n)); CK(dalloc(ctx->isl.ncont, n)); CK(dalloc(ctx->isl.cmask, 4 * (size_t)n));
  CK(dalloc(ctx->isl.cursor, n)); CK(dalloc(ctx->isl.place, n)); CK(dalloc(ctx->isl.blocal, n));
  P.planes = ctx->planes;
  P.nplanes = (uint32_t)hplanes.size();
  P.flags = ctx->flags;
  ctx->members_missing = std::find(shape, shape + n, (uint8_t)PEHIP_UNION) != shape + n;
  if (!ctx->members_missing) launch_aabb(ctx->bodies(), ctx->stream);  // unions: after their members
  ctx->nc = 0;
  ctx->order_pending = false;
  return sync_check(ctx);
}

pehip_status pe_amd::check_params(const pehip_params* p, std::string& why) {
  if (p->relaxation_model != PEHIP_INELASTIC_FRICTIONLESS &&
      p->relaxation_model != PEHIP_APPROX_INELASTIC_COULOMB_DECOUPLING &&
      p->relaxation_model != PEHIP_APPROX_INELASTIC_COULOMB_ORTHO &&
      p->relaxation_model != PEHIP_INELASTIC_COULOMB_DECOUPLING &&
      p->relaxation_model != PEHIP_INELASTIC_COULOMB_ORTHO &&
      p->relaxation_model != PEHIP_INELASTIC_GMD) {
    why = "Unsupported relaxation model.";  // HCSITS.h:1926
    return PEHIP_NOT_SUPPORTED;
  }
  if (!(p->erp >= 0 && p->erp <= 1)) { why = "Error reduction parameter out of range."; return PEHIP_INVALID_ARG; }
  if (p->sweep_mode != PEHIP_SWEEP_COLOURED && p->sweep_mode != PEHIP_SWEEP_REFERENCE_ORDER) {
    why = "Invalid sweep mode.";
    return PEHIP_INVALID_ARG;
  }
  if (p->damping > 1 || p->damping < 0) { why = "Invalid damping."; return PEHIP_INVALID_ARG; }
  if (p->sweep_schedule < PEHIP_SCHEDULE_AUTO || p->sweep_schedule > PEHIP_SCHEDULE_ISLANDS) {
    why = "Invalid sweep schedule.";
    return PEHIP_INVALID_ARG;
  }
  return PEHIP_OK;
}

extern "C" {

void pehip_default_params(pehip_params* p) {
  p->max_iterations = 100;
  p->relaxation_model = PEHIP_APPROX_INELASTIC_COULOMB_DECOUPLING;
  p->erp = 0.7;
  p->abort_threshold = 0.0;
  p->gravity[0] = 0.0; p->gravity[1] = 0.0; p->gravity[2] = 0.0;
  p->damping = 1.0;
  p->relaxation_param = 0.9;
  p->max_sub_iterations = 20;
  p->sweep_mode = PEHIP_SWEEP_COLOURED;
  p->exact_trig = 0;
  p->sweep_schedule = PEHIP_SCHEDULE_AUTO;
}


pehip_status pehip_create(const pehip_params* params, int device, pehip_ctx** out) {
  if (!out) { g_create_error = "out is NULL"; return PEHIP_INVALID_ARG; }
  *out = nullptr;
  pehip_params p;
  if (params) p = *params; else pehip_default_params(&p);
  std::string why;
  pehip_status st = check_params(&p, why);
  if (st != PEHIP_OK) { g_create_error = why; return st; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_error = "no HIP device available";
    return PEHIP_HIP_ERROR;
  }
  if (device < 0 || device >= ndev) { g_create_error = "invalid device ordinal"; return PEHIP_INVALID_ARG; }
  pehip_ctx* ctx = new pehip_ctx();
  ctx->prm = p;
  ctx->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    ctx->ncu = (uint32_t)ncu;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&ctx->hstage, 4096, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&ctx->status, 4 * kStWords) != hipSuccess || hipMalloc((void**)&ctx->maxpen, 64) != hipSuccess) {
    g_create_error = "HIP initialisation failed";
    pehip_destroy(ctx);
    return PEHIP_HIP_ERROR;
  }
  ctx->flags = ctx->status + kStFlags;
  ctx->total = ctx->status + kStTotal;
  ctx->col.remaining = ctx->status + kStRemaining;
  ctx->col.hist = ctx->status + kStHist;
  if (hipMalloc((void**)&ctx->deltas, sizeof(unsigned long long) * (p.max_iterations + 1)) != hipSuccess) {
    g_create_error = "HIP allocation failed";
    pehip_destroy(ctx);
    return PEHIP_OUT_OF_MEMORY;
  }
  for (auto& e : ctx->ev) hipEventCreate(&e);
  ctx->colour_off.reserve(kMaxColours + 1);
  {
    std::string why;
    if (ctx->carve_contacts(nullptr, kScratchContacts) > kScratchBytes) {
      g_create_error = "static scratch smaller than the contact layout";
      pehip_destroy(ctx);
      return PEHIP_HIP_ERROR;
    }
    const pehip_status ss = scratch_address(device, &ctx->scratch, why);
    if (ss != PEHIP_OK) {
      g_create_error = why;
      pehip_destroy(ctx);
      return ss;
    }
  }
  *out = ctx;
  return PEHIP_OK;
}

pehip_status pehip_destroy(pehip_ctx* ctx) {
  if (!ctx) return PEHIP_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  pe_adjoint_free(ctx);
  ctx->free_bodies();
  ctx->free_contacts();
  dfree(ctx->csf); dfree(ctx->cdf); dfree(ctx->status); dfree(ctx->maxpen); dfree(ctx->deltas);
  ctx->flags = ctx->total = ctx->col.remaining = ctx->col.hist = nullptr;
  if (ctx->hstage) hipHostFree(ctx->hstage);
  for (auto& e : ctx->ev) if (e) hipEventDestroy(e);
  for (auto& e : ctx->launch_ev) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return PEHIP_OK;
}

pehip_status pehip_set_params(pehip_ctx* ctx, const pehip_params* params) {
  if (!ctx || !params) return PEHIP_INVALID_ARG;
  std::string why;
  pehip_status st = check_params(params, why);
  if (st != PEHIP_OK) return fail(ctx, st, why);
  if (params->max_iterations != ctx->prm.max_iterations) {
    hipSetDevice(ctx->device);
    CK(hipStreamSynchronize(ctx->stream));
    // the sweep-delta array is replaced only once the new one exists (a failure keeps the old
    // parameters and the array sized for them)
    unsigned long long* deltas = nullptr;
    CK(dalloc(deltas, (size_t)params->max_iterations + 1));
    dfree(ctx->deltas);
    ctx->deltas = deltas;
  }
  ctx->prm = *params;
  return PEHIP_OK;
}

pehip_status pehip_get_params(const pehip_ctx* ctx, pehip_params* params) {
  if (!ctx || !params) return PEHIP_INVALID_ARG;
  *params = ctx->prm;
  return PEHIP_OK;
}

pehip_status pehip_set_materials(pehip_ctx* ctx, uint32_t n, const double* csf_pair, const double* cdf_pair) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n == 0 || !csf_pair || !cdf_pair) return fail(ctx, PEHIP_INVALID_ARG, "empty material table");
  for (size_t i = 0; i < (size_t)n * n; ++i)
    if (!(csf_pair[i] >= 0) || !(cdf_pair[i] >= 0))
      return fail(ctx, PEHIP_INVALID_ARG, "Invalid coefficient of friction!");  // Materials.cpp:203-209
  // uploaded bodies and members index the tables by their material IDs, all below the current
  // count (pe's materials are only ever added, Materials.cpp:175-247)
  if (ctx->n && n < ctx->nmat)
    return fail(ctx, PEHIP_INVALID_ARG, "material table cannot shrink below the bodies' materials");
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  // the new tables replace the old ones only once both are on the device
  const size_t nn = (size_t)n * n;
  double *csf = nullptr, *cdf = nullptr;
  hipError_t e = dalloc(csf, nn);
  if (e == hipSuccess) e = dalloc(cdf, nn);
  if (e == hipSuccess) e = hipMemcpy(csf, csf_pair, sizeof(double) * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(cdf, cdf_pair, sizeof(double) * nn, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dfree(csf); dfree(cdf);
    return fail(ctx, e == hipErrorOutOfMemory ? PEHIP_OUT_OF_MEMORY : PEHIP_HIP_ERROR,
                std::string("material tables: ") + hipGetErrorString(e));
  }
  dfree(ctx->csf); dfree(ctx->cdf);
  ctx->csf = csf;
  ctx->cdf = cdf;
  ctx->nmat = n;
  return PEHIP_OK;
}

pehip_status pehip_upload_bodies(pehip_ctx* ctx, uint32_t n, const uint8_t* shape, const uint64_t* sid,
                                 const uint32_t* material, const uint8_t* fixed, const double* param,
                                 const double* pos, const double* q, const double* v, const double* w,
                                 const double* inv_mass, const double* inertia_body,
                                 const double* inv_inertia_body, const double* plane_normal,
                                 const double* plane_d) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n == 0) return fail(ctx, PEHIP_INVALID_ARG, "no bodies");
  if (n >= kFixedBit) return fail(ctx, PEHIP_INVALID_ARG, "too many bodies");
  if (!shape || !sid || !material || !fixed || !param || !pos || !q || !v || !w || !inv_mass || !inertia_body ||
      !inv_inertia_body || !plane_normal || !plane_d)
    return fail(ctx, PEHIP_INVALID_ARG, "NULL body array");
  if (ctx->nmat == 0) return fail(ctx, PEHIP_INVALID_ARG, "materials must be set before bodies");
  std::vector<uint8_t> fx(fixed, fixed + n);
  std::vector<uint32_t> hplanes;
  for (uint32_t i = 0; i < n; ++i) {
    if (shape[i] < PEHIP_SPHERE || shape[i] > PEHIP_UNION) return fail(ctx, PEHIP_INVALID_ARG, "invalid shape");
    if (material[i] >= ctx->nmat) return fail(ctx, PEHIP_INVALID_ARG, "Invalid material ID");
    if (shape[i] == PEHIP_PLANE) {
      hplanes.push_back(i);
      fx[i] = 1;
      if (inv_mass[i] != 0) return fail(ctx, PEHIP_INVALID_ARG, "plane with non-zero inverse mass");
    } else {
      const double* p = param + 3 * (size_t)i;
      if (shape[i] == PEHIP_SPHERE && !(p[0] > 0)) return fail(ctx, PEHIP_INVALID_ARG, "Invalid sphere radius");
      if (shape[i] == PEHIP_BOX && !(p[0] > 0 && p[1] > 0 && p[2] > 0))
        return fail(ctx, PEHIP_INVALID_ARG, "Invalid side length");
      if (shape[i] == PEHIP_CAPSULE && !(p[0] > 0 && p[1] > 0))
        return fail(ctx, PEHIP_INVALID_ARG, "Invalid capsule radius or length");
      if (shape[i] == PEHIP_UNION && !(p[1] > 0 && p[0] >= p[1] && std::isfinite(p[0])))
        return fail(ctx, PEHIP_INVALID_ARG, "Invalid union bounding radius");
      if (shape[i] == PEHIP_UNION && fx[i]) return fail(ctx, PEHIP_INVALID_ARG, "fixed unions are not supported");
      if (fx[i] && inv_mass[i] != 0) return fail(ctx, PEHIP_INVALID_ARG, "fixed body with non-zero inverse mass");
    }
  }
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  pe_adjoint_free(ctx);
  ctx->free_bodies();
  // a failure from here on leaves no bodies (never a body count over arrays that do not exist)
  const pehip_status st = upload_body_arrays(ctx, n, fx, hplanes, shape, sid, material, param, pos, q, v, w,
                                             inv_mass, inertia_body, inv_inertia_body, plane_normal, plane_d);
  if (st != PEHIP_OK) ctx->free_bodies();
  return st;
}

pehip_status pehip_upload_members(pehip_ctx* ctx, uint32_t nm, const uint32_t* owner, const uint64_t* sid,
                                  const uint32_t* material, const double* radius, const double* rpos,
                                  const double* gpos) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (ctx->n == 0) return fail(ctx, PEHIP_INVALID_ARG, "members need their union bodies uploaded first");
  if (nm == 0 || !owner || !sid || !material || !radius || !rpos || !gpos)
    return fail(ctx, PEHIP_INVALID_ARG, "NULL or empty member array");
  std::vector<uint32_t> start((size_t)ctx->n + 1, 0);
  for (uint32_t k = 0; k < nm; ++k) {
    if (owner[k] >= ctx->n || ctx->h_shape[owner[k]] != PEHIP_UNION)
      return fail(ctx, PEHIP_INVALID_ARG, "member owner is not a union body");
    if (k && owner[k] < owner[k - 1]) return fail(ctx, PEHIP_INVALID_ARG, "members not grouped by union in body order");
    if (material[k] >= ctx->nmat) return fail(ctx, PEHIP_INVALID_ARG, "Invalid material ID");
    if (!(radius[k] > 0)) return fail(ctx, PEHIP_INVALID_ARG, "Invalid sphere radius");
    for (int j = 0; j < 3; ++j)
      if (!std::isfinite(rpos[3 * (size_t)k + j]) || !std::isfinite(gpos[3 * (size_t)k + j]))
        return fail(ctx, PEHIP_INVALID_ARG, "non-finite member position");
    ++start[owner[k] + 1];
  }
  for (uint32_t i = 0; i < ctx->n; ++i)
    if (ctx->h_shape[i] == PEHIP_UNION && start[i + 1] == 0) return fail(ctx, PEHIP_INVALID_ARG, "union without members");
  for (uint32_t i = 0; i < ctx->n; ++i) start[i + 1] += start[i];
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  ctx->free_members();
  // a failure below leaves the unions without members: steps refuse to run until an upload succeeds
  const pehip_status st = [&]() -> pehip_status {
    const size_t D = sizeof(double);
    CK(dalloc(ctx->mstart, (size_t)ctx->n + 1)); CK(dalloc(ctx->mowner, nm)); CK(dalloc(ctx->mmat, nm));
    CK(dalloc(ctx->msid, nm)); CK(dalloc(ctx->mrad, nm)); CK(dalloc(ctx->mrpos, 3 * (size_t)nm));
    CK(dalloc(ctx->mpos, 3 * (size_t)nm));
    CK(hipMemcpy(ctx->mstart, start.data(), 4 * start.size(), hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->mowner, owner, 4 * (size_t)nm, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->mmat, material, 4 * (size_t)nm, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->msid, sid, 8 * (size_t)nm, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->mrad, radius, D * nm, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->mrpos, rpos, 3 * D * nm, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->mpos, gpos, 3 * D * nm, hipMemcpyHostToDevice));
    ctx->nm = nm;
    ctx->members_missing = false;
    launch_aabb(ctx->bodies(), ctx->stream);
    ctx->nc = 0;
    return sync_check(ctx);
  }();
  if (st != PEHIP_OK) {
    ctx->free_members();
    ctx->members_missing = true;
  }
  return st;
}

pehip_status pehip_download_members(pehip_ctx* ctx, uint32_t nm, double* gpos) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (nm != ctx->nm || (nm && !gpos)) return fail(ctx, PEHIP_INVALID_ARG, "member count mismatch");
  if (nm == 0) return PEHIP_OK;
  hipSetDevice(ctx->device);
  CK(hipMemcpyAsync(gpos, ctx->mpos, 3 * sizeof(double) * nm, hipMemcpyDeviceToHost, ctx->stream));
  return sync_check(ctx);
}

pehip_status pehip_upload_state(pehip_ctx* ctx, uint32_t n, const double* pos, const double* q, const double* v,
                                const double* w, const double* plane_normal, const double* plane_d) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n != ctx->n) return fail(ctx, PEHIP_INVALID_ARG, "body count mismatch");
  if (!pos || !q || !v || !w || !plane_normal || !plane_d) return fail(ctx, PEHIP_INVALID_ARG, "NULL state array");
  if (ctx->members_missing) return fail(ctx, PEHIP_INVALID_ARG, "union members not uploaded");
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  std::vector<double> hR(9 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i) {
    qd qq;
    for (int k = 0; k < 4; ++k) qq.v[k] = q[4 * (size_t)i + k];
    const m33 m = quat_to_R(qq);
    for (int k = 0; k < 9; ++k) hR[9 * (size_t)i + k] = m.a[k];
  }
  const size_t D = sizeof(double);
  if (ctx->nm) CKS(move_members(ctx, pos, q, hR));
  CK(hipMemcpy(ctx->pos, pos, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->q, q, 4 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->R, hR.data(), 9 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->v, v, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->w, w, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->pnormal, plane_normal, 3 * D * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->pd, plane_d, D * n, hipMemcpyHostToDevice));
  launch_aabb(ctx->bodies(), ctx->stream);
  return sync_check(ctx);
}

pehip_status pehip_set_forces(pehip_ctx* ctx, uint32_t n, const double* force, const double* torque) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n != ctx->n || !force || !torque) return fail(ctx, PEHIP_INVALID_ARG, "force arrays do not match the bodies");
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  CK(hipMemcpy(ctx->force, force, 3 * sizeof(double) * n, hipMemcpyHostToDevice));
  CK(hipMemcpy(ctx->torque, torque, 3 * sizeof(double) * n, hipMemcpyHostToDevice));
  return PEHIP_OK;
}

pehip_status pehip_set_contact_order(pehip_ctx* ctx, uint32_t n, const uint32_t* body1, const uint32_t* body2) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n && (!body1 || !body2)) return fail(ctx, PEHIP_INVALID_ARG, "NULL order arrays");
  ctx->order_b1.assign(body1, body1 + n);
  ctx->order_b2.assign(body2, body2 + n);
  ctx->order_pending = true;
  return PEHIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// The step.
// ---------------------------------------------------------------------------------------------
static void mark(pehip_ctx* ctx, int k) {
  if (ctx->profiling) hipEventRecord(ctx->ev[k], ctx->stream);
}

static pehip_status record_launch_event(pehip_ctx* ctx, size_t k) {
  while (ctx->launch_ev.size() <= k) {
    hipEvent_t e;
    CK(hipEventCreate(&e));
    ctx->launch_ev.push_back(e);
  }
  CK(hipEventRecord(ctx->launch_ev[k], ctx->stream));
  return PEHIP_OK;
}

// HashGrids::findContacts equivalent: keys, sort, cell ranges, count, scan, write.
static pehip_status detect(pehip_ctx* ctx) {
  const hipStream_t s = ctx->stream;
  Bodies B = ctx->bodies();
  fill_u32(ctx->flags, 0u, 1, s);
  launch_cell_keys(B, ctx->grid, ctx->bp, s);
  radix_sort_pairs(ctx->bp.keys, ctx->bp.vals, ctx->bp.keys_tmp, ctx->bp.vals_tmp, ctx->n, ctx->key_bits,
                   ctx->sort_scratch, s);
  launch_cell_ranges(ctx->n, ctx->grid, ctx->bp, s);
  launch_sorted_boxes(B, ctx->bp, s);
  launch_pairs(B, ctx->grid, ctx->bp, ctx->raw, ctx->csf, ctx->cdf, ctx->nmat, false, ctx->spheres_only, s);
  const uint32_t ngrid = ctx->n - ctx->bp.nplanes;
  exclusive_scan_u32(ctx->bp.counts, ctx->bp.offsets, ngrid, ctx->total, ctx->scan_scratch, s);
  CK(hipMemcpyAsync(ctx->hstage, ctx->status, 4 * (kStTotal + 1), hipMemcpyDeviceToHost, s));
  CKS(sync_check(ctx));
  const uint32_t nc = ctx->hstage[kStTotal], fl = ctx->hstage[kStFlags];
  if (fl & ERR_NONFINITE) return fail(ctx, PEHIP_NONFINITE, "non-finite body position or bounding box");
  if (fl & ERR_LEVEL_OVERFLOW) return fail(ctx, PEHIP_INVALID_ARG, "body larger than the coarsest grid cell");
  CKS(ensure_contact_capacity(ctx, nc));
  if (nc) {
    launch_pairs_place(B, ctx->bp, ctx->raw, ctx->csf, ctx->cdf, ctx->nmat, s);
    if (fl & ERR_STAGE_FULL)
      launch_pairs(B, ctx->grid, ctx->bp, ctx->raw, ctx->csf, ctx->cdf, ctx->nmat, true, ctx->spheres_only, s);
  }
  ctx->nc = nc;
  return PEHIP_OK;
}

// Contact-graph colouring -> ctx->perm (colour-major) and ctx->colour_off.
static pehip_status colour(pehip_ctx* ctx) {
  const hipStream_t s = ctx->stream;
  const uint32_t nc = ctx->nc;
  Bodies B = ctx->bodies();
  launch_adjacency(B, ctx->raw, nc, ctx->col, ctx->scan_scratch, s);
  launch_colour_init(ctx->raw, B, nc, ctx->col, s);
  // Rounds in batches, the first as long as the last step's colouring, each batch followed by the
  // histogram and one read-back of (remaining, histogram, flags): a step whose colouring needs no
  // more rounds than the last one's synchronises once.  A round after the last is a no-op (every
  // contact is coloured), so the colours do not depend on the batching.
  uint32_t round = 0, batch = std::max(ctx->colour_rounds, 4u);
  for (;;) {
    for (uint32_t k = 0; k < batch; ++k) launch_colour_round(ctx->raw, B, nc, ctx->col, round++, ctx->flags, s);
    launch_colour_hist(nc, ctx->col, s);
    CK(hipMemcpyAsync(ctx->hstage, ctx->status, 4 * kStWords, hipMemcpyDeviceToHost, s));
    CKS(sync_check(ctx));
    if (ctx->hstage[kStRemaining] == 0) break;
    if (round > 4u * 4096u) return fail(ctx, PEHIP_HIP_ERROR, "colouring did not terminate");
    batch = 4;
  }
  // the rounds this colouring needed, and one more: the next step's first batch
  ctx->colour_rounds = ctx->hstage[kStRemaining + 1] + 2;
  if (ctx->hstage[kStFlags] & ERR_COLOURS) return fail(ctx, PEHIP_NOT_SUPPORTED, "more than 128 colours needed");
  const uint32_t* hist = ctx->hstage + kStHist;
  uint32_t ncol = 0;
  for (uint32_t k = 0; k < kMaxColours; ++k)
    if (hist[k]) ncol = k + 1;
  ctx->colour_off.assign(ncol + 1, 0);
  for (uint32_t k = 0; k < ncol; ++k) ctx->colour_off[k + 1] = ctx->colour_off[k] + hist[k];
  // stable sort of contact indices by colour (colours < 128: one 8-bit pass, straight into perm)
  iota_u32(ctx->sort_vals_tmp, nc, s);
  radix_sort_pairs_8bit(ctx->col.colour, ctx->sort_vals_tmp, ctx->sort_keys, ctx->perm, nc, ctx->csort_scratch, s);
  ctx->stats.num_colours = ncol;
  return PEHIP_OK;
}

// Reference order: map the requested (body1, body2) sequence onto the detected contacts.
static pehip_status reference_perm(pehip_ctx* ctx) {
  const uint32_t nc = ctx->nc;
  if (!ctx->order_pending) return fail(ctx, PEHIP_INVALID_ARG, "reference-order mode needs pehip_set_contact_order");
  ctx->order_pending = false;
  if (ctx->order_b1.size() != nc)
    return fail(ctx, PEHIP_INVALID_ARG, "contact order does not match the detected contact set (count " +
                                            std::to_string(ctx->order_b1.size()) + " vs " + std::to_string(nc) + ")");
  std::vector<uint32_t> b1(nc), b2(nc);
  if (nc) {
    // the stream is non-blocking: wait for the narrowphase write pass before reading
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(b1.data(), ctx->raw.b1, 4 * (size_t)nc, hipMemcpyDeviceToHost));
    CK(hipMemcpy(b2.data(), ctx->raw.b2, 4 * (size_t)nc, hipMemcpyDeviceToHost));
  }
  std::unordered_map<uint64_t, std::vector<uint32_t>> byPair;
  for (uint32_t c = 0; c < nc; ++c) byPair[((uint64_t)b1[c] << 32) | b2[c]].push_back(c);
  std::unordered_map<uint64_t, size_t> used;
  std::vector<uint32_t> perm(nc);
  for (uint32_t k = 0; k < nc; ++k) {
    const uint64_t key = ((uint64_t)ctx->order_b1[k] << 32) | ctx->order_b2[k];
    auto it = byPair.find(key);
    size_t& u = used[key];
    if (it == byPair.end() || u >= it->second.size())
      return fail(ctx, PEHIP_INVALID_ARG, "contact order does not match the detected contact set (pair " +
                                              std::to_string(ctx->order_b1[k]) + "," + std::to_string(ctx->order_b2[k]) + ")");
    perm[k] = it->second[u++];
  }
  if (nc) CK(hipMemcpy(ctx->perm, perm.data(), 4 * (size_t)nc, hipMemcpyHostToDevice));
  ctx->colour_off.assign(2, 0);
  ctx->colour_off[1] = nc;
  ctx->stats.num_colours = 0;
  return PEHIP_OK;
}

// Island sweeps (PEHIP_SCHEDULE_ISLANDS): the contact graph's islands are found and packed into
// units on the device; one launch runs every small and large island's sweeps, and the islands over
// kIslandLarge contacts go to the dataflow sweeps (over their slots only) first.
static pehip_status island_sweeps(pehip_ctx* ctx, const Bodies& B, double dtinv, bool exact, bool* swept) {
  const pehip_params& P = ctx->prm;
  const hipStream_t s = ctx->stream;
  const uint32_t nc = ctx->nc;
  *swept = false;
  ColourOffsets O;
  O.ncol = (uint32_t)ctx->colour_off.size() - 1;
  for (uint32_t k = 0; k <= O.ncol; ++k) O.off[k] = ctx->colour_off[k];
  launch_island_prepare(ctx->sol, nc, O, B, ctx->col.deg, ctx->isl, s);
  CK(hipMemcpyAsync(ctx->hstage, ctx->isl.hdr + ISL_SMALL_UNITS, 12, hipMemcpyDeviceToHost, s));
  CKS(sync_check(ctx));
  const uint32_t su = ctx->hstage[0], lu = ctx->hstage[1], dfc = ctx->hstage[2];
  if (dfc && (dfc > kDataflowMaxContacts || ctx->dataflow_refused))
    return fail(ctx, PEHIP_NOT_SUPPORTED, "island sweeps: " + std::to_string(dfc) +
                                              " contacts in islands over 256 contacts need the dataflow sweeps, " +
                                              (ctx->dataflow_refused ? ctx->dataflow_refusal
                                                                     : std::string("which take at most 65536")));
  launch_island_assign(ctx->sol, nc, O, B, ctx->col.deg, ctx->isl, su, lu, dfc, ctx->cscan_scratch, s);
  size_t nev = 0;
  if (dfc) {
    DataflowBuffers D;
    D.perm = ctx->perm; D.adj_start = ctx->col.adj_start; D.adj = ctx->col.adj; D.deg = ctx->col.deg;
    D.inv = ctx->df_inv; D.rank1 = ctx->df_rank1; D.rank2 = ctx->df_rank2; D.done = ctx->df_done;
    D.deltas = ctx->deltas; D.flags = ctx->flags; D.spin_limit = kDataflowSpinLimit;
    fill_u32(ctx->flags, 0u, 1, s);
    launch_dataflow_prepare(ctx->sol, nc, B, D, s);
    if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
    DataflowLaunchInfo info;
    if (launch_sweep_dataflow(P.relaxation_model, exact, ctx->sol, dfc, P.max_iterations, B, dtinv,
                              (int)P.max_sub_iterations, D, ctx->ncu, s, &info, ctx->isl.dslots) != SWEEP_OK) {
      const std::string why = std::string("dataflow sweep launch refused: occupancy query ") +
                              hipGetErrorName(info.occupancy_err) + ", " + std::to_string(info.per_cu) +
                              " workgroups per CU, " + std::to_string(ctx->ncu) + " CUs, grid " +
                              std::to_string(info.grid) + ", cooperative launch " + hipGetErrorName(info.launch_err);
      (void)hipGetLastError();  // the refusal is not a sticky error
      return fail(ctx, PEHIP_HIP_ERROR, why);
    }
    if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
  }
  if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
  launch_sweep_islands(P.relaxation_model, exact, ctx->sol, ctx->isl, su, lu, P.max_iterations, B,
                       (int)P.max_sub_iterations, ctx->deltas, s);
  if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
  ctx->n_launch_events = ctx->profiling ? nev : 0;
  if (dfc) {
    CK(hipMemcpyAsync(ctx->hstage, ctx->flags, 4, hipMemcpyDeviceToHost, s));
    CKS(sync_check(ctx));
    if (ctx->hstage[0] & ERR_TIMEOUT) return fail(ctx, PEHIP_HIP_ERROR, "dataflow sweep timed out");
    ++ctx->dataflow_steps;
  }
  *swept = true;
  return PEHIP_OK;
}

pehip_status pe_one_step(pehip_ctx* ctx, double dt) {
  const hipStream_t s = ctx->stream;
  const double dtinv = 1.0 / dt;
  const pehip_params& P = ctx->prm;
  const bool exact = P.exact_trig != 0 || P.sweep_mode == PEHIP_SWEEP_REFERENCE_ORDER;
  Bodies B = ctx->bodies();
  mark(ctx, 0);
  CKS(detect(ctx));
  mark(ctx, 1);
  const d3 g = mk3(P.gravity[0], P.gravity[1], P.gravity[2]);
  launch_predict(B, g, dt, s);
  const uint32_t nc = ctx->nc;
  uint32_t iters = 0;
  int32_t schedule_used = -1;
  bool dataflow_ran = false;
  double last_delta = 0;
  fill_u64(ctx->maxpen, 0ull, 2, s);
  if (nc) {
    if (P.sweep_mode == PEHIP_SWEEP_REFERENCE_ORDER) CKS(reference_perm(ctx));
    else CKS(colour(ctx));
  } else {
    ctx->order_pending = false;
    ctx->stats.num_colours = 0;
    ctx->colour_off.clear();
  }
  mark(ctx, 2);
  if (nc) {
    launch_setup(B, ctx->raw, ctx->perm, nc, ctx->sol, P.erp, dtinv, ctx->maxpen, s);
    fill_u64(ctx->deltas, 0ull, (size_t)P.max_iterations + 1, s);
  }
  mark(ctx, 3);
  if (nc && P.max_iterations) {
    if (P.sweep_mode == PEHIP_SWEEP_REFERENCE_ORDER) {
      launch_sweep_serial(P.relaxation_model, exact, ctx->sol, nc, P.max_iterations, B, dtinv, (int)P.max_sub_iterations,
                          P.abort_threshold, ctx->deltas, s);
      CK(hipMemcpyAsync(ctx->hstage, ctx->deltas + P.max_iterations, 8, hipMemcpyDeviceToHost, s));
      CKS(sync_check(ctx));
      iters = (uint32_t)(*(unsigned long long*)ctx->hstage);
    } else if (P.sweep_schedule == PEHIP_SCHEDULE_SINGLE_BLOCK ||
               (P.sweep_schedule == PEHIP_SCHEDULE_AUTO && P.abort_threshold > 0 &&
                P.relaxation_model != PEHIP_INELASTIC_COULOMB_DECOUPLING &&
                P.relaxation_model != PEHIP_INELASTIC_GMD && nc <= kSingleBlockMaxContacts)) {
      ColourOffsets O;
      O.ncol = (uint32_t)ctx->colour_off.size() - 1;
      for (uint32_t k = 0; k <= O.ncol; ++k) O.off[k] = ctx->colour_off[k];
      if (ctx->profiling) CKS(record_launch_event(ctx, 0));
      launch_sweep_block(P.relaxation_model, exact, ctx->sol, O, P.max_iterations, B, dtinv,
                         (int)P.max_sub_iterations, P.abort_threshold, ctx->deltas, s);
      if (ctx->profiling) CKS(record_launch_event(ctx, 1));
      ctx->n_launch_events = ctx->profiling ? 2 : 0;
      CK(hipMemcpyAsync(ctx->hstage, ctx->deltas + P.max_iterations, 8, hipMemcpyDeviceToHost, s));
      CKS(sync_check(ctx));
      iters = (uint32_t)(*(unsigned long long*)ctx->hstage);
      schedule_used = PEHIP_SCHEDULE_SINGLE_BLOCK;
    } else {
      // All sweeps in one cooperative launch (same Gauss-Seidel order per body as the colours).
      // AUTO falls back to the colour launches (the same bits) when the device refuses the
      // cooperative launch, e.g. while another process holds the device's cooperative-launch
      // resources; nothing has run then.  An explicit DATAFLOW schedule reports the refusal.
      bool swept = false;
      if (P.abort_threshold <= 0 && P.sweep_schedule == PEHIP_SCHEDULE_ISLANDS) {
        CKS(island_sweeps(ctx, B, dtinv, exact, &swept));
        if (swept) {
          iters = P.max_iterations;
          schedule_used = PEHIP_SCHEDULE_ISLANDS;
        }
      }
      if (!swept && P.abort_threshold <= 0 &&
          (P.sweep_schedule == PEHIP_SCHEDULE_DATAFLOW ||
           (P.sweep_schedule == PEHIP_SCHEDULE_AUTO && nc <= kDataflowMaxContacts && !ctx->dataflow_refused))) {
        DataflowBuffers D;
        D.perm = ctx->perm; D.adj_start = ctx->col.adj_start; D.adj = ctx->col.adj; D.deg = ctx->col.deg;
        D.inv = ctx->df_inv; D.rank1 = ctx->df_rank1; D.rank2 = ctx->df_rank2; D.done = ctx->df_done;
        D.deltas = ctx->deltas; D.flags = ctx->flags; D.spin_limit = kDataflowSpinLimit;
        fill_u32(ctx->flags, 0u, 1, s);
        launch_dataflow_prepare(ctx->sol, nc, B, D, s);
        if (ctx->profiling) CKS(record_launch_event(ctx, 0));
        DataflowLaunchInfo info;
        if (launch_sweep_dataflow(P.relaxation_model, exact, ctx->sol, nc, P.max_iterations, B, dtinv,
                                  (int)P.max_sub_iterations, D, ctx->ncu, s, &info) != SWEEP_OK) {
          const std::string why = std::string("dataflow sweep launch refused: occupancy query ") +
                                  hipGetErrorName(info.occupancy_err) + ", " + std::to_string(info.per_cu) +
                                  " workgroups per CU, " + std::to_string(ctx->ncu) + " CUs, grid " +
                                  std::to_string(info.grid) + ", cooperative launch " +
                                  hipGetErrorName(info.launch_err);
          (void)hipGetLastError();  // the refusal is not a sticky error
          if (P.sweep_schedule == PEHIP_SCHEDULE_DATAFLOW) return fail(ctx, PEHIP_HIP_ERROR, why);
          ctx->dataflow_refused = true;
          ctx->dataflow_refusal = why;
          std::fprintf(stderr, "pe_amd: %s; this context uses colour launches from now on\n", why.c_str());
        } else {
          if (ctx->profiling) CKS(record_launch_event(ctx, 1));
          ctx->n_launch_events = ctx->profiling ? 2 : 0;
          // its timeout flag is read with the step's statistics (one synchronisation per step
          // fewer): a timed-out step still fails, after the integration it no longer matters for
          dataflow_ran = true;
          iters = P.max_iterations;
          schedule_used = PEHIP_SCHEDULE_DATAFLOW;
          ++ctx->dataflow_steps;
          swept = true;
        }
      }
      if (!swept) {
        const uint32_t ncol = (uint32_t)ctx->colour_off.size() - 1;
        size_t nev = 0;
        schedule_used = PEHIP_SCHEDULE_COLOUR_LAUNCHES;
        // pe's abort test (HCSITS.h:1935-1946 semantics: stop after the first sweep whose impulse
        // change is below the threshold) runs on the device after every sweep; the sweeps queued
        // behind a stop return at once, and the host looks at the flag every kAbortPoll sweeps.
        const bool abort_on = P.abort_threshold > 0;
        uint32_t* stop = ctx->flags + 4;
        if (abort_on) fill_u32(stop, 0u, 2, s);
        iters = P.max_iterations;
        for (uint32_t it = 0; it < P.max_iterations; ++it) {
          for (uint32_t k = 0; k < ncol; ++k) {
            if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
            launch_sweep(P.relaxation_model, exact, ctx->sol, ctx->colour_off[k], ctx->colour_off[k + 1], B, dtinv,
                         (int)P.max_sub_iterations, ctx->deltas + it, abort_on ? stop : nullptr, ctx->ncu, s);
            if (ctx->profiling) CKS(record_launch_event(ctx, nev++));
          }
          if (!abort_on) continue;
          launch_abort_check(ctx->deltas, it, P.abort_threshold, stop, s);
          if ((it + 1) % kAbortPoll == 0 || it + 1 == P.max_iterations) {
            CK(hipMemcpyAsync(ctx->hstage, stop, 8, hipMemcpyDeviceToHost, s));
            CKS(sync_check(ctx));
            if (ctx->hstage[0]) {
              iters = ctx->hstage[1];
              break;
            }
          }
        }
        // events of the launches that returned at once (after the stop) are not sweep launches
        ctx->n_launch_events = std::min(nev, (size_t)2 * ncol * iters);
      }
    }
  }
  mark(ctx, 4);
  launch_integrate(B, dt, P.damping, exact, s);
  mark(ctx, 5);
  // stats of this step
  CK(hipMemcpyAsync(ctx->hstage, ctx->maxpen, 16, hipMemcpyDeviceToHost, s));
  if (iters) CK(hipMemcpyAsync(ctx->hstage + 4, ctx->deltas + (iters - 1), 8, hipMemcpyDeviceToHost, s));
  if (dataflow_ran) CK(hipMemcpyAsync(ctx->hstage + 6, ctx->flags, 4, hipMemcpyDeviceToHost, s));
  CKS(sync_check(ctx));
  if (dataflow_ran && (ctx->hstage[6] & ERR_TIMEOUT)) return fail(ctx, PEHIP_HIP_ERROR, "dataflow sweep timed out");
  double mp;
  std::memcpy(&mp, ctx->hstage, 8);
  unsigned long long nfx;
  std::memcpy(&nfx, ctx->hstage + 2, 8);
  if (iters) std::memcpy(&last_delta, ctx->hstage + 4, 8);
  ctx->stats.num_contacts = nc;
  ctx->stats.schedule_used = schedule_used;
  ctx->stats.max_penetration = mp;
  ctx->stats.iterations = iters;
  ctx->stats.last_delta = last_delta;
  ctx->stats.fixed_contacts = nc ? nfx : 0;
  // Compulsory bytes of one sweep: per contact its record (r1, r2, n, t, o, dist, mu, Dn^-1,
  // D^-1 sym, Dto^-1 sym, Dto, p read+write: 36 doubles + 2 indices = 296 B) and per
  // non-fixed contact end the body's v, w, m^-1, Iw^-1 sym and dv, dw read+write (200 B).
  const double sweep_bytes = nc ? (double)iters * (296.0 * nc + 200.0 * (2.0 * nc - (double)nfx)) : 0.0;
  ctx->stats.sweep_bytes = sweep_bytes;
  if (ctx->profiling && P.sweep_mode == PEHIP_SWEEP_COLOURED && nc && iters) {
    double ms = 0;
    for (size_t k = 0; k + 1 < ctx->n_launch_events; k += 2) {
      float t = 0;
      hipEventElapsedTime(&t, ctx->launch_ev[k], ctx->launch_ev[k + 1]);
      ms += t;
    }
    ctx->stats.prof_sweep_launches += ctx->n_launch_events / 2;
    ctx->stats.prof_sweep_kernel_ms += ms;
    ctx->stats.prof_sweep_bytes += sweep_bytes;
  }
  if (ctx->profiling) ++ctx->stats.prof_steps;
  if (ctx->profiling) {
    float t[5];
    for (int k = 0; k < 5; ++k) hipEventElapsedTime(&t[k], ctx->ev[k], ctx->ev[k + 1]);
    ctx->stats.t_detection_ms = t[0];
    ctx->stats.t_colouring_ms = t[1];
    ctx->stats.t_setup_ms = t[2];
    ctx->stats.t_solve_ms = t[3];
    ctx->stats.t_integrate_ms = t[4];
  }
  return PEHIP_OK;
}

extern "C" {

pehip_status pehip_step(pehip_ctx* ctx, double dt, uint32_t nsteps) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (!(dt > 0)) return fail(ctx, PEHIP_INVALID_ARG, "Invalid time step size");  // World.cpp:333-334
  if (ctx->n == 0) return fail(ctx, PEHIP_INVALID_ARG, "no bodies uploaded");
  if (ctx->members_missing) return fail(ctx, PEHIP_INVALID_ARG, "union members not uploaded");
  hipSetDevice(ctx->device);
  for (uint32_t k = 0; k < nsteps; ++k) CKS(pe_one_step(ctx, dt));
  return PEHIP_OK;
}

pehip_status pehip_step_noalloc(pehip_ctx* ctx, double dt, uint32_t nsteps) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (!(dt > 0)) return fail(ctx, PEHIP_INVALID_ARG, "Invalid time step size");
  if (ctx->n == 0) return fail(ctx, PEHIP_INVALID_ARG, "no bodies uploaded");
  if (ctx->members_missing) return fail(ctx, PEHIP_INVALID_ARG, "union members not uploaded");
  if (ctx->prm.sweep_mode != PEHIP_SWEEP_COLOURED)
    return fail(ctx, PEHIP_NOT_SUPPORTED, "pehip_step_noalloc runs the coloured sweep mode only");
  if (ctx->profiling) return fail(ctx, PEHIP_NOT_SUPPORTED, "pehip_step_noalloc does not profile");
  hipSetDevice(ctx->device);
  if (!ctx->contacts_in_scratch) {
    CK(hipStreamSynchronize(ctx->stream));
    ctx->unbind_contacts();
    ctx->carve_contacts(static_cast<char*>(ctx->scratch), kScratchContacts);
    ctx->ccap = kScratchContacts;
    ctx->contacts_in_scratch = true;
  }
  ctx->no_alloc = true;
  pehip_status st = PEHIP_OK;
  for (uint32_t k = 0; k < nsteps && st == PEHIP_OK; ++k) st = pe_one_step(ctx, dt);
  ctx->no_alloc = false;
  return st;
}

pehip_status pehip_detect(pehip_ctx* ctx) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (ctx->n == 0) return fail(ctx, PEHIP_INVALID_ARG, "no bodies uploaded");
  if (ctx->members_missing) return fail(ctx, PEHIP_INVALID_ARG, "union members not uploaded");
  hipSetDevice(ctx->device);
  CKS(detect(ctx));
  return sync_check(ctx);
}

pehip_status pehip_download_state(pehip_ctx* ctx, uint32_t n, double* pos, double* q, double* v, double* w,
                                  double* plane_normal, double* plane_d) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (n != ctx->n) return fail(ctx, PEHIP_INVALID_ARG, "body count mismatch");
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  const size_t D = sizeof(double);
  if (pos) CK(hipMemcpy(pos, ctx->pos, 3 * D * n, hipMemcpyDeviceToHost));
  if (q) CK(hipMemcpy(q, ctx->q, 4 * D * n, hipMemcpyDeviceToHost));
  if (v) CK(hipMemcpy(v, ctx->v, 3 * D * n, hipMemcpyDeviceToHost));
  if (w) CK(hipMemcpy(w, ctx->w, 3 * D * n, hipMemcpyDeviceToHost));
  if (plane_normal) CK(hipMemcpy(plane_normal, ctx->pnormal, 3 * D * n, hipMemcpyDeviceToHost));
  if (plane_d) CK(hipMemcpy(plane_d, ctx->pd, D * n, hipMemcpyDeviceToHost));
  return PEHIP_OK;
}

pehip_status pehip_get_contacts(pehip_ctx* ctx, size_t cap, size_t* n, uint32_t* body1, uint32_t* body2,
                                double* pos, double* normal, double* dist, double* mu) {
  if (!ctx || !n) return PEHIP_INVALID_ARG;
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  *n = ctx->nc;
  const size_t m = std::min<size_t>(cap, ctx->nc);
  if (!m) return PEHIP_OK;
  if (body1) CK(hipMemcpy(body1, ctx->raw.b1, 4 * m, hipMemcpyDeviceToHost));
  if (body2) CK(hipMemcpy(body2, ctx->raw.b2, 4 * m, hipMemcpyDeviceToHost));
  if (pos) CK(hipMemcpy(pos, ctx->raw.pos, 24 * m, hipMemcpyDeviceToHost));
  if (normal) CK(hipMemcpy(normal, ctx->raw.nrm, 24 * m, hipMemcpyDeviceToHost));
  if (dist) CK(hipMemcpy(dist, ctx->raw.dist, 8 * m, hipMemcpyDeviceToHost));
  if (mu) CK(hipMemcpy(mu, ctx->raw.mu, 8 * m, hipMemcpyDeviceToHost));
  return PEHIP_OK;
}

pehip_status pehip_get_sweep_order(pehip_ctx* ctx, size_t cap, size_t* n, uint32_t* contact_index,
                                   uint32_t* ncolours, uint32_t* colour_offsets) {
  if (!ctx || !n) return PEHIP_INVALID_ARG;
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  *n = ctx->nc;
  const size_t m = std::min<size_t>(cap, ctx->nc);
  if (m && contact_index) CK(hipMemcpy(contact_index, ctx->perm, 4 * m, hipMemcpyDeviceToHost));
  const uint32_t ncol = ctx->colour_off.empty() ? 0u : (uint32_t)ctx->colour_off.size() - 1;
  if (ncolours) *ncolours = ncol;
  if (colour_offsets)
    for (uint32_t k = 0; k <= ncol && !ctx->colour_off.empty(); ++k) colour_offsets[k] = ctx->colour_off[k];
  return PEHIP_OK;
}

pehip_status pehip_get_impulses(pehip_ctx* ctx, size_t cap, size_t* n, double* p) {
  if (!ctx || !n) return PEHIP_INVALID_ARG;
  hipSetDevice(ctx->device);
  CK(hipStreamSynchronize(ctx->stream));
  const uint32_t nc = ctx->nc;
  *n = nc;
  if (!nc || !p) return PEHIP_OK;
  std::vector<uint32_t> perm(nc);
  std::vector<double> ps(3 * (size_t)nc);
  CK(hipMemcpy(perm.data(), ctx->perm, 4 * (size_t)nc, hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; ++k)
    CK(hipMemcpy(ps.data() + (size_t)k * nc, ctx->sol.f + (size_t)(CF_P + k) * ctx->sol.cap, 8 * (size_t)nc,
                 hipMemcpyDeviceToHost));
  for (uint32_t sidx = 0; sidx < nc; ++sidx) {
    const uint32_t c = perm[sidx];
    if (c < cap)
      for (int k = 0; k < 3; ++k) p[3 * (size_t)c + k] = ps[(size_t)k * nc + sidx];
  }
  return PEHIP_OK;
}

pehip_status pehip_get_stats(pehip_ctx* ctx, pehip_stats* out) {
  if (!ctx || !out) return PEHIP_INVALID_ARG;
  hipSetDevice(ctx->device);
  if (ctx->n) {
    const uint32_t nb = launch_kinetic(ctx->bodies(), ctx->ke_partial, ctx->stream);
    std::vector<double> part(nb);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(part.data(), ctx->ke_partial, 8 * (size_t)nb, hipMemcpyDeviceToHost));
    double e = 0;
    for (double x : part) e += x;
    ctx->stats.kinetic_energy = e;
  }
  *out = ctx->stats;
  return PEHIP_OK;
}

pehip_status pehip_get_step_stats(pehip_ctx* ctx, pehip_stats* out) {
  if (!ctx || !out) return PEHIP_INVALID_ARG;
  *out = ctx->stats;
  return PEHIP_OK;
}

pehip_status pehip_set_profiling(pehip_ctx* ctx, int enable) {
  if (!ctx) return PEHIP_INVALID_ARG;
  ctx->profiling = enable != 0;
  if (ctx->profiling) {
    ctx->stats.prof_steps = 0;
    ctx->stats.prof_sweep_launches = 0;
    ctx->stats.prof_sweep_kernel_ms = 0;
    ctx->stats.prof_sweep_bytes = 0;
  }
  return PEHIP_OK;
}

pehip_status pehip_eval_trig(pehip_ctx* ctx, int fn, int exact, size_t n, const double* x, const double* y,
                             double* out) {
  if (!ctx) return PEHIP_INVALID_ARG;
  if (fn < 0 || fn > 5 || !x || !out || (fn == 2 && !y)) return fail(ctx, PEHIP_INVALID_ARG, "invalid trig request");
  if (!n) return PEHIP_OK;
  hipSetDevice(ctx->device);
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  const pehip_status st = [&]() -> pehip_status {
    CK(dalloc(dx, n));
    CK(dalloc(dy, n));
    CK(dalloc(dout, n));
    CK(hipMemcpy(dx, x, 8 * n, hipMemcpyHostToDevice));
    CK(hipMemcpy(dy, y ? y : x, 8 * n, hipMemcpyHostToDevice));
    launch_eval_trig(fn, exact != 0, n, dx, dy, dout, ctx->stream);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(out, dout, 8 * n, hipMemcpyDeviceToHost));
    return sync_check(ctx);
  }();
  dfree(dx); dfree(dy); dfree(dout);
  return st;
}

void* pehip_stream(pehip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

const char* pehip_last_error(const pehip_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

const char* pehip_version(void) { return kVersion; }

}  // extern "C"
