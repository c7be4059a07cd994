// Host side of the batched-scene C-ABI: scene planning (per-scene hash grid, contact
// slots, processing order), the persistent-kernel launches (optionally replayed from a
// captured hipGraph), contact-capacity growth, and the autotuner.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batch.hpp"

using namespace pe_amd;

namespace {

template <class T>
hipError_t balloc(T*& p, size_t n) {
  p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)&p, n * sizeof(T));
}
template <class T>
void bfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

std::string g_batch_create_error;

}  // namespace

struct pehip_batch {
  pehip_params prm;
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  std::string err;
  pehip_batch_tuning tun{256, 2, 1, 1};

  // bodies (all scenes, concatenated)
  uint64_t n = 0;
  uint8_t *shape = nullptr, *fixed = nullptr;
  uint32_t* mat = nullptr;
  uint64_t* sid = nullptr;
  double *param = nullptr, *pos = nullptr, *q = nullptr, *R = nullptr, *v = nullptr, *w = nullptr;
  double *pnormal = nullptr, *pd = nullptr, *aabb = nullptr, *inv_mass = nullptr, *Ib = nullptr, *Ibinv = nullptr;
  double *force = nullptr, *torque = nullptr, *sb = nullptr, *dvw = nullptr;
  uint32_t *bucket = nullptr, *sorted = nullptr, *count = nullptr;

  // materials
  uint32_t nmat = 0;
  double *csf = nullptr, *cdf = nullptr;

  // scenes
  uint32_t nscenes = 0;
  std::vector<SceneInfo> hinfo;
  std::vector<uint32_t> hplane_idx;
  SceneInfo* info = nullptr;
  uint32_t *order = nullptr, *plane_idx = nullptr, *colour_off = nullptr, *queue = nullptr, *any_flags = nullptr;
  SceneOut* out = nullptr;
  bool scene_mu = false;
  uint32_t capacity_retries = 0;
  uint32_t launches = 0;

  // contacts
  uint64_t ccap = 0;
  RawContacts raw{};
  SolverContacts sol{};
  uint32_t *colour = nullptr, *perm = nullptr;

  // captured step graph
  hipGraphExec_t gexec = nullptr;
  double g_dt = 0;
  pehip_batch_tuning g_tun{};
  pehip_params g_prm{};
  uint32_t* hstage = nullptr;  // pinned

  Bodies bodies() const {
    Bodies B;
    B.n = (uint32_t)n; B.shape = shape; B.fixed = fixed; B.mat = mat; B.sid = sid; B.param = param;
    B.pos = pos; B.q = q; B.R = R; B.v = v; B.w = w; B.pnormal = pnormal; B.pd = pd; B.aabb = aabb;
    B.inv_mass = inv_mass; B.Ib = Ib; B.Ibinv = Ibinv; B.force = force; B.torque = torque;
    B.sb = sb; B.dvw = dvw;
    return B;
  }
  void drop_graph() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
  }
  void free_contacts() {
    bfree(raw.b1); bfree(raw.b2); bfree(raw.pos); bfree(raw.nrm); bfree(raw.dist); bfree(raw.mu);
    bfree(sol.b1); bfree(sol.b2); bfree(sol.f); bfree(colour); bfree(perm);
    ccap = 0;
    drop_graph();
  }
  void free_all() {
    free_contacts();
    bfree(shape); bfree(fixed); bfree(mat); bfree(sid); bfree(param); bfree(pos); bfree(q); bfree(R);
    bfree(v); bfree(w); bfree(pnormal); bfree(pd); bfree(aabb); bfree(inv_mass); bfree(Ib); bfree(Ibinv);
    bfree(force); bfree(torque); bfree(sb); bfree(dvw); bfree(bucket); bfree(sorted); bfree(count);
    bfree(info); bfree(order); bfree(plane_idx); bfree(colour_off); bfree(out);
    n = 0;
    nscenes = 0;
  }
};

#define BK(call)                                                                 \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      b->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
      return e_ == hipErrorOutOfMemory ? PEHIP_OUT_OF_MEMORY : PEHIP_HIP_ERROR;  \
    }                                                                            \
  } while (0)

#define BKS(st)                    \
  do {                             \
    pehip_status s_ = (st);        \
    if (s_ != PEHIP_OK) return s_; \
  } while (0)

static pehip_status bfail(pehip_batch* b, pehip_status st, const std::string& msg) {
  b->err = msg;
  return st;
}

// Contact slots of every scene from hinfo[].ccap; (re)allocates the contact arrays.
static pehip_status layout_contacts_into(pehip_batch* b);
static pehip_status layout_contacts(pehip_batch* b) {
  BK(hipStreamSynchronize(b->stream));
  b->free_contacts();
  // a failure leaves the batch without scenes (never scenes over contact arrays that do not exist)
  const pehip_status st = layout_contacts_into(b);
  if (st != PEHIP_OK) b->free_all();
  return st;
}
static pehip_status layout_contacts_into(pehip_batch* b) {
  uint64_t c = 0;
  for (auto& s : b->hinfo) {
    s.c0 = (uint32_t)c;
    c += s.ccap;
  }
  if (c >= 0xFFFFFFF0ull) return bfail(b, PEHIP_OUT_OF_MEMORY, "contact capacity exceeds 2^32 slots");
  const size_t cap = std::max<uint64_t>(c, 1);
  BK(balloc(b->raw.b1, cap)); BK(balloc(b->raw.b2, cap)); BK(balloc(b->raw.pos, 3 * cap));
  BK(balloc(b->raw.nrm, 3 * cap)); BK(balloc(b->raw.dist, cap)); BK(balloc(b->raw.mu, cap));
  BK(balloc(b->sol.b1, cap)); BK(balloc(b->sol.b2, cap)); BK(balloc(b->sol.f, (size_t)CF_COUNT * cap));
  BK(balloc(b->colour, cap)); BK(balloc(b->perm, cap));
  b->raw.cap = cap;
  b->sol.cap = cap;
  b->ccap = c;
  BK(hipMemcpy(b->info, b->hinfo.data(), sizeof(SceneInfo) * b->nscenes, hipMemcpyHostToDevice));
  return PEHIP_OK;
}

static BatchArgs make_args(const pehip_batch* b, double dt) {
  BatchArgs A;
  A.nscenes = b->nscenes;
  A.info = b->info;
  A.order = b->order;
  A.plane_idx = b->plane_idx;
  A.csf = b->csf;
  A.cdf = b->cdf;
  A.nmat = b->nmat;
  A.scene_mu = b->scene_mu ? 1 : 0;
  A.queue = b->queue;
  A.bucket = b->bucket;
  A.sorted = b->sorted;
  A.count = b->count;
  A.colour = b->colour;
  A.perm = b->perm;
  A.colour_off = b->colour_off;
  A.out = b->out;
  A.any_flags = b->any_flags;
  A.gravity = mk3(b->prm.gravity[0], b->prm.gravity[1], b->prm.gravity[2]);
  A.dt = dt;
  A.dtinv = 1.0 / dt;
  A.erp = b->prm.erp;
  A.damping = b->prm.damping;
  A.thr = b->prm.abort_threshold;
  A.iters = (int)b->prm.max_iterations;
  A.max_sub = (int)b->prm.max_sub_iterations;
  return A;
}

static uint32_t grid_of(const pehip_batch* b, const pehip_batch_tuning& t) {
  const uint64_t g = (uint64_t)b->ncu * t.blocks_per_cu;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, b->nscenes));
}

static void enqueue_launch(pehip_batch* b, const BatchArgs& A, uint32_t add_steps) {
  launch_batch_begin(b->queue, add_steps, b->stream);
  launch_batch(b->prm.relaxation_model, b->prm.exact_trig != 0, b->bodies(), b->raw, b->sol, A, grid_of(b, b->tun),
               b->tun.threads, b->stream);
  b->launches += 2;
}

static bool same_params(const pehip_params& a, const pehip_params& c) { return std::memcmp(&a, &c, sizeof(a)) == 0; }

static pehip_status capture_graph(pehip_batch* b, double dt) {
  b->drop_graph();
  const BatchArgs A = make_args(b, dt);
  hipGraph_t g = nullptr;
  BK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  launch_batch_begin(b->queue, b->tun.steps_per_launch, b->stream);
  launch_batch(b->prm.relaxation_model, b->prm.exact_trig != 0, b->bodies(), b->raw, b->sol, A, grid_of(b, b->tun),
               b->tun.threads, b->stream);
  BK(hipStreamEndCapture(b->stream, &g));
  const hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  BK(e);
  b->g_dt = dt;
  b->g_tun = b->tun;
  b->g_prm = b->prm;
  return PEHIP_OK;
}

static pehip_status read_flags(pehip_batch* b, uint32_t* flags) {
  BK(hipMemcpyAsync(b->hstage, b->any_flags, 4, hipMemcpyDeviceToHost, b->stream));
  BK(hipGetLastError());
  BK(hipStreamSynchronize(b->stream));
  *flags = b->hstage[0];
  return PEHIP_OK;
}

// Runs nsteps for every scene; grows the contact slots of scenes that ran out and lets them
// catch up.
static pehip_status batch_run(pehip_batch* b, double dt, uint32_t nsteps) {
  b->launches = 0;
  fill_u32(b->any_flags, 0u, 1, b->stream);
  const uint32_t spl = std::max<uint32_t>(1, b->tun.steps_per_launch);
  uint32_t remaining = nsteps;
  const BatchArgs A = make_args(b, dt);
  if (b->tun.use_graph && remaining >= spl) {
    if (!b->gexec || b->g_dt != dt || std::memcmp(&b->g_tun, &b->tun, sizeof(b->tun)) != 0 ||
        !same_params(b->g_prm, b->prm))
      BKS(capture_graph(b, dt));
    while (remaining >= spl) {
      BK(hipGraphLaunch(b->gexec, b->stream));
      b->launches += 2;
      remaining -= spl;
    }
  }
  while (remaining) {
    const uint32_t a = std::min(spl, remaining);
    enqueue_launch(b, A, a);
    remaining -= a;
  }
  uint32_t flags = 0;
  BKS(read_flags(b, &flags));
  for (int round = 0; (flags & ERR_CAPACITY) && round < 16; ++round) {
    std::vector<SceneOut> o(b->nscenes);
    BK(hipMemcpy(o.data(), b->out, sizeof(SceneOut) * b->nscenes, hipMemcpyDeviceToHost));
    for (uint32_t s = 0; s < b->nscenes; ++s)
      if (o[s].flags & ERR_CAPACITY) {
        const uint64_t need = std::max<uint64_t>(o[s].nc, b->hinfo[s].ccap);
        b->hinfo[s].ccap = (uint32_t)std::min<uint64_t>(need + need / 2 + 64, 0x7FFFFFFFull);
      }
    ++b->capacity_retries;
    BKS(layout_contacts(b));
    const BatchArgs A2 = make_args(b, dt);
    fill_u32(b->any_flags, 0u, 1, b->stream);
    enqueue_launch(b, A2, 0);
    BKS(read_flags(b, &flags));
  }
  if (flags & ERR_CAPACITY) return bfail(b, PEHIP_CAPACITY_RETRY, "contact capacity could not be grown enough");
  if (flags & ERR_NONFINITE) return bfail(b, PEHIP_NONFINITE, "non-finite body position or bounding box");
  if (flags & ERR_LEVEL_OVERFLOW) return bfail(b, PEHIP_INVALID_ARG, "body larger than its scene's grid cell");
  if (flags & ERR_COLOURS) return bfail(b, PEHIP_NOT_SUPPORTED, "more than 128 colours needed");
  if (flags & ERR_INTERNAL) return bfail(b, PEHIP_HIP_ERROR, "internal error: inconsistent sweep order");
  return PEHIP_OK;
}

extern "C" {

pehip_status pehip_batch_create(const pehip_params* params, int device, pehip_batch** outp) {
  if (!outp) { g_batch_create_error = "out is NULL"; return PEHIP_INVALID_ARG; }
  *outp = nullptr;
  pehip_params p;
  if (params) p = *params; else pehip_default_params(&p);
  std::string why;
  const pehip_status st = check_params(&p, why);
  if (st != PEHIP_OK) { g_batch_create_error = why; return st; }
  if (p.sweep_mode != PEHIP_SWEEP_COLOURED) {
    g_batch_create_error = "batched scenes run in coloured mode only";
    return PEHIP_NOT_SUPPORTED;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_batch_create_error = "no HIP device available";
    return PEHIP_HIP_ERROR;
  }
  if (device < 0 || device >= ndev) { g_batch_create_error = "invalid device ordinal"; return PEHIP_INVALID_ARG; }
  pehip_batch* b = new pehip_batch();
  b->prm = p;
  b->device = device;
  int ncu = 0;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&b->hstage, 4096, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&b->queue, 64) != hipSuccess || hipMalloc((void**)&b->any_flags, 64) != hipSuccess) {
    g_batch_create_error = "HIP initialisation failed";
    pehip_batch_destroy(b);
    return PEHIP_HIP_ERROR;
  }
  b->ncu = ncu > 0 ? ncu : 256;
  *outp = b;
  return PEHIP_OK;
}

pehip_status pehip_batch_destroy(pehip_batch* b) {
  if (!b) return PEHIP_OK;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  b->free_all();
  bfree(b->csf); bfree(b->cdf); bfree(b->queue); bfree(b->any_flags);
  if (b->hstage) hipHostFree(b->hstage);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return PEHIP_OK;
}

pehip_status pehip_batch_set_params(pehip_batch* b, const pehip_params* params) {
  if (!b || !params) return PEHIP_INVALID_ARG;
  std::string why;
  const pehip_status st = check_params(params, why);
  if (st != PEHIP_OK) return bfail(b, st, why);
  if (params->sweep_mode != PEHIP_SWEEP_COLOURED)
    return bfail(b, PEHIP_NOT_SUPPORTED, "batched scenes run in coloured mode only");
  b->prm = *params;
  return PEHIP_OK;
}

pehip_status pehip_batch_set_materials(pehip_batch* b, uint32_t n, const double* csf_pair, const double* cdf_pair) {
  if (!b) return PEHIP_INVALID_ARG;
  if (n == 0 || !csf_pair || !cdf_pair) return bfail(b, PEHIP_INVALID_ARG, "empty material table");
  if (b->n && n < b->nmat) return bfail(b, PEHIP_INVALID_ARG, "material table cannot shrink below the bodies' materials");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  // the new tables replace the old ones only once both are on the device
  const size_t nn = (size_t)n * n;
  double *csf = nullptr, *cdf = nullptr;
  hipError_t e = balloc(csf, nn);
  if (e == hipSuccess) e = balloc(cdf, nn);
  if (e == hipSuccess) e = hipMemcpy(csf, csf_pair, sizeof(double) * nn, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(cdf, cdf_pair, sizeof(double) * nn, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    bfree(csf);
    bfree(cdf);
    return bfail(b, e == hipErrorOutOfMemory ? PEHIP_OUT_OF_MEMORY : PEHIP_HIP_ERROR,
                 std::string("material tables: ") + hipGetErrorString(e));
  }
  bfree(b->csf);
  bfree(b->cdf);
  b->csf = csf;
  b->cdf = cdf;
  b->nmat = n;
  b->drop_graph();
  return PEHIP_OK;
}

pehip_status pehip_batch_upload(pehip_batch* b, uint32_t nscenes, const uint32_t* scene_offset, const uint8_t* shape,
                                const uint64_t* sid, const uint32_t* material, const uint8_t* fixed,
                                const double* param, const double* pos, const double* q, const double* v,
                                const double* w, const double* inv_mass, const double* inertia_body,
                                const double* inv_inertia_body, const double* plane_normal, const double* plane_d,
                                double contacts_per_body) {
  if (!b) return PEHIP_INVALID_ARG;
  if (nscenes == 0 || !scene_offset) return bfail(b, PEHIP_INVALID_ARG, "no scenes");
  if (!shape || !sid || !material || !fixed || !param || !pos || !q || !v || !w || !inv_mass || !inertia_body ||
      !inv_inertia_body || !plane_normal || !plane_d)
    return bfail(b, PEHIP_INVALID_ARG, "NULL body array");
  if (b->nmat == 0) return bfail(b, PEHIP_INVALID_ARG, "materials must be set before bodies");
  if (!(contacts_per_body > 0)) return bfail(b, PEHIP_INVALID_ARG, "contacts_per_body must be positive");
  if (scene_offset[0] != 0) return bfail(b, PEHIP_INVALID_ARG, "scene_offset[0] must be 0");
  const uint64_t n = scene_offset[nscenes];
  if (n == 0 || n >= kFixedBit) return bfail(b, PEHIP_INVALID_ARG, "invalid body count");
  std::vector<SceneInfo> info(nscenes);
  std::vector<uint32_t> pidx;
  std::vector<uint8_t> fx(fixed, fixed + n);
  for (uint32_t s = 0; s < nscenes; ++s) {
    const uint32_t lo = scene_offset[s], hi = scene_offset[s + 1];
    if (hi < lo) return bfail(b, PEHIP_INVALID_ARG, "scene offsets must not decrease");
    if (hi - lo > kBatchMaxBodies) return bfail(b, PEHIP_NOT_SUPPORTED, "more than 4096 bodies in a scene");
    SceneInfo& S = info[s];
    std::memset(&S, 0, sizeof(S));
    S.body0 = lo;
    S.nbody = hi - lo;
    S.plane0 = (uint32_t)pidx.size();
    double mx = 0;
    double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = lo; i < hi; ++i) {
      if (shape[i] < PEHIP_SPHERE || shape[i] > PEHIP_CAPSULE) return bfail(b, PEHIP_INVALID_ARG, "invalid shape");
      if (material[i] >= b->nmat) return bfail(b, PEHIP_INVALID_ARG, "Invalid material ID");
      const double* p = param + 3 * (size_t)i;
      if (shape[i] == PEHIP_PLANE) {
        pidx.push_back(i);
        fx[i] = 1;
        if (inv_mass[i] != 0) return bfail(b, PEHIP_INVALID_ARG, "plane with non-zero inverse mass");
        continue;
      }
      if (shape[i] == PEHIP_SPHERE && !(p[0] > 0)) return bfail(b, PEHIP_INVALID_ARG, "Invalid sphere radius");
      if (shape[i] == PEHIP_BOX && !(p[0] > 0 && p[1] > 0 && p[2] > 0))
        return bfail(b, PEHIP_INVALID_ARG, "Invalid side length");
      if (shape[i] == PEHIP_CAPSULE && !(p[0] > 0 && p[1] > 0))
        return bfail(b, PEHIP_INVALID_ARG, "Invalid capsule radius or length");
      if (fx[i] && inv_mass[i] != 0) return bfail(b, PEHIP_INVALID_ARG, "fixed body with non-zero inverse mass");
      const double sz = shape[i] == PEHIP_SPHERE    ? 2.0 * (p[0] + 1e-8)
                        : shape[i] == PEHIP_CAPSULE ? p[1] + 2.0 * (p[0] + 1e-8)
                                                    : std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) + 2e-8;
      mx = std::max(mx, sz);
      for (int k = 0; k < 3; ++k) {
        blo[k] = std::min(blo[k], pos[3 * (size_t)i + k]);
        bhi[k] = std::max(bhi[k], pos[3 * (size_t)i + k]);
      }
    }
    S.nplane = (uint32_t)pidx.size() - S.plane0;
    if (S.nplane > kBatchMaxPlanes) return bfail(b, PEHIP_NOT_SUPPORTED, "more than 16 planes in a scene");
    if (mx == 0) { mx = 1.0; blo[0] = blo[1] = blo[2] = 0; bhi[0] = bhi[1] = bhi[2] = 1; }
    S.span = mx * 1.001;
    S.inv = 1.0 / S.span;
    uint32_t l[3];
    for (int k = 0; k < 3; ++k) {
      const double cells = std::ceil((bhi[k] - blo[k]) * S.inv) + 2.0;
      uint32_t e = 2;
      while (e < 10 && (double)(1u << e) < cells) ++e;
      l[k] = e;
    }
    while (l[0] + l[1] + l[2] > 12) {
      int m = 0;
      for (int k = 1; k < 3; ++k) if (l[k] > l[m]) m = k;
      --l[m];
    }
    S.lx = l[0]; S.ly = l[1]; S.lz = l[2];
    S.ccap = (uint32_t)std::min<double>(std::ceil(contacts_per_body * S.nbody) + 64.0, 0x7FFFFFFF);
    S.mu_s = S.mu_d = 0;
  }
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  b->free_all();
  // a failure from here on leaves the batch without scenes (never a body count over arrays that
  // do not exist)
  const pehip_status st = [&]() -> pehip_status {
    b->n = n;
    b->nscenes = nscenes;
    b->hinfo = info;
    b->hplane_idx = pidx;
    b->scene_mu = false;
    b->capacity_retries = 0;
    std::vector<double> hR(9 * (size_t)n);
    for (uint64_t i = 0; i < n; ++i) {
      qd qq;
      for (int k = 0; k < 4; ++k) qq.v[k] = q[4 * (size_t)i + k];
      const m33 m = quat_to_R(qq);
      for (int k = 0; k < 9; ++k) hR[9 * (size_t)i + k] = m.a[k];
    }
    // largest scenes first (ties by index): the persistent grid ends on the small ones
    std::vector<uint32_t> ord(nscenes);
    for (uint32_t s = 0; s < nscenes; ++s) ord[s] = s;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) { return info[a].nbody > info[c].nbody; });
    BK(balloc(b->shape, n)); BK(balloc(b->fixed, n)); BK(balloc(b->mat, n)); BK(balloc(b->sid, n));
    BK(balloc(b->param, 3 * n)); BK(balloc(b->pos, 3 * n)); BK(balloc(b->q, 4 * n)); BK(balloc(b->R, 9 * n));
    BK(balloc(b->v, 3 * n)); BK(balloc(b->w, 3 * n)); BK(balloc(b->pnormal, 3 * n)); BK(balloc(b->pd, n));
    BK(balloc(b->aabb, 6 * n)); BK(balloc(b->inv_mass, n)); BK(balloc(b->Ib, 9 * n)); BK(balloc(b->Ibinv, 9 * n));
    BK(balloc(b->force, 3 * n)); BK(balloc(b->torque, 3 * n)); BK(balloc(b->sb, (size_t)kSB * n));
    BK(balloc(b->dvw, (size_t)kDVW * n)); BK(balloc(b->bucket, n)); BK(balloc(b->sorted, n)); BK(balloc(b->count, n));
    BK(balloc(b->info, nscenes)); BK(balloc(b->order, nscenes)); BK(balloc(b->plane_idx, pidx.size()));
    BK(balloc(b->colour_off, (size_t)nscenes * (kMaxColours + 1))); BK(balloc(b->out, nscenes));
    const size_t D = sizeof(double);
    BK(hipMemcpy(b->shape, shape, n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->fixed, fx.data(), n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->mat, material, 4 * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->sid, sid, 8 * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->param, param, 3 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->pos, pos, 3 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->q, q, 4 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->R, hR.data(), 9 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->v, v, 3 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->w, w, 3 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->pnormal, plane_normal, 3 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->pd, plane_d, D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->inv_mass, inv_mass, D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->Ib, inertia_body, 9 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->Ibinv, inv_inertia_body, 9 * D * n, hipMemcpyHostToDevice));
    BK(hipMemcpy(b->order, ord.data(), 4 * (size_t)nscenes, hipMemcpyHostToDevice));
    if (!pidx.empty()) BK(hipMemcpy(b->plane_idx, pidx.data(), 4 * pidx.size(), hipMemcpyHostToDevice));
    BK(hipMemset(b->force, 0, 3 * D * n));
    BK(hipMemset(b->torque, 0, 3 * D * n));
    BK(hipMemset(b->dvw, 0, kDVW * D * n));
    BK(hipMemset(b->sb, 0, kSB * D * n));
    BK(hipMemset(b->out, 0, sizeof(SceneOut) * nscenes));
    BK(hipMemset(b->colour_off, 0, 4 * (size_t)nscenes * (kMaxColours + 1)));
    BK(hipMemset(b->queue, 0, 64));
    BKS(layout_contacts(b));
    launch_aabb(b->bodies(), b->stream);
    BK(hipGetLastError());
    BK(hipStreamSynchronize(b->stream));
    return PEHIP_OK;
  }();
  if (st != PEHIP_OK) b->free_all();
  return st;
}

pehip_status pehip_batch_set_scene_friction(pehip_batch* b, uint32_t nscenes, const double* static_mu,
                                            const double* dynamic_mu) {
  if (!b) return PEHIP_INVALID_ARG;
  if (nscenes != b->nscenes) return bfail(b, PEHIP_INVALID_ARG, "scene count mismatch");
  if (!static_mu || !dynamic_mu) {
    b->scene_mu = false;
  } else {
    for (uint32_t s = 0; s < nscenes; ++s) {
      if (!(static_mu[s] >= 0) || !(dynamic_mu[s] >= 0)) return bfail(b, PEHIP_INVALID_ARG, "negative friction");
      b->hinfo[s].mu_s = static_mu[s];
      b->hinfo[s].mu_d = dynamic_mu[s];
    }
    b->scene_mu = true;
  }
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  BK(hipMemcpy(b->info, b->hinfo.data(), sizeof(SceneInfo) * b->nscenes, hipMemcpyHostToDevice));
  b->drop_graph();
  return PEHIP_OK;
}

pehip_status pehip_batch_upload_state(pehip_batch* b, uint64_t n, const double* pos, const double* q, const double* v,
                                      const double* w, const double* plane_normal, const double* plane_d) {
  if (!b) return PEHIP_INVALID_ARG;
  if (n != b->n) return bfail(b, PEHIP_INVALID_ARG, "body count mismatch");
  if (!pos || !q || !v || !w || !plane_normal || !plane_d) return bfail(b, PEHIP_INVALID_ARG, "NULL state array");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  std::vector<double> hR(9 * (size_t)n);
  for (uint64_t i = 0; i < n; ++i) {
    qd qq;
    for (int k = 0; k < 4; ++k) qq.v[k] = q[4 * (size_t)i + k];
    const m33 m = quat_to_R(qq);
    for (int k = 0; k < 9; ++k) hR[9 * (size_t)i + k] = m.a[k];
  }
  const size_t D = sizeof(double);
  BK(hipMemcpy(b->pos, pos, 3 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->q, q, 4 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->R, hR.data(), 9 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->v, v, 3 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->w, w, 3 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->pnormal, plane_normal, 3 * D * n, hipMemcpyHostToDevice));
  BK(hipMemcpy(b->pd, plane_d, D * n, hipMemcpyHostToDevice));
  launch_aabb(b->bodies(), b->stream);
  BK(hipGetLastError());
  BK(hipStreamSynchronize(b->stream));
  return PEHIP_OK;
}

pehip_status pehip_batch_step(pehip_batch* b, double dt, uint32_t nsteps) {
  if (!b) return PEHIP_INVALID_ARG;
  if (!(dt > 0)) return bfail(b, PEHIP_INVALID_ARG, "Invalid time step size");  // World.cpp:333-334
  if (b->n == 0) return bfail(b, PEHIP_INVALID_ARG, "no scenes uploaded");
  if (nsteps == 0) return PEHIP_OK;
  hipSetDevice(b->device);
  return batch_run(b, dt, nsteps);
}

pehip_status pehip_batch_download_state(pehip_batch* b, uint64_t n, double* pos, double* q, double* v, double* w,
                                        double* plane_normal, double* plane_d) {
  if (!b) return PEHIP_INVALID_ARG;
  if (n != b->n) return bfail(b, PEHIP_INVALID_ARG, "body count mismatch");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  const size_t D = sizeof(double);
  if (pos) BK(hipMemcpy(pos, b->pos, 3 * D * n, hipMemcpyDeviceToHost));
  if (q) BK(hipMemcpy(q, b->q, 4 * D * n, hipMemcpyDeviceToHost));
  if (v) BK(hipMemcpy(v, b->v, 3 * D * n, hipMemcpyDeviceToHost));
  if (w) BK(hipMemcpy(w, b->w, 3 * D * n, hipMemcpyDeviceToHost));
  if (plane_normal) BK(hipMemcpy(plane_normal, b->pnormal, 3 * D * n, hipMemcpyDeviceToHost));
  if (plane_d) BK(hipMemcpy(plane_d, b->pd, D * n, hipMemcpyDeviceToHost));
  return PEHIP_OK;
}

pehip_status pehip_batch_scene_stats(pehip_batch* b, uint32_t nscenes, uint32_t* contacts, uint32_t* colours,
                                     uint32_t* iterations, double* max_penetration, uint32_t* steps_done) {
  if (!b) return PEHIP_INVALID_ARG;
  if (nscenes != b->nscenes) return bfail(b, PEHIP_INVALID_ARG, "scene count mismatch");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  std::vector<SceneOut> o(nscenes);
  BK(hipMemcpy(o.data(), b->out, sizeof(SceneOut) * nscenes, hipMemcpyDeviceToHost));
  for (uint32_t s = 0; s < nscenes; ++s) {
    if (contacts) contacts[s] = o[s].nc;
    if (colours) colours[s] = o[s].ncol;
    if (iterations) iterations[s] = o[s].iters;
    if (max_penetration) {
      const unsigned long long bits = o[s].counters[0];
      double d;
      std::memcpy(&d, &bits, 8);
      max_penetration[s] = d;
    }
    if (steps_done) steps_done[s] = o[s].steps;
  }
  return PEHIP_OK;
}

pehip_status pehip_batch_get_contacts(pehip_batch* b, uint32_t scene, size_t cap, size_t* n, uint32_t* body1,
                                      uint32_t* body2, double* pos, double* normal, double* dist, double* mu) {
  if (!b) return PEHIP_INVALID_ARG;
  if (scene >= b->nscenes) return bfail(b, PEHIP_INVALID_ARG, "invalid scene");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  SceneOut o;
  BK(hipMemcpy(&o, b->out + scene, sizeof(o), hipMemcpyDeviceToHost));
  const size_t nc = o.nc, m = std::min(cap, nc), c0 = b->hinfo[scene].c0;
  if (n) *n = nc;
  const size_t D = sizeof(double);
  if (m) {
    if (body1) BK(hipMemcpy(body1, b->raw.b1 + c0, 4 * m, hipMemcpyDeviceToHost));
    if (body2) BK(hipMemcpy(body2, b->raw.b2 + c0, 4 * m, hipMemcpyDeviceToHost));
    if (pos) BK(hipMemcpy(pos, b->raw.pos + 3 * c0, 3 * D * m, hipMemcpyDeviceToHost));
    if (normal) BK(hipMemcpy(normal, b->raw.nrm + 3 * c0, 3 * D * m, hipMemcpyDeviceToHost));
    if (dist) BK(hipMemcpy(dist, b->raw.dist + c0, D * m, hipMemcpyDeviceToHost));
    if (mu) BK(hipMemcpy(mu, b->raw.mu + c0, D * m, hipMemcpyDeviceToHost));
  }
  return PEHIP_OK;
}

pehip_status pehip_batch_get_sweep_order(pehip_batch* b, uint32_t scene, size_t cap, size_t* n,
                                         uint32_t* contact_index, uint32_t* ncolours, uint32_t* colour_offsets) {
  if (!b) return PEHIP_INVALID_ARG;
  if (scene >= b->nscenes) return bfail(b, PEHIP_INVALID_ARG, "invalid scene");
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  SceneOut o;
  BK(hipMemcpy(&o, b->out + scene, sizeof(o), hipMemcpyDeviceToHost));
  const size_t nc = o.nc, m = std::min(cap, nc), c0 = b->hinfo[scene].c0;
  if (n) *n = nc;
  if (ncolours) *ncolours = o.ncol;
  if (m && contact_index) {
    BK(hipMemcpy(contact_index, b->perm + c0, 4 * m, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < m; ++k) contact_index[k] -= (uint32_t)c0;
  }
  if (colour_offsets)
    BK(hipMemcpy(colour_offsets, b->colour_off + (size_t)scene * (kMaxColours + 1), 4 * ((size_t)o.ncol + 1),
                 hipMemcpyDeviceToHost));
  return PEHIP_OK;
}

pehip_status pehip_batch_get_stats(pehip_batch* b, pehip_batch_stats* st) {
  if (!b || !st) return PEHIP_INVALID_ARG;
  std::memset(st, 0, sizeof(*st));
  st->nscenes = b->nscenes;
  st->num_bodies = b->n;
  st->contact_capacity = b->ccap;
  st->capacity_retries = b->capacity_retries;
  st->launches = b->launches;
  if (!b->nscenes) return PEHIP_OK;
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  std::vector<SceneOut> o(b->nscenes);
  BK(hipMemcpy(o.data(), b->out, sizeof(SceneOut) * b->nscenes, hipMemcpyDeviceToHost));
  uint64_t mn = ~0ull;
  double bytes = 0;
  for (uint32_t s = 0; s < b->nscenes; ++s) {
    st->num_contacts += o[s].nc;
    st->max_colours = std::max(st->max_colours, o[s].ncol);
    mn = std::min<uint64_t>(mn, o[s].steps);
    bytes += (double)o[s].iters * (296.0 * o[s].nc + 200.0 * (2.0 * o[s].nc - (double)o[s].counters[1]));
  }
  st->steps_done_min = mn;
  st->sweep_bytes = bytes;
  return PEHIP_OK;
}

pehip_status pehip_batch_get_tuning(pehip_batch* b, pehip_batch_tuning* t) {
  if (!b || !t) return PEHIP_INVALID_ARG;
  *t = b->tun;
  return PEHIP_OK;
}

pehip_status pehip_batch_set_tuning(pehip_batch* b, const pehip_batch_tuning* t) {
  if (!b || !t) return PEHIP_INVALID_ARG;
  if (t->threads != 64 && t->threads != 128 && t->threads != 256 && t->threads != 512)
    return bfail(b, PEHIP_INVALID_ARG, "threads must be 64, 128, 256 or 512");
  if (t->blocks_per_cu < 1 || t->blocks_per_cu > 16) return bfail(b, PEHIP_INVALID_ARG, "blocks_per_cu must be 1..16");
  if (t->steps_per_launch < 1) return bfail(b, PEHIP_INVALID_ARG, "steps_per_launch must be >= 1");
  b->tun = *t;
  return PEHIP_OK;
}

pehip_status pehip_batch_autotune(pehip_batch* b, double dt, uint32_t trial_steps, pehip_batch_tuning* best) {
  if (!b) return PEHIP_INVALID_ARG;
  if (!(dt > 0)) return bfail(b, PEHIP_INVALID_ARG, "Invalid time step size");
  if (b->n == 0) return bfail(b, PEHIP_INVALID_ARG, "no scenes uploaded");
  if (trial_steps == 0) trial_steps = 1;
  hipSetDevice(b->device);
  BK(hipStreamSynchronize(b->stream));
  const uint64_t n = b->n;
  // snapshot of everything a step changes
  struct Buf { void* dev; size_t bytes; void* save; };
  std::vector<Buf> bufs = {{b->pos, 24 * n, nullptr}, {b->q, 32 * n, nullptr}, {b->R, 72 * n, nullptr},
                           {b->v, 24 * n, nullptr}, {b->w, 24 * n, nullptr}, {b->pnormal, 24 * n, nullptr},
                           {b->pd, 8 * n, nullptr}, {b->aabb, 48 * n, nullptr}, {b->force, 24 * n, nullptr},
                           {b->torque, 24 * n, nullptr}, {b->out, sizeof(SceneOut) * b->nscenes, nullptr},
                           {b->queue, 64, nullptr}};
  pehip_status result = PEHIP_OK;
  for (auto& x : bufs) {
    if (hipMalloc(&x.save, x.bytes) != hipSuccess) { result = bfail(b, PEHIP_OUT_OF_MEMORY, "autotune snapshot"); break; }
    hipMemcpyAsync(x.save, x.dev, x.bytes, hipMemcpyDeviceToDevice, b->stream);
  }
  const pehip_batch_tuning orig = b->tun;
  pehip_batch_tuning win = orig;
  double best_ms = 1e300;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const uint32_t threads[4] = {64, 128, 256, 512};
  const uint32_t bpcs[4] = {1, 2, 4, 8};
  const uint32_t spls[2] = {1, trial_steps};
  for (int ti = 0; ti < 4 && result == PEHIP_OK; ++ti)
    for (int bi = 0; bi < 4 && result == PEHIP_OK; ++bi)
      for (int si = 0; si < 2 && result == PEHIP_OK; ++si) {
        if (si == 1 && trial_steps == 1) continue;
        pehip_batch_tuning t{threads[ti], bpcs[bi], spls[si], 0};
        b->tun = t;
        for (auto& x : bufs) hipMemcpyAsync(x.dev, x.save, x.bytes, hipMemcpyDeviceToDevice, b->stream);
        hipEventRecord(e0, b->stream);
        result = batch_run(b, dt, trial_steps);
        hipEventRecord(e1, b->stream);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (result == PEHIP_OK && ms < best_ms) {
          best_ms = ms;
          win = t;
        }
      }
  for (auto& x : bufs) {
    if (x.save) {
      hipMemcpyAsync(x.dev, x.save, x.bytes, hipMemcpyDeviceToDevice, b->stream);
    }
  }
  hipStreamSynchronize(b->stream);
  for (auto& x : bufs) if (x.save) hipFree(x.save);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  win.use_graph = orig.use_graph;
  b->tun = result == PEHIP_OK ? win : orig;
  if (best) *best = b->tun;
  return result;
}

void* pehip_batch_stream(pehip_batch* b) { return b ? (void*)b->stream : nullptr; }

const char* pehip_batch_last_error(const pehip_batch* b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

}  // extern "C"
