// Batched scenes: device data of many independent pe worlds advanced by one persistent
// kernel, one workgroup per scene (internal to the library).
#pragma once

#include <string>

#include "../../include/pe_hip.h"
#include "kernels.hpp"

namespace pe_amd {

constexpr uint32_t kBatchMaxBodies = 4096;  // bodies per scene (LDS claim table)
constexpr uint32_t kBatchMaxPlanes = 16;
constexpr uint32_t kBatchTable = 4096;      // buckets of a scene's hash grid (LDS)
constexpr uint32_t kBatchMaxThreads = 512;

// Host-computed description of one scene.
struct SceneInfo {
  uint32_t body0, nbody;    // bodies [body0, body0 + nbody) of the batch arrays
  uint32_t plane0, nplane;  // planes at plane_idx[plane0 .. plane0 + nplane)
  uint32_t c0, ccap;        // contact slots [c0, c0 + ccap)
  uint32_t lx, ly, lz;      // log2 of the bucket-table dimensions (lx + ly + lz <= 12)
  uint32_t pad;
  double span, inv;         // cell edge (>= every grid body's AABB extent) and 1 / span
  double mu_s, mu_d;        // per-scene friction (BatchArgs::scene_mu)
};

// Per-scene results of the last step.
struct SceneOut {
  uint32_t nc, ncol, iters, flags, steps;
  uint32_t pad;
  unsigned long long counters[2];  // max penetration bits, contacts with a fixed body
  double delta;                    // impulse change of the last sweep
};

struct BatchArgs {
  uint32_t nscenes;
  const SceneInfo* info;
  const uint32_t* order;      // processing order of the scenes (largest first)
  const uint32_t* plane_idx;
  const double* csf;
  const double* cdf;
  uint32_t nmat;
  int scene_mu;
  uint32_t* queue;            // [0] next position in order, [1] step target
  uint32_t* bucket;           // per body
  uint32_t* sorted;           // per body: grid bodies of a scene sorted by bucket
  uint32_t* count;            // per body: contacts emitted, then their offsets
  uint32_t* colour;           // per contact slot
  uint32_t* perm;             // per contact slot: sweep slot -> raw contact
  uint32_t* colour_off;       // per scene: kMaxColours + 1
  SceneOut* out;
  uint32_t* any_flags;
  d3 gravity;
  double dt, dtinv, erp, damping, thr;
  int iters, max_sub;
};

// Validation of pehip_params shared by single-world contexts and batches.
pehip_status check_params(const pehip_params* p, std::string& why);

void launch_batch(int model, bool exact_trig, const Bodies& B, const RawContacts& R, const SolverContacts& C,
                  const BatchArgs& A, uint32_t grid, uint32_t threads, hipStream_t s);
// queue[0] = 0, queue[1] += add_steps (a kernel, so that captured graphs hold no memset).
void launch_batch_begin(uint32_t* queue, uint32_t add_steps, hipStream_t s);

}  // namespace pe_amd
