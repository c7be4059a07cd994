// Launchers of the HIP kernels of the time step (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace pe_amd {

// Hash-grid hierarchy: level l has cubic cells of edge span[l] (ascending); every
// level uses the same power-of-two periodic bucket table of dimx*dimy*dimz cells.
constexpr int kMaxLevels = 8;
struct GridParams {
  int nlevels;
  double span[kMaxLevels];
  double inv[kMaxLevels];
  uint32_t dimx, dimy, dimz;
  uint32_t table;  // dimx*dimy*dimz
};

// Raw contacts as emitted by the narrowphase (detection order).
struct RawContacts {
  size_t cap;
  uint32_t* b1;
  uint32_t* b2;
  double* pos;   // 3*cap
  double* nrm;   // 3*cap
  double* dist;
  double* mu;
};

// Error flags raised by kernels (read back once per step).
enum : uint32_t {
  ERR_NONFINITE = 1u,
  ERR_LEVEL_OVERFLOW = 2u,
  ERR_CAPACITY = 4u,
  ERR_COLOURS = 8u,
  ERR_INTERNAL = 16u,
  ERR_TIMEOUT = 32u,
  ERR_STAGE_FULL = 64u   // a body emitted more contacts than the count pass stages (not an error)
};

struct BroadphaseBuffers {
  uint32_t* keys;      // n
  uint32_t* vals;      // n
  uint32_t* keys_tmp;  // n
  uint32_t* vals_tmp;  // n
  uint32_t* cell_start;  // nlevels*table
  uint32_t* cell_end;    // nlevels*table
  const uint32_t* planes;  // plane body indices
  uint32_t nplanes;
  uint32_t* counts;    // n+1 per sorted slot
  uint32_t* offsets;   // n+1
  uint32_t* flags;     // error flags
  double* saabb;       // 6 * n: the grid bodies' AABBs in sorted-slot order, field-major (saabb[k * n + s])
  uint8_t* sfixed;     // n: fixed flags in sorted-slot order
  uint32_t stride;     // n (field stride of saabb)
  // The count pass keeps the first kStage contacts a body emits, in emission order, at
  // slot * kStage + k (geometry ids, point, normal, distance); a body that emits more raises
  // ERR_STAGE_FULL and is written by the write pass instead.
  uint32_t* st_b1;     // kStage * n
  uint32_t* st_b2;     // kStage * n
  double* st_pos;      // 3 * kStage * n
  double* st_nrm;      // 3 * kStage * n
  double* st_dist;     // kStage * n
};
constexpr uint32_t kStage = 8;

// Scratch of the device-wide primitives: sized by scan_scratch_words / sort_scratch_words.
size_t scan_scratch_words(size_t n);
size_t sort_scratch_words(size_t n);
// Exclusive prefix sum of n u32 values; writes the total to *total (device pointer).
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total,
                        uint32_t* scratch, hipStream_t s);
// One stable pass on the low 8 key bits from (keys_in, vals_in) into (keys_out, vals_out)
// (keys below 256: a complete sort; scratch as for radix_sort_pairs).
void radix_sort_pairs_8bit(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                           size_t n, uint32_t* scratch, hipStream_t s);
// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits; the result
// ends in keys/vals (tmp arrays are ping-pong space).
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                      size_t n, int bits, uint32_t* scratch, hipStream_t s);

void launch_aabb(const Bodies& B, hipStream_t s);
void launch_predict(const Bodies& B, d3 g, double dt, hipStream_t s);
void launch_integrate(const Bodies& B, double dt, double damping, bool exact_trig, hipStream_t s);
void launch_eval_trig(int fn, bool exact, size_t n, const double* x, const double* y, double* out, hipStream_t s);
uint32_t launch_kinetic(const Bodies& B, double* partial, hipStream_t s);

// Broadphase: cell keys + sort + cell ranges.
void launch_cell_keys(const Bodies& B, const GridParams& G, BroadphaseBuffers& P, hipStream_t s);
void launch_cell_ranges(uint32_t n, const GridParams& G, BroadphaseBuffers& P, hipStream_t s);
// The grid bodies' AABBs and fixed flags gathered into sorted-slot order (after the sort).
void launch_sorted_boxes(const Bodies& B, BroadphaseBuffers& P, hipStream_t s);
// Pair enumeration + narrowphase; count pass (write = false) fills P.counts, write pass emits
// into R at P.offsets.  spheres_only: the scene holds spheres and planes only (a narrowphase of
// those two shapes, and smaller kernels).
void launch_pairs(const Bodies& B, const GridParams& G, BroadphaseBuffers& P, const RawContacts& R,
                  const double* csf, const double* cdf, uint32_t nmat, bool write, bool spheres_only, hipStream_t s);
// After the count pass and the scan: the staged contacts of bodies with at most kStage of them
// placed in R at P.offsets with their bodies and friction (the write pass then covers the bodies
// with more).
void launch_pairs_place(const Bodies& B, const BroadphaseBuffers& P, const RawContacts& R, const double* csf,
                        const double* cdf, uint32_t nmat, hipStream_t s);

// Colouring of the contact conflict graph.
struct ColourBuffers {
  uint32_t* deg;        // n bodies
  uint32_t* adj_start;  // n+1
  uint32_t* adj_cursor; // n
  uint32_t* adj;        // 2*nc
  uint32_t* colour;     // nc
  uint32_t* prio;       // nc
  uint32_t* cround;     // nc: round in which the contact was coloured
  uint32_t* remaining;  // 2: contacts still uncoloured, the last round that coloured one
  uint32_t* hist;       // 128
};
constexpr uint32_t kMaxColours = 128;
// word offsets in the context's status block
constexpr uint32_t kStFlags = 0, kStTotal = 16, kStRemaining = 18, kStHist = 32, kStWords = kStHist + kMaxColours;
constexpr uint32_t kUncoloured = 0xFFFFFFFFu;
void launch_adjacency(const Bodies& B, const RawContacts& R, uint32_t nc, ColourBuffers& C,
                      uint32_t* scan_scratch, hipStream_t s);
void launch_colour_init(const RawContacts& R, const Bodies& B, uint32_t nc, ColourBuffers& C, hipStream_t s);
void launch_colour_round(const RawContacts& R, const Bodies& B, uint32_t nc, ColourBuffers& C, uint32_t round,
                         uint32_t* flags, hipStream_t s);
void fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
void fill_u64(unsigned long long* p, unsigned long long value, size_t n, hipStream_t s);
void iota_u32(uint32_t* p, size_t n, hipStream_t s);
void launch_colour_hist(uint32_t nc, ColourBuffers& C, hipStream_t s);

// Colour boundaries of the solver order, by value (a kernel argument).
struct ColourOffsets {
  uint32_t ncol;
  uint32_t off[kMaxColours + 1];
};

// Contact caching (HCSITS.h:1741-1818) in solver order perm[0..nc).
void launch_setup(const Bodies& B, const RawContacts& R, const uint32_t* perm, uint32_t nc,
                  SolverContacts& C, double erp, double dtinv, unsigned long long* maxpen,
                  hipStream_t s);
// Dataflow sweeps: all K sweeps in one cooperative launch, every (sweep, contact) waiting only
// for the earlier updates of its own bodies.
struct DataflowBuffers {
  const uint32_t* perm;       // solver slot -> raw contact
  const uint32_t* adj_start;  // body -> contacts (raw), from the colouring
  const uint32_t* adj;
  const uint32_t* deg;        // contacts per non-fixed body
  uint32_t* inv;              // raw contact -> solver slot
  uint32_t* rank1;            // per slot: position among its body 1's slots
  uint32_t* rank2;
  uint32_t* done;             // per body: updates done
  unsigned long long* deltas; // per sweep: L-inf impulse change
  uint32_t* flags;            // ERR_TIMEOUT
  uint32_t spin_limit;
};
enum pehip_sweep_status { SWEEP_OK = 0, SWEEP_LAUNCH_FAILED = 1 };
// What a failed dataflow launch saw: the HIP error of the occupancy query and of the launch,
// the resident workgroups per compute unit and the grid asked for.
struct DataflowLaunchInfo {
  hipError_t occupancy_err = hipSuccess, launch_err = hipSuccess;
  int per_cu = 0;
  uint32_t grid = 0;
};
// The per-body ranks and cleared counters the dataflow sweeps wait on (three small kernels) ...
void launch_dataflow_prepare(const SolverContacts& C, uint32_t nc, const Bodies& B, const DataflowBuffers& D,
                             hipStream_t s);
// ... then all `iters` sweeps in one cooperative launch.
// slots: the slots each sweep relaxes, ascending (nullptr: all nc slots); nc: their number.  A slot
// list must hold every slot of each body it touches (whole islands).
pehip_sweep_status launch_sweep_dataflow(int model, bool exact_trig, const SolverContacts& C, uint32_t nc,
                                         uint32_t iters, const Bodies& B, double dtinv, int maxSub,
                                         const DataflowBuffers& D, uint32_t ncu, hipStream_t s,
                                         DataflowLaunchInfo* info, const uint32_t* slots = nullptr);

// Island sweeps.  The non-fixed bodies joined by contacts form islands (connected components of
// the contact graph without the fixed bodies); no body takes part in two islands, so each island's
// Gauss-Seidel sweeps are independent of every other's.  An island of up to kIslandSmall contacts
// is relaxed by one wavefront with one lane per contact (several small islands share a wavefront),
// one of up to kIslandLarge contacts by one workgroup, all K sweeps in a row with the bodies'
// records and velocity corrections in LDS; larger islands go to the dataflow sweeps.  Within an
// island the contacts are relaxed colour by colour in the coloured order, so every body sees its
// updates in the same order as with the colour launches: the same bits.
constexpr uint32_t kIslandSmall = 64;          // contacts per small unit (one wavefront)
constexpr uint32_t kIslandSmallBodies = 128;   // non-fixed bodies per small unit
constexpr uint32_t kIslandLarge = 256;         // contacts per large unit (one workgroup)
constexpr uint32_t kIslandLargeBodies = 512;   // non-fixed bodies per large unit
constexpr uint32_t kIslandClasses = 7;         // small-island size classes: 1, 2, 3-4, ..., 33-64
constexpr uint32_t kIslandNone = 0xFFFFFFFFu;
// header words (IslandBuffers::hdr)
enum : uint32_t {
  ISL_CLASS_COUNT = 0,     // [7] small islands per size class
  ISL_CLASS_BASE = 8,      // [7] first small unit of each class
  ISL_SMALL_UNITS = 16,
  ISL_LARGE_UNITS = 17,
  ISL_DATAFLOW_CONTACTS = 18,
  ISL_HDR_WORDS = 32
};
struct IslandBuffers {
  // per body
  uint32_t* parent;   // union-find forest over the non-fixed bodies (a root is its set's smallest index)
  uint32_t* ncont;    // per root: the island's contacts
  uint32_t* cmask;    // per root: 4 words, the colours present in the island
  uint32_t* cursor;   // per root: lanes handed out
  uint32_t* place;    // per root: kind << 30 | class << 24 | position (kind 0 small, 1 large, 2 dataflow)
  uint32_t* blocal;   // per non-fixed body with contacts: its index in its unit's body table
  // per contact slot, and per unit (bounds from island_unit_bounds)
  uint32_t* sroot;    // per slot: root of its island
  uint32_t* dflag;    // per slot: 1 if its island goes to the dataflow sweeps
  uint32_t* dpos;     // per slot: exclusive scan of dflag
  uint32_t* dslots;   // the dataflow islands' slots, ascending
  uint32_t* sslot;    // small units: kIslandSmall lanes each, slot or kIslandNone
  uint32_t* sround;   // small units: each lane's round (rank of its colour within its island)
  uint32_t* sbody;    // small units: kIslandSmallBodies bodies each
  uint32_t* snb;      // small units: bodies
  uint32_t* srounds;  // small units: rounds per sweep
  uint32_t* lslot;    // large units: kIslandLarge lanes each
  uint32_t* lround;
  uint32_t* lbody;    // large units: kIslandLargeBodies bodies each
  uint32_t* lnb;
  uint32_t* lrounds;
  uint32_t* hdr;      // ISL_HDR_WORDS words
};
// Upper bounds of the unit counts for nc contacts: small units (every small island of size m takes
// at most 2m - 1 lanes of its class, plus one partly filled unit per class) and large units.
inline size_t island_small_units_bound(size_t nc) { return (2 * nc) / kIslandSmall + kIslandClasses + 1; }
inline size_t island_large_units_bound(size_t nc) { return nc / (kIslandSmall + 1) + 1; }
// Islands of the coloured order: union-find, sizes, colour masks, units, bodies.  After it,
// hdr[ISL_SMALL_UNITS], hdr[ISL_LARGE_UNITS] and hdr[ISL_DATAFLOW_CONTACTS] hold the counts (the
// host reads them before launch_island_assign).
void launch_island_prepare(const SolverContacts& C, uint32_t nc, const ColourOffsets& O, const Bodies& B,
                           const uint32_t* deg, IslandBuffers& I, hipStream_t s);
// Slots and bodies into their units, and the dataflow islands' slot list.
void launch_island_assign(const SolverContacts& C, uint32_t nc, const ColourOffsets& O, const Bodies& B,
                          const uint32_t* deg, IslandBuffers& I, uint32_t small_units, uint32_t large_units,
                          uint32_t dataflow_contacts, uint32_t* scan_scratch, hipStream_t s);
// All `iters` sweeps of the small and large units (one launch).
void launch_sweep_islands(int model, bool exact_trig, const SolverContacts& C, const IslandBuffers& I,
                          uint32_t small_units, uint32_t large_units, uint32_t iters, const Bodies& B, int maxSub,
                          unsigned long long* deltas, hipStream_t s);

// One colour of one sweep; ncu: compute units of the device (sizes the lanes per contact).  A
// non-null stop makes the launch return at once once stop[0] is set (see launch_abort_check).
void launch_sweep(int model, bool exact_trig, const SolverContacts& C, uint32_t begin, uint32_t end,
                  const Bodies& B, double dtinv, int maxSub, unsigned long long* delta, const uint32_t* stop,
                  uint32_t ncu, hipStream_t s);
// pe's abort test on the device: sets stop[0] = 1 and stop[1] = it + 1 if sweep it changed no
// impulse by thr or more and no earlier sweep did (stop[0..1] cleared before the first sweep).
void launch_abort_check(const unsigned long long* deltas, uint32_t it, double thr, uint32_t* stop, hipStream_t s);
// All `iters` sweeps in one workgroup, colour by colour with a barrier between colours; after
// each sweep deltas[it] = its L-inf impulse change, and the sweeps stop after the first one
// below thr > 0.  deltas[iters] receives the number of sweeps run.
void launch_sweep_block(int model, bool exact_trig, const SolverContacts& C, const ColourOffsets& O, uint32_t iters,
                        const Bodies& B, double dtinv, int maxSub, double thr, unsigned long long* deltas,
                        hipStream_t s);
// All sweeps, strictly sequential, one wavefront.  deltas[it] receives the L-inf impulse
// change of sweep it; deltas[iters] the number of sweeps run (stops early below thr > 0).
void launch_sweep_serial(int model, bool exact_trig, const SolverContacts& C, uint32_t nc, uint32_t iters,
                         const Bodies& B, double dtinv, int maxSub, double thr,
                         unsigned long long* deltas, hipStream_t s);

}  // namespace pe_amd
