// batch.hip -- batched domains and the solve server (hlbm_batch_* in include/hlbm.h).
//
// Host code.  Solves live in slices of two population arenas (A/B), one flag
// arena and one bounce-back mask arena, carved by a first-fit allocator against
// an HBM budget.  A chunk is `chunk_steps` time steps of batch_kernel (one BGK
// and one MRT launch per step, captured once as a HIP graph), each walking its
// part of the tile table of all running solves.
// Admission runs the collide-only kernel on the new slice; retirement runs the
// stream-only kernel and macro_variables, leaving rho/u in the solve's idle
// arena half until hlbm_batch_result copies them out.
//
// hlbm_batch_run enqueues and returns: up to two chunks are in flight on the
// main stream, so the host's admission / retirement work for the next chunk
// overlaps the device's current one.  Pinned staging buffers (domain table,
// tile table, obstacle lists) come in two sets, one per chunk in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/hlbm.h"
#include "common.hpp"
#include "free_list.hpp"
#include "launch.hpp"
#include "util_kernels.hpp"

using namespace hlbm;

namespace {

int bfail(int code, const std::string& m) { return fail(code, m); }

double now_s() {
  using namespace std::chrono;
  static const steady_clock::time_point t0 = steady_clock::now();
  return duration<double>(steady_clock::now() - t0).count();
}

struct Solve {
  int64_t id = 0;
  hlbm_solve_desc desc{};
  std::vector<hlbm_obstacle> obs;
  int state = HLBM_SOLVE_QUEUED;
  int64_t steps_done = 0;
  int fused_left = 0;
  int parity = 0;
  uint32_t N = 0;
  uint64_t stride = 0, f_off = 0, f_len = 0, flag_off = 0, flag_len = 0;
  BatchDomain dom{};
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool retiring = false;      // retirement enqueued; DONE once `ev` has completed
  hipEvent_t ev = nullptr;
  int res_arena = 0;          // arena whose slice holds rho (N), u (3N) and the statistics (7), as doubles
  bool staged = false;        // rho, u and the statistics also go to the pinned read-back pool at stage_off
  uint64_t stage_off = 0;     // (doubles)
  std::vector<double> rho, u, f;
  std::vector<double> init_f; // initial populations (FP64 node layout), empty = init_rho * w
  double stats[7] = {0, 0, 0, 0, 0, 0, 0};
};

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct hlbm_batch {
  hlbm_batch_desc d{};
  size_t esize = 8;
  uint64_t arena_elems = 0, flag_bytes = 0;
  void* arena[2] = {nullptr, nullptr};
  uint8_t* flags = nullptr;
  uint32_t* bbm = nullptr;  // per-node bounce-back masks, same offsets as the flags
  FreeList fl_pop, fl_flag;
  int64_t next_id = 1;
  std::map<int64_t, Solve> solves;
  std::deque<int64_t> queue;
  std::vector<int64_t> running;
  BatchDomain* d_doms = nullptr;
  BatchTile* d_tiles = nullptr;
  uint32_t* d_ntiles = nullptr;
  // pinned staging, one set per chunk in flight
  BatchDomain* h_doms_set[2] = {nullptr, nullptr};
  BatchTile* h_tiles_set[2] = {nullptr, nullptr};
  uint32_t* h_ntiles_set[2] = {nullptr, nullptr};
  ObstacleDev* h_obs[2] = {nullptr, nullptr};
  ObstacleDev* d_obs[2] = {nullptr, nullptr};
  size_t obs_cap = 0, obs_used[2] = {0, 0};
  uint64_t max_doms = 0, max_tiles = 0;
  double* f_tmp = nullptr;  // one FP64 population plane (keep_populations, initial populations)
  // pinned read-back pool: a retiring solve's rho, u and statistics are copied
  // here on the compute stream, so that polling and hlbm_batch_result never
  // wait on a device queue (solves that find no room take the copy stream)
  double* h_stage = nullptr;
  FreeList fl_stage;
  double* d_part = nullptr; // simulation_state partial sums (retirement, one reduction at a time on `stream`)
  uint64_t max_nodes = 0;
  unsigned grid[2] = {8u, 8u};  // workgroups of the BGK and MRT launches (graph captured with these)
  hipStream_t stream = nullptr;       // chunks, admission, retirement
  hipStream_t copy_stream = nullptr;  // result read-back
  hipStream_t side = nullptr;         // the MRT solves' chain of a chunk (the BGK chain runs on `stream`)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t graph = nullptr;
  struct ChunkEvents {
    hipEvent_t start = nullptr, end = nullptr;
    bool pending = false;
  } cev[2];
  int64_t run_index = 0;  // chunks enqueued
  std::vector<int64_t> retiring;
  bool tiles_dirty = true;
  int xcd_split = 1;
  int sr = 0;  // 16-bit storage rounds stochastically
  int64_t chunks = 0, n_done = 0;
  int64_t relocated = 0;  // arena pairs replaced by the placement check
  uint64_t peak_bytes = 0;
  double node_updates = 0, gpu_ms = 0;
};

static StepParams single_params(hlbm_batch* b, const Solve& s, int parity) {
  StepParams p{};
  p.nx = s.dom.nx;
  p.ny = s.dom.ny;
  p.nz = s.dom.nz;
  for (int k = 0; k < 3; k++) p.per[k] = s.dom.per[k];
  p.N = s.N;
  p.stride = s.stride;
  p.lay = default_layout(s.dom.nx, s.dom.ny, s.dom.nz);  // batch solves keep the default layout
  p.ln[0] = s.dom.nx;
  p.ln[1] = s.dom.ny;
  p.ln[2] = s.dom.nz;
  p.fin = (const char*)b->arena[parity] + s.f_off * b->esize;
  p.fout = (char*)b->arena[parity ^ 1] + s.f_off * b->esize;
  p.flags = b->flags + s.flag_off;
  p.any_wall = 1;
  p.bbm = b->bbm + s.flag_off;  // built at admission
  p.pair32 = 1;                 // checked at submission
  p.sr = b->sr;
  p.seed = mix32((uint32_t)s.steps_done * 0x9E3779B9u + 1u);
  p.wall_bb = s.dom.wall_bb;
  p.plane_bb = 1;
  p.cav_on[0] = s.dom.cav_on[0];
  p.cav_on[1] = s.dom.cav_on[1];
  p.cav_gather = 1;
  p.mrt_custom = s.dom.mrt_custom;
  for (int k = 0; k < 5; k++) {
    p.cav[0][k] = s.dom.cav[0][k];
    p.cav[1][k] = s.dom.cav[1][k];
  }
  p.tau = s.dom.tau;
  p.inv_tau = s.dom.inv_tau;
  for (int k = 0; k < 3; k++) p.sig[k] = s.dom.sig[k];
  for (int i = 0; i < Q; i++) p.ef[i] = s.dom.ef[i];
  for (int k = 0; k < 10; k++) p.mrt[k] = s.dom.mrt[k];
  set_fp32_constants(p);
  p.parts = PART_ALL;
  p.gx = (s.dom.nx + 127) / 128;
  p.gy = (s.dom.ny + 1) / 2;
  p.gz = s.dom.nz;
  p.xcd_remap = 1;
  return p;
}

static hipError_t single_launch(hlbm_batch* b, const Solve& s, int mode, int parity) {
  StepParams p = single_params(b, s, parity);
  const dim3 block(128, 2, 1), grid((unsigned)p.gx * p.gy * p.gz, 1, 1);
  const int coll = s.dom.coll;
  if (b->d.storage == HLBM_F64) return launch_step_f64(coll, mode, 1, grid, block, p, b->stream);
  return launch_step_lowp(b->d.storage, coll, mode, 1, grid, block, p, b->stream);
}

static hipError_t chunk_launch(hlbm_batch* b, int coll, int step, hipStream_t st) {
  BatchParams bp;
  bp.doms = b->d_doms;
  bp.tiles = b->d_tiles;
  bp.ntiles = b->d_ntiles;
  bp.arena[0] = b->arena[0];
  bp.arena[1] = b->arena[1];
  bp.flags = b->flags;
  bp.bbm = b->bbm;
  bp.step = step;
  bp.xcd_split = b->xcd_split;
  bp.sr = b->sr;
  if (b->d.storage == HLBM_F64) return launch_batch_f64(coll, b->grid[coll], bp, st);
  return launch_batch_lowp(b->d.storage, coll, b->grid[coll], bp, st);
}

// The steps of a chunk.  BGK and MRT solves never share a domain, so the two
// collision models run as independent chains, BGK on `stream` and MRT on the
// side stream, joined only at the end: each chain's per-step tail (its last
// workgroups) overlaps the other chain's work.  Both streams have the default
// priority.
static hipError_t chunk_sequence(hlbm_batch* b, int nsteps) {
  hipError_t e = hipEventRecord(b->ev_fork, b->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_fork, 0);
  for (int k = 0; k < nsteps && e == hipSuccess; k++) {
    e = chunk_launch(b, COLL_BGK, k, b->stream);
    if (e == hipSuccess) e = chunk_launch(b, COLL_MRT, k, b->side);
  }
  if (e == hipSuccess) e = hipEventRecord(b->ev_join, b->side);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev_join, 0);
  return e;
}

// Launch sizes follow the work: one workgroup per work unit (a persistent grid
// of resident workgroups walking the units ran the 16-bit kernel 8-13 % slower),
// with a quarter of headroom so that a mix that grows or shrinks a little keeps
// its captured graph; more units than workgroups, or fewer than half, resize.
// Returns whether a size changed.
static bool size_grids(hlbm_batch* b, uint32_t nbgk, uint32_t nmrt) {
  const int es = (int)b->esize;
  const uint64_t units[2] = {(uint64_t)nbgk * batch_units_per_tile(es, COLL_BGK),
                             (uint64_t)nmrt * batch_units_per_tile(es, COLL_MRT)};
  bool changed = false;
  for (int k = 0; k < 2; k++) {
    const uint64_t need = std::max<uint64_t>(8, round_up(units[k], 8));
    if (need > b->grid[k] || 2 * need < b->grid[k]) {
      b->grid[k] = (unsigned)std::min<uint64_t>(round_up(need + need / 4, 8), 0x7FFFFFF8u);
      changed = true;
    }
  }
  return changed;
}

static int build_graph(hlbm_batch* b) {
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return bfail(HLBM_ERR_HIP, "graph capture begin failed");
  if (chunk_sequence(b, b->d.chunk_steps) != hipSuccess) {
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    return bfail(HLBM_ERR_HIP, "batch kernel launch failed during capture");
  }
  if (hipStreamEndCapture(b->stream, &g) != hipSuccess) return bfail(HLBM_ERR_HIP, "graph capture end failed");
  hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(HLBM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  return HLBM_OK;
}

extern "C" {

double hlbm_clock(void) { return now_s(); }

// Placement check of the population arenas.  The same batch runs at distinct
// speed levels (for 200 served 16-bit solves: 58.8 % against 63.2-63.5 % of
// 8 TB/s) depending on the HBM regions the two arenas occupy, and a level holds
// for a whole region: probing consecutive 2.5 GB stretches of one arena pair
// gave one level throughout, or one level below an offset and the other above
// it.  So the arena pair is timed with the streaming pattern of a 16-bit step
// (launch_placement_probe) over the stretches solves fill first (first fit from
// offset 0), up to three fresh pairs are timed the same way, and the fastest
// pair is kept when it is more than 2 % faster.  Rejected pairs stay allocated
// until the search ends, so that each try is handed a region of its own, and a
// try needs the pair's size plus 4 GiB free.  Arenas under 8 GiB are not
// checked; HLBM_PLACEMENT=0 turns the check off, HLBM_PLACEMENT=force takes
// every candidate (exercises the move), HLBM_PLACEMENT_LOG prints each pair's
// time and the rate of every stretch.  No solve exists yet, so the move is a
// change of two pointers.
static int place_arenas(hlbm_batch* b) {
  const char* pe = getenv("HLBM_PLACEMENT");
  const bool force = pe && std::strcmp(pe, "force") == 0;
  if (pe && !force && atoi(pe) == 0) return HLBM_OK;
  const size_t abytes = b->arena_elems * b->esize;
  if (abytes < ((size_t)8 << 30) && !force) return HLBM_OK;
  const uint32_t n = (uint32_t)std::min<uint64_t>(64u << 20, abytes / 38 / 256 * 256);  // elements per probe plane
  const size_t region = (size_t)38 * n;                                                 // bytes of one probe
  const int nreg = (int)std::min<size_t>(abytes / std::max<size_t>(region, 1), 6);
  if (n == 0 || nreg < 1) return HLBM_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return bfail(HLBM_ERR_HIP, "placement events");
  }
  hipStream_t st = b->stream;
  auto pass = [&](void* A, void* B) {
    for (int r = 0; r < nreg; r++)
      if (launch_placement_probe((const char*)A + r * region, (char*)B + r * region, n, st) != hipSuccess) return false;
    return true;
  };
  // one untimed pass, then two timed ones
  auto time_pair = [&](void* A, void* B, float* ms) {
    return pass(A, B) && hipEventRecord(e0, st) == hipSuccess && pass(A, B) && pass(A, B) &&
           hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
           hipEventElapsedTime(ms, e0, e1) == hipSuccess;
  };
  const bool log = getenv("HLBM_PLACEMENT_LOG") != nullptr;
  auto report = [&](const char* what, void* A, void* B, float ms) {
    if (!log) return;
    std::string line;
    // every stretch of the pair, timed alone (GB/s)
    for (size_t off = 0; off + region <= abytes; off += region) {
      float t = 0.f;
      if (!(launch_placement_probe((const char*)A + off, (char*)B + off, n, st) == hipSuccess &&
            hipEventRecord(e0, st) == hipSuccess &&
            launch_placement_probe((const char*)A + off, (char*)B + off, n, st) == hipSuccess &&
            hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&t, e0, e1) == hipSuccess))
        break;
      line += " " + std::to_string((int)(2.0 * region / 1e6 / t));
    }
    std::fprintf(stderr, "placement %s %p %p: %.3f ms over %d stretches; stretches:%s\n", what, A, B, ms, nreg,
                 line.c_str());
  };
  int rc = HLBM_OK;
  float best = 0.f;
  std::vector<void*> rejected;
  if (!time_pair(b->arena[0], b->arena[1], &best)) rc = bfail(HLBM_ERR_HIP, "placement probe");
  if (rc == HLBM_OK) report("current", b->arena[0], b->arena[1], best);
  for (int t = 0; t < 3 && rc == HLBM_OK; t++) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < 2 * abytes + ((size_t)4 << 30)) {
      if (log) std::fprintf(stderr, "placement: %zu bytes free, no further candidate\n", fr);
      break;
    }
    void* nb[2] = {nullptr, nullptr};
    if (hipMalloc(&nb[0], abytes) != hipSuccess || hipMalloc(&nb[1], abytes) != hipSuccess) {
      for (void* q : nb)
        if (q) (void)hipFree(q);
      (void)hipGetLastError();
      break;
    }
    float ms = 0.f;
    if (!time_pair(nb[0], nb[1], &ms)) {
      rejected.push_back(nb[0]);
      rejected.push_back(nb[1]);
      rc = bfail(HLBM_ERR_HIP, "placement probe");
      break;
    }
    report("candidate", nb[0], nb[1], ms);
    if (force || ms < 0.98f * best) {
      rejected.push_back(b->arena[0]);
      rejected.push_back(b->arena[1]);
      b->arena[0] = nb[0];
      b->arena[1] = nb[1];
      best = ms;
      b->relocated++;
    } else {
      rejected.push_back(nb[0]);
      rejected.push_back(nb[1]);
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess && rc == HLBM_OK) rc = bfail(HLBM_ERR_HIP, "placement sync");
  for (void* q : rejected) (void)hipFree(q);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int hlbm_batch_create(const hlbm_batch_desc* desc, hlbm_batch** out) {
  if (!desc || !out) return bfail(HLBM_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (desc->storage < HLBM_F64 || desc->storage > HLBM_BF16) return bfail(HLBM_ERR_ARG, "bad storage");
  if (desc->chunk_steps < 1 || desc->chunk_steps > 100000) return bfail(HLBM_ERR_ARG, "bad chunk_steps");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(HLBM_ERR_NODEVICE, "no HIP device visible");
  hlbm_batch* b = new hlbm_batch();
  b->d = *desc;
  b->esize = desc->storage == HLBM_F64 ? 8 : (desc->storage == HLBM_F32 ? 4 : 2);
  if (const char* xs = getenv("HLBM_BATCH_XCD_SPLIT")) b->xcd_split = atoi(xs) ? 1 : 0;
  b->sr = (desc->storage == HLBM_F16 || desc->storage == HLBM_BF16) && !desc->round_nearest ? 1 : 0;
  // budget = 2 arenas of 19 planes + flags + masks: per node 38*esize + 1 + 4 bytes
  const uint64_t per_node = 38 * b->esize + 5;
  b->max_nodes = desc->budget_bytes / per_node;
  if (b->max_nodes < 4096) {
    delete b;
    return bfail(HLBM_ERR_ARG, "budget too small");
  }
  b->arena_elems = b->max_nodes * 19;
  b->flag_bytes = b->max_nodes + 4096 * 256;
  b->max_doms = 16384;
  b->max_tiles = b->max_nodes / BATCH_TILE_NODES + b->max_doms;
  auto cleanup = [&](const char* what) {
    hlbm_batch_destroy(b);
    return bfail(HLBM_ERR_NOMEM, std::string("batch allocation failed: ") + what);
  };
  if (hipMalloc(&b->arena[0], b->arena_elems * b->esize) != hipSuccess) return cleanup("arena A");
  if (hipMalloc(&b->arena[1], b->arena_elems * b->esize) != hipSuccess) return cleanup("arena B");
  if (hipMalloc((void**)&b->flags, b->flag_bytes) != hipSuccess) return cleanup("flags");
  if (hipMalloc((void**)&b->bbm, b->flag_bytes * 4) != hipSuccess) return cleanup("masks");
  if (hipMalloc((void**)&b->d_doms, b->max_doms * sizeof(BatchDomain)) != hipSuccess) return cleanup("domains");
  if (hipMalloc((void**)&b->d_tiles, b->max_tiles * sizeof(BatchTile)) != hipSuccess) return cleanup("tiles");
  if (hipMalloc((void**)&b->d_ntiles, 2 * sizeof(uint32_t)) != hipSuccess) return cleanup("tile count");
  if (hipMalloc((void**)&b->d_part, 7 * sizeof(double) * (size_t)reduce_parts(0xFFFFFFFFu)) != hipSuccess)
    return cleanup("reduction partials");
  for (int k = 0; k < 2; k++) {
    if (hipHostMalloc((void**)&b->h_doms_set[k], b->max_doms * sizeof(BatchDomain)) != hipSuccess ||
        hipHostMalloc((void**)&b->h_tiles_set[k], b->max_tiles * sizeof(BatchTile)) != hipSuccess ||
        hipHostMalloc((void**)&b->h_ntiles_set[k], 2 * sizeof(uint32_t)) != hipSuccess)
      return cleanup("pinned");
    if (hipEventCreate(&b->cev[k].start) != hipSuccess || hipEventCreate(&b->cev[k].end) != hipSuccess)
      return cleanup("events");
    b->h_ntiles_set[k][0] = b->h_ntiles_set[k][1] = 0;
  }
  {
    const uint64_t stage = std::min<uint64_t>((uint64_t)256 << 20, 4 * b->max_nodes + 7 * 64);  // doubles, <= 2 GiB
    if (hipHostMalloc((void**)&b->h_stage, stage * sizeof(double)) != hipSuccess) return cleanup("read-back pool");
    b->fl_stage.init(stage);
  }
  if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->copy_stream) != hipSuccess ||
      hipStreamCreate(&b->side) != hipSuccess)
    return cleanup("stream");
  if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
    return cleanup("events");
  if (hipMemcpy(b->d_ntiles, b->h_ntiles_set[0], 2 * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup("tile count init");
  {
    const int rc = place_arenas(b);
    if (rc) {
      std::string m = last_error();
      hlbm_batch_destroy(b);
      return bfail(rc, m);
    }
  }
  b->fl_pop.init(b->arena_elems);
  b->fl_flag.init(b->flag_bytes);
  if (desc->use_graph) {
    int rc = build_graph(b);
    if (rc) {
      std::string m = last_error();
      hlbm_batch_destroy(b);
      return bfail(rc, m);
    }
  }
  *out = b;
  return HLBM_OK;
}

int hlbm_batch_destroy(hlbm_batch* b) {
  if (!b) return HLBM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->side) (void)hipStreamSynchronize(b->side);
  if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (void* p : {b->arena[0], b->arena[1], (void*)b->flags, (void*)b->bbm, (void*)b->d_doms, (void*)b->d_tiles,
                  (void*)b->d_ntiles, (void*)b->f_tmp, (void*)b->d_part, (void*)b->d_obs[0], (void*)b->d_obs[1]})
    if (p) (void)hipFree(p);
  for (int k = 0; k < 2; k++) {
    for (void* p : {(void*)b->h_doms_set[k], (void*)b->h_tiles_set[k], (void*)b->h_ntiles_set[k], (void*)b->h_obs[k]})
      if (p) (void)hipHostFree(p);
    if (k == 0 && b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->cev[k].start) (void)hipEventDestroy(b->cev[k].start);
    if (b->cev[k].end) (void)hipEventDestroy(b->cev[k].end);
  }
  for (auto& kv : b->solves)
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
  delete b;
  return HLBM_OK;
}

int hlbm_batch_submit(hlbm_batch* b, const hlbm_solve_desc* s, int64_t* id) {
  if (!b || !s || !id) return bfail(HLBM_ERR_ARG, "NULL argument");
  for (int k = 0; k < 3; k++)
    if (s->nodes[k] < 2) return bfail(HLBM_ERR_ARG, "nodes must be >= 2");
  if (s->steps < 0) return bfail(HLBM_ERR_ARG, "steps must be >= 0");
  if (!(s->tau > 0.5)) return bfail(HLBM_ERR_ARG, "tau must be > 0.5 (nu = (tau - 0.5)/3 > 0)");
  if (s->collision != HLBM_BGK && s->collision != HLBM_MRT) return bfail(HLBM_ERR_ARG, "bad collision");
  if ((s->cavity_on[0] || s->cavity_on[1]) && s->periodic[2])
    return bfail(HLBM_ERR_ARG, "a moving wall needs a closed z axis");
  if (s->n_obstacles > 0 && (!s->obstacles || !(s->dx > 0.0))) return bfail(HLBM_ERR_ARG, "obstacles need dx > 0");
  // hippoLBM bounces walls on a periodic face off ghost nodes that hold the
  // previous step's populations (bounce_back.hpp:163-185, streaming.cu:107-111);
  // single domains reproduce that (special_kernel), the batch kernel does not:
  // obstacles whose candidate box reaches a periodic face are refused
  if (s->wall_bb && s->n_obstacles > 0) {
    for (int k = 0; k < s->n_obstacles; k++) {
      const hlbm_obstacle& o = s->obstacles[k];
      for (int d = 0; d < 3; d++) {
        if (!s->periodic[d]) continue;
        const double lo = o.kind == 0 ? o.p0[d] - o.radius : o.p0[d];
        const double hi = o.kind == 0 ? o.p0[d] + o.radius : o.p1[d];
        if (o.kind == 2 || int(lo / s->dx) <= 0 || int(hi / s->dx) >= s->nodes[d] - 1)
          return bfail(HLBM_ERR_UNSUPPORTED,
                       "batch solves take no wall_bounce_back obstacle that reaches a periodic face; use a domain");
      }
    }
  }
  const uint64_t N = (uint64_t)s->nodes[0] * s->nodes[1] * s->nodes[2];
  const uint64_t stride = round_up(N, 128);
  if (stride * 19 > b->arena_elems || N > b->max_nodes) return bfail(HLBM_ERR_NOMEM, "solve can never fit the budget");
  if ((stride + N) * b->esize > 0xFFFFFFFFull)
    return bfail(HLBM_ERR_UNSUPPORTED, "solve too large for the batch kernel's 32-bit offsets; use a domain");
  Solve v;
  v.id = b->next_id++;
  v.desc = *s;
  v.desc.obstacles = nullptr;
  if (s->n_obstacles > 0) v.obs.assign(s->obstacles, s->obstacles + s->n_obstacles);
  v.desc.init_f = nullptr;
  if (s->init_f) v.init_f.assign(s->init_f, s->init_f + (size_t)19 * N);
  v.N = (uint32_t)N;
  v.stride = stride;
  v.f_len = stride * 19;
  v.flag_len = round_up(N, 256);
  v.t_submit = now_s();
  BatchDomain& d = v.dom;
  d.nx = s->nodes[0];
  d.ny = s->nodes[1];
  d.nz = s->nodes[2];
  for (int k = 0; k < 3; k++) d.per[k] = s->periodic[k] ? 1 : 0;
  d.coll = s->collision == HLBM_MRT ? COLL_MRT : COLL_BGK;
  d.wall_bb = s->wall_bb && s->n_obstacles > 0 ? 1 : 0;
  d.any_wall = s->n_obstacles > 0 ? 1 : 0;
  d.mrt_custom = s->mrt_custom;
  d.inv_nx = 1.0 / (double)d.nx;
  d.inv_nxy = 1.0 / ((double)d.nx * (double)d.ny);
  d.tau = s->tau;
  d.inv_tau = 1 / s->tau;
  for (int i = 0; i < Q; i++) d.ef[i] = CX(i) * s->fext[0] + CY(i) * s->fext[1] + CZ(i) * s->fext[2];
  d.has_force = 0;
  for (int i = 0; i < Q; i++)
    if (d.ef[i] != 0.0) d.has_force = 1;
  for (int k = 0; k < 3; k++) d.f3[k] = 3.0f * (float)s->fext[k];
  bgk_fp32_constants(d.tau, d.k1, d.kw);
  {
    const RestShift sh = rest_shift(s->collision == HLBM_MRT, s->mrt_custom, s->mrt_rates);
    for (int k = 0; k < 3; k++) d.sig[k] = (float)sh.s[k];
  }
  for (int k = 0; k < 10; k++) d.mrt[k] = s->mrt_rates[k];
  mrt_fp32_rates(d.mrt_custom, d.mrt, d.inv_tau, d.mrtf);
  const double L = d.nz;  // cavity.hpp:58-63
  for (int side = 0; side < 2; side++) {
    d.cav_on[side] = s->cavity_on[side] ? 1 : 0;
    const double uxx = s->cavity_u[side][0] * (1 + 0.5 / (L - 1));
    const double uyy = s->cavity_u[side][1] * (1 + 0.5 / (L - 1));
    const double uzz = s->cavity_u[side][2] * (1 + 0.5 / (L - 1));
    for (int j = 0; j < 5; j++) {
      const int q = Faces::in[4 + side][j];
      d.cav[side][j] = 6. * WD(q) * (CX(q) * uxx + CY(q) * uyy + CZ(q) * uzz);
      d.cavf[side][j] = (float)d.cav[side][j];
    }
  }
  *id = v.id;
  b->solves.emplace(v.id, std::move(v));
  b->queue.push_back(*id);
  return HLBM_OK;
}

// the rest state of a solve's shifted storage (RestShift)
static RestShift solve_shift(const Solve& s) {
  return rest_shift(s.desc.collision == HLBM_MRT, s.desc.mrt_custom, s.desc.mrt_rates);
}

static int admit_one(hlbm_batch* b, Solve& s) {
  hipStream_t st = b->stream;
  if (hipMemsetAsync(b->flags + s.flag_off, 0, s.flag_len, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "memset");
  if (!s.obs.empty()) {
    std::vector<ObstacleDev> od;
    const int nn[3] = {s.dom.nx, s.dom.ny, s.dom.nz};
    const double dx = s.desc.dx;
    for (const hlbm_obstacle& o : s.obs) {
      ObstacleDev v;
      if (obstacle_to_dev(o, nn, dx, v)) od.push_back(v);
    }
    if (!od.empty()) {
      // the obstacle list goes through this chunk's pinned staging set
      const int set = (int)(b->run_index & 1);
      if (b->obs_used[set] + od.size() > b->obs_cap) {
        // grow both sets (rare): wait until no queued copy reads the old ones
        if (hipStreamSynchronize(st) != hipSuccess) return bfail(HLBM_ERR_HIP, "sync");
        const size_t cap = std::max<size_t>(2 * b->obs_cap, b->obs_used[set] + od.size() + 1024);
        for (int k = 0; k < 2; k++) {
          ObstacleDev *h = nullptr, *dv = nullptr;
          if (hipHostMalloc((void**)&h, cap * sizeof(ObstacleDev)) != hipSuccess ||
              hipMalloc((void**)&dv, cap * sizeof(ObstacleDev)) != hipSuccess)
            return bfail(HLBM_ERR_NOMEM, "obstacle staging");
          if (b->h_obs[k]) (void)hipHostFree(b->h_obs[k]);
          if (b->d_obs[k]) (void)hipFree(b->d_obs[k]);
          b->h_obs[k] = h;
          b->d_obs[k] = dv;
        }
        b->obs_cap = cap;
        b->obs_used[0] = b->obs_used[1] = 0;
      }
      ObstacleDev* hs = b->h_obs[set] + b->obs_used[set];
      ObstacleDev* ds = b->d_obs[set] + b->obs_used[set];
      std::memcpy(hs, od.data(), od.size() * sizeof(ObstacleDev));
      b->obs_used[set] += od.size();
      hipError_t e = hipMemcpyAsync(ds, hs, od.size() * sizeof(ObstacleDev), hipMemcpyHostToDevice, st);
      if (e == hipSuccess)
        e = launch_rasterize(b->flags + s.flag_off, s.dom.nx, s.dom.ny, s.dom.nz, dx, ds, (int)od.size(),
                             default_layout(s.dom.nx, s.dom.ny, s.dom.nz), st);
      if (e != hipSuccess) return bfail(HLBM_ERR_HIP, "rasterize");
    }
  }
  if (launch_build_bbm(b->flags + s.flag_off, s.dom.nx, s.dom.ny, s.dom.nz, s.dom.per, b->bbm + s.flag_off,
                       default_layout(s.dom.nx, s.dom.ny, s.dom.nz), st) !=
      hipSuccess)
    return bfail(HLBM_ERR_HIP, "bounce-back mask");
  void* fa = (char*)b->arena[0] + s.f_off * b->esize;
  if (s.init_f.empty()) {
    int lo[3] = {0, 0, 0}, hi[3] = {s.dom.nx - 1, s.dom.ny - 1, s.dom.nz - 1};
    if (launch_set_distribution(b->d.storage, fa, s.stride, default_layout(s.dom.nx, s.dom.ny, s.dom.nz), lo, hi,
                                s.desc.init_rho,
                                solve_shift(s), st) !=
        hipSuccess)
      return bfail(HLBM_ERR_HIP, "set_distribution");
  } else {
    // host populations, one FP64 plane at a time through f_tmp (stream-ordered),
    // converted to the storage form (shifted g = f - w below FP64)
    if (!b->f_tmp && hipMalloc((void**)&b->f_tmp, b->max_nodes * 8) != hipSuccess) return bfail(HLBM_ERR_NOMEM, "tmp");
    for (int i = 0; i < Q; i++) {
      void* plane = (char*)fa + (size_t)i * s.stride * b->esize;
      if (hipMemcpyAsync(b->f_tmp, s.init_f.data() + (size_t)i * s.N, (size_t)s.N * 8, hipMemcpyHostToDevice, st) !=
              hipSuccess ||
          launch_convert(b->d.storage, true, plane, b->f_tmp, s.N, i, solve_shift(s), st) != hipSuccess)
        return bfail(HLBM_ERR_HIP, "initial populations");
    }
  }
  s.parity = 0;
  if (s.desc.steps > 0) {
    if (single_launch(b, s, MODE_COLLIDE, 0) != hipSuccess) return bfail(HLBM_ERR_HIP, "collide");
    s.parity = 1;
    s.steps_done = 1;
    s.fused_left = s.desc.steps - 1;
  } else {
    s.fused_left = 0;
  }
  s.state = HLBM_SOLVE_RUNNING;
  s.t_admit = now_s();
  b->node_updates += s.desc.steps > 0 ? (double)s.N : 0.0;
  return HLBM_OK;
}

// Enqueue a finished solve's read-out: stream-only step (the reference's
// post-stream state), then macro_variables into the idle arena half of its
// slice (rho at the slice start, u after it).  The solve is DONE when the
// recorded event completes; hlbm_batch_result copies the fields out.
static int retire_one(hlbm_batch* b, Solve& s) {
  hipStream_t st = b->stream;
  int par = s.parity;
  if (s.desc.steps > 0) {
    if (single_launch(b, s, MODE_STREAM, par) != hipSuccess) return bfail(HLBM_ERR_HIP, "stream");
    par ^= 1;
  }
  const void* fcur = (const char*)b->arena[par] + s.f_off * b->esize;
  s.res_arena = par ^ 1;
  double* rho = (double*)((char*)b->arena[s.res_arena] + s.f_off * b->esize);  // 4N doubles <= 19N * esize bytes
  // walls keep rho = 0 (macro_variables never writes them)
  if (hipMemsetAsync(rho, 0, (size_t)s.N * 8, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "memset");
  if (launch_macro(b->d.storage, fcur, s.stride, s.N, b->flags + s.flag_off, rho, rho + s.N, st) != hipSuccess)
    return bfail(HLBM_ERR_HIP, "macro");
  // simulation_state of this solve alone (FP64), after rho and u: the failure check
  if (launch_reduce(rho, rho + s.N, b->flags + s.flag_off, s.N, b->d_part, rho + 4 * (size_t)s.N, st) != hipSuccess)
    return bfail(HLBM_ERR_HIP, "simulation_state");
  if (s.desc.keep_populations) {  // test and I/O path: read everything back now
    if (!b->f_tmp && hipMalloc((void**)&b->f_tmp, b->max_nodes * 8) != hipSuccess) return bfail(HLBM_ERR_NOMEM, "tmp");
    s.f.resize((size_t)19 * s.N);
    for (int i = 0; i < Q; i++) {
      const void* plane = (const char*)fcur + (size_t)i * s.stride * b->esize;
      if (launch_convert(b->d.storage, false, (void*)plane, b->f_tmp, s.N, i, solve_shift(s), st) != hipSuccess ||
          hipMemcpyAsync(s.f.data() + (size_t)i * s.N, b->f_tmp, s.N * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        return bfail(HLBM_ERR_HIP, "populations out");
    }
    if (hipStreamSynchronize(st) != hipSuccess) return bfail(HLBM_ERR_HIP, "sync");
  }
  {
    const uint64_t nd = 4 * (uint64_t)s.N + 7;
    s.staged = b->fl_stage.alloc(nd, &s.stage_off);
    if (s.staged && hipMemcpyAsync(b->h_stage + s.stage_off, rho, nd * sizeof(double), hipMemcpyDeviceToHost, st) !=
                        hipSuccess)
      return bfail(HLBM_ERR_HIP, "read-back");
  }
  if (!s.ev && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) return bfail(HLBM_ERR_HIP, "event");
  if (hipEventRecord(s.ev, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "event");
  s.retiring = true;
  b->retiring.push_back(s.id);
  return HLBM_OK;
}

// Retirements whose device work has completed become DONE (wait = block on the oldest).
static int poll_retiring(hlbm_batch* b, bool wait, int* finished) {
  std::vector<int64_t> keep;
  for (size_t k = 0; k < b->retiring.size(); k++) {
    Solve& s = b->solves[b->retiring[k]];
    hipError_t q = (wait && k == 0) ? hipEventSynchronize(s.ev) : hipEventQuery(s.ev);
    if (q == hipSuccess) {
      if (s.staged) {
        std::memcpy(s.stats, b->h_stage + s.stage_off + 4 * (size_t)s.N, sizeof(s.stats));
      } else {
        const double* dstat = (const double*)((const char*)b->arena[s.res_arena] + s.f_off * b->esize) + 4 * (size_t)s.N;
        if (hipMemcpyAsync(s.stats, dstat, sizeof(s.stats), hipMemcpyDeviceToHost, b->copy_stream) != hipSuccess ||
            hipStreamSynchronize(b->copy_stream) != hipSuccess)
          return bfail(HLBM_ERR_HIP, "statistics copy");
      }
      bool finite = true;
      for (double v : s.stats) finite = finite && std::isfinite(v);
      s.retiring = false;
      s.state = finite ? HLBM_SOLVE_DONE : HLBM_SOLVE_FAILED;
      s.t_done = now_s();
      b->n_done++;
      if (finished) (*finished)++;
    } else if (q == hipErrorNotReady) {
      keep.push_back(b->retiring[k]);
    } else {
      return bfail(HLBM_ERR_HIP, std::string("retirement: ") + hipGetErrorString(q));
    }
  }
  b->retiring.swap(keep);
  return HLBM_OK;
}

// Account the device time of a completed chunk.
static int finish_chunk(hlbm_batch* b, int set) {
  auto& c = b->cev[set];
  if (!c.pending) return HLBM_OK;
  if (hipEventSynchronize(c.end) != hipSuccess) return bfail(HLBM_ERR_HIP, "chunk sync");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c.start, c.end);
  b->gpu_ms += ms;
  b->chunks++;
  c.pending = false;
  return HLBM_OK;
}

int hlbm_batch_run(hlbm_batch* b, int32_t* n_finished) {
  if (!b) return bfail(HLBM_ERR_ARG, "NULL batch");
  int finished = 0;
  // this chunk uses staging set `set`; the chunk that last used it (two calls
  // ago) must be complete before the set is rewritten
  const int set = (int)(b->run_index & 1);
  int rc = finish_chunk(b, set);
  if (rc) return rc;
  b->obs_used[set] = 0;
  rc = poll_retiring(b, false, &finished);
  if (rc) return rc;
  // admission (first fit, in arrival order, backfilling smaller solves)
  for (auto it = b->queue.begin(); it != b->queue.end();) {
    Solve& s = b->solves[*it];
    if (b->running.size() >= b->max_doms) break;
    uint64_t foff, goff;
    if (!b->fl_pop.alloc(s.f_len, &foff)) {
      ++it;
      continue;
    }
    if (!b->fl_flag.alloc(s.flag_len, &goff)) {
      b->fl_pop.release(foff, s.f_len);
      ++it;
      continue;
    }
    s.f_off = foff;
    s.flag_off = goff;
    s.dom.f_off = foff;
    s.dom.stride = s.stride;
    s.dom.flag_off = goff;
    rc = admit_one(b, s);
    if (rc) return rc;
    b->running.push_back(s.id);
    it = b->queue.erase(it);
    b->tiles_dirty = true;
    b->peak_bytes = std::max<uint64_t>(b->peak_bytes, b->fl_pop.used * b->esize * 2 + b->fl_flag.used * 5);
  }
  // per-domain chunk lengths and the tile table (BGK domains' tiles first)
  const int K = b->d.chunk_steps;
  BatchDomain* h_doms = b->h_doms_set[set];
  BatchTile* h_tiles = b->h_tiles_set[set];
  uint32_t* h_ntiles = b->h_ntiles_set[set];
  uint32_t nt = 0, nbgk = 0;
  for (size_t k = 0; k < b->running.size(); k++) {
    Solve& s = b->solves[b->running[k]];
    s.dom.active_steps = std::min(K, s.fused_left);
    s.dom.parity = s.parity;
    s.dom.t0 = (uint32_t)s.steps_done;
    h_doms[k] = s.dom;
  }
  if (b->tiles_dirty) {
    for (int pass = 0; pass < 2; pass++) {
      for (size_t k = 0; k < b->running.size(); k++) {
        const Solve& s = b->solves[b->running[k]];
        if (s.dom.coll != (pass == 0 ? COLL_BGK : COLL_MRT)) continue;
        for (uint32_t first = 0; first < s.N; first += BATCH_TILE_NODES) h_tiles[nt++] = BatchTile{(uint32_t)k, first};
      }
      if (pass == 0) nbgk = nt;
    }
  }
  hipStream_t st = b->stream;
  if (!b->running.empty()) {
    if (hipMemcpyAsync(b->d_doms, h_doms, b->running.size() * sizeof(BatchDomain), hipMemcpyHostToDevice, st) !=
        hipSuccess)
      return bfail(HLBM_ERR_HIP, "upload domains");
    if (b->tiles_dirty) {
      h_ntiles[0] = nbgk;
      h_ntiles[1] = nt;
      if (hipMemcpyAsync(b->d_tiles, h_tiles, (size_t)nt * sizeof(BatchTile), hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(b->d_ntiles, h_ntiles, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, st) != hipSuccess)
        return bfail(HLBM_ERR_HIP, "upload tiles");
      b->tiles_dirty = false;
      if (size_grids(b, nbgk, nt - nbgk) && b->d.use_graph) {
        // the captured launches carry the old sizes: let the chunk that may still
        // replay them finish, then capture the chunk again
        if (hipStreamSynchronize(st) != hipSuccess) return bfail(HLBM_ERR_HIP, "sync");
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        const int rc2 = build_graph(b);
        if (rc2) return rc2;
      }
    }
    int steps_needed = 0;
    for (int64_t id : b->running) steps_needed = std::max(steps_needed, b->solves[id].dom.active_steps);
    // the events also mark when this call's staging set is free again
    auto& c = b->cev[set];
    if (hipEventRecord(c.start, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "event");
    if (steps_needed > 0) {
      if (b->graph && steps_needed == K) {
        if (hipGraphLaunch(b->graph, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "graph launch");
      } else {
        if (chunk_sequence(b, steps_needed) != hipSuccess) return bfail(HLBM_ERR_HIP, "batch kernel launch");
      }
    }
    if (hipEventRecord(c.end, st) != hipSuccess) return bfail(HLBM_ERR_HIP, "event");
    c.pending = true;
    b->run_index++;
    for (int64_t id : b->running) {
      Solve& s = b->solves[id];
      const int a = s.dom.active_steps;
      s.parity ^= (a & 1);
      s.fused_left -= a;
      s.steps_done += a;
      b->node_updates += (double)a * s.N;
    }
  }
  // retirement of the solves this chunk finishes: enqueued behind it
  std::vector<int64_t> keep;
  for (int64_t id : b->running) {
    Solve& s = b->solves[id];
    if (s.fused_left == 0) {
      rc = retire_one(b, s);
      if (rc) return rc;
      b->tiles_dirty = true;
    } else {
      keep.push_back(id);
    }
  }
  b->running.swap(keep);
  // nothing left to run: wait for the outstanding read-outs instead of spinning
  if (b->running.empty() && b->queue.empty() && !b->retiring.empty()) {
    rc = poll_retiring(b, true, &finished);
    if (rc) return rc;
  }
  if (n_finished) *n_finished = finished;
  return HLBM_OK;
}

int hlbm_batch_status(hlbm_batch* b, int64_t id, int32_t* state, int64_t* steps_done) {
  if (!b) return bfail(HLBM_ERR_ARG, "NULL batch");
  auto it = b->solves.find(id);
  if (it == b->solves.end()) return bfail(HLBM_ERR_ARG, "unknown solve id");
  if (it->second.retiring) {
    const int rc = poll_retiring(b, false, nullptr);
    if (rc) return rc;
  }
  if (state) *state = it->second.state;
  if (steps_done) *steps_done = it->second.steps_done;
  return HLBM_OK;
}

int hlbm_batch_times(hlbm_batch* b, int64_t id, double* t_submit, double* t_admit, double* t_done) {
  if (!b) return bfail(HLBM_ERR_ARG, "NULL batch");
  auto it = b->solves.find(id);
  if (it == b->solves.end()) return bfail(HLBM_ERR_ARG, "unknown solve id");
  if (t_submit) *t_submit = it->second.t_submit;
  if (t_admit) *t_admit = it->second.t_admit;
  if (t_done) *t_done = it->second.t_done;
  return HLBM_OK;
}

int hlbm_batch_stats(hlbm_batch* b, int64_t id, double out[7]) {
  if (!b || !out) return bfail(HLBM_ERR_ARG, "NULL argument");
  auto it = b->solves.find(id);
  if (it == b->solves.end()) return bfail(HLBM_ERR_ARG, "unknown solve id");
  if (it->second.retiring) {
    const int rc = poll_retiring(b, false, nullptr);
    if (rc) return rc;
  }
  const Solve& s = it->second;
  if (s.state != HLBM_SOLVE_DONE && s.state != HLBM_SOLVE_FAILED) return bfail(HLBM_ERR_ARG, "solve not finished");
  for (int k = 0; k < 7; k++) out[k] = s.stats[k];
  return HLBM_OK;
}

int hlbm_batch_result(hlbm_batch* b, int64_t id, double* rho, double* u, double* f) {
  if (!b) return bfail(HLBM_ERR_ARG, "NULL batch");
  auto it = b->solves.find(id);
  if (it == b->solves.end()) return bfail(HLBM_ERR_ARG, "unknown solve id");
  Solve& s = it->second;
  if (s.state != HLBM_SOLVE_DONE && s.state != HLBM_SOLVE_FAILED) return bfail(HLBM_ERR_ARG, "solve not finished");
  if (f && !s.desc.keep_populations) return bfail(HLBM_ERR_ARG, "populations were not kept (keep_populations = 0)");
  if (s.staged) {
    // the retirement's read-back has completed (the solve is finished)
    const double* h = b->h_stage + s.stage_off;
    if (rho) std::memcpy(rho, h, (size_t)s.N * 8);
    if (u) std::memcpy(u, h + s.N, 3 * (size_t)s.N * 8);
    b->fl_stage.release(s.stage_off, 4 * (uint64_t)s.N + 7);
  } else {
    // the read-out waits for nothing on the main stream: the fields are final
    const double* dres = (const double*)((const char*)b->arena[s.res_arena] + s.f_off * b->esize);
    if ((rho && hipMemcpyAsync(rho, dres, (size_t)s.N * 8, hipMemcpyDeviceToHost, b->copy_stream) != hipSuccess) ||
        (u && hipMemcpyAsync(u, dres + s.N, 3 * (size_t)s.N * 8, hipMemcpyDeviceToHost, b->copy_stream) != hipSuccess) ||
        hipStreamSynchronize(b->copy_stream) != hipSuccess)
      return bfail(HLBM_ERR_HIP, "result copy");
  }
  if (f) std::memcpy(f, s.f.data(), s.f.size() * 8);
  b->fl_pop.release(s.f_off, s.f_len);
  b->fl_flag.release(s.flag_off, s.flag_len);
  if (s.ev) (void)hipEventDestroy(s.ev);
  b->solves.erase(it);
  return HLBM_OK;
}

int hlbm_batch_info_get(hlbm_batch* b, hlbm_batch_info* out) {
  if (!b || !out) return bfail(HLBM_ERR_ARG, "NULL argument");
  // a consistent snapshot: chunks in flight are waited for and accounted
  for (int k = 0; k < 2; k++) {
    const int rc = finish_chunk(b, (int)((b->run_index + k) & 1));
    if (rc) return rc;
  }
  const int rc = poll_retiring(b, false, nullptr);
  if (rc) return rc;
  out->queued = (int64_t)b->queue.size();
  out->running = (int64_t)b->running.size();
  out->done = b->n_done;
  out->chunks = b->chunks;
  out->bytes_in_use = b->fl_pop.used * b->esize * 2 + b->fl_flag.used * 5;
  out->budget_bytes = b->d.budget_bytes;
  out->node_updates = b->node_updates;
  out->gpu_ms = b->gpu_ms;
  out->peak_bytes_in_use = b->peak_bytes;
  out->relocated = b->relocated;
  return HLBM_OK;
}

}  // extern "C"
