// api.hip -- the C ABI of include/hlbm.h.
//
// Host code only (kernels live in step_*.hip and util_kernels.hip).  Compiled
// with -ffp-contract=off: the few FP64 constants prepared here (c_i.F, the
// cavity coefficients, bounding boxes) must round as in hippoLBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hlbm.h"
#include "common.hpp"
#include "launch.hpp"
#include "util_kernels.hpp"

using namespace hlbm;

namespace {

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(HLBM_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// PRE: f^T as the reference holds it (natural layout, post-BC).
// POST (A/B): f*^{T-1}, post-collision, streaming pending.
// AA_PENDING (AA): natural layout after an odd step, boundary program pending.
// AA_SWAPPED (AA): post-collision populations in the opposite slots of each node.
enum Phase { PHASE_PRE = 0, PHASE_POST = 1, PHASE_AA_PENDING = 2, PHASE_AA_SWAPPED = 3 };

size_t elem_size(int storage) { return storage == HLBM_F64 ? 8 : (storage == HLBM_F32 ? 4 : 2); }

}  // namespace

struct hlbm_domain {
  hlbm_domain_desc d;
  int nx = 0, ny = 0, nz = 0;
  uint32_t N = 0;
  uint64_t stride = 0;
  size_t esize = 8;
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;
  int phase = PHASE_PRE;
  int64_t t = 0;
  uint8_t* flags = nullptr;
  double* m0 = nullptr;
  double* m1 = nullptr;
  double* red_part = nullptr;
  double* red_out = nullptr;
  double* tmp = nullptr;  // one FP64 population plane, conversions
  bool any_wall = false;
  uint32_t* bbm = nullptr;  // bounce-back mask of the bulk kernel, built on demand
  void* oflow = nullptr;    // outflow values captured before in-place AA steps
  uint32_t oflow_n = 0;     // nodes per captured plane
  bool bbm_dirty = true;
  // ghost flags of the periodic axes' first ghost layer (padded grid, see
  // ghost_count); allocated when some axis is periodic
  uint8_t* gflags = nullptr;
  // nodes next to a wall across a periodic face (special_kernel)
  uint32_t nspec = 0, spec_cap = 0;
  uint32_t* spec_idx = nullptr;
  uint32_t* spec_mask = nullptr;
  uint32_t* spec_count = nullptr;
  void* spec_val = nullptr;      // previous-step values, C type, [i * nspec + k]
  void* spec_scratch = nullptr;  // the same size, for the autotuner's steps
  std::vector<uint32_t> spec_host_idx, spec_host_mask;
  uint64_t plan_serial = 1, spec_checked = 0;
  StepParams plan{};
  dim3 block{128, 2, 1};
  int xcd_remap = 1;
  int npt = 1;
  int force_wide = 0;
  int variant = 0;  // bulk-kernel variant (hlbm_set_launch)
  int boundary_first = 1;  // program-plane kernel before the bulk kernel (measured 1-2 % faster)
  // the program-plane kernel on a second stream beside the bulk kernel
  int overlap = 1;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int device = 0;
  uint64_t bytes = 0;
  NodeLayout lay{};  // node layout in HBM (hlbm_set_layout; default x fastest)
};

// nodes along the k-th axis of the domain's layout (0 = fastest)
static int along(const hlbm_domain* h, int k) {
  const int n[3] = {h->nx, h->ny, h->nz};
  return n[h->lay.ord[k]];
}

static bool default_layout_p(const hlbm_domain* h) {
  return h->lay.ord[0] == 0 && h->lay.ord[1] == 1 && h->lay.ord[2] == 2;
}

// Host copies between the ABI's node order (x fastest) and the domain's layout.
template <class T>
static void to_layout(const hlbm_domain* h, const T* src, T* dst) {
  size_t k = 0;
  for (int z = 0; z < h->nz; z++)
    for (int y = 0; y < h->ny; y++)
      for (int x = 0; x < h->nx; x++, k++) dst[node_at(h->lay, x, y, z)] = src[k];
}
template <class T>
static void from_layout(const hlbm_domain* h, const T* src, T* dst) {
  size_t k = 0;
  for (int z = 0; z < h->nz; z++)
    for (int y = 0; y < h->ny; y++)
      for (int x = 0; x < h->nx; x++, k++) dst[k] = src[node_at(h->lay, x, y, z)];
}

// bulk-kernel tiles: block.x nodes along the layout's fastest axis by
// block.y * npt rows along the next one; the grid walks the slowest axis
static int eff_npt(const hlbm_domain* h) { return default_layout_p(h) ? h->npt : 1; }

static unsigned tiles_y(const hlbm_domain* h) {
  const unsigned rows = h->block.y * eff_npt(h);
  return (along(h, 1) + rows - 1) / rows;
}

static dim3 grid_for(const hlbm_domain* h) {
  const unsigned gx = (along(h, 0) + h->block.x - 1) / h->block.x;
  return dim3(gx * tiles_y(h) * along(h, 2), 1, 1);
}

// The rest state the domain's shifted storage is relative to (RestShift).
static RestShift dom_shift(const hlbm_domain* h) {
  return rest_shift(h->d.collision == HLBM_MRT, h->d.mrt_custom, h->d.mrt_rates);
}

static StepParams base_params(hlbm_domain* h) {
  StepParams p = h->plan;
  p.nx = h->nx;
  p.ny = h->ny;
  p.nz = h->nz;
  for (int d = 0; d < 3; d++) p.per[d] = h->d.periodic[d] ? 1 : 0;
  p.N = h->N;
  p.stride = h->stride;
  p.flags = h->flags;
  p.m0 = h->m0;
  p.m1 = h->m1;
  p.tau = h->d.tau;
  p.inv_tau = 1 / h->d.tau;
  for (int d = 0; d < 3; d++) p.F[d] = h->d.fext[d];
  // c_i . F in the reference's order: ex*Fx + ey*Fy + ez*Fz (bgk.hpp:57, mrt.hpp:188)
  for (int i = 0; i < Q; i++) p.ef[i] = CX(i) * h->d.fext[0] + CY(i) * h->d.fext[1] + CZ(i) * h->d.fext[2];
  {
    const RestShift sh = dom_shift(h);
    for (int k = 0; k < 3; k++) p.sig[k] = (float)sh.s[k];
  }
  p.mrt_custom = h->d.mrt_custom;
  p.lay = h->lay;
  for (int k = 0; k < 3; k++) p.ln[k] = along(h, k);
  p.gx = (int)((along(h, 0) + h->block.x - 1) / h->block.x);
  p.gy = (int)tiles_y(h);
  p.gz = along(h, 2);
  p.xcd_remap = h->xcd_remap;
  for (int k = 0; k < 10; k++) p.mrt[k] = h->d.mrt_rates[k];
  set_fp32_constants(p);  // after tau, c_i . F, the cavity terms and the MRT rates
  if (!h->any_wall) p.wall_bb = 0;
  p.any_wall = h->any_wall ? 1 : 0;
  p.bbm = h->bbm;
  p.nspec = p.wall_bb ? h->nspec : 0;
  p.spec_idx = h->spec_idx;
  p.spec_mask = h->spec_mask;
  p.spec_val = h->spec_val;
  p.force_wide = h->force_wide || h->variant == 1 ? 1 : 0;
  p.boundary_first = h->boundary_first;
  p.pair32 = ((uint64_t)h->stride + h->N) * h->esize < 0xFFFFFFFFull ? 1 : 0;
  // 16-bit storage: stochastic rounding keyed on (node, time step)
  p.sr = (h->d.storage == HLBM_F16 || h->d.storage == HLBM_BF16) && !h->d.round_nearest ? 1 : 0;
  p.seed = mix32((uint32_t)h->t * 0x9E3779B9u + 1u);
  return p;
}

static bool any_periodic(const hlbm_domain* h) {
  return h->d.periodic[0] || h->d.periodic[1] || h->d.periodic[2];
}

static size_t compute_size(const hlbm_domain* h) { return h->d.storage == HLBM_F64 ? 8 : 4; }

// Nodes next to a wall across a periodic face (k_find_special), listed after
// the bounce-back mask is built.  The listed values are written by the next
// collision kernel; the flags only change with the state at a pre-collision
// phase (see materialize_for_change), so no listed value is read before.
static int build_special(hlbm_domain* h, hipStream_t s) {
  h->nspec = 0;
  h->spec_host_idx.clear();
  h->spec_host_mask.clear();
  h->spec_checked = 0;
  if (!any_periodic(h) || !h->gflags) return HLBM_OK;
  if (!h->spec_idx) {
    uint64_t cap = 0;
    const uint64_t n[3] = {(uint64_t)h->nx, (uint64_t)h->ny, (uint64_t)h->nz};
    for (int d = 0; d < 3; d++)
      if (h->d.periodic[d]) cap += 2 * n[(d + 1) % 3] * n[(d + 2) % 3];
    cap = std::min<uint64_t>(cap, h->N);
    if (hipMalloc((void**)&h->spec_idx, cap * 4) != hipSuccess ||
        hipMalloc((void**)&h->spec_mask, cap * 4) != hipSuccess ||
        hipMalloc((void**)&h->spec_count, 4) != hipSuccess)
      return fail(HLBM_ERR_NOMEM, "periodic-face wall list");
    h->spec_cap = (uint32_t)cap;
    h->bytes += cap * 8 + 4;
  }
  const int per[3] = {h->d.periodic[0] ? 1 : 0, h->d.periodic[1] ? 1 : 0, h->d.periodic[2] ? 1 : 0};
  uint32_t count = 0;
  HIP_TRY(hipMemsetAsync(h->spec_count, 0, 4, s));
  HIP_TRY(launch_find_special(h->flags, h->gflags, h->nx, h->ny, h->nz, per, h->bbm, h->spec_idx, h->spec_mask,
                              h->spec_count, h->spec_cap, h->lay, s));
  HIP_TRY(hipMemcpyAsync(&count, h->spec_count, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (count > h->spec_cap) return fail(HLBM_ERR_HIP, "periodic-face wall list overflow");
  if (count == 0) return HLBM_OK;
  h->spec_host_idx.resize(count);
  h->spec_host_mask.resize(count);
  HIP_TRY(hipMemcpy(h->spec_host_idx.data(), h->spec_idx, (size_t)count * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(h->spec_host_mask.data(), h->spec_mask, (size_t)count * 4, hipMemcpyDeviceToHost));
  for (void** q : {&h->spec_val, &h->spec_scratch})
    if (*q) {
      (void)hipFree(*q);
      *q = nullptr;
    }
  const size_t vbytes = (size_t)count * Q * compute_size(h);
  if (hipMalloc(&h->spec_val, vbytes) != hipSuccess || hipMalloc(&h->spec_scratch, vbytes) != hipSuccess)
    return fail(HLBM_ERR_NOMEM, "periodic-face wall values");
  HIP_TRY(hipMemset(h->spec_val, 0, vbytes));
  h->nspec = count;
  return HLBM_OK;
}

// The previous-step values stand for the ghost's populations only while the
// operators between streaming and the next collision treat the node and its
// ghost alike.  Plane operators that are functions of the node's own
// populations (lid_driven_cavity, neumann, pressure_inlet) do, on planes of
// non-periodic axes, which hold the ghost too (plane lists span the All box,
// grid_region.hpp:148-193).  Refused: post_bounce_back faces (the ghost
// restores its own stale saved values, bounce_back.hpp:115-140), program planes
// of periodic axes (the ghost across the face is not on them), start_iteration
// boxes (applied to the ghost by its own position), an outflow reading the node
// as its first interior layer, and a wall ghost over a fluid image.
static int check_special(hlbm_domain* h) {
  if (!h->nspec || h->spec_checked == h->plan_serial) return HLBM_OK;
  const StepParams& P = h->plan;
  int closed = 0, open = 0;  // planes of the non-periodic / periodic axes
  for (int pl = 0; pl < 6; pl++) (h->d.periodic[plane_dim(pl)] ? open : closed) |= 1 << pl;
  const int n[3] = {h->nx, h->ny, h->nz};
  for (size_t k = 0; k < h->spec_host_idx.size(); k++) {
    if (h->spec_host_mask[k] >> 31)
      return fail(HLBM_ERR_UNSUPPORTED,
                  "wall_bounce_back: a wall across a periodic face whose ghost node, rasterized at its own "
                  "position, is a wall while its periodic image is fluid (hippoLBM then streams the wall's own "
                  "populations, which this engine does not keep)");
    const uint32_t idx = h->spec_host_idx[k];
    int c[3];
    {
      const uint32_t n0 = (uint32_t)along(h, 0), n01 = n0 * (uint32_t)along(h, 1);
      c[h->lay.ord[2]] = (int)(idx / n01);
      c[h->lay.ord[1]] = (int)(idx % n01 / n0);
      c[h->lay.ord[0]] = (int)(idx % n0);
    }
    int planes = 0;
    for (int pl = 0; pl < 6; pl++) {
      const int d = plane_dim(pl);
      if (c[d] == (plane_side(pl) ? n[d] - 1 : 0)) planes |= 1 << pl;
    }
    bool outflow_reads = false;
    for (int j = 0; j < P.nprog; j++) {
      if (P.prog[j].kind != PK_OUTFLOW) continue;
      const int d = plane_dim(P.prog[j].plane);
      outflow_reads |= c[d] == (plane_side(P.prog[j].plane) ? n[d] - 2 : 1);
    }
    if ((P.plane_bb && (planes & closed)) || (planes & P.prog_planes & open) || P.nset || outflow_reads)
      return fail(HLBM_ERR_UNSUPPORTED,
                  "wall_bounce_back: a node next to a wall across a periodic face lies on a pre/post_bounce_back "
                  "face, on a boundary-operator plane of a periodic axis, under start_iteration set_distribution "
                  "or next to an outflow plane");
  }
  h->spec_checked = h->plan_serial;
  return HLBM_OK;
}

// The bulk kernel's bounce-back mask, rebuilt after the flags changed.
static int ensure_bbm(hlbm_domain* h, hipStream_t s) {
  if (!h->bbm_dirty && h->bbm) return check_special(h);
  if (!h->bbm) {
    if (hipMalloc((void**)&h->bbm, (size_t)h->N * 4) != hipSuccess) return fail(HLBM_ERR_NOMEM, "bounce-back mask");
    h->bytes += (uint64_t)h->N * 4;
  }
  const int per[3] = {h->d.periodic[0] ? 1 : 0, h->d.periodic[1] ? 1 : 0, h->d.periodic[2] ? 1 : 0};
  hipError_t e = launch_build_bbm(h->flags, h->nx, h->ny, h->nz, per, h->bbm, h->lay, s);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, std::string("bounce-back mask: ") + hipGetErrorString(e));
  const int rc = build_special(h, s);
  if (rc) return rc;
  h->bbm_dirty = false;
  return check_special(h);
}

// One step's kernels.  With program planes and a bulk kernel, the program-plane
// kernel runs on the domain's side stream, forked from and joined back into s:
// it is latency-bound on few nodes, and the two kernels touch disjoint nodes
// (A/B) or disjoint slots (AA).  The AA outflow capture runs first, on s.
static hipError_t run_kernels(hlbm_domain* h, StepParams& p, int mode, hipStream_t s) {
  const int coll = h->d.collision == HLBM_MRT ? COLL_MRT : COLL_BGK;
  auto go = [&](hipStream_t st) {
    return h->d.storage == HLBM_F64 ? launch_step_f64(coll, mode, h->npt, grid_for(h), h->block, p, st)
                                    : launch_step_lowp(h->d.storage, coll, mode, h->npt, grid_for(h), h->block, p, st);
  };
  const bool split = h->overlap && p.prog_planes && mode != MODE_COLLIDE && mode != MODE_AA_FIX;
  if (split && !h->side) {
    if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
      h->overlap = 0;  // keep one stream
  }
  if (!split || !h->side) {
    p.parts = PART_ALL;
    return go(s);
  }
  hipError_t e = hipSuccess;
  if (p.oflow) {
    p.parts = PART_CAPTURE;
    e = go(s);
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev_fork, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(h->side, h->ev_fork, 0);
  if (e == hipSuccess) {
    p.parts = PART_BOUNDARY;
    e = go(h->side);
  }
  if (e == hipSuccess) {
    p.parts = PART_BULK;
    e = go(s);
  }
  if (e == hipSuccess) e = hipEventRecord(h->ev_join, h->side);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, h->ev_join, 0);
  return e;
}

static int launch(hlbm_domain* h, int mode, bool write_macro, hipStream_t s, void* out = nullptr,
                  bool skip_program = false) {
  if (h->any_wall && h->plan.wall_bb) {
    const int rc = ensure_bbm(h, s);
    if (rc) return rc;
  }
  StepParams p = base_params(h);
  const bool aa = mode_aa(mode);
  p.fin = h->buf[h->cur];
  p.fout = aa ? (out ? out : h->buf[h->cur]) : h->buf[h->cur ^ 1];
  p.write_macro = write_macro ? 1 : 0;
  p.skip_program = skip_program ? 1 : 0;
  const bool in_place = mode == MODE_AA_EVEN || mode == MODE_AA_ODD;
  p.oflow = in_place ? h->oflow : nullptr;
  p.oflow_n = h->oflow_n;
  const hipError_t e = run_kernels(h, p, mode, s);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, std::string("step kernel launch: ") + hipGetErrorString(e));
  if (!aa) h->cur ^= 1;
  return HLBM_OK;
}

// Bring the state to the reference's post-stream/post-BC form f^T.
static int materialize(hlbm_domain* h, hipStream_t s) {
  if (h->phase == PHASE_POST) {
    int rc = launch(h, MODE_STREAM, false, s);
    if (rc) return rc;
    h->phase = PHASE_PRE;
  } else if (h->phase == PHASE_AA_PENDING) {
    if (h->plan.prog_planes) {
      int rc = launch(h, MODE_AA_FIX, false, s);
      if (rc) return rc;
    }
    h->phase = PHASE_PRE;
  } else if (h->phase == PHASE_AA_SWAPPED) {
    // streaming a swapped AA state in place would race; stream into a second
    // buffer, which then becomes the domain's buffer (the old one is freed)
    void* nb = nullptr;
    const size_t fbytes = (size_t)h->stride * Q * h->esize;
    if (hipMalloc(&nb, fbytes) != hipSuccess) return fail(HLBM_ERR_NOMEM, "AA read-out buffer");
    int rc = launch(h, MODE_AA_STREAM, false, s, nb);
    if (rc) {
      (void)hipFree(nb);
      return rc;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
      (void)hipFree(nb);
      return fail(HLBM_ERR_HIP, "AA read-out sync");
    }
    (void)hipFree(h->buf[h->cur]);
    h->buf[h->cur] = nb;
    h->phase = PHASE_PRE;
  }
  return HLBM_OK;
}

// Flags and the boundary plan change only at a pre-collision phase: the pending
// streaming of the last step belongs to the old ones (and the periodic-face
// wall list, rebuilt with the flags, is filled by the next collision).
static int materialize_for_change(hlbm_domain* h) { return h->phase == PHASE_PRE ? HLBM_OK : materialize(h, 0); }

extern "C" {

const char* hlbm_last_error(void) { return last_error().c_str(); }

const char* hlbm_version(void) { return "hippolbm-mi355x 0.1 (gfx950)"; }

int hlbm_device_count(int32_t* count) {
  if (!count) return fail(HLBM_ERR_ARG, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return HLBM_OK;
}

int hlbm_domain_create(const hlbm_domain_desc* desc, hlbm_domain** out) {
  if (!desc || !out) return fail(HLBM_ERR_ARG, "desc/out is NULL");
  *out = nullptr;
  for (int d = 0; d < 3; d++)
    if (desc->nodes[d] < 2) return fail(HLBM_ERR_ARG, "nodes must be >= 2 on every axis");
  const uint64_t N64 = (uint64_t)desc->nodes[0] * desc->nodes[1] * desc->nodes[2];
  if (N64 >= (1ull << 31)) return fail(HLBM_ERR_ARG, "domain too large for one subdomain (>= 2^31 nodes)");
  if (desc->storage < HLBM_F64 || desc->storage > HLBM_BF16) return fail(HLBM_ERR_ARG, "bad storage");
  if (desc->collision != HLBM_BGK && desc->collision != HLBM_MRT) return fail(HLBM_ERR_ARG, "bad collision");
  if (desc->pattern != HLBM_PATTERN_AB && desc->pattern != HLBM_PATTERN_AA) return fail(HLBM_ERR_ARG, "bad pattern");
  // tau <= 1/2 means a zero or negative viscosity, nu = (tau - 1/2)/3: the update is unstable
  if (!(desc->tau > 0.5)) return fail(HLBM_ERR_ARG, "tau must be > 0.5 (nu = (tau - 0.5)/3 > 0)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(HLBM_ERR_NODEVICE, "no HIP device visible");

  hlbm_domain* h = new hlbm_domain();
  h->d = *desc;
  h->nx = desc->nodes[0];
  h->ny = desc->nodes[1];
  h->nz = desc->nodes[2];
  h->N = (uint32_t)N64;
  h->lay = default_layout(h->nx, h->ny, h->nz);
  h->esize = elem_size(desc->storage);
  const uint64_t align = 256 / h->esize;
  h->stride = (N64 + align - 1) / align * align;
  if (const char* pad = getenv("HLBM_PAD")) h->stride += (uint64_t)strtoull(pad, nullptr, 10);
  if (const char* xr = getenv("HLBM_XCD_REMAP")) h->xcd_remap = atoi(xr);
  if (const char* np = getenv("HLBM_NPT")) h->npt = atoi(np) == 2 ? 2 : 1;
  // the 64-bit-address variant otherwise runs only on domains above 2^32 bytes per plane
  if (const char* fw = getenv("HLBM_FORCE_WIDE")) h->force_wide = atoi(fw) ? 1 : 0;
  if (const char* bf = getenv("HLBM_BOUNDARY_FIRST")) h->boundary_first = atoi(bf) ? 1 : 0;
  if (const char* ov = getenv("HLBM_OVERLAP")) h->overlap = atoi(ov) ? 1 : 0;
  (void)hipGetDevice(&h->device);
  const char* blk = getenv("HLBM_BLOCK");
  if (blk) {
    int bx = 0, by = 0;
    if (sscanf(blk, "%dx%d", &bx, &by) == 2 && bx > 0 && by > 0 && bx * by <= 256 && (bx * by) % 64 == 0)
      h->block = dim3(bx, by, 1);
  }
  const size_t fbytes = (size_t)h->stride * Q * h->esize;
  auto cleanup = [&](const std::string& what) {
    for (void* p : {h->buf[0], h->buf[1], (void*)h->flags, (void*)h->m0, (void*)h->m1, (void*)h->red_part,
                    (void*)h->red_out, (void*)h->tmp, (void*)h->gflags})
      if (p) (void)hipFree(p);
    delete h;
    return fail(HLBM_ERR_NOMEM, "device allocation failed: " + what);
  };
  const bool aa = desc->pattern == HLBM_PATTERN_AA;
  if (hipMalloc(&h->buf[0], fbytes) != hipSuccess) return cleanup("populations A");
  if (!aa && hipMalloc(&h->buf[1], fbytes) != hipSuccess) return cleanup("populations B");
  if (hipMalloc((void**)&h->flags, N64) != hipSuccess) return cleanup("flags");
  if (hipMalloc((void**)&h->m0, N64 * 8) != hipSuccess) return cleanup("rho");
  if (hipMalloc((void**)&h->m1, N64 * 24) != hipSuccess) return cleanup("u");
  if (hipMalloc((void**)&h->red_part, 7 * 8 * 1024) != hipSuccess) return cleanup("reduction");
  if (hipMalloc((void**)&h->red_out, 7 * 8) != hipSuccess) return cleanup("reduction");
  const size_t gbytes = ghost_count(h->nx, h->ny, h->nz);
  if (any_periodic(h) && (hipMalloc((void**)&h->gflags, gbytes) != hipSuccess || hipMemset(h->gflags, 0, gbytes) != hipSuccess))
    return cleanup("ghost flags");
  h->bytes = (aa ? 1 : 2) * fbytes + N64 * (1 + 8 + 24) + 7 * 8 * 1025 + (h->gflags ? gbytes : 0);
  if (hipMemset(h->buf[0], 0, fbytes) != hipSuccess || (!aa && hipMemset(h->buf[1], 0, fbytes) != hipSuccess) ||
      hipMemset(h->flags, 0, N64) != hipSuccess || hipMemset(h->m0, 0, N64 * 8) != hipSuccess ||
      hipMemset(h->m1, 0, N64 * 24) != hipSuccess)
    return cleanup("memset");
  if (hipDeviceSynchronize() != hipSuccess) return cleanup("sync");
  *out = h;
  // zero populations in the storage's own encoding (shifted storage holds minus the rest state)
  if (h->d.storage != HLBM_F64) {
    int lo[3] = {0, 0, 0}, hi[3] = {h->nx - 1, h->ny - 1, h->nz - 1};
    const RestShift sh = dom_shift(h);
    if (launch_set_distribution(h->d.storage, h->buf[0], h->stride, h->lay, lo, hi, 0.0, sh, 0) != hipSuccess ||
        (!aa && launch_set_distribution(h->d.storage, h->buf[1], h->stride, h->lay, lo, hi, 0.0, sh, 0) !=
                    hipSuccess) ||
        hipDeviceSynchronize() != hipSuccess) {
      hlbm_domain_destroy(h);
      *out = nullptr;
      return fail(HLBM_ERR_HIP, "initial fill failed");
    }
  }
  return HLBM_OK;
}

int hlbm_domain_destroy(hlbm_domain* h) {
  if (!h) return HLBM_OK;
  (void)hipDeviceSynchronize();
  for (void* p : {h->buf[0], h->buf[1], (void*)h->flags, (void*)h->m0, (void*)h->m1, (void*)h->red_part,
                  (void*)h->red_out, (void*)h->tmp, (void*)h->bbm, h->oflow, (void*)h->gflags, (void*)h->spec_idx,
                  (void*)h->spec_mask, (void*)h->spec_count, h->spec_val, h->spec_scratch})
    if (p) (void)hipFree(p);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->side) (void)hipStreamDestroy(h->side);
  delete h;
  return HLBM_OK;
}

int hlbm_domain_bytes(const hlbm_domain* h, uint64_t* bytes) {
  if (!h || !bytes) return fail(HLBM_ERR_ARG, "NULL argument");
  *bytes = h->bytes + (h->tmp ? (uint64_t)h->N * 8 : 0) + (uint64_t)h->nspec * Q * compute_size(h) * 2;
  return HLBM_OK;
}

int hlbm_set_params(hlbm_domain* h, double tau, const double fext[3]) {
  if (!h || !fext) return fail(HLBM_ERR_ARG, "NULL argument");
  if (!(tau > 0.5)) return fail(HLBM_ERR_ARG, "tau must be > 0.5 (nu = (tau - 0.5)/3 > 0)");
  h->d.tau = tau;
  for (int d = 0; d < 3; d++) h->d.fext[d] = fext[d];
  return HLBM_OK;
}

// Compile the three operator lists into the kernel's boundary plan.
int hlbm_set_bcs(hlbm_domain* h, const hlbm_bc* list, int32_t n) {
  if (!h || (n > 0 && !list) || n < 0) return fail(HLBM_ERR_ARG, "bad bc list");
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  StepParams P{};
  bool pre_bb = false, post_bb = false;
  bool cav_after_pre[2] = {false, false};
  double cav_u[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int last_stage = HLBM_STAGE_START;
  for (int k = 0; k < n; k++) {
    const hlbm_bc& b = list[k];
    if (b.stage < last_stage)
      return fail(HLBM_ERR_ARG, "bc list must be ordered start_iteration, pre_stream, post_stream, boundary");
    last_stage = b.stage;
    if (b.stage == HLBM_STAGE_START) {
      if (b.kind != HLBM_BC_SET_DISTRIBUTION)
        return fail(HLBM_ERR_UNSUPPORTED, "only set_distribution may run in start_iteration");
      if (P.nset >= MAX_SET) return fail(HLBM_ERR_UNSUPPORTED, "at most 4 set_distribution operators per step");
      int* B = P.set_box[P.nset];
      const int nn[3] = {h->nx, h->ny, h->nz};
      bool inside = true;
      for (int d = 0; d < 3; d++) {
        B[d] = 0;
        B[3 + d] = nn[d] - 1;
        if (b.plane == 1) {  // bounded: int(b/dx) box restricted to the grid (set_distribution.cu:84-95)
          if (!(h->d.dx > 0.0)) return fail(HLBM_ERR_ARG, "desc.dx must be > 0 for a bounded set_distribution");
          const int a = int(b.bounds[d] / h->d.dx), c = int(b.bounds[3 + d] / h->d.dx);
          if (c < 0 || a > nn[d] - 1) inside = false;
          B[d] = std::max(a, 0);
          B[3 + d] = std::min(c, nn[d] - 1);
        }
      }
      if (inside) P.set_val[P.nset++] = b.U[0];
    } else if (b.stage == HLBM_STAGE_PRE_STREAM) {
      switch (b.kind) {
        case HLBM_BC_PRE_BOUNCE_BACK:
          pre_bb = true;
          break;
        case HLBM_BC_WALL_BOUNCE_BACK:
          P.wall_bb = 1;
          break;
        case HLBM_BC_CAVITY_Z_0:
        case HLBM_BC_CAVITY_Z_L: {
          const int side = b.kind == HLBM_BC_CAVITY_Z_L ? 1 : 0;
          // a cavity before pre_bounce_back acts on buffers that pre_bounce_back then overwrites
          if (!pre_bb) break;
          if (cav_after_pre[side]) return fail(HLBM_ERR_UNSUPPORTED, "more than one cavity operator on one z face");
          cav_after_pre[side] = true;
          for (int d = 0; d < 3; d++) cav_u[side][d] = b.U[d];
          break;
        }
        default:
          return fail(HLBM_ERR_UNSUPPORTED, "only pre_bounce_back, wall_bounce_back and cavity_z_* may run before streaming");
      }
    } else if (b.stage == HLBM_STAGE_POST_STREAM || b.stage == HLBM_STAGE_BOUNDARY) {
      switch (b.kind) {
        case HLBM_BC_POST_BOUNCE_BACK:
          // post_bounce_back first in the post-stream lists: the restored values
          // (bounce-back + moving wall) are what the gather already produces
          if (!post_bb && P.nprog == 0) P.cav_gather = 1;
          post_bb = true;
          // The gather already applied on-node bounce-back (+ moving wall); a
          // post_bounce_back re-applies it only after an operator changed the node.
          if (pre_bb && P.nprog > 0) {
            if (P.nprog >= MAX_PROG) return fail(HLBM_ERR_UNSUPPORTED, "too many boundary operators");
            P.prog[P.nprog].kind = PK_REBB;
            P.prog[P.nprog].plane = -1;
            P.nprog++;
          }
          break;
        case HLBM_BC_LID_DRIVEN_CAVITY:
        case HLBM_BC_NEUMANN:
          if (b.plane < 0 || b.plane > 5) return fail(HLBM_ERR_ARG, "bad plane");
          if (P.nprog >= MAX_PROG) return fail(HLBM_ERR_UNSUPPORTED, "too many boundary operators");
          P.prog[P.nprog].kind = b.kind == HLBM_BC_LID_DRIVEN_CAVITY ? PK_LID : PK_NEUMANN;
          P.prog[P.nprog].plane = b.plane;
          for (int d = 0; d < 3; d++) P.prog[P.nprog].U[d] = b.U[d];
          P.prog_planes |= 1 << b.plane;
          P.nprog++;
          break;
        case HLBM_BC_PRESSURE_INLET:
        case HLBM_BC_OUTFLOW: {
          if (b.plane < 0 || b.plane > 5) return fail(HLBM_ERR_ARG, "bad plane");
          const int dim = plane_dim(b.plane);
          if (h->d.periodic[dim]) return fail(HLBM_ERR_UNSUPPORTED, "pressure_inlet / outflow need a non-periodic axis");
          if (b.kind == HLBM_BC_PRESSURE_INLET && !(b.U[0] > 0.0))
            return fail(HLBM_ERR_ARG, "pressure_inlet density (U[0]) must be > 0");
          const int n_axis = dim == 0 ? h->nx : (dim == 1 ? h->ny : h->nz);
          if (b.kind == HLBM_BC_OUTFLOW && n_axis < 3) return fail(HLBM_ERR_UNSUPPORTED, "outflow needs >= 3 nodes");
          if (P.nprog >= MAX_PROG) return fail(HLBM_ERR_UNSUPPORTED, "too many boundary operators");
          P.prog[P.nprog].kind = b.kind == HLBM_BC_PRESSURE_INLET ? PK_PRESSURE : PK_OUTFLOW;
          P.prog[P.nprog].plane = b.plane;
          for (int d = 0; d < 3; d++) P.prog[P.nprog].U[d] = b.U[d];
          P.prog_planes |= 1 << b.plane;
          P.nprog++;
          break;
        }
        default:
          return fail(HLBM_ERR_UNSUPPORTED,
                      "only post_bounce_back, lid_driven_cavity, neumann, pressure_inlet and outflow may run after "
                      "streaming");
      }
    } else {
      return fail(HLBM_ERR_ARG, "bad stage");
    }
  }
  P.plane_bb = (pre_bb && post_bb) ? 1 : 0;
  // the moving-wall term exists only through the saved/restored buffers of a non-periodic z axis
  const double L = h->nz;  // domain_size_ z (cavity.cu:69, cavity.hpp:58-62)
  for (int side = 0; side < 2; side++) {
    P.cav_on[side] = (P.plane_bb && cav_after_pre[side] && !h->d.periodic[2]) ? 1 : 0;
    const double uxx = cav_u[side][0] * (1 + 0.5 / (L - 1));
    const double uyy = cav_u[side][1] * (1 + 0.5 / (L - 1));
    const double uzz = cav_u[side][2] * (1 + 0.5 / (L - 1));
    for (int j = 0; j < 5; j++) {
      const int q = Faces::in[4 + side][j];
      P.cav[side][j] = 6. * WD(q) * (CX(q) * uxx + CY(q) * uyy + CZ(q) * uzz);
    }
  }
  // outflow entries carry their ordinal (the slot of their captured values)
  int n_out = 0;
  for (int k = 0; k < P.nprog; k++)
    if (P.prog[k].kind == PK_OUTFLOW) P.prog[k].U[0] = (double)n_out++;
  // outflow reads the first interior layer as the gather produces it: no earlier
  // entry may have changed nodes of that layer (entries on planes of another
  // axis), nor a re-applied bounce-back added a moving-wall term there
  for (int k = 0; k < P.nprog; k++) {
    if (P.prog[k].kind != PK_OUTFLOW) continue;
    for (int j = 0; j < k; j++) {
      const ProgEntry& e = P.prog[j];
      if (e.kind == PK_REBB ? (P.cav_on[0] || P.cav_on[1]) : plane_dim(e.plane) != plane_dim(P.prog[k].plane))
        return fail(HLBM_ERR_UNSUPPORTED, "outflow must precede operators that change its interior layer");
    }
  }
  bool has_rebb = false;
  for (int k = 0; k < P.nprog; k++) has_rebb |= P.prog[k].kind == PK_REBB;
  if (has_rebb) {
    for (int dim = 0; dim < 3; dim++)
      if (!h->d.periodic[dim]) P.prog_planes |= dim == 2 ? 3 : (dim == 1 ? 12 : 48);
  }
  // in-place AA steps capture the outflow values first (the step overwrites the interior layer)
  if (h->oflow) {
    (void)hipFree(h->oflow);
    h->oflow = nullptr;
    h->oflow_n = 0;
  }
  if (n_out > 0 && h->d.pattern == HLBM_PATTERN_AA) {
    const uint64_t pl = std::max({(uint64_t)h->nx * h->ny, (uint64_t)h->ny * h->nz, (uint64_t)h->nx * h->nz});
    const size_t csize = h->d.storage == HLBM_F64 ? 8 : 4;
    if (hipMalloc(&h->oflow, (size_t)n_out * 5 * pl * csize) != hipSuccess) {
      h->oflow = nullptr;
      return fail(HLBM_ERR_NOMEM, "outflow capture buffer");
    }
    h->oflow_n = (uint32_t)pl;
  }
  h->plan = P;
  h->plan_serial++;
  return HLBM_OK;
}

int hlbm_init_obst(hlbm_domain* h) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  HIP_TRY(hipMemset(h->flags, 0, h->N));
  if (h->gflags) HIP_TRY(hipMemset(h->gflags, 0, ghost_count(h->nx, h->ny, h->nz)));
  HIP_TRY(hipDeviceSynchronize());
  h->any_wall = false;
  h->bbm_dirty = true;
  return HLBM_OK;
}

static int refresh_any_wall(hlbm_domain* h) {
  h->bbm_dirty = true;
  std::vector<uint8_t> fl(h->N);
  HIP_TRY(hipMemcpy(fl.data(), h->flags, h->N, hipMemcpyDeviceToHost));
  h->any_wall = std::any_of(fl.begin(), fl.end(), [](uint8_t v) { return v != 0; });
  return HLBM_OK;
}

int hlbm_add_obstacles(hlbm_domain* h, const hlbm_obstacle* list, int32_t n) {
  if (!h || (n > 0 && !list) || n < 0) return fail(HLBM_ERR_ARG, "bad obstacle list");
  if (n == 0) return HLBM_OK;
  if (!(h->d.dx > 0.0)) return fail(HLBM_ERR_ARG, "desc.dx must be > 0 to rasterize obstacles");
  const double dx = h->d.dx;
  const int nn[3] = {h->nx, h->ny, h->nz};
  const int per[3] = {h->d.periodic[0] ? 1 : 0, h->d.periodic[1] ? 1 : 0, h->d.periodic[2] ? 1 : 0};
  std::vector<ObstacleDev> obs;
  for (int k = 0; k < n; k++) {
    if (list[k].kind < 0 || list[k].kind > 2)
      return fail(HLBM_ERR_ARG, "obstacle kind must be 0 (ball), 1 (wall) or 2 (quadric)");
    ObstacleDev od;
    if (obstacle_to_dev(list[k], nn, dx, od, per)) obs.push_back(od);
  }
  if (obs.empty()) return HLBM_OK;
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  ObstacleDev* dobs = nullptr;
  HIP_TRY(hipMalloc(&dobs, obs.size() * sizeof(ObstacleDev)));
  hipError_t e = hipMemcpy(dobs, obs.data(), obs.size() * sizeof(ObstacleDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_rasterize(h->flags, h->nx, h->ny, h->nz, dx, dobs, (int)obs.size(), h->lay, 0);
  if (e == hipSuccess && h->gflags)
    e = launch_rasterize_ghost(h->gflags, h->nx, h->ny, h->nz, per, dx, dobs, (int)obs.size(), 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(dobs);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, std::string("rasterize: ") + hipGetErrorString(e));
  return refresh_any_wall(h);
}

// reference layout <-> node layout index helpers
static inline size_t ref_index(const hlbm_domain* h, int x, int y, int z) {
  const size_t Lx = h->nx + 4, Ly = h->ny + 4;
  return ((size_t)(z + 2) * Ly + (y + 2)) * Lx + (x + 2);
}

static size_t ref_count(const hlbm_domain* h) { return (size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4); }

int hlbm_set_flags(hlbm_domain* h, const int32_t* obst, int32_t layout) {
  if (!h || !obst) return fail(HLBM_ERR_ARG, "NULL argument");
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  std::vector<uint8_t> fl(h->N);
  size_t k = 0;
  for (int z = 0; z < h->nz; z++)
    for (int y = 0; y < h->ny; y++)
      for (int x = 0; x < h->nx; x++, k++) {
        const int32_t v = layout == HLBM_LAYOUT_REF ? obst[ref_index(h, x, y, z)] : obst[k];
        if (v != -1 && v != -2) return fail(HLBM_ERR_ARG, "obstacle values must be -1 (fluid) or -2 (wall)");
        fl[k] = v == -2 ? 1 : 0;
      }
  if (!default_layout_p(h)) {
    std::vector<uint8_t> lf(h->N);
    to_layout(h, fl.data(), lf.data());
    fl.swap(lf);
  }
  HIP_TRY(hipMemcpy(h->flags, fl.data(), h->N, hipMemcpyHostToDevice));
  if (h->gflags) {
    const int per[3] = {h->d.periodic[0] ? 1 : 0, h->d.periodic[1] ? 1 : 0, h->d.periodic[2] ? 1 : 0};
    if (layout == HLBM_LAYOUT_REF) {
      // the first ghost layer as given (the reference's obstacle field over the All box)
      std::vector<uint8_t> g(ghost_count(h->nx, h->ny, h->nz), 0);
      size_t j = 0;
      for (int z = -1; z <= h->nz; z++)
        for (int y = -1; y <= h->ny; y++)
          for (int x = -1; x <= h->nx; x++, j++) {
            const int32_t v = obst[ref_index(h, x, y, z)];
            if (v != -1 && v != -2) return fail(HLBM_ERR_ARG, "obstacle values must be -1 (fluid) or -2 (wall)");
            g[j] = v == -2 ? 1 : 0;
          }
      HIP_TRY(hipMemcpy(h->gflags, g.data(), g.size(), hipMemcpyHostToDevice));
    } else {
      // node layout carries no ghosts: they take their periodic images' flags
      HIP_TRY(launch_ghost_images(h->gflags, h->flags, h->nx, h->ny, h->nz, per, h->lay, 0));
      HIP_TRY(hipDeviceSynchronize());
    }
  }
  h->any_wall = std::any_of(fl.begin(), fl.end(), [](uint8_t v) { return v != 0; });
  h->bbm_dirty = true;
  return HLBM_OK;
}

int hlbm_get_flags(hlbm_domain* h, int32_t* obst, int32_t layout) {
  if (!h || !obst) return fail(HLBM_ERR_ARG, "NULL argument");
  std::vector<uint8_t> fl(h->N);
  HIP_TRY(hipMemcpy(fl.data(), h->flags, h->N, hipMemcpyDeviceToHost));
  if (!default_layout_p(h)) {
    std::vector<uint8_t> nf(h->N);
    from_layout(h, fl.data(), nf.data());
    fl.swap(nf);
  }
  size_t k = 0;
  for (int z = 0; z < h->nz; z++)
    for (int y = 0; y < h->ny; y++)
      for (int x = 0; x < h->nx; x++, k++) {
        const int32_t v = fl[k] ? -2 : -1;
        if (layout == HLBM_LAYOUT_REF)
          obst[ref_index(h, x, y, z)] = v;
        else
          obst[k] = v;
      }
  return HLBM_OK;
}

int hlbm_set_distribution(hlbm_domain* h, double value, const double* bounds) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  int rc = materialize(h, 0);
  if (rc) return rc;
  int lo[3] = {0, 0, 0}, hi[3] = {h->nx - 1, h->ny - 1, h->nz - 1};
  if (bounds) {
    if (!(h->d.dx > 0.0)) return fail(HLBM_ERR_ARG, "desc.dx must be > 0 for a bounded set_distribution");
    const int nn[3] = {h->nx, h->ny, h->nz};
    for (int d = 0; d < 3; d++) {
      const int a = int(bounds[d] / h->d.dx), b = int(bounds[3 + d] / h->d.dx);
      if (b < 0 || a > nn[d] - 1) return HLBM_OK;  // no intersection (set_distribution.cu:94-95)
      lo[d] = std::max(a, 0);
      hi[d] = std::min(b, nn[d] - 1);
    }
  }
  HIP_TRY(launch_set_distribution(h->d.storage, h->buf[h->cur], h->stride, h->lay, lo, hi, value, dom_shift(h),
                                  0));
  HIP_TRY(hipDeviceSynchronize());
  return HLBM_OK;
}

int hlbm_init_double_couette(hlbm_domain* h, int32_t dim, const double U[3]) {
  if (!h || !U) return fail(HLBM_ERR_ARG, "NULL argument");
  if (dim < 0 || dim > 2) return fail(HLBM_ERR_ARG, "dim must be 0 (x), 1 (y) or 2 (z)");
  int rc = materialize(h, 0);
  if (rc) return rc;
  const int n = dim == 0 ? h->nx : (dim == 1 ? h->ny : h->nz);
  // dU = Uc / (0.5 * (domain_size[dim] - 1)) (init_double_couette.cu:85-107)
  double dU[3];
  for (int d = 0; d < 3; d++) dU[d] = U[d] / (0.5 * (n - 1));
  HIP_TRY(launch_double_couette(h->d.storage, h->buf[h->cur], h->stride, h->nx, h->ny, h->nz, dim, dU, U,
                                dom_shift(h), h->lay, 0));
  HIP_TRY(hipDeviceSynchronize());
  return HLBM_OK;
}

static int ensure_tmp(hlbm_domain* h) {
  if (!h->tmp) HIP_TRY(hipMalloc((void**)&h->tmp, (size_t)h->N * 8));
  return HLBM_OK;
}

int hlbm_set_populations_f64(hlbm_domain* h, const double* f, int32_t layout) {
  if (!h || !f) return fail(HLBM_ERR_ARG, "NULL argument");
  int rc = ensure_tmp(h);
  if (rc) return rc;
  std::vector<double> plane, lplane;
  const size_t Nall = ref_count(h);
  char* base = (char*)h->buf[h->cur];
  for (int i = 0; i < Q; i++) {
    const double* src = f + (size_t)i * h->N;
    if (layout == HLBM_LAYOUT_REF) {
      plane.resize(h->N);
      size_t k = 0;
      for (int z = 0; z < h->nz; z++)
        for (int y = 0; y < h->ny; y++)
          for (int x = 0; x < h->nx; x++, k++) plane[k] = f[(size_t)i * Nall + ref_index(h, x, y, z)];
      src = plane.data();
    }
    if (!default_layout_p(h)) {
      lplane.resize(h->N);
      to_layout(h, src, lplane.data());
      src = lplane.data();
    }
    void* dst = base + (size_t)i * h->stride * h->esize;
    if (h->d.storage == HLBM_F64) {
      HIP_TRY(hipMemcpy(dst, src, (size_t)h->N * 8, hipMemcpyHostToDevice));
    } else {
      HIP_TRY(hipMemcpy(h->tmp, src, (size_t)h->N * 8, hipMemcpyHostToDevice));
      HIP_TRY(launch_convert(h->d.storage, true, dst, h->tmp, h->N, i, dom_shift(h), 0));
    }
  }
  HIP_TRY(hipDeviceSynchronize());
  h->phase = PHASE_PRE;
  return HLBM_OK;
}

static void fill_periodic_ghosts(const hlbm_domain* h, double* fref, int ncomp) {
  // single-rank update_ghost: every periodic shift copies the opposite Real slab
  // into the ghost slab (operator_ghost_manager.hpp:65-114, make_domain.hpp:213-255)
  const int n[3] = {h->nx, h->ny, h->nz};
  const int L[3] = {h->nx + 4, h->ny + 4, h->nz + 4};
  const size_t Nall = (size_t)L[0] * L[1] * L[2];
  for (int s = 0; s < 27; s++) {
    int rel[3], t = s;
    for (int d = 0; d < 3; d++) {
      const int v = t % 3;
      rel[d] = v - 1;
      t = (t - v) / 3;
    }
    if (!rel[0] && !rel[1] && !rel[2]) continue;
    bool exists = true;
    for (int d = 0; d < 3; d++)
      if (rel[d] && !h->d.periodic[d]) exists = false;
    if (!exists) continue;
    int lo[3], hi[3], sh[3];
    for (int d = 0; d < 3; d++) {
      if (rel[d] == -1) {
        lo[d] = 0;
        hi[d] = 1;
        sh[d] = n[d];
      } else if (rel[d] == 1) {
        lo[d] = L[d] - 2;
        hi[d] = L[d] - 1;
        sh[d] = -n[d];
      } else {
        lo[d] = 2;
        hi[d] = L[d] - 3;
        sh[d] = 0;
      }
    }
    for (int c = 0; c < ncomp; c++)
      for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
          for (int x = lo[0]; x <= hi[0]; x++) {
            const size_t dst = ((size_t)z * L[1] + y) * L[0] + x;
            const size_t src = ((size_t)(z + sh[2]) * L[1] + (y + sh[1])) * L[0] + (x + sh[0]);
            fref[(size_t)c * Nall + dst] = fref[(size_t)c * Nall + src];
          }
  }
}

int hlbm_get_populations_f64(hlbm_domain* h, double* f, int32_t layout) {
  if (!h || !f) return fail(HLBM_ERR_ARG, "NULL argument");
  int rc = materialize(h, 0);
  if (rc) return rc;
  rc = ensure_tmp(h);
  if (rc) return rc;
  std::vector<double> plane(layout == HLBM_LAYOUT_REF ? h->N : 0), lplane(default_layout_p(h) ? 0 : h->N);
  const size_t Nall = ref_count(h);
  char* base = (char*)h->buf[h->cur];
  for (int i = 0; i < Q; i++) {
    double* out = layout == HLBM_LAYOUT_REF ? plane.data() : f + (size_t)i * h->N;
    double* dst = default_layout_p(h) ? out : lplane.data();
    void* srcp = base + (size_t)i * h->stride * h->esize;
    if (h->d.storage == HLBM_F64) {
      HIP_TRY(hipMemcpy(dst, srcp, (size_t)h->N * 8, hipMemcpyDeviceToHost));
    } else {
      HIP_TRY(launch_convert(h->d.storage, false, srcp, h->tmp, h->N, i, dom_shift(h), 0));
      HIP_TRY(hipMemcpy(dst, h->tmp, (size_t)h->N * 8, hipMemcpyDeviceToHost));
    }
    if (!default_layout_p(h)) from_layout(h, lplane.data(), out);
    if (layout == HLBM_LAYOUT_REF) {
      size_t k = 0;
      for (int z = 0; z < h->nz; z++)
        for (int y = 0; y < h->ny; y++)
          for (int x = 0; x < h->nx; x++, k++) f[(size_t)i * Nall + ref_index(h, x, y, z)] = plane[k];
    }
  }
  if (layout == HLBM_LAYOUT_REF) fill_periodic_ghosts(h, f, Q);
  return HLBM_OK;
}

int hlbm_ghost_update_ref(const hlbm_domain* h, double* f_ref) {
  if (!h || !f_ref) return fail(HLBM_ERR_ARG, "NULL argument");
  fill_periodic_ghosts(h, f_ref, Q);
  return HLBM_OK;
}

int hlbm_step(hlbm_domain* h, int32_t nsteps, uint32_t flags, void* stream) {
  if (!h || nsteps < 0) return fail(HLBM_ERR_ARG, "bad step arguments");
  hipStream_t s = (hipStream_t)stream;
  const bool aa = h->d.pattern == HLBM_PATTERN_AA;
  for (int k = 0; k < nsteps; k++) {
    const bool wm = (flags & HLBM_STEP_WRITE_MACRO) && k == nsteps - 1;
    if (aa) {
      // even: own slots -> opposite slots (program first if the state is pending);
      // odd: gather from the swapped slots, push to the neighbours
      const bool even = h->phase != PHASE_AA_SWAPPED;
      int rc = even ? launch(h, MODE_AA_EVEN, wm, s, nullptr, h->phase == PHASE_PRE) : launch(h, MODE_AA_ODD, wm, s);
      if (rc) return rc;
      h->phase = even ? PHASE_AA_SWAPPED : PHASE_AA_PENDING;
    } else {
      const int mode = h->phase == PHASE_PRE ? MODE_COLLIDE : MODE_FUSED;
      int rc = launch(h, mode, wm, s);
      if (rc) return rc;
      h->phase = PHASE_POST;
    }
    h->t++;
  }
  return HLBM_OK;
}

int hlbm_timestep(const hlbm_domain* h, int64_t* t) {
  if (!h || !t) return fail(HLBM_ERR_ARG, "NULL argument");
  *t = h->t;
  return HLBM_OK;
}

int hlbm_synchronize(hlbm_domain* h) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  HIP_TRY(hipDeviceSynchronize());
  return HLBM_OK;
}

static int copy_macro_out(hlbm_domain* h, double* rho, double* u, int32_t layout) {
  const size_t N = h->N;
  std::vector<double> r(N), v(3 * N);
  HIP_TRY(hipMemcpy(r.data(), h->m0, N * 8, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(v.data(), h->m1, 3 * N * 8, hipMemcpyDeviceToHost));
  if (!default_layout_p(h)) {
    std::vector<double> t(N);
    from_layout(h, r.data(), t.data());
    r.swap(t);
    for (int c = 0; c < 3; c++) {
      from_layout(h, v.data() + c * N, t.data());
      std::memcpy(v.data() + c * N, t.data(), N * 8);
    }
  }
  if (layout != HLBM_LAYOUT_REF) {
    if (rho) std::memcpy(rho, r.data(), N * 8);
    if (u) std::memcpy(u, v.data(), 3 * N * 8);
    return HLBM_OK;
  }
  const size_t Nall = ref_count(h);
  size_t k = 0;
  for (int z = 0; z < h->nz; z++)
    for (int y = 0; y < h->ny; y++)
      for (int x = 0; x < h->nx; x++, k++) {
        const size_t j = ref_index(h, x, y, z);
        if (rho) rho[j] = r[k];
        if (u)
          for (int c = 0; c < 3; c++) u[c * Nall + j] = v[c * N + k];
      }
  return HLBM_OK;
}

int hlbm_macro(hlbm_domain* h, double* rho, double* u, int32_t layout) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  int rc = materialize(h, 0);
  if (rc) return rc;
  HIP_TRY(launch_macro(h->d.storage, h->buf[h->cur], h->stride, h->N, h->flags, h->m0, h->m1, 0));
  HIP_TRY(hipDeviceSynchronize());
  if (rho || u) return copy_macro_out(h, rho, u, layout);
  return HLBM_OK;
}

int hlbm_get_macro(hlbm_domain* h, double* rho, double* u, int32_t layout) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  HIP_TRY(hipDeviceSynchronize());
  return copy_macro_out(h, rho, u, layout);
}

int hlbm_reduce_stats(hlbm_domain* h, double out[7]) {
  if (!h || !out) return fail(HLBM_ERR_ARG, "NULL argument");
  HIP_TRY(launch_reduce(h->m0, h->m1, h->flags, h->N, h->red_part, h->red_out, 0));
  HIP_TRY(hipMemcpy(out, h->red_out, 7 * 8, hipMemcpyDeviceToHost));
  return HLBM_OK;
}

int hlbm_plane_profile(hlbm_domain* h, int32_t dim, double scale, double* sum, int32_t* count, double* vmin,
                       double* vmax) {
  if (!h || !sum || !count || !vmin || !vmax) return fail(HLBM_ERR_ARG, "NULL argument");
  if (dim < 0 || dim > 2) return fail(HLBM_ERR_ARG, "dim must be 0 (x), 1 (y) or 2 (z)");
  const int n[3] = {h->nx, h->ny, h->nz};
  const size_t parts = profile_parts(h->nx, h->ny, h->nz, dim);
  ProfileStat* dev = nullptr;
  HIP_TRY(hipMalloc((void**)&dev, (parts + (size_t)n[dim]) * sizeof(ProfileStat)));
  std::vector<ProfileStat> host(n[dim]);
  hipError_t e =
      launch_plane_profile(h->m1, h->flags, h->N, h->nx, h->ny, h->nz, dim, scale, dev, dev + parts, h->lay, 0);
  if (e == hipSuccess)
    e = hipMemcpy(host.data(), dev + parts, (size_t)n[dim] * sizeof(ProfileStat), hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, hipGetErrorString(e));
  for (int p = 0; p < n[dim]; p++) {
    sum[p] = host[p].sum;
    count[p] = host[p].count;
    vmin[p] = host[p].vmin;
    vmax[p] = host[p].vmax;
  }
  return HLBM_OK;
}

int hlbm_plane_profile_noalloc(hlbm_domain* h, int32_t dim, double scale, hlbm_plane_stat* out, void* stream) {
  if (!h || !out) return fail(HLBM_ERR_ARG, "NULL argument");
  if (dim < 0 || dim > 2) return fail(HLBM_ERR_ARG, "dim must be 0 (x), 1 (y) or 2 (z)");
  if (profile_parts(h->nx, h->ny, h->nz, dim) > (size_t)HLBM_PROFILE_SCRATCH_COLUMNS)
    return fail(HLBM_ERR_ARG, "plane profile: more columns than HLBM_PROFILE_SCRATCH_COLUMNS");
  ProfileStat* part = nullptr;
  HIP_TRY(profile_scratch(&part));
  HIP_TRY(launch_plane_profile(h->m1, h->flags, h->N, h->nx, h->ny, h->nz, dim, scale, part, (ProfileStat*)out, h->lay,
                               (hipStream_t)stream));
  return HLBM_OK;
}

int hlbm_line_velocity(hlbm_domain* h, const int32_t start[3], int32_t axis, int32_t n, double scale, double* u) {
  if (!h || !start || !u) return fail(HLBM_ERR_ARG, "NULL argument");
  if (axis < 0 || axis > 2 || n < 1) return fail(HLBM_ERR_ARG, "axis must be 0..2 and n >= 1");
  double* dev = nullptr;
  HIP_TRY(hipMalloc((void**)&dev, (size_t)n * 3 * 8));
  const int st[3] = {start[0], start[1], start[2]};
  hipError_t e = launch_line_velocity(h->m1, h->N, h->nx, h->ny, h->nz, st, axis, n, scale, dev, h->lay, 0);
  if (e == hipSuccess) e = hipMemcpy(u, dev, (size_t)n * 3 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, hipGetErrorString(e));
  return HLBM_OK;
}

int hlbm_set_collision(hlbm_domain* h, int32_t collision, const double* mrt_rates) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  if (collision != HLBM_BGK && collision != HLBM_MRT) return fail(HLBM_ERR_ARG, "collision must be BGK or MRT");
  // the pending streaming belongs to the old collision; shifted storage then
  // moves to the new rest state
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  const RestShift old = dom_shift(h);
  h->d.collision = collision;
  h->d.mrt_custom = mrt_rates ? 1 : 0;
  for (int k = 0; k < 10; k++) h->d.mrt_rates[k] = mrt_rates ? mrt_rates[k] : 0.0;
  const RestShift now = dom_shift(h);
  RestShift delta;
  bool same = true;
  for (int k = 0; k < 3; k++) {
    delta.s[k] = old.s[k] - now.s[k];
    same = same && delta.s[k] == 0.0;
  }
  if (!same && h->d.storage != HLBM_F64) {
    HIP_TRY(launch_reshift(h->d.storage, h->buf[h->cur], h->stride, h->N, delta, 0));
    HIP_TRY(hipDeviceSynchronize());
  }
  return HLBM_OK;
}

int hlbm_set_launch(hlbm_domain* h, int32_t block_x, int32_t block_y, int32_t npt, int32_t xcd_remap,
                    int32_t variant) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  if (block_x < 1 || block_y < 1 || block_x * block_y > 256 || (block_x * block_y) % 64 != 0)
    return fail(HLBM_ERR_ARG, "block_x*block_y must be a multiple of 64 and at most 256");
  if (npt != 1 && npt != 2) return fail(HLBM_ERR_ARG, "npt must be 1 or 2");
  if (variant != 0 && variant != 1) return fail(HLBM_ERR_ARG, "variant must be 0 or 1");
  h->block = dim3((unsigned)block_x, (unsigned)block_y, 1);
  h->variant = variant;
  h->npt = variant == 1 ? 1 : npt;
  h->xcd_remap = xcd_remap ? 1 : 0;
  return HLBM_OK;
}

int hlbm_set_layout(hlbm_domain* h, const int32_t order[3]) {
  if (!h || !order) return fail(HLBM_ERR_ARG, "NULL argument");
  const int o[3] = {order[0], order[1], order[2]};
  bool seen[3] = {false, false, false};
  for (int k = 0; k < 3; k++) {
    if (o[k] < 0 || o[k] > 2 || seen[o[k]]) return fail(HLBM_ERR_ARG, "order must be a permutation of 0, 1, 2");
    seen[o[k]] = true;
  }
  if (o[0] == h->lay.ord[0] && o[1] == h->lay.ord[1]) return HLBM_OK;
  // the pending streaming belongs to the old layout; the periodic-face wall list
  // and the bounce-back mask are rebuilt for the new one
  {
    const int rc = materialize_for_change(h);
    if (rc) return rc;
  }
  {
    const int rc = ensure_tmp(h);
    if (rc) return rc;
  }
  const int n[3] = {h->nx, h->ny, h->nz};
  const NodeLayout from = h->lay, to = make_layout(n, o);
  // every per-node array through the FP64 scratch plane, one plane at a time
  auto move = [&](void* a, size_t esize) -> hipError_t {
    hipError_t e = launch_permute(h->tmp, a, esize, n, from, to, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(a, h->tmp, (size_t)h->N * esize, hipMemcpyDeviceToDevice, 0);
    return e;
  };
  for (int i = 0; i < Q; i++) HIP_TRY(move((char*)h->buf[h->cur] + (size_t)i * h->stride * h->esize, h->esize));
  HIP_TRY(move(h->flags, 1));
  HIP_TRY(move(h->m0, 8));
  for (int c = 0; c < 3; c++) HIP_TRY(move(h->m1 + (size_t)c * h->N, 8));
  HIP_TRY(hipDeviceSynchronize());
  h->lay = to;
  h->bbm_dirty = true;
  return HLBM_OK;
}

int hlbm_get_layout(const hlbm_domain* h, int32_t order[3]) {
  if (!h || !order) return fail(HLBM_ERR_ARG, "NULL argument");
  for (int k = 0; k < 3; k++) order[k] = h->lay.ord[k];
  return HLBM_OK;
}

namespace {
struct TuneKey {
  int nx, ny, nz, storage, collision, pattern, prog_planes, wall_bb, device;
  bool operator<(const TuneKey& o) const {
    const int a[9] = {nx, ny, nz, storage, collision, pattern, prog_planes, wall_bb, device};
    const int b[9] = {o.nx, o.ny, o.nz, o.storage, o.collision, o.pattern, o.prog_planes, o.wall_bb, o.device};
    return std::lexicographical_compare(a, a + 9, b, b + 9);
  }
};
struct TuneVal {
  int bx, by, npt, remap, variant;
  double ms_best, ms_default;
  int order[3];
};
std::mutex tune_mu;
std::map<TuneKey, TuneVal> tune_cache;
}  // namespace

// Time `reps` steps with the domain's current launch shape on scratch memory.
static int time_shape(hlbm_domain* h, void* scratch, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                      double* ms) {
  const bool aa = h->d.pattern == HLBM_PATTERN_AA;
  StepParams p = base_params(h);
  p.write_macro = 0;
  p.skip_program = 0;
  p.spec_val = h->spec_scratch;  // the listed previous-step values belong to the real state
  auto one = [&](int k) -> hipError_t {
    int mode;
    if (aa) {
      mode = (k & 1) ? MODE_AA_ODD : MODE_AA_EVEN;
      p.fin = scratch;
      p.fout = scratch;
    } else {
      mode = MODE_FUSED;
      p.fin = h->buf[h->cur];
      p.fout = h->buf[h->cur ^ 1];  // overwritten by the next real step anyway
    }
    return run_kernels(h, p, mode, s);
  };
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2 && e == hipSuccess; k++) e = one(k);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int k = 0; k < reps && e == hipSuccess; k++) e = one(k);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e != hipSuccess) return fail(HLBM_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
  *ms = (double)t / reps;
  return HLBM_OK;
}

// Population placement (domains with 256 MiB of populations or more).  The
// same kernel on the same data runs up to ~10 % apart depending on where its
// population buffers land in HBM (measured per allocation; stable for the
// allocation's lifetime).  Up to `tries` fresh buffers are timed with the
// current launch shape; the state moves to one that is faster by more than 2 %.
// A/B: the step reads the current buffer and writes the other one, which the
// next step overwrites anyway, so both pairs are timed as they are.  AA: the
// step works in place, so the current lattice is saved first and restored
// after its timing, and candidates are timed on copies.
static int tune_placement(hlbm_domain* h, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int reps, int tries,
                          bool force, int* moved) {
  *moved = 0;
  const bool aa = h->d.pattern == HLBM_PATTERN_AA;
  const size_t fbytes = (size_t)h->stride * Q * h->esize;
  if (fbytes < ((size_t)256 << 20)) return HLBM_OK;
  const int nbuf = aa ? 1 : 2;
  auto copy = [&](void* dst, const void* src) {
    return hipMemcpyAsync(dst, src, fbytes, hipMemcpyDeviceToDevice, s) == hipSuccess;
  };
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < 2 * fbytes + ((size_t)4 << 30)) return HLBM_OK;
  void* save = nullptr;  // AA: the state while the lattice itself is timed
  if (aa) {
    if (hipMalloc(&save, fbytes) != hipSuccess) return HLBM_OK;
    if (!copy(save, h->buf[h->cur])) {
      (void)hipFree(save);
      return fail(HLBM_ERR_HIP, "autotune: placement copy");
    }
  }
  // rejected candidates stay allocated until the search ends, so that each
  // try is handed a region of its own rather than the one just freed
  std::vector<void*> rejected;
  double best = 0.0;
  int rc = time_shape(h, aa ? h->buf[h->cur] : nullptr, reps, s, e0, e1, &best);
  if (rc == HLBM_OK && aa && (!copy(h->buf[h->cur], save) || hipStreamSynchronize(s) != hipSuccess))
    rc = fail(HLBM_ERR_HIP, "autotune: placement restore");
  for (int t = 0; t < tries && rc == HLBM_OK; t++) {
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < (size_t)(nbuf + 1) * fbytes + ((size_t)4 << 30)) break;
    void* nb[2] = {nullptr, nullptr};
    bool ok = true;
    for (int k = 0; k < nbuf && ok; k++) ok = hipMalloc(&nb[k], fbytes) == hipSuccess;
    if (!ok) {
      for (void* q : nb)
        if (q) (void)hipFree(q);
      break;
    }
    // A/B: either buffer may hold state, depending on the phase
    for (int k = 0; k < nbuf && ok; k++) ok = copy(nb[aa ? 0 : k], aa ? save : h->buf[k]);
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
      for (void* q : nb)
        if (q) (void)hipFree(q);
      rc = fail(HLBM_ERR_HIP, "autotune: placement copy");
      break;
    }
    double ms = 0.0;
    void* old[2] = {h->buf[0], h->buf[1]};
    if (aa) {
      rc = time_shape(h, nb[0], reps, s, e0, e1, &ms);
      if (rc == HLBM_OK && (!copy(nb[0], save) || hipStreamSynchronize(s) != hipSuccess))
        rc = fail(HLBM_ERR_HIP, "autotune: placement restore");
    } else {
      h->buf[0] = nb[0];
      h->buf[1] = nb[1];
      rc = time_shape(h, nullptr, reps, s, e0, e1, &ms);
      h->buf[0] = old[0];
      h->buf[1] = old[1];
    }
    if (rc == HLBM_OK && (force || ms < 0.98 * best)) {
      best = ms;
      (*moved)++;
      if (aa) {
        (void)hipFree(h->buf[h->cur]);
        h->buf[h->cur] = nb[0];
      } else {
        (void)hipFree(old[0]);
        (void)hipFree(old[1]);
        h->buf[0] = nb[0];
        h->buf[1] = nb[1];
      }
    } else {
      for (void* q : nb)
        if (q) rejected.push_back(q);
    }
  }
  for (void* q : rejected) (void)hipFree(q);
  if (save) (void)hipFree(save);
  return rc;
}

static int pack_order(const hlbm_domain* h) { return 100 * h->lay.ord[0] + 10 * h->lay.ord[1] + h->lay.ord[2]; }

// Node layouts worth trying: the current one, and for every axis that holds
// planes of the boundary program, the one that puts that axis last (its planes
// become contiguous); the fastest axis must fill at least half a wave.
static std::vector<std::array<int, 3>> layout_candidates(const hlbm_domain* h) {
  std::vector<std::array<int, 3>> c{{h->lay.ord[0], h->lay.ord[1], h->lay.ord[2]}};
  const int n[3] = {h->nx, h->ny, h->nz};
  for (int a = 0; a < 3; a++) {
    const int bits = a == 2 ? 3 : (a == 1 ? 12 : 48);
    if (!(h->plan.prog_planes & bits)) continue;
    std::array<int, 3> o{};
    int k = 0;
    for (int b = 0; b < 3; b++)
      if (b != a) o[k++] = b;
    o[2] = a;
    if (n[o[0]] < 32) continue;
    if (std::find(c.begin(), c.end(), o) == c.end()) c.push_back(o);
  }
  return c;
}

// Time each candidate layout with the default launch shape and keep the fastest.
static int choose_layout(hlbm_domain* h, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const auto cand = layout_candidates(h);
  if (cand.size() < 2) return HLBM_OK;
  const dim3 keep_block = h->block;
  const int keep_npt = h->npt, keep_remap = h->xcd_remap, keep_variant = h->variant;
  hlbm_set_launch(h, 128, 2, 1, 1, 0);
  const bool aa = h->d.pattern == HLBM_PATTERN_AA;
  const size_t fbytes = (size_t)h->stride * Q * h->esize;
  double best = 1e300;
  std::array<int, 3> pick = cand[0];
  int rc = HLBM_OK;
  for (const auto& o : cand) {
    rc = hlbm_set_layout(h, o.data());
    if (rc) break;
    void* scratch = nullptr;  // AA steps work in place: time them on a copy
    if (aa) {
      if (hipMalloc(&scratch, fbytes) != hipSuccess) {
        rc = fail(HLBM_ERR_NOMEM, "autotune scratch populations");
        break;
      }
      if (hipMemcpyAsync(scratch, h->buf[h->cur], fbytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
        (void)hipFree(scratch);
        rc = fail(HLBM_ERR_HIP, "autotune scratch copy");
        break;
      }
    }
    double ms = 0.0;
    rc = time_shape(h, scratch, reps, s, e0, e1, &ms);
    if (scratch) (void)hipFree(scratch);
    if (rc) break;
    if (ms < best) {
      best = ms;
      pick = o;
    }
  }
  if (rc == HLBM_OK) rc = hlbm_set_layout(h, pick.data());
  h->block = keep_block;
  h->npt = keep_npt;
  h->xcd_remap = keep_remap;
  h->variant = keep_variant;
  return rc;
}

int hlbm_autotune(hlbm_domain* h, int32_t reps, void* stream, hlbm_tune_info* out) {
  if (!h) return fail(HLBM_ERR_ARG, "NULL domain");
  if (reps < 1) reps = 1;
  hipStream_t s = (hipStream_t)stream;
  if (h->any_wall && h->plan.wall_bb) {
    const int rc = ensure_bbm(h, s);
    if (rc) return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(HLBM_ERR_HIP, "autotune events");
  }
  auto done = [&](int rc) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  };
  // placement first: it belongs to this domain's allocation, not to its shape
  int moved = 0;
  {
    // HLBM_PLACEMENT: 0 = off, "force" = always move (exercises the relocation)
    const char* pe = getenv("HLBM_PLACEMENT");
    const bool force = pe && std::strcmp(pe, "force") == 0;
    if (!(pe && !force && atoi(pe) == 0)) {
      const int rc = tune_placement(h, s, e0, e1, reps, 3, force, &moved);
      if (rc) return done(rc);
    }
  }
  const TuneKey key{h->nx, h->ny, h->nz, h->d.storage, h->d.collision, h->d.pattern, h->plan.prog_planes,
                    h->any_wall && h->plan.wall_bb ? 1 : 0, h->device};
  {
    std::lock_guard<std::mutex> g(tune_mu);
    auto it = tune_cache.find(key);
    if (it != tune_cache.end()) {
      const TuneVal v = it->second;
      const int rc = hlbm_set_layout(h, v.order);
      if (rc) return done(rc);
      hlbm_set_launch(h, v.bx, v.by, v.npt, v.remap, v.variant);
      if (out)
        *out = hlbm_tune_info{v.bx, v.by, v.npt, v.remap, 0, v.variant, moved, pack_order(h), v.ms_best, v.ms_default};
      return done(HLBM_OK);
    }
  }
  {
    const int rc = choose_layout(h, reps, s, e0, e1);
    if (rc) return done(rc);
  }
  void* scratch = nullptr;
  if (h->d.pattern == HLBM_PATTERN_AA) {
    const size_t fbytes = (size_t)h->stride * Q * h->esize;
    if (hipMalloc(&scratch, fbytes) != hipSuccess) return done(fail(HLBM_ERR_NOMEM, "autotune scratch populations"));
    if (hipMemcpyAsync(scratch, h->buf[h->cur], fbytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      (void)hipFree(scratch);
      return done(fail(HLBM_ERR_HIP, "autotune scratch copy"));
    }
  }
  const dim3 keep_block = h->block;
  const int keep_npt = h->npt, keep_remap = h->xcd_remap, keep_variant = h->variant;
  static const int shapes[][2] = {{128, 2}, {64, 1}, {64, 2}, {64, 4}, {128, 1}, {256, 1}};
  TuneVal best{128, 2, 1, 1, 0, 1e300, 0.0, {h->lay.ord[0], h->lay.ord[1], h->lay.ord[2]}};
  int rc = HLBM_OK, n = 0;
  {
    // short steps: time enough of them that each candidate's measurement spans
    // about 5 ms (at most 500 steps), else launch jitter picks the shape
    double ms = 0.0;
    rc = time_shape(h, scratch, reps, s, e0, e1, &ms);
    if (rc == HLBM_OK && ms > 0.0 && reps * ms < 5.0) reps = std::min(500, (int)std::ceil(5.0 / ms));
  }
  for (int variant = 0; variant <= 1 && rc == HLBM_OK; variant++) {
    for (const auto& sh : shapes) {
      // a layout other than the default runs one node per thread
      for (int npt = 1; npt <= (variant || !default_layout_p(h) ? 1 : 2) && rc == HLBM_OK; npt++) {
        for (int remap = 1; remap >= 0 && rc == HLBM_OK; remap--) {
          hlbm_set_launch(h, sh[0], sh[1], npt, remap, variant);
          double ms = 0.0;
          rc = time_shape(h, scratch, reps, s, e0, e1, &ms);
          if (rc) break;
          n++;
          if (sh[0] == 128 && sh[1] == 2 && npt == 1 && remap == 1 && variant == 0) best.ms_default = ms;
          if (ms < best.ms_best)
            best = TuneVal{sh[0], sh[1], npt, remap, variant, ms, best.ms_default, {best.order[0], best.order[1],
                                                                                   best.order[2]}};
        }
      }
      if (rc) break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (scratch) (void)hipFree(scratch);
  if (rc) {
    h->block = keep_block;
    h->npt = keep_npt;
    h->xcd_remap = keep_remap;
    h->variant = keep_variant;
    return rc;
  }
  hlbm_set_launch(h, best.bx, best.by, best.npt, best.remap, best.variant);
  {
    std::lock_guard<std::mutex> g(tune_mu);
    tune_cache[key] = best;
  }
  if (out) *out = hlbm_tune_info{best.bx, best.by, best.npt, best.remap, n, best.variant, moved, pack_order(h), best.ms_best,
                                 best.ms_default};
  return HLBM_OK;
}

int hlbm_copy_bench(uint64_t bytes, int32_t reps, void* stream, double* ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return fail(HLBM_ERR_ARG, "bad copy benchmark arguments");
  bytes &= ~(uint64_t)15;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t s = (hipStream_t)stream;
  int rc = HLBM_OK;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    rc = fail(HLBM_ERR_NOMEM, "copy benchmark allocation failed");
  } else if (hipMemset(a, 1, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = fail(HLBM_ERR_HIP, "copy benchmark setup failed");
  } else {
    hipError_t e = launch_copy(b, a, bytes, s);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy(b, a, bytes, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      rc = fail(HLBM_ERR_HIP, std::string("copy benchmark: ") + hipGetErrorString(e));
    else
      *ms_per_copy = (double)ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
