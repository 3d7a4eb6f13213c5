// capi.hip -- C-ABI of the gfx950 UVLM hot path (include/uvlm_gpu.h).
//
// A context owns one HIP stream (default priority), the device buffers of the
// stateless calls and the device-resident SimpleSimulator state.  Buffers
// grow on demand and are reused; host inputs of the simulation calls go
// through pinned staging buffers so the step stays asynchronous.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/uvlm_gpu.h"
#include "batch.h"
#include "farfield.h"
#include "internal.h"
#include "mlp.h"
#include "step.h"

using namespace uvlm;

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

struct Staging {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};

}  // namespace

struct uvlm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  CutOff cut{};  // core cut-off of the segment law (default: the reference's)
  // stateless scratch
  Buf t_a, t_b, t_c, t_out, t_recs, t_partial, t_A, t_x, t_ipiv, t_info, t_work;
  // simulation
  bool sim = false;
  int num = 0, R = 0, C = 0, N = 0, cap = 0, capg = 0;
  unsigned flags = 0;
  int wstart = 0, wrows = 0, gstart = 0, wgrows = 0;
  int max_steps = 0;
  Buf wing, gamma, gamma_prev, wake, wgam, cp, nrm, area, recs, chains, partial, A,
      rhs, ipiv, info, uw, vm, mid, uflow, vedge, pos, pst, k1, k2, k3, k4,
      vprev, loads, ref_area, work, tmp;
  bool has_vprev = false;
  int vprev_rows = 0;
  int loads_count = 0;
  double pairs_eval = 0.0, pairs_ref = 0.0;
  Staging st_wing, st_vm, st_ve, st_area;
  // measurement hook
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used, ev_free;
  std::vector<double> ev_pairs;
  // batched cases
  uvlm::BatchEngine* batch = nullptr;
  double batch_pairs = 0.0;
  // far-field split of the wake sum in the advection stages
  uvlm::FarField* ff = nullptr;
  uvlm::FarField* ffb = nullptr;  // the bound sheets as sources (wake nodes downstream)
  double ff_theta = 0.0, ff_tol = 0.0, ff_last_err = -1.0;
  double ff_reach = 0.0, ff_theta_min = 0.0;  // graded criterion (reach 0: off)
  int ff_audit_every = 0, ff_counter = 0, ff_audits = 0;
  bool ff_off = false;
  // the split's sources are rebuilt only when the wake changed
  long wake_version = 0, ff_version = -1;
  // the solve's split call also serves the loads' edge midpoints: wake
  // velocity there (ffm) for this wake and wing state
  long wing_version = 0, ffm_wake = -1, ffm_wing = -1;
  Buf wtg, wvel;
  Buf ffm;
  Buf ffg, ffv, ffx, ffxs, ffstat, ffbox, bnear, ffbnd;
  // kept-inverse solver (UVLM_SOLVE_F64_INV); one for the simulation, one
  // for the stateless uvlm_solve
  uvlm::InvSolver* inv = nullptr;
  uvlm::InvSolver* t_inv = nullptr;
  int inv_last_sweeps = 0;
  // refinement sweeps of the mixed-precision solves in the simulation
  long long ir_solves = 0, ir_sweeps = 0;
  int ir_last = 0;
  // Sources of the last stateless sum, packed on the device.  The host copy
  // detects an identical next call (same kind, record layout, shape and
  // bytes), which then skips the upload and the packing: a caller that
  // queries one lattice target by target pays for its targets only.
  struct SourceCache {
    int kind = -1;  // 0 lattice rings, 1 lattice chains, 2 free rings
    int layout = 0;
    size_t a = 0, b = 0, cc = 0;
    std::vector<double> nodes, gamma;
  } src;
  Buf src_recs;
  long long src_hits = 0, src_uploads = 0;
  // far-field correction network
  uvlm::MlpTrainer* mlp = nullptr;
  Buf mw, mx, mv, my, mg, mdev;
};

namespace {

int fail(uvlm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                    \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(ctx, UVLM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int ensure(uvlm_ctx* c, Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (bytes <= b.cap) return UVLM_OK;
  if (b.p) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = 2 * bytes;  // doubling: O(log) reallocations for a growing wake
  if (hipMalloc(&b.p, want) != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(c, UVLM_ERR_OOM, "device allocation of %zu bytes failed", want);
  }
  b.cap = want;
  return UVLM_OK;
}

void release(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

template <typename T>
T* P(Buf& b) {
  return reinterpret_cast<T*>(b.p);
}

// host -> device through a pinned staging buffer, stream-ordered
int stage_upload(uvlm_ctx* c, Staging& st, void* dst, const void* src,
                 size_t bytes) {
  if (bytes == 0) return UVLM_OK;
  if (st.pending) {
    HIPCHK(c, hipEventSynchronize(st.ev));
    st.pending = false;
  }
  if (bytes > st.cap) {
    if (st.host) HIPCHK(c, hipHostFree(st.host));
    st.host = nullptr;
    st.cap = 0;
    HIPCHK(c, hipHostMalloc(&st.host, bytes, hipHostMallocDefault));
    st.cap = bytes;
  }
  if (!st.ev) HIPCHK(c, hipEventCreateWithFlags(&st.ev, hipEventDisableTiming));
  std::memcpy(st.host, src, bytes);
  HIPCHK(c, hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(st.ev, c->stream));
  st.pending = true;
  return UVLM_OK;
}

void free_staging(Staging& st) {
  if (st.pending && st.ev) (void)hipEventSynchronize(st.ev);
  if (st.host) (void)hipHostFree(st.host);
  if (st.ev) (void)hipEventDestroy(st.ev);
  st = Staging{};
}

int check_launch(uvlm_ctx* c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(c, UVLM_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
  return UVLM_OK;
}

int prec_of(unsigned flags) { return (int)(flags & 0x0f); }

#define TRY(expr)              \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != UVLM_OK) return rc_; \
  } while (0)

int set_dev(uvlm_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device));
  return UVLM_OK;
}

bool default_cut(const uvlm_ctx* c) { return c->cut.eps == kCutoff; }

bool src_same(const uvlm_ctx* c, int kind, int layout, size_t a, size_t b, size_t cc,
              const double* nodes, size_t nn, const double* gamma, size_t ng) {
  const auto& s = c->src;
  return s.kind == kind && s.layout == layout && s.a == a && s.b == b && s.cc == cc &&
         s.nodes.size() == nn && s.gamma.size() == ng &&
         std::memcmp(s.nodes.data(), nodes, sizeof(double) * nn) == 0 &&
         std::memcmp(s.gamma.data(), gamma, sizeof(double) * ng) == 0;
}
void src_store(uvlm_ctx* c, int kind, int layout, size_t a, size_t b, size_t cc,
               const double* nodes, size_t nn, const double* gamma, size_t ng) {
  auto& s = c->src;
  s.kind = kind;
  s.layout = layout;
  s.a = a;
  s.b = b;
  s.cc = cc;
  s.nodes.assign(nodes, nodes + nn);
  s.gamma.assign(gamma, gamma + ng);
}

int check_prec(uvlm_ctx* c, unsigned flags) {
  const int p = prec_of(flags);
  if (p != kF64Exact && p != kF64Fast && p != kF32C)
    return fail(c, UVLM_ERR_ARG, "unsupported precision flag %u", flags & 0x0fu);
  if (p == kF32C && !default_cut(c))
    return fail(c, UVLM_ERR_ARG, "UVLM_PREC_F32C runs with the default cut-off only");
  return UVLM_OK;
}

// Launch the N-body sum of targets against recs [0, nb+nw) and return the
// plan (partial sums in c->partial or b).
int nbody(uvlm_ctx* c, Buf& partial, const double* targets, int nt,
          const RingRec* recs, int nb, int nw, unsigned flags, int seg,
          ChunkPlan* plan_out) {
  ChunkPlan plan = plan_chunks(nt, nb, nw, (flags & UVLM_SEQUENTIAL) != 0);
  TRY(ensure(c, partial, sizeof(double) * 3 * (size_t)std::max(1, plan.total()) *
                             (size_t)std::max(1, nt)));
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (c->prof && nt > 0 && nb + nw > 0) {
    if (!c->ev_free.empty()) {
      ev = c->ev_free.back();
      c->ev_free.pop_back();
    } else {
      HIPCHK(c, hipEventCreate(&ev.first));
      HIPCHK(c, hipEventCreate(&ev.second));
    }
    HIPCHK(c, hipEventRecord(ev.first, c->stream));
  }
  launch_nbody(c->stream, prec_of(flags), seg, targets, nt, recs, plan,
               P<double>(partial), nt, c->cut);
  if (ev.first) {
    HIPCHK(c, hipEventRecord(ev.second, c->stream));
    c->ev_used.push_back(ev);
    const double segs = (seg == kSegRingChord) ? 2.0 : 4.0;
    c->ev_pairs.push_back(segs * (double)nt * (double)(nb + nw));
  }
  *plan_out = plan;
  return check_launch(c);
}

// N-body of targets against the simulation's sources (bound and/or wake):
// ring records in F64_EXACT, segment chains in the fast modes.  Counts the
// evaluated segment pairs and returns the chunk plan for the combine.
int sum_sources(uvlm_ctx* c, const double* targets, int nt, bool bound, bool wake,
                ChunkPlan* plan_out) {
  const int prec = prec_of(c->flags);
  const bool seq = (c->flags & UVLM_SEQUENTIAL) != 0;
  if (prec == kF64Exact) {
    const int N = c->N, Nw = c->num * c->wgrows * c->C;
    const RingRec* base = P<RingRec>(c->recs) + (bound ? 0 : N);
    TRY(nbody(c, c->partial, targets, nt, base, bound ? N : 0, wake ? Nw : 0, c->flags,
              kSegVoring, plan_out));
    c->pairs_eval += 4.0 * nt * (double)((bound ? N : 0) + (wake ? Nw : 0));
    return UVLM_OK;
  }
  const int eb = chain_elems(c->num, c->R, c->C);
  const int ew = chain_elems(c->num, c->wgrows, c->C);
  const int nb = bound ? eb : 0, nw = wake ? ew : 0;
  ChunkPlan plan = plan_chunks(nt, nb, nw, seq, chain_slots(prec));
  TRY(ensure(c, c->partial, sizeof(double) * 3 * (size_t)std::max(1, plan.total()) *
                                (size_t)std::max(1, nt)));
  const double segs = (bound ? chain_segments(c->num, c->R, c->C) : 0.0) +
                      (wake ? chain_segments(c->num, c->wgrows, c->C) : 0.0);
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (c->prof && nt > 0 && nb + nw > 0) {
    if (!c->ev_free.empty()) {
      ev = c->ev_free.back();
      c->ev_free.pop_back();
    } else {
      HIPCHK(c, hipEventCreate(&ev.first));
      HIPCHK(c, hipEventCreate(&ev.second));
    }
    HIPCHK(c, hipEventRecord(ev.first, c->stream));
  }
  launch_nbody_chain(c->stream, prec, targets, nt,
                     P<ChainRec>(c->chains) + (bound ? 0 : eb), plan,
                     P<double>(c->partial), nt, c->cut);
  if (ev.first) {
    HIPCHK(c, hipEventRecord(ev.second, c->stream));
    c->ev_used.push_back(ev);
    c->ev_pairs.push_back(segs * nt);
  }
  c->pairs_eval += segs * nt;
  *plan_out = plan;
  return check_launch(c);
}

// The wing's few targets (collocation points for the RHS, edge midpoints
// for the loads) go through the split with a tighter criterion than the
// wake's own nodes: their velocities set the circulations and the loads
// directly, and they cost little.
constexpr double kWingThetaFactor = 2.0;
// Tags of the wings' target sets in the split (their Morton order is kept
// between steps; the wings' points keep their arrangement as they move).
constexpr int kOrderWingTargets = 1, kOrderEdgeMidpoints = 2;
// Audits of wakes of at least kAuditFull nodes sum every stride-th node
// exactly, stride = max(kAuditStride, nodes / kAuditNodes): at most an
// eighth of the cost of the exact stage, and ~kAuditNodes sums for the
// largest wakes.
constexpr int kAuditFull = 32768, kAuditStride = 8, kAuditNodes = 8192;
// With the split on, the bound sheets act on the wake's nodes exactly in
// the first kBoundRows node rows of each wake sheet (the wake at the wings)
// and through each wing's moment expansion beyond, where a node is farther
// than kBoundTheta wing radii from the wing's centre (else that wing is
// summed exactly at the node): relative error ~kBoundTheta^-4 of the bound
// sheets' velocity at nodes that move wake acting on the wings like a
// dipole.
constexpr int kBoundRows = 24;
constexpr double kBoundTheta = 4.0;

// The near/far split applies: set, not switched off by an audit, a wake.
bool ff_active(const uvlm_ctx* c) {
  return c->ff && c->ff_theta > 0.0 && !c->ff_off && c->wgrows > 0;
}

// Far-field sources from the current wake (positions gathered into pos,
// circulations into ffg) unless already built for this wake state.
int ff_current(uvlm_ctx* c) {
  c->ff->set_f32c(prec_of(c->flags) == kF32C);
  if (c->ff_version == c->wake_version) return UVLM_OK;
  const int C1 = c->C + 1, rows = c->wrows, V = c->num * rows * C1;
  TRY(ensure(c, c->pos, sizeof(double) * 3 * (size_t)V));
  TRY(ensure(c, c->ffg, sizeof(double) * (size_t)c->num * c->wgrows * c->C));
  launch_gather_rows(c->stream, P<double>(c->wake), c->cap, c->wstart, c->num, rows, C1,
                     P<double>(c->pos));
  launch_gather_rows_w(c->stream, P<double>(c->wgam), c->capg, c->gstart, c->num, c->wgrows,
                       c->C, P<double>(c->ffg));
  std::string e;
  const int rc = c->ff->build(c->stream, P<double>(c->pos), P<double>(c->ffg), c->num,
                              c->wgrows, c->C, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  c->ff_version = c->wake_version;
  return UVLM_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

// ---------------------------------------------------------------------------
// Scratch of uvlm_induced_velocity_noalloc: a file-scope device array (each
// device holds its own copy until the process ends).  Host code never uses
// its name as an address: the device address of each device's copy is looked
// up once per device (hipGetSymbolAddress, with that device current) and
// kept in g_scratch_addr.
constexpr size_t kScratchRecBytes = size_t(16) << 20;   // source records
constexpr size_t kScratchPartBytes = size_t(48) << 20;  // chunk partial sums
constexpr size_t kScratchBytes = kScratchRecBytes + kScratchPartBytes;
static __device__ __align__(256) unsigned char g_scratch[kScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchDevices] = {};

// Device address of this device's scratch (device `dev` current); looked up
// on the first call per device.
int scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kScratchDevices) return UVLM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return UVLM_ERR_HIP;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return UVLM_OK;
}
}  // namespace

int uvlm_create(int device, uvlm_ctx** out) {
  if (!out) return UVLM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return UVLM_ERR_HIP;
  }
  if (device < 0 || device >= n) return UVLM_ERR_ARG;
  uvlm_ctx* c = new (std::nothrow) uvlm_ctx();
  if (!c) return UVLM_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return UVLM_ERR_HIP;
  }
  c->cut = make_cutoff(kCutoff);
  void* scratch = nullptr;  // set-up of uvlm_induced_velocity_noalloc's scratch
  if (scratch_address(device, &scratch) != UVLM_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return UVLM_ERR_HIP;
  }
  *out = c;
  return UVLM_OK;
}

int uvlm_destroy(uvlm_ctx* c) {
  if (!c) return UVLM_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  delete c->batch;
  c->batch = nullptr;
  delete c->ff;
  c->ff = nullptr;
  delete c->ffb;
  c->ffb = nullptr;
  delete c->inv;
  c->inv = nullptr;
  delete c->t_inv;
  c->t_inv = nullptr;
  delete c->mlp;
  c->mlp = nullptr;
  Buf* bufs[] = {&c->t_a, &c->t_b, &c->t_c, &c->t_out, &c->t_recs, &c->t_partial,
                 &c->t_A, &c->t_x, &c->t_ipiv, &c->t_info, &c->t_work, &c->wing,
                 &c->gamma, &c->gamma_prev, &c->wake, &c->wgam, &c->cp, &c->nrm,
                 &c->area, &c->recs, &c->chains, &c->partial, &c->A, &c->rhs, &c->ipiv,
                 &c->info, &c->uw, &c->vm, &c->mid, &c->uflow, &c->vedge, &c->pos,
                 &c->pst, &c->k1, &c->k2, &c->k3, &c->k4, &c->vprev, &c->loads,
                 &c->ref_area, &c->work, &c->tmp, &c->ffg, &c->ffv, &c->ffx, &c->ffstat,
                 &c->mw, &c->mx, &c->mv, &c->my, &c->mg, &c->mdev, &c->ffm, &c->ffbox, &c->ffxs};
  for (Buf* b : bufs) release(*b);
  free_staging(c->st_wing);
  free_staging(c->st_vm);
  free_staging(c->st_ve);
  free_staging(c->st_area);
  for (auto& e : c->ev_used) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto& e : c->ev_free) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return UVLM_OK;
}

const char* uvlm_last_error(const uvlm_ctx* c) {
  return c ? c->err.c_str() : "no context";
}

int uvlm_set_cutoff(uvlm_ctx* c, double eps) {
  if (!c) return UVLM_ERR_ARG;
  if (!(eps > 0.0) || !(eps < 1e-2))
    return fail(c, UVLM_ERR_ARG, "cut-off %g outside (0, 1e-2)", eps);
  if (eps != kCutoff) {
    if (c->sim && prec_of(c->flags) == kF32C)
      return fail(c, UVLM_ERR_ARG, "UVLM_PREC_F32C runs with the default cut-off only");
    if (c->ff && c->ff_theta > 0.0)
      return fail(c, UVLM_ERR_ARG, "the far-field split runs with the default cut-off only");
    if (c->batch)
      return fail(c, UVLM_ERR_ARG, "batches run with the default cut-off only");
  }
  c->cut = make_cutoff(eps);
  return UVLM_OK;
}

int uvlm_source_cache_stats(uvlm_ctx* c, long long* uploads, long long* hits) {
  if (!c) return UVLM_ERR_ARG;
  if (uploads) *uploads = c->src_uploads;
  if (hits) *hits = c->src_hits;
  return UVLM_OK;
}

int uvlm_sync(uvlm_ctx* c) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

// ---------------------------------------------------------------------------
int uvlm_biot_savart(uvlm_ctx* c, const double* start, const double* end,
                     const double* pos, size_t n, unsigned flags, double* out) {
  if (!c || (n && (!start || !end || !pos || !out))) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  TRY(check_prec(c, flags));
  if (n == 0) return UVLM_OK;
  const size_t bytes = sizeof(double) * 3 * n;
  TRY(ensure(c, c->t_a, bytes));
  TRY(ensure(c, c->t_b, bytes));
  TRY(ensure(c, c->t_c, bytes));
  TRY(ensure(c, c->t_out, bytes));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, start, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_b.p, end, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_c.p, pos, bytes, hipMemcpyHostToDevice, c->stream));
  launch_segments(c->stream, prec_of(flags) == kF64Exact ? kF64Exact : kF64Fast,
                  P<double>(c->t_a), P<double>(c->t_b), P<double>(c->t_c), (int)n,
                  c->cut, P<double>(c->t_out));
  TRY(check_launch(c));
  HIPCHK(c, hipMemcpyAsync(out, c->t_out.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_induced_velocity(uvlm_ctx* c, const double* targets, size_t nt,
                          const double* nodes, const double* gamma, size_t num,
                          size_t rows, size_t cols, unsigned flags, double* out) {
  if (!c) return UVLM_ERR_ARG;
  if (nt && (!targets || !out)) return fail(c, UVLM_ERR_ARG, "null target/out");
  TRY(set_dev(c));
  TRY(check_prec(c, flags));
  if (nt > (size_t)INT32_MAX / 4) return fail(c, UVLM_ERR_ARG, "too many targets");
  const size_t nr = num * rows * cols;
  if (nt == 0) return UVLM_OK;
  if (nr == 0) {
    std::memset(out, 0, sizeof(double) * 3 * nt);
    return UVLM_OK;
  }
  if (!nodes || !gamma) return fail(c, UVLM_ERR_ARG, "null nodes/gamma");
  const size_t nn = num * (rows + 1) * (cols + 1);
  const int prec = prec_of(flags);
  const int kind = prec == kF64Exact ? 0 : 1;
  const int ne = kind ? chain_elems((int)num, (int)rows, (int)cols) : 0;
  TRY(ensure(c, c->t_a, sizeof(double) * 3 * nt));
  TRY(ensure(c, c->t_out, sizeof(double) * 3 * nt));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, targets, sizeof(double) * 3 * nt,
                           hipMemcpyHostToDevice, c->stream));
  if (src_same(c, kind, prec, num, rows, cols, nodes, 3 * nn, gamma, nr)) {
    c->src_hits += 1;
  } else {
    c->src.kind = -1;  // invalid until the new set is packed
    TRY(ensure(c, c->t_b, sizeof(double) * 3 * nn));
    TRY(ensure(c, c->t_c, sizeof(double) * nr));
    TRY(ensure(c, c->src_recs, kind ? sizeof(ChainRec) * ne : sizeof(RingRec) * nr));
    HIPCHK(c, hipMemcpyAsync(c->t_b.p, nodes, sizeof(double) * 3 * nn,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->t_c.p, gamma, sizeof(double) * nr,
                             hipMemcpyHostToDevice, c->stream));
    if (kind == 0)
      launch_pack_lattice(c->stream, P<double>(c->t_b), (int)rows + 1, 0,
                          P<double>(c->t_c), (int)rows, 0, (int)num, (int)rows,
                          (int)cols, P<RingRec>(c->src_recs));
    else
      launch_pack_chains(c->stream, P<double>(c->t_b), (int)rows + 1, 0, P<double>(c->t_c),
                         (int)rows, 0, (int)num, (int)rows, (int)cols,
                         P<ChainRec>(c->src_recs), prec == kF32C);
    TRY(check_launch(c));
    src_store(c, kind, prec, num, rows, cols, nodes, 3 * nn, gamma, nr);
    c->src_uploads += 1;
  }
  ChunkPlan plan;
  if (kind == 0) {
    TRY(nbody(c, c->t_partial, P<double>(c->t_a), (int)nt, P<RingRec>(c->src_recs), 0,
              (int)nr, flags, kSegVoring, &plan));
  } else {
    plan = plan_chunks((int)nt, 0, ne, (flags & UVLM_SEQUENTIAL) != 0);
    TRY(ensure(c, c->t_partial,
               sizeof(double) * 3 * (size_t)std::max(1, plan.total()) * nt));
    launch_nbody_chain(c->stream, prec, P<double>(c->t_a), (int)nt,
                       P<ChainRec>(c->src_recs), plan, P<double>(c->t_partial), (int)nt,
                       c->cut);
    TRY(check_launch(c));
  }
  launch_sum_chunks(c->stream, P<double>(c->t_partial), (int)nt, (int)nt,
                    plan.total(), P<double>(c->t_out));
  TRY(check_launch(c));
  HIPCHK(c, hipMemcpyAsync(out, c->t_out.p, sizeof(double) * 3 * nt,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_induced_velocity_noalloc(uvlm_ctx* c, const double* targets, size_t nt,
                                  const double* nodes, const double* gamma, size_t num,
                                  size_t rows, size_t cols, unsigned flags, double* out,
                                  void* stream) {
  if (!c) return UVLM_ERR_ARG;
  if (nt && (!targets || !out)) return fail(c, UVLM_ERR_ARG, "null target/out");
  TRY(check_prec(c, flags));
  if (nt == 0) return UVLM_OK;
  const size_t nr = num * rows * cols;
  if (nr && (!nodes || !gamma)) return fail(c, UVLM_ERR_ARG, "null nodes/gamma");
  const int prec = prec_of(flags);
  const bool seq = (flags & UVLM_SEQUENTIAL) != 0;
  // sizes first: nothing is launched for a call that does not fit
  const size_t rec_bytes = prec == kF64Exact
                               ? sizeof(RingRec) * nr
                               : sizeof(ChainRec) * (size_t)2 * num * (rows + 1) * (cols + 1);
  if (nt > (size_t)INT32_MAX / 4 || nr > (size_t)INT32_MAX / 4 || rec_bytes > kScratchRecBytes)
    return fail(c, UVLM_ERR_CAPACITY,
                "noalloc: %zu source records (%zu bytes) exceed the %zu-byte scratch", nr,
                rec_bytes, kScratchRecBytes);
  ChunkPlan plan{};
  int ne = 0;
  if (nr) {
    ne = prec == kF64Exact ? (int)nr : chain_elems((int)num, (int)rows, (int)cols);
    plan = plan_chunks((int)nt, 0, ne, seq);
  }
  const size_t part_bytes = sizeof(double) * 3 * (size_t)std::max(1, plan.total()) * nt;
  if (part_bytes > kScratchPartBytes)
    return fail(c, UVLM_ERR_CAPACITY,
                "noalloc: %zu targets x %d chunks need %zu bytes of partial sums (at most %zu)",
                nt, plan.total(), part_bytes, kScratchPartBytes);
  TRY(set_dev(c));
  void* scratch = nullptr;
  if (scratch_address(c->device, &scratch) != UVLM_OK)
    return fail(c, UVLM_ERR_HIP, "noalloc: scratch address lookup failed");
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
  void* recs = scratch;
  double* part = reinterpret_cast<double*>(static_cast<unsigned char*>(scratch) +
                                           kScratchRecBytes);
  if (nr) {
    if (prec == kF64Exact) {
      launch_pack_lattice(s, nodes, (int)rows + 1, 0, gamma, (int)rows, 0, (int)num,
                          (int)rows, (int)cols, static_cast<RingRec*>(recs));
      launch_nbody(s, prec, kSegVoring, targets, (int)nt, static_cast<const RingRec*>(recs),
                   plan, part, (int)nt, c->cut);
    } else {
      launch_pack_chains(s, nodes, (int)rows + 1, 0, gamma, (int)rows, 0, (int)num, (int)rows,
                         (int)cols, static_cast<ChainRec*>(recs), prec == kF32C);
      launch_nbody_chain(s, prec, targets, (int)nt, static_cast<const ChainRec*>(recs), plan,
                         part, (int)nt, c->cut);
    }
  }
  // no sources: zero chunks sum to zero (a kernel, not a memset)
  launch_sum_chunks(s, part, (int)nt, (int)nt, nr ? plan.total() : 0, out);
  return check_launch(c);
}

int uvlm_ring_induced_velocity(uvlm_ctx* c, const double* targets, size_t nt,
                               const double* rings, const double* gamma,
                               size_t n_rings, unsigned flags, double* out) {
  if (!c) return UVLM_ERR_ARG;
  if (nt && (!targets || !out)) return fail(c, UVLM_ERR_ARG, "null target/out");
  TRY(set_dev(c));
  TRY(check_prec(c, flags));
  if (nt == 0) return UVLM_OK;
  if (n_rings == 0) {
    std::memset(out, 0, sizeof(double) * 3 * nt);
    return UVLM_OK;
  }
  if (!rings || !gamma) return fail(c, UVLM_ERR_ARG, "null rings/gamma");
  int seg = kSegRingPerSeg;
  if (flags & UVLM_CHORDWISE_ONLY)
    seg = kSegRingChord;
  else if (flags & UVLM_SCALE_PER_RING)
    seg = kSegRingPerRing;
  TRY(ensure(c, c->t_a, sizeof(double) * 3 * nt));
  TRY(ensure(c, c->t_out, sizeof(double) * 3 * nt));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, targets, sizeof(double) * 3 * nt,
                           hipMemcpyHostToDevice, c->stream));
  if (src_same(c, 2, 0, n_rings, 0, 0, rings, 12 * n_rings, gamma, n_rings)) {
    c->src_hits += 1;
  } else {
    c->src.kind = -1;
    TRY(ensure(c, c->t_b, sizeof(double) * 12 * n_rings));
    TRY(ensure(c, c->t_c, sizeof(double) * n_rings));
    TRY(ensure(c, c->src_recs, sizeof(RingRec) * n_rings));
    HIPCHK(c, hipMemcpyAsync(c->t_b.p, rings, sizeof(double) * 12 * n_rings,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->t_c.p, gamma, sizeof(double) * n_rings,
                             hipMemcpyHostToDevice, c->stream));
    launch_pack_rings(c->stream, P<double>(c->t_b), P<double>(c->t_c), (int)n_rings,
                      P<RingRec>(c->src_recs));
    TRY(check_launch(c));
    src_store(c, 2, 0, n_rings, 0, 0, rings, 12 * n_rings, gamma, n_rings);
    c->src_uploads += 1;
  }
  ChunkPlan plan;
  TRY(nbody(c, c->t_partial, P<double>(c->t_a), (int)nt, P<RingRec>(c->src_recs), 0,
            (int)n_rings, flags, seg, &plan));
  launch_sum_chunks(c->stream, P<double>(c->t_partial), (int)nt, (int)nt,
                    plan.total(), P<double>(c->t_out));
  TRY(check_launch(c));
  HIPCHK(c, hipMemcpyAsync(out, c->t_out.p, sizeof(double) * 3 * nt,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_influence_matrix(uvlm_ctx* c, const double* nodes, size_t num,
                          size_t rows, size_t cols, const double* cpos,
                          const double* normals, unsigned flags, double* A) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  TRY(check_prec(c, flags));
  const size_t N = num * rows * cols;
  if (N == 0) return UVLM_OK;
  if (!nodes || !cpos || !normals || !A) return fail(c, UVLM_ERR_ARG, "null argument");
  const size_t nn = num * (rows + 1) * (cols + 1);
  TRY(ensure(c, c->t_a, sizeof(double) * 3 * nn));
  TRY(ensure(c, c->t_b, sizeof(double) * 3 * N));
  TRY(ensure(c, c->t_c, sizeof(double) * 3 * N));
  TRY(ensure(c, c->t_recs, sizeof(RingRec) * N));
  TRY(ensure(c, c->t_A, sizeof(double) * N * N));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, nodes, sizeof(double) * 3 * nn,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_b.p, cpos, sizeof(double) * 3 * N,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_c.p, normals, sizeof(double) * 3 * N,
                           hipMemcpyHostToDevice, c->stream));
  launch_pack_lattice(c->stream, P<double>(c->t_a), (int)rows + 1, 0, nullptr, 0,
                      0, (int)num, (int)rows, (int)cols, P<RingRec>(c->t_recs));
  if (prec_of(flags) == kF64Exact)
    launch_matrix(c->stream, P<RingRec>(c->t_recs), P<double>(c->t_b), P<double>(c->t_c),
                  (int)N, c->cut, P<double>(c->t_A));
  else
    launch_matrix_fast(c->stream, P<RingRec>(c->t_recs), P<double>(c->t_b), P<double>(c->t_c),
                       (int)N, c->cut, P<double>(c->t_A));
  TRY(check_launch(c));
  // device matrix is column-major: A^T row-major == A column-major
  double* tmp = new (std::nothrow) double[N * N];
  if (!tmp) return fail(c, UVLM_ERR_OOM, "host buffer");
  hipError_t e = hipMemcpyAsync(tmp, c->t_A.p, sizeof(double) * N * N,
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    delete[] tmp;
    return fail(c, UVLM_ERR_HIP, "copy back: %s", hipGetErrorString(e));
  }
  for (size_t L = 0; L < N; ++L)
    for (size_t K = 0; K < N; ++K) A[K * N + L] = tmp[L * N + K];
  delete[] tmp;
  return UVLM_OK;
}

int uvlm_ring_influence_matrix(uvlm_ctx* c, const double* rings, size_t n,
                               const double* targets, const double* normals, double* A) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  if (n == 0) return UVLM_OK;
  if (!rings || !targets || !normals || !A) return fail(c, UVLM_ERR_ARG, "null argument");
  if (n > 46000) return fail(c, UVLM_ERR_ARG, "too many rings");
  TRY(ensure(c, c->t_a, sizeof(double) * 12 * n));
  TRY(ensure(c, c->t_b, sizeof(double) * 3 * n));
  TRY(ensure(c, c->t_c, sizeof(double) * 3 * n));
  TRY(ensure(c, c->t_out, sizeof(double) * n));
  TRY(ensure(c, c->t_recs, sizeof(RingRec) * n));
  TRY(ensure(c, c->t_A, sizeof(double) * n * n));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, rings, sizeof(double) * 12 * n, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_b.p, targets, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_c.p, normals, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                           c->stream));
  // Gamma = 1 for every ring: the per-ring scale is then exact
  std::vector<double> ones(n, 1.0);
  HIPCHK(c, hipMemcpyAsync(c->t_out.p, ones.data(), sizeof(double) * n,
                           hipMemcpyHostToDevice, c->stream));
  launch_pack_rings(c->stream, P<double>(c->t_a), P<double>(c->t_out), (int)n,
                    P<RingRec>(c->t_recs));
  launch_ring_matrix(c->stream, P<RingRec>(c->t_recs), P<double>(c->t_b), P<double>(c->t_c),
                     (int)n, c->cut, P<double>(c->t_A));
  TRY(check_launch(c));
  std::vector<double> tmp(n * n);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), c->t_A.p, sizeof(double) * n * n,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t L = 0; L < n; ++L)
    for (size_t K = 0; K < n; ++K) A[K * n + L] = tmp[L * n + K];
  return UVLM_OK;
}

int uvlm_solve(uvlm_ctx* c, size_t n, const double* A, const double* rhs,
               double* x, int mode, double tol, int max_iter, int* iters_out) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  if (iters_out) *iters_out = 0;
  if (n == 0) return UVLM_OK;
  if (!A || !rhs || !x) return fail(c, UVLM_ERR_ARG, "null argument");
  if (n > 46000) return fail(c, UVLM_ERR_ARG, "matrix too large");
  TRY(ensure(c, c->t_A, sizeof(double) * n * n));
  TRY(ensure(c, c->t_x, sizeof(double) * 2 * n));
  TRY(ensure(c, c->t_ipiv, dense_lu_work_bytes((int)n)));
  TRY(ensure(c, c->t_info, sizeof(int) * 4));
  // row-major host A -> column-major device A: transpose on the host
  double* tmp = new (std::nothrow) double[n * n];
  if (!tmp) return fail(c, UVLM_ERR_OOM, "host buffer");
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) tmp[j * n + i] = A[i * n + j];
  hipError_t e = hipMemcpy(c->t_A.p, tmp, sizeof(double) * n * n, hipMemcpyHostToDevice);
  delete[] tmp;
  if (e != hipSuccess) return fail(c, UVLM_ERR_HIP, "upload: %s", hipGetErrorString(e));
  double* db = P<double>(c->t_x);
  double* dx = db + n;
  HIPCHK(c, hipMemcpyAsync(db, rhs, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (mode == UVLM_SOLVE_F64) {
    if (dense_lu_factor(c->stream, P<double>(c->t_A), (int)n, c->t_ipiv.p) < 0)
      return fail(c, UVLM_ERR_HIP, "LU factorisation: %s", hipGetErrorString(hipGetLastError()));
    dense_lu_solve(c->stream, P<double>(c->t_A), (int)n, c->t_ipiv.p, db, dx);
    TRY(check_launch(c));
    int info = 0;
    HIPCHK(c, hipMemcpyAsync(&info, dense_lu_info(c->t_ipiv.p), sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (info != 0) return fail(c, UVLM_ERR_SINGULAR, "Matrix is not invertible (zero pivot at column %d)", info - 1);
    return UVLM_OK;
  }
  if (mode == UVLM_SOLVE_F64_FULLPIV) {
    TRY(ensure(c, c->t_work, dense_fullpiv_work_bytes((int)n)));
    dense_fullpiv_solve(c->stream, P<double>(c->t_A), (int)n, db, dx, c->t_work.p,
                        c->t_work.cap, P<int>(c->t_info));
    TRY(check_launch(c));
    int rank = 0;
    HIPCHK(c, hipMemcpyAsync(&rank, c->t_info.p, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (iters_out) *iters_out = rank;
    return UVLM_OK;
  }
  if (mode == UVLM_SOLVE_F64_INV) {
    if (!c->t_inv && !(c->t_inv = new (std::nothrow) InvSolver()))
      return fail(c, UVLM_ERR_OOM, "host allocation");
    c->t_inv->invalidate();  // an unrelated matrix: build its inverse afresh
    int it = 0;
    std::string msg;
    const int rc = c->t_inv->solve(c->stream, P<double>(c->t_A), (int)n, db, dx, false,
                                   tol > 0 ? tol : 1e-14, max_iter > 0 ? max_iter : 30, &it,
                                   &msg);
    if (iters_out) *iters_out = it;
    if (rc != UVLM_OK) return fail(c, rc, "%s", msg.c_str());
    HIPCHK(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return UVLM_OK;
  }
  if (mode != UVLM_SOLVE_F32_IR && mode != UVLM_SOLVE_BF16_IR)
    return fail(c, UVLM_ERR_ARG, "unknown solve mode %d", mode);
  TRY(ensure(c, c->t_work, dense_mixed_work_bytes((int)n)));
  int it = 0;
  const int rc = dense_mixed_solve(c->stream, P<double>(c->t_A), (int)n, db, dx, mode,
                                   tol > 0 ? tol : 1e-14, max_iter > 0 ? max_iter : 30,
                                   c->t_work.p, c->t_work.cap, &it);
  TRY(check_launch(c));
  if (iters_out) *iters_out = it;
  HIPCHK(c, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (rc < 0) return fail(c, UVLM_ERR_ARG, "workspace too small");
  if (rc > 0) return fail(c, UVLM_ERR_NOCONV, "refinement did not converge in %d sweeps", it);
  return UVLM_OK;
}

// ---------------------------------------------------------------------------
int uvlm_sim_init(uvlm_ctx* c, size_t num, size_t rows, size_t cols,
                  size_t max_steps, unsigned flags) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  TRY(check_prec(c, flags));
  if (num == 0 || rows == 0 || cols == 0)
    return fail(c, UVLM_ERR_ARG, "empty lattice");
  if (cols & 1) return fail(c, UVLM_ERR_ARG, "columns must be even.");
  const size_t N = num * rows * cols;
  const size_t cap = max_steps + 2;
  const size_t nw_max = num * cap * cols;
  if (N > 46000 || nw_max + N > (size_t)INT32_MAX / 8)
    return fail(c, UVLM_ERR_ARG, "lattice too large");
  c->num = (int)num;
  c->R = (int)rows;
  c->C = (int)cols;
  c->N = (int)N;
  c->cap = (int)cap;
  c->capg = (int)cap;
  c->max_steps = (int)max_steps;
  c->flags = flags;
  c->wstart = c->cap;
  c->wrows = 0;
  c->gstart = c->capg;
  c->wgrows = 0;
  c->has_vprev = false;
  c->vprev_rows = 0;
  c->loads_count = 0;
  c->pairs_eval = c->pairs_ref = 0.0;
  c->wake_version += 1;
  if (c->inv) c->inv->invalidate();
  const size_t nn = num * (rows + 1) * (cols + 1);
  const size_t E = num * rows * (2 * cols + 1);
  const size_t V = num * cap * (cols + 1);
  TRY(ensure(c, c->wing, sizeof(double) * 3 * nn));
  TRY(ensure(c, c->gamma, sizeof(double) * N));
  TRY(ensure(c, c->gamma_prev, sizeof(double) * N));
  TRY(ensure(c, c->wake, sizeof(double) * 3 * V));
  TRY(ensure(c, c->wgam, sizeof(double) * num * cap * cols));
  TRY(ensure(c, c->cp, sizeof(double) * 3 * N));
  TRY(ensure(c, c->nrm, sizeof(double) * 3 * N));
  TRY(ensure(c, c->area, sizeof(double) * N));
  TRY(ensure(c, c->recs, sizeof(RingRec) * (N + nw_max)));
  if (prec_of(flags) != kF64Exact)
    TRY(ensure(c, c->chains, sizeof(ChainRec) * ((size_t)chain_elems((int)num, (int)rows, (int)cols) +
                                                 (size_t)chain_elems((int)num, (int)cap, (int)cols))));
  TRY(ensure(c, c->A, sizeof(double) * N * N));
  TRY(ensure(c, c->rhs, sizeof(double) * N));
  TRY(ensure(c, c->ipiv, dense_lu_work_bytes(N)));
  TRY(ensure(c, c->info, sizeof(int) * 4));
  TRY(ensure(c, c->uw, sizeof(double) * 3 * N));
  TRY(ensure(c, c->vm, sizeof(double) * 3 * N));
  TRY(ensure(c, c->mid, sizeof(double) * 3 * E));
  TRY(ensure(c, c->uflow, sizeof(double) * 3 * E));
  TRY(ensure(c, c->vedge, sizeof(double) * 3 * E));
  TRY(ensure(c, c->ref_area, sizeof(double) * num));
  TRY(ensure(c, c->loads, sizeof(double) * 2 * num * (max_steps + 2)));
  HIPCHK(c, hipMemsetAsync(c->gamma.p, 0, sizeof(double) * N, c->stream));
  HIPCHK(c, hipMemsetAsync(c->gamma_prev.p, 0, sizeof(double) * N, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sim = true;
  return UVLM_OK;
}

int uvlm_sim_set_wing(uvlm_ctx* c, const double* nodes) {
  if (!c || !nodes) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  const size_t nn = (size_t)c->num * (c->R + 1) * (c->C + 1);
  HIPCHK(c, hipMemcpyAsync(c->gamma_prev.p, c->gamma.p, sizeof(double) * c->N,
                           hipMemcpyDeviceToDevice, c->stream));
  TRY(stage_upload(c, c->st_wing, c->wing.p, nodes, sizeof(double) * 3 * nn));
  launch_pack_lattice(c->stream, P<double>(c->wing), c->R + 1, 0, P<double>(c->gamma),
                      c->R, 0, c->num, c->R, c->C, P<RingRec>(c->recs));
  c->wing_version += 1;
  return check_launch(c);
}

namespace {
// Move the live rows of a ring buffer to its end when no headroom is left.
int compact_rows(uvlm_ctx* c, Buf& buf, int width, int cap, int& start, int rows) {
  if (start > 0 || rows == 0) {
    if (rows == 0) start = cap;
    return UVLM_OK;
  }
  const size_t row_bytes = sizeof(double) * (size_t)width;
  const size_t bytes = row_bytes * (size_t)c->num * rows;
  TRY(ensure(c, c->tmp, bytes));
  for (int n = 0; n < c->num; ++n)
    HIPCHK(c, hipMemcpyAsync((char*)c->tmp.p + row_bytes * (size_t)n * rows,
                             (char*)buf.p + row_bytes * ((size_t)n * cap + start),
                             row_bytes * rows, hipMemcpyDeviceToDevice, c->stream));
  const int ns = cap - rows;
  for (int n = 0; n < c->num; ++n)
    HIPCHK(c, hipMemcpyAsync((char*)buf.p + row_bytes * ((size_t)n * cap + ns),
                             (char*)c->tmp.p + row_bytes * (size_t)n * rows,
                             row_bytes * rows, hipMemcpyDeviceToDevice, c->stream));
  start = ns;
  return UVLM_OK;
}

// Wake sources from nodes (ring buffer or a compact stage copy): ring
// records in F64_EXACT, segment chains in the fast modes.
void pack_wake(uvlm_ctx* c, const double* nodes, int sheet_rows, int row_off) {
  if (prec_of(c->flags) == kF64Exact)
    launch_pack_lattice(c->stream, nodes, sheet_rows, row_off, P<double>(c->wgam),
                        c->capg, c->gstart, c->num, c->wgrows, c->C,
                        P<RingRec>(c->recs) + c->N);
  else
    launch_pack_chains(c->stream, nodes, sheet_rows, row_off, P<double>(c->wgam),
                       c->capg, c->gstart, c->num, c->wgrows, c->C,
                       P<ChainRec>(c->chains) + chain_elems(c->num, c->R, c->C),
                       prec_of(c->flags) == kF32C);
}
}  // namespace

int uvlm_sim_shed(uvlm_ctx* c, size_t step) {
  if (!c) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  if (c->wrows + 1 > c->cap)
    return fail(c, UVLM_ERR_CAPACITY, "wake capacity of %d rows exhausted", c->cap);
  TRY(compact_rows(c, c->wake, 3 * (c->C + 1), c->cap, c->wstart, c->wrows));
  c->wstart -= 1;
  c->wrows += 1;
  int grow = -1;
  if (step > 0) {
    if (c->wgrows + 1 > c->capg)
      return fail(c, UVLM_ERR_CAPACITY, "wake capacity of %d rows exhausted", c->capg);
    TRY(compact_rows(c, c->wgam, c->C, c->capg, c->gstart, c->wgrows));
    c->gstart -= 1;
    c->wgrows += 1;
    grow = c->gstart;
  }
  if (c->wgrows + 1 != c->wrows)
    return fail(c, UVLM_ERR_STATE, "wake shape out of step: %d node rows, %d gamma rows",
                c->wrows, c->wgrows);
  launch_shed(c->stream, P<double>(c->wing), P<double>(c->gamma), c->num, c->R, c->C,
              P<double>(c->wake), c->cap, c->wstart, P<double>(c->wgam), c->capg, grow);
  pack_wake(c, P<double>(c->wake), c->cap, c->wstart);
  c->wake_version += 1;
  return check_launch(c);
}

int uvlm_sim_solve(uvlm_ctx* c, const double* vmorph_cp, const double ff[3],
                   int solve_mode, double* gamma_out) {
  if (!c || !vmorph_cp || !ff) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  const int N = c->N;
  const int Nw = c->num * c->wgrows * c->C;
  launch_geometry(c->stream, P<double>(c->wing), c->num, c->R, c->C, P<double>(c->cp),
                  P<double>(c->nrm), P<double>(c->area));
  TRY(stage_upload(c, c->st_vm, c->vm.p, vmorph_cp, sizeof(double) * 3 * N));
  // wake velocity at the collocation points (through the split when set;
  // one split call also takes the loads' edge midpoints: the wake does not
  // change between the solve and the loads)
  if (ff_active(c)) {
    TRY(ff_current(c));
    const int E = c->num * c->R * (2 * c->C + 1);
    const size_t nb = sizeof(double) * 3 * (size_t)(N + E);
    TRY(ensure(c, c->wtg, nb));
    TRY(ensure(c, c->wvel, nb));
    TRY(ensure(c, c->ffm, sizeof(double) * 3 * (size_t)E));
    HIPCHK(c, hipMemcpyAsync(c->wtg.p, c->cp.p, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice,
                             c->stream));
    launch_edge_midpoints(c->stream, P<double>(c->wing), c->num, c->R, c->C,
                          P<double>(c->wtg) + 3 * (size_t)N);
    std::string e;
    const int rc = c->ff->velocity(c->stream, P<double>(c->wtg), N + E,
                                   kWingThetaFactor * c->ff_theta, P<double>(c->wvel), &e,
                                   nullptr, true, kOrderWingTargets);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
    HIPCHK(c, hipMemcpyAsync(c->uw.p, c->wvel.p, sizeof(double) * 3 * N,
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->ffm.p, P<double>(c->wvel) + 3 * (size_t)N,
                             sizeof(double) * 3 * E, hipMemcpyDeviceToDevice, c->stream));
    c->ffm_wake = c->wake_version;
    c->ffm_wing = c->wing_version;
  } else {
    ChunkPlan plan;
    TRY(sum_sources(c, P<double>(c->cp), N, false, true, &plan));
    launch_sum_chunks(c->stream, P<double>(c->partial), N, N, plan.total(),
                      P<double>(c->uw));
  }
  launch_rhs(c->stream, P<double>(c->uw), P<double>(c->vm), P<double>(c->nrm), N, ff,
             P<double>(c->rhs));
  if (prec_of(c->flags) == kF64Exact)
    launch_matrix(c->stream, P<RingRec>(c->recs), P<double>(c->cp), P<double>(c->nrm), N,
                  c->cut, P<double>(c->A));
  else
    launch_matrix_fast(c->stream, P<RingRec>(c->recs), P<double>(c->cp), P<double>(c->nrm), N,
                       c->cut, P<double>(c->A));
  TRY(check_launch(c));
  c->pairs_eval += 4.0 * N * (double)N;
  c->pairs_ref += 4.0 * N * (double)N + 4.0 * N * (double)Nw;
  if (solve_mode == UVLM_SOLVE_F64) {
    if (dense_lu_factor(c->stream, P<double>(c->A), N, c->ipiv.p) < 0)
      return fail(c, UVLM_ERR_HIP, "LU factorisation: %s", hipGetErrorString(hipGetLastError()));
    dense_lu_solve(c->stream, P<double>(c->A), N, c->ipiv.p, P<double>(c->rhs),
                   P<double>(c->gamma));
  } else if (solve_mode == UVLM_SOLVE_F64_FULLPIV) {
    TRY(ensure(c, c->work, dense_fullpiv_work_bytes(N)));
    dense_fullpiv_solve(c->stream, P<double>(c->A), N, P<double>(c->rhs), P<double>(c->gamma),
                        c->work.p, c->work.cap, nullptr);
  } else if (solve_mode == UVLM_SOLVE_F32_IR || solve_mode == UVLM_SOLVE_BF16_IR) {
    TRY(ensure(c, c->work, dense_mixed_work_bytes(N)));
    int it = 0;
    const int rc = dense_mixed_solve(c->stream, P<double>(c->A), N, P<double>(c->rhs),
                                     P<double>(c->gamma), solve_mode, 1e-14, 30,
                                     c->work.p, c->work.cap, &it);
    if (rc < 0) return fail(c, UVLM_ERR_ARG, "workspace too small");
    if (rc > 0) return fail(c, UVLM_ERR_NOCONV, "refinement did not converge");
    c->ir_solves += 1;
    c->ir_sweeps += it;
    c->ir_last = it;
  } else if (solve_mode == UVLM_SOLVE_F64_INV) {
    // warm start from the previous step's circulations (zero after init)
    if (!c->inv && !(c->inv = new (std::nothrow) InvSolver()))
      return fail(c, UVLM_ERR_OOM, "host allocation");
    std::string msg;
    const int rc = c->inv->solve(c->stream, P<double>(c->A), N, P<double>(c->rhs),
                                 P<double>(c->gamma), true, 1e-14, 40, &c->inv_last_sweeps,
                                 &msg);
    if (rc != UVLM_OK) return fail(c, rc, "%s", msg.c_str());
  } else {
    return fail(c, UVLM_ERR_ARG, "unknown solve mode %d", solve_mode);
  }
  launch_set_gamma(c->stream, P<double>(c->gamma), N, P<RingRec>(c->recs));
  if (prec_of(c->flags) != kF64Exact)
    launch_pack_chains(c->stream, P<double>(c->wing), c->R + 1, 0, P<double>(c->gamma), c->R,
                       0, c->num, c->R, c->C, P<ChainRec>(c->chains),
                       prec_of(c->flags) == kF32C);
  TRY(check_launch(c));
  if (gamma_out) {
    int info = 0;
    if (solve_mode == UVLM_SOLVE_F64)
      HIPCHK(c, hipMemcpyAsync(&info, dense_lu_info(c->ipiv.p), sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHK(c, hipMemcpyAsync(gamma_out, c->gamma.p, sizeof(double) * N,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (info != 0)
      return fail(c, UVLM_ERR_SINGULAR, "Matrix is not invertible (zero pivot at column %d)",
                  info - 1);
  }
  return UVLM_OK;
}

int uvlm_sim_loads(uvlm_ctx* c, const double* vmorph_edges, const double ff[3],
                   double dt, const double* ref_area) {
  if (!c || !vmorph_edges || !ff || !ref_area) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  if (c->loads_count >= c->max_steps + 2)
    return fail(c, UVLM_ERR_CAPACITY, "load history full");
  const int N = c->N;
  const int Nw = c->num * c->wgrows * c->C;
  const int E = c->num * c->R * (2 * c->C + 1);
  launch_edge_midpoints(c->stream, P<double>(c->wing), c->num, c->R, c->C,
                        P<double>(c->mid));
  TRY(stage_upload(c, c->st_ve, c->vedge.p, vmorph_edges, sizeof(double) * 3 * E));
  TRY(stage_upload(c, c->st_area, c->ref_area.p, ref_area, sizeof(double) * c->num));
  ChunkPlan plan;
  if (ff_active(c)) {  // bound sheets exactly, the wake through the split
    TRY(sum_sources(c, P<double>(c->mid), E, true, false, &plan));
    if (c->ffm_wake != c->wake_version || c->ffm_wing != c->wing_version) {
      TRY(ff_current(c));
      TRY(ensure(c, c->ffm, sizeof(double) * 3 * (size_t)E));
      std::string e;
      const int rc = c->ff->velocity(c->stream, P<double>(c->mid), E,
                                     kWingThetaFactor * c->ff_theta, P<double>(c->ffm), &e,
                                     nullptr, true, kOrderEdgeMidpoints);
      if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
    }
    launch_ff_combine(c->stream, P<double>(c->partial), E, plan.cb, P<double>(c->ffm), E, ff, 0,
                      0, P<double>(c->uflow));
  } else {
    TRY(sum_sources(c, P<double>(c->mid), E, true, true, &plan));
    launch_flow_velocity(c->stream, P<double>(c->partial), E, E, plan.cb, plan.cw, ff,
                         P<double>(c->uflow), 0, 0);
  }
  const double Q = std::sqrt((ff[0] * ff[0] + ff[1] * ff[1]) + ff[2] * ff[2]);
  launch_loads(c->stream, P<double>(c->wing), P<double>(c->gamma),
               P<double>(c->gamma_prev), P<double>(c->nrm), P<double>(c->area),
               P<double>(c->uflow), P<double>(c->vedge), c->num, c->R, c->C, dt, Q,
               P<double>(c->ref_area),
               P<double>(c->loads) + (size_t)2 * c->num * c->loads_count);
  TRY(check_launch(c));
  c->loads_count += 1;
  const double L = (double)c->num * (4.0 * c->R * c->C - c->C);
  c->pairs_ref += 4.0 * L * (double)(N + Nw);
  return UVLM_OK;
}

int uvlm_sim_advect(uvlm_ctx* c, int scheme, size_t ignore_steps, const double ff[3],
                    double dt) {
  if (!c || !ff) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  if (scheme < 0 || scheme > 4) return fail(c, UVLM_ERR_ARG, "unknown scheme %d", scheme);
  TRY(set_dev(c));
  const int N = c->N;
  const int Nw = c->num * c->wgrows * c->C;
  const int C1 = c->C + 1;
  const int rows = c->wrows;
  const int V = c->num * rows * C1;
  if (V == 0) return UVLM_OK;
  const size_t vb = sizeof(double) * 3 * (size_t)V;
  TRY(ensure(c, c->pos, vb));
  TRY(ensure(c, c->pst, vb));
  TRY(ensure(c, c->k1, vb));
  launch_gather_rows(c->stream, P<double>(c->wake), c->cap, c->wstart, c->num, rows, C1,
                     P<double>(c->pos));
  const int te_period = rows * C1;
  auto exact_stage = [&](const double* targets, double* out) -> int {
    ChunkPlan plan;
    TRY(sum_sources(c, targets, V, true, true, &plan));
    launch_flow_velocity(c->stream, P<double>(c->partial), V, V, plan.cb, plan.cw, ff, out,
                         te_period, C1);
    return check_launch(c);
  };
  // Far-field stage: bound sheets exactly, the wake (whose nodes are the
  // targets) through the near/far split; every ff_audit_every steps the
  // first stage is also summed exactly, compared, and the split switched
  // off when the difference exceeds ff_tol of the largest velocity.
  const bool use_ff = ff_active(c);
  bool audit = false;
  if (use_ff) {
    audit = c->ff_audit_every > 0 && (c->ff_counter % c->ff_audit_every) == 0;
    ++c->ff_counter;
    TRY(ensure(c, c->ffg, sizeof(double) * (size_t)c->num * c->wgrows * c->C));
    TRY(ensure(c, c->ffv, vb));
    TRY(ensure(c, c->ffstat, 2 * sizeof(unsigned long long)));
    if (audit) TRY(ensure(c, c->ffx, vb));
    launch_gather_rows_w(c->stream, P<double>(c->wgam), c->capg, c->gstart, c->num, c->wgrows,
                         c->C, P<double>(c->ffg));
  }
  FFFocus focus;
  if (use_ff && c->ff_reach > 0.0) {  // bounding box of the wings' nodes
    TRY(ensure(c, c->ffbox, 6 * sizeof(double)));
    launch_ff_bbox(c->stream, P<double>(c->wing), c->num * (c->R + 1) * (c->C + 1),
                   P<double>(c->ffbox));
    focus.box = P<double>(c->ffbox);
    focus.reach = c->ff_reach;
    focus.theta_min = c->ff_theta_min;
  }
  // bound sheets as far-field sources for the wake's downstream nodes
  const bool bound_split = use_ff && rows > kBoundRows;
  if (bound_split) {
    if (!c->ffb && !(c->ffb = new (std::nothrow) FarField()))
      return fail(c, UVLM_ERR_OOM, "far-field state");
    std::string e;
    const int rc = c->ffb->build(c->stream, P<double>(c->wing), P<double>(c->gamma), c->num,
                                 c->R, c->C, &e);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
    TRY(ensure(c, c->bnear, sizeof(double) * 3 * (size_t)c->num * kBoundRows * C1));
    TRY(ensure(c, c->ffbnd, vb));
  }
  auto ff_stage = [&](const double* targets, double* out) -> int {
    ChunkPlan plan;
    std::string e;
    int rc;
    const double* bnd;  // taken after sum_sources, which may reallocate c->partial
    int bcb;
    if (bound_split) {
      const int Vn = c->num * kBoundRows * C1;
      launch_gather_rows(c->stream, targets, rows, 0, c->num, kBoundRows, C1, P<double>(c->bnear));
      TRY(sum_sources(c, P<double>(c->bnear), Vn, true, false, &plan));
      rc = c->ffb->root_velocity(c->stream, targets, V, rows * C1, C1, kBoundRows, kBoundTheta,
                                 P<double>(c->partial), Vn, plan.cb, P<double>(c->ffbnd), &e);
      if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
      bnd = P<double>(c->ffbnd);
      bcb = 1;
    } else {
      TRY(sum_sources(c, targets, V, true, false, &plan));
      bnd = P<double>(c->partial);
      bcb = plan.cb;
    }
    if (targets == P<double>(c->pos)) {  // the current wake: built for the RHS already
      TRY(ff_current(c));
    } else {  // an intermediate stage's positions
      c->ff->set_f32c(prec_of(c->flags) == kF32C);
      rc = c->ff->build(c->stream, targets, P<double>(c->ffg), c->num, c->wgrows, c->C, &e);
      if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
      c->ff_version = -1;
    }
    rc = c->ff->velocity(c->stream, targets, V, c->ff_theta, P<double>(c->ffv), &e, &focus, true);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
    launch_ff_combine(c->stream, bnd, V, bcb, P<double>(c->ffv), V, ff, te_period, C1, out);
    return check_launch(c);
  };
  bool first = true;
  auto stage = [&](const double* targets, Buf& out) -> int {
    if (!use_ff) return exact_stage(targets, P<double>(out));
    TRY(ff_stage(targets, P<double>(out)));
    if (audit && first) {
      first = false;
      // Large wakes are audited on every stride-th node (at least
      // kAuditStride, about kAuditNodes nodes; the subset moves along with
      // each audit) and keep the split's velocities.
      const int stride = V >= kAuditFull ? std::max(kAuditStride, V / kAuditNodes) : 1;
      if (stride == 1) {
        TRY(exact_stage(targets, P<double>(c->ffx)));
        launch_ff_maxdiff(c->stream, P<double>(out), P<double>(c->ffx), V,
                          P<unsigned long long>(c->ffstat), targets, &focus);
      } else {
        const int off = c->ff_audits % stride, Vs = (V - off + stride - 1) / stride;
        TRY(ensure(c, c->ffxs, sizeof(double) * 3 * (size_t)Vs));
        launch_gather_strided(c->stream, targets, off, stride, Vs, P<double>(c->ffxs));
        ChunkPlan plan;
        TRY(sum_sources(c, P<double>(c->ffxs), Vs, true, true, &plan));
        launch_flow_velocity(c->stream, P<double>(c->partial), Vs, Vs, plan.cb, plan.cw, ff,
                             P<double>(c->ffx), 0, 0);
        launch_ff_maxdiff(c->stream, P<double>(out), P<double>(c->ffx), Vs,
                          P<unsigned long long>(c->ffstat), P<double>(c->ffxs), &focus, off,
                          stride, te_period, C1);
      }
      unsigned long long h[2] = {0, 0};
      HIPCHK(c, hipMemcpyAsync(h, c->ffstat.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      double e, m;
      std::memcpy(&e, &h[0], sizeof(double));
      std::memcpy(&m, &h[1], sizeof(double));
      c->ff_last_err = m > 0.0 ? e / m : e;
      ++c->ff_audits;
      if (c->ff_last_err > c->ff_tol) c->ff_off = true;
      // a fully audited stage continues with the exact velocities
      if (stride == 1)
        HIPCHK(c, hipMemcpyAsync(out.p, c->ffx.p, vb, hipMemcpyDeviceToDevice, c->stream));
    }
    return UVLM_OK;
  };
  const size_t total = (size_t)V;
  const int ign = (int)std::min<size_t>(ignore_steps, (size_t)INT32_MAX);
  double* k2 = nullptr;
  double* k3 = nullptr;
  double* k4 = nullptr;
  int ref_stages = 1;
  TRY(stage(P<double>(c->pos), c->k1));  // wake recs were packed by shed
  if (scheme == UVLM_ADVECT_RK2) {
    TRY(ensure(c, c->k2, vb));
    launch_stage_pos(c->stream, P<double>(c->pos), P<double>(c->k1), dt, 1.0, total,
                     P<double>(c->pst));
    pack_wake(c, P<double>(c->pst), rows, 0);
    TRY(stage(P<double>(c->pst), c->k2));
    k2 = P<double>(c->k2);
    ref_stages = 2;
  } else if (scheme == UVLM_ADVECT_RK4 || scheme == UVLM_ADVECT_RK4_CLASSIC) {
    TRY(ensure(c, c->k2, vb));
    TRY(ensure(c, c->k3, vb));
    launch_stage_pos(c->stream, P<double>(c->pos), P<double>(c->k1), dt, 2.0, total,
                     P<double>(c->pst));
    pack_wake(c, P<double>(c->pst), rows, 0);
    TRY(stage(P<double>(c->pst), c->k2));
    launch_stage_pos(c->stream, P<double>(c->pos), P<double>(c->k2), dt, 2.0, total,
                     P<double>(c->pst));
    pack_wake(c, P<double>(c->pst), rows, 0);
    TRY(stage(P<double>(c->pst), c->k3));
    k2 = P<double>(c->k2);
    k3 = P<double>(c->k3);
    ref_stages = 4;  // the reference also evaluates the unused k4 stage
    if (scheme == UVLM_ADVECT_RK4_CLASSIC) {
      TRY(ensure(c, c->k4, vb));
      launch_stage_pos(c->stream, P<double>(c->pos), P<double>(c->k3), dt, 1.0, total,
                       P<double>(c->pst));
      pack_wake(c, P<double>(c->pst), rows, 0);
      TRY(stage(P<double>(c->pst), c->k4));
      k4 = P<double>(c->k4);
    }
  }
  const double* vprev = nullptr;
  if (scheme == UVLM_ADVECT_AB2 && c->has_vprev) vprev = P<double>(c->vprev);
  launch_update(c->stream, scheme, P<double>(c->pos), P<double>(c->k1), k2, k3, k4, vprev,
                c->vprev_rows, ign, dt, c->num, rows, C1, P<double>(c->wake), c->cap,
                c->wstart);
  c->wake_version += 1;
  TRY(check_launch(c));
  if (scheme == UVLM_ADVECT_AB2) {
    TRY(ensure(c, c->vprev, vb));
    HIPCHK(c, hipMemcpyAsync(c->vprev.p, c->k1.p, vb, hipMemcpyDeviceToDevice, c->stream));
    c->has_vprev = true;
    c->vprev_rows = rows;
  }
  c->pairs_ref += 4.0 * ref_stages * V * (double)(N + Nw);
  return UVLM_OK;
}

int uvlm_sim_erase_oldest_wake(uvlm_ctx* c) {
  if (!c) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  if (c->wrows > 0) c->wrows -= 1;
  if (c->wgrows > 0) c->wgrows -= 1;
  c->wake_version += 1;
  return UVLM_OK;
}

int uvlm_sim_wake_shape(uvlm_ctx* c, size_t* node_rows, size_t* gamma_rows) {
  if (!c) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  if (node_rows) *node_rows = (size_t)c->wrows;
  if (gamma_rows) *gamma_rows = (size_t)c->wgrows;
  return UVLM_OK;
}

int uvlm_sim_get_wake(uvlm_ctx* c, double* nodes, double* gamma) {
  if (!c) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  const int C1 = c->C + 1;
  for (int n = 0; n < c->num; ++n) {
    if (nodes && c->wrows)
      HIPCHK(c, hipMemcpyAsync(nodes + (size_t)3 * n * c->wrows * C1,
                               P<double>(c->wake) + (size_t)3 * ((size_t)n * c->cap + c->wstart) * C1,
                               sizeof(double) * 3 * c->wrows * C1, hipMemcpyDeviceToHost,
                               c->stream));
    if (gamma && c->wgrows)
      HIPCHK(c, hipMemcpyAsync(gamma + (size_t)n * c->wgrows * c->C,
                               P<double>(c->wgam) + ((size_t)n * c->capg + c->gstart) * c->C,
                               sizeof(double) * c->wgrows * c->C, hipMemcpyDeviceToHost,
                               c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_sim_set_wake(uvlm_ctx* c, size_t node_rows, size_t gamma_rows,
                      const double* nodes, const double* gamma,
                      const double* wing_gamma) {
  if (!c) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  if ((int)node_rows > c->cap || (int)gamma_rows > c->capg)
    return fail(c, UVLM_ERR_CAPACITY, "wake larger than capacity");
  if (!(gamma_rows + 1 == node_rows || (node_rows == 0 && gamma_rows == 0)))
    return fail(c, UVLM_ERR_ARG, "inconsistent wake shape: %zu node rows, %zu gamma rows",
                node_rows, gamma_rows);
  const int C1 = c->C + 1;
  c->wrows = (int)node_rows;
  c->wstart = c->cap - c->wrows;
  c->wake_version += 1;
  c->wgrows = (int)gamma_rows;
  c->gstart = c->capg - c->wgrows;
  for (int n = 0; n < c->num; ++n) {
    if (c->wrows)
      HIPCHK(c, hipMemcpyAsync(P<double>(c->wake) + (size_t)3 * ((size_t)n * c->cap + c->wstart) * C1,
                               nodes + (size_t)3 * n * c->wrows * C1,
                               sizeof(double) * 3 * c->wrows * C1, hipMemcpyHostToDevice,
                               c->stream));
    if (c->wgrows)
      HIPCHK(c, hipMemcpyAsync(P<double>(c->wgam) + ((size_t)n * c->capg + c->gstart) * c->C,
                               gamma + (size_t)n * c->wgrows * c->C,
                               sizeof(double) * c->wgrows * c->C, hipMemcpyHostToDevice,
                               c->stream));
  }
  if (wing_gamma)
    HIPCHK(c, hipMemcpyAsync(c->gamma.p, wing_gamma, sizeof(double) * c->N,
                             hipMemcpyHostToDevice, c->stream));
  if (c->inv) c->inv->reset_history();
  c->has_vprev = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_sim_get_gamma(uvlm_ctx* c, double* gamma) {
  if (!c || !gamma) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  HIPCHK(c, hipMemcpyAsync(gamma, c->gamma.p, sizeof(double) * c->N, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_sim_get_loads(uvlm_ctx* c, double* out, size_t* count_inout) {
  if (!c || !count_inout) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  TRY(set_dev(c));
  const size_t n = std::min(*count_inout, (size_t)c->loads_count);
  if (n && out)
    HIPCHK(c, hipMemcpyAsync(out, c->loads.p, sizeof(double) * 2 * c->num * n,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *count_inout = (size_t)c->loads_count;
  return UVLM_OK;
}

int uvlm_profile_enable(uvlm_ctx* c, int enable) {
  if (!c) return UVLM_ERR_ARG;
  c->prof = enable != 0;
  if (c->ff) c->ff->set_profile(c->prof);
  return UVLM_OK;
}

int uvlm_profile_read_near(uvlm_ctx* c, size_t* launches, double* kernel_ms, double* pairs) {
  if (!c) return UVLM_ERR_ARG;
  size_t n = 0;
  double ms = 0.0, pr = 0.0;
  if (c->ff) {
    TRY(set_dev(c));
    std::string e;
    const int rc = c->ff->profile_read(c->stream, &n, &ms, &pr, &e);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  }
  if (launches) *launches = n;
  if (kernel_ms) *kernel_ms = ms;
  if (pairs) *pairs = pr;
  return UVLM_OK;
}

int uvlm_profile_read(uvlm_ctx* c, size_t* launches, double* kernel_ms, double* pairs) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double ms = 0.0, pr = 0.0;
  for (size_t k = 0; k < c->ev_used.size(); ++k) {
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev_used[k].first, c->ev_used[k].second));
    ms += t;
    pr += c->ev_pairs[k];
    c->ev_free.push_back(c->ev_used[k]);
  }
  if (launches) *launches = c->ev_used.size();
  if (kernel_ms) *kernel_ms = ms;
  if (pairs) *pairs = pr;
  c->ev_used.clear();
  c->ev_pairs.clear();
  return UVLM_OK;
}

int uvlm_sim_pair_counts(uvlm_ctx* c, double* evaluated, double* reference) {
  if (!c) return UVLM_ERR_ARG;
  if (evaluated) *evaluated = c->pairs_eval;
  if (reference) *reference = c->pairs_ref;
  return UVLM_OK;
}

// ---------------------------------------------------------------------------
int uvlm_farfield_velocity(uvlm_ctx* c, const double* targets, size_t nt, const double* nodes,
                           const double* gamma, size_t num, size_t rows, size_t cols,
                           double theta, double* out) {
  if (!c || (nt && (!targets || !out)) || !nodes || !gamma) return UVLM_ERR_ARG;
  if (!(theta > 1.0)) return fail(c, UVLM_ERR_ARG, "theta must exceed 1");
  if (!default_cut(c))
    return fail(c, UVLM_ERR_ARG, "the far-field split runs with the default cut-off only");
  if (nt > (1u << 26) || num * (rows + 1) * (cols + 1) > (1u << 26))
    return fail(c, UVLM_ERR_ARG, "problem too large");
  TRY(set_dev(c));
  if (nt == 0) return UVLM_OK;
  if (!c->ff) {
    c->ff = new (std::nothrow) FarField();
    if (c->ff) c->ff->set_profile(c->prof);
  }
  if (!c->ff) return fail(c, UVLM_ERR_OOM, "far-field state");
  const size_t nb = 3 * num * (rows + 1) * (cols + 1), gb = num * rows * cols;
  TRY(ensure(c, c->t_a, sizeof(double) * 3 * nt));
  TRY(ensure(c, c->t_b, sizeof(double) * nb));
  TRY(ensure(c, c->t_c, sizeof(double) * std::max<size_t>(gb, 1)));
  TRY(ensure(c, c->t_out, sizeof(double) * 3 * nt));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, targets, sizeof(double) * 3 * nt, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->t_b.p, nodes, sizeof(double) * nb, hipMemcpyHostToDevice,
                           c->stream));
  if (gb)
    HIPCHK(c, hipMemcpyAsync(c->t_c.p, gamma, sizeof(double) * gb, hipMemcpyHostToDevice,
                             c->stream));
  std::string e;
  c->ff_version = -1;  // the simulation's sources are replaced
  c->ff->set_f32c(false);
  const int rc = c->ff->wake_velocity(c->stream, P<double>(c->t_b), P<double>(c->t_c), (int)num,
                                      (int)rows, (int)cols, P<double>(c->t_a), (int)nt, theta,
                                      P<double>(c->t_out), &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  HIPCHK(c, hipMemcpyAsync(out, c->t_out.p, sizeof(double) * 3 * nt, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

namespace {

int mlp_upload(uvlm_ctx* c, const float* w, const float* X, const float* V, size_t n) {
  if (!c->mlp && !(c->mlp = new (std::nothrow) MlpTrainer()))
    return fail(c, UVLM_ERR_OOM, "network state");
  TRY(ensure(c, c->mw, sizeof(float) * UVLM_MLP_PARAMS));
  TRY(ensure(c, c->mx, sizeof(float) * kMlpIn * std::max<size_t>(n, 1)));
  TRY(ensure(c, c->my, sizeof(float) * kMlpOut * std::max<size_t>(n, 1)));
  HIPCHK(c, hipMemcpyAsync(c->mw.p, w, sizeof(float) * UVLM_MLP_PARAMS, hipMemcpyHostToDevice,
                           c->stream));
  if (n)
    HIPCHK(c, hipMemcpyAsync(c->mx.p, X, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice,
                             c->stream));
  if (V) {
    TRY(ensure(c, c->mv, sizeof(float) * kMlpOut * std::max<size_t>(n, 1)));
    if (n)
      HIPCHK(c, hipMemcpyAsync(c->mv.p, V, sizeof(float) * kMlpOut * n, hipMemcpyHostToDevice,
                               c->stream));
  }
  return UVLM_OK;
}

}  // namespace

static_assert(UVLM_MLP_PARAMS == MlpShape::total && UVLM_MLP_IN == kMlpIn &&
                  UVLM_MLP_OUT == kMlpOut,
              "network layout of the C-ABI");

int uvlm_farfield_set_mlp(uvlm_ctx* c, const float* weights, double ymax) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  if (!c->ff) {
    c->ff = new (std::nothrow) FarField();
    if (c->ff) c->ff->set_profile(c->prof);
  }
  if (!c->ff) return fail(c, UVLM_ERR_OOM, "far-field state");
  if (weights && !(ymax > 0.0)) return fail(c, UVLM_ERR_ARG, "ymax must be positive");
  std::string e;
  const int rc = c->ff->set_mlp(c->stream, weights, weights != nullptr, (float)ymax, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  return UVLM_OK;
}

int uvlm_farfield_mlp_stats(uvlm_ctx* c, unsigned long long* pairs, unsigned long long* trips) {
  if (!c) return UVLM_ERR_ARG;
  unsigned long long p = 0, t = 0;
  if (c->ff) {
    TRY(set_dev(c));
    std::string e;
    const int rc = c->ff->mlp_stats(c->stream, &p, &t, &e);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  }
  if (pairs) *pairs = p;
  if (trips) *trips = t;
  return UVLM_OK;
}

namespace {
int copy_samples(uvlm_ctx* c, const std::vector<float>& f, const std::vector<float>& r,
                 float* feats, float* resid, size_t* count) {
  const size_t k = std::min(*count, f.size() / kMlpIn);
  if (k) {
    std::memcpy(feats, f.data(), sizeof(float) * kMlpIn * k);
    std::memcpy(resid, r.data(), sizeof(float) * kMlpOut * k);
  }
  *count = k;
  (void)c;
  return UVLM_OK;
}
}  // namespace

int uvlm_farfield_target_order(uvlm_ctx* c, const double* targets, size_t nt,
                               unsigned* keys, int* order) {
  if (!c || (nt && (!targets || !keys || !order))) return UVLM_ERR_ARG;
  if (nt > (1u << 26)) return fail(c, UVLM_ERR_ARG, "problem too large");
  TRY(set_dev(c));
  if (nt == 0) return UVLM_OK;
  if (!c->ff) {
    c->ff = new (std::nothrow) FarField();
    if (c->ff) c->ff->set_profile(c->prof);
  }
  if (!c->ff) return fail(c, UVLM_ERR_OOM, "far-field state");
  const size_t padded = (nt + 63) / 64 * 64;
  TRY(ensure(c, c->t_a, sizeof(double) * 3 * nt));
  TRY(ensure(c, c->t_out, std::max(sizeof(unsigned) * nt, sizeof(int) * padded)));
  TRY(ensure(c, c->t_b, sizeof(int) * padded));
  HIPCHK(c, hipMemcpyAsync(c->t_a.p, targets, sizeof(double) * 3 * nt, hipMemcpyHostToDevice,
                           c->stream));
  std::string e;
  const int rc = c->ff->morton_order(c->stream, P<double>(c->t_a), (int)nt, (int)(padded / 64),
                                     P<int>(c->t_b), P<unsigned>(c->t_out), &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  HIPCHK(c, hipMemcpyAsync(keys, c->t_out.p, sizeof(unsigned) * nt,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(order, c->t_b.p, sizeof(int) * padded, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_farfield_samples(uvlm_ctx* c, const double* targets, size_t nt, const double* nodes,
                          const double* gamma, size_t num, size_t rows, size_t cols,
                          double theta, unsigned seed, float* feats, float* resid,
                          size_t* count) {
  if (!c || !count || (*count && (!feats || !resid))) return UVLM_ERR_ARG;
  if (*count > (1u << 22)) return fail(c, UVLM_ERR_ARG, "too many samples");
  // the split builds the interaction lists the samples are drawn from
  const size_t want = *count;
  *count = 0;
  TRY(set_dev(c));
  if (nt == 0 || want == 0) return UVLM_OK;
  std::vector<double> tmp(3 * nt);
  TRY(uvlm_farfield_velocity(c, targets, nt, nodes, gamma, num, rows, cols, theta, tmp.data()));
  std::vector<float> f, r;
  std::string e;
  const int rc = c->ff->sample_residuals(c->stream, P<double>(c->t_a), (int)nt, (int)want, seed,
                                         &f, &r, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  *count = want;
  return copy_samples(c, f, r, feats, resid, count);
}

namespace {
// Samples of the split at theta drawn from the simulation's current wake.
int sim_samples(uvlm_ctx* c, double theta, size_t want, unsigned seed, std::vector<float>* f,
                std::vector<float>* r) {
  f->clear();
  r->clear();
  if (c->wgrows == 0 || want == 0) return UVLM_OK;
  if (!c->ff) {
    c->ff = new (std::nothrow) FarField();
    if (c->ff) c->ff->set_profile(c->prof);
  }
  if (!c->ff) return fail(c, UVLM_ERR_OOM, "far-field state");
  const int C1 = c->C + 1, rows = c->wrows, V = c->num * rows * C1;
  const size_t vb = sizeof(double) * 3 * (size_t)V;
  TRY(ensure(c, c->pos, vb));
  TRY(ensure(c, c->ffv, vb));
  TRY(ensure(c, c->ffg, sizeof(double) * (size_t)c->num * c->wgrows * c->C));
  launch_gather_rows(c->stream, P<double>(c->wake), c->cap, c->wstart, c->num, rows, C1,
                     P<double>(c->pos));
  launch_gather_rows_w(c->stream, P<double>(c->wgam), c->capg, c->gstart, c->num, c->wgrows,
                       c->C, P<double>(c->ffg));
  // the lists of the advection call: graded criterion (when set) and the
  // per-target far test
  FFFocus focus;
  if (c->ff_reach > 0.0) {
    TRY(ensure(c, c->ffbox, 6 * sizeof(double)));
    launch_ff_bbox(c->stream, P<double>(c->wing), c->num * (c->R + 1) * (c->C + 1),
                   P<double>(c->ffbox));
    focus.box = P<double>(c->ffbox);
    focus.reach = c->ff_reach;
    focus.theta_min = c->ff_theta_min;
  }
  std::string e;
  c->ff_version = -1;
  c->ff->set_f32c(prec_of(c->flags) == kF32C);
  int rc = c->ff->build(c->stream, P<double>(c->pos), P<double>(c->ffg), c->num, c->wgrows, c->C,
                        &e);
  if (rc == UVLM_OK)
    rc = c->ff->velocity(c->stream, P<double>(c->pos), V, theta, P<double>(c->ffv), &e, &focus,
                         true);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  rc = c->ff->sample_residuals(c->stream, P<double>(c->pos), V, (int)want, seed, f, r, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  return UVLM_OK;
}
}  // namespace

int uvlm_sim_farfield_samples(uvlm_ctx* c, double theta, unsigned seed, float* feats,
                              float* resid, size_t* count) {
  if (!c || !count || (*count && (!feats || !resid))) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  if (!(theta > 1.0)) return fail(c, UVLM_ERR_ARG, "theta must exceed 1");
  if (*count > (1u << 22)) return fail(c, UVLM_ERR_ARG, "too many samples");
  const size_t want = *count;
  *count = 0;
  TRY(set_dev(c));
  std::vector<float> f, r;
  TRY(sim_samples(c, theta, want, seed, &f, &r));
  *count = want;
  return copy_samples(c, f, r, feats, resid, count);
}

int uvlm_sim_farfield_learn(uvlm_ctx* c, double theta, size_t max_samples, size_t iters,
                            size_t batch, float lr, unsigned seed, float* weights,
                            float* last_loss) {
  if (!c || !weights || max_samples == 0 || iters == 0 || batch == 0) return UVLM_ERR_ARG;
  if (!c->sim) return fail(c, UVLM_ERR_STATE, "uvlm_sim_init not called");
  if (!(theta > 1.0)) return fail(c, UVLM_ERR_ARG, "theta must exceed 1");
  if (max_samples > (1u << 22) || iters > (1u << 20) || batch > (1u << 16))
    return fail(c, UVLM_ERR_ARG, "training problem too large");
  if (!(lr > 0.0f)) return fail(c, UVLM_ERR_ARG, "learning rate must be positive");
  TRY(set_dev(c));
  std::vector<float> f, r;
  TRY(sim_samples(c, theta, max_samples, seed, &f, &r));
  const size_t n = f.size() / kMlpIn;
  if (n == 0) return fail(c, UVLM_ERR_STATE, "no far pairs to learn from");
  float ymax = 0.0f;
  for (float v : r) ymax = std::max(ymax, std::fabs(v));
  TRY(mlp_upload(c, weights, f.data(), r.data(), n));
  std::string e;
  std::vector<float> lh;
  int rc = c->mlp->train(c->stream, P<float>(c->mw), P<float>(c->mx), P<float>(c->mv), (int)n,
                         (int)iters, (int)batch, lr, seed, &lh, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  HIPCHK(c, hipMemcpyAsync(weights, c->mw.p, sizeof(float) * UVLM_MLP_PARAMS,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (last_loss) *last_loss = lh.empty() ? 0.0f : lh.back();
  rc = c->ff->set_mlp(c->stream, weights, true, 2.0f * ymax, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  return UVLM_OK;
}

int uvlm_sim_set_farfield(uvlm_ctx* c, double theta, int audit_every, double tol) {
  if (!c) return UVLM_ERR_ARG;
  if (theta > 0.0 && !(theta > 1.0)) return fail(c, UVLM_ERR_ARG, "theta must exceed 1");
  if (theta > 0.0 && !default_cut(c))
    return fail(c, UVLM_ERR_ARG, "the far-field split runs with the default cut-off only");
  if (theta > 0.0 && !c->ff) {
    c->ff = new (std::nothrow) FarField();
    if (c->ff) c->ff->set_profile(c->prof);
  }
  if (theta > 0.0 && !c->ff) return fail(c, UVLM_ERR_OOM, "far-field state");
  c->ff_theta = theta > 0.0 ? theta : 0.0;
  c->ff_audit_every = audit_every > 0 ? audit_every : 0;
  c->ff_tol = tol > 0.0 ? tol : 1e-3;
  c->ff_counter = 0;
  c->ff_off = false;
  return UVLM_OK;
}

int uvlm_sim_set_farfield_focus(uvlm_ctx* c, double reach, double theta_min) {
  if (!c) return UVLM_ERR_ARG;
  if (reach > 0.0 && !(theta_min > 1.0))
    return fail(c, UVLM_ERR_ARG, "theta_min must exceed 1");
  c->ff_reach = reach > 0.0 ? reach : 0.0;
  c->ff_theta_min = theta_min;
  return UVLM_OK;
}

int uvlm_sim_farfield_stats(uvlm_ctx* c, double* last_audit_err, int* audits, int* active) {
  if (!c) return UVLM_ERR_ARG;
  if (last_audit_err) *last_audit_err = c->ff_last_err;
  if (audits) *audits = c->ff_audits;
  if (active) *active = (c->ff_theta > 0.0 && !c->ff_off) ? 1 : 0;
  return UVLM_OK;
}

// ---------------------------------------------------------------------------

int uvlm_mlp_forward(uvlm_ctx* c, const float* w, const float* feats, size_t n, float* y,
                     int mode) {
  if (!c || !w || (n && (!feats || !y))) return UVLM_ERR_ARG;
  if (mode != UVLM_MLP_FP32 && mode != UVLM_MLP_BF16)
    return fail(c, UVLM_ERR_ARG, "unknown network mode %d", mode);
  if (n > (1u << 24)) return fail(c, UVLM_ERR_ARG, "too many samples");
  TRY(set_dev(c));
  if (n == 0) return UVLM_OK;
  TRY(mlp_upload(c, w, feats, nullptr, n));
  std::string e;
  if (mode == UVLM_MLP_FP32) {
    const int rc = c->mlp->forward(c->stream, P<float>(c->mw), P<float>(c->mx), (int)n,
                                   P<float>(c->my), &e);
    if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  } else {
    TRY(ensure(c, c->mdev, mlp_dev_bytes()));
    launch_mlp_to_dev(c->stream, P<float>(c->mw), c->mdev.p);
    launch_mlp_infer(c->stream, mlp_dev_view(c->mdev.p), P<float>(c->mx), (int)n,
                     P<float>(c->my));
    TRY(check_launch(c));
  }
  HIPCHK(c, hipMemcpyAsync(y, c->my.p, sizeof(float) * kMlpOut * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return UVLM_OK;
}

int uvlm_mlp_grad(uvlm_ctx* c, const float* w, const float* feats, const float* targets,
                  size_t n, float* grad, float* loss) {
  if (!c || !w || !feats || !targets || !grad || n == 0) return UVLM_ERR_ARG;
  if (n > (1u << 20)) return fail(c, UVLM_ERR_ARG, "too many samples");
  TRY(set_dev(c));
  TRY(mlp_upload(c, w, feats, targets, n));
  TRY(ensure(c, c->mg, sizeof(float) * (UVLM_MLP_PARAMS + 4)));
  std::string e;
  const int rc = c->mlp->grad(c->stream, P<float>(c->mw), P<float>(c->mx), P<float>(c->mv),
                              (int)n, P<float>(c->mg), P<float>(c->mg) + UVLM_MLP_PARAMS, &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  TRY(check_launch(c));
  HIPCHK(c, hipMemcpyAsync(grad, c->mg.p, sizeof(float) * UVLM_MLP_PARAMS, hipMemcpyDeviceToHost,
                           c->stream));
  float l = 0.0f;
  HIPCHK(c, hipMemcpyAsync(&l, P<float>(c->mg) + UVLM_MLP_PARAMS, sizeof(float),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  return UVLM_OK;
}

int uvlm_mlp_train(uvlm_ctx* c, float* w, const float* feats, const float* targets, size_t n,
                   size_t iters, size_t batch, float lr, unsigned seed, float* losses) {
  if (!c || !w || !feats || !targets || n == 0 || iters == 0 || batch == 0) return UVLM_ERR_ARG;
  if (n > (1u << 24) || iters > (1u << 20) || batch > (1u << 16))
    return fail(c, UVLM_ERR_ARG, "training problem too large");
  if (!(lr > 0.0f)) return fail(c, UVLM_ERR_ARG, "learning rate must be positive");
  TRY(set_dev(c));
  TRY(mlp_upload(c, w, feats, targets, n));
  std::string e;
  std::vector<float> lh;
  const int rc = c->mlp->train(c->stream, P<float>(c->mw), P<float>(c->mx), P<float>(c->mv),
                               (int)n, (int)iters, (int)batch, lr, seed, losses ? &lh : nullptr,
                               &e);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  HIPCHK(c, hipMemcpyAsync(w, c->mw.p, sizeof(float) * UVLM_MLP_PARAMS, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (losses) std::copy(lh.begin(), lh.end(), losses);
  return UVLM_OK;
}

int uvlm_sim_refine_stats(uvlm_ctx* c, long long* solves, long long* sweeps, int* last_sweeps) {
  if (!c) return UVLM_ERR_ARG;
  if (solves) *solves = c->ir_solves;
  if (sweeps) *sweeps = c->ir_sweeps;
  if (last_sweeps) *last_sweeps = c->ir_last;
  return UVLM_OK;
}

int uvlm_sim_inverse_stats(uvlm_ctx* c, long long* sweeps, long long* refreshes,
                           long long* rebuilds, int* last_sweeps) {
  if (!c) return UVLM_ERR_ARG;
  if (sweeps) *sweeps = c->inv ? c->inv->sweeps_total() : 0;
  if (refreshes) *refreshes = c->inv ? c->inv->refreshes() : 0;
  if (rebuilds) *rebuilds = c->inv ? c->inv->rebuilds() : 0;
  if (last_sweeps) *last_sweeps = c->inv_last_sweeps;
  return UVLM_OK;
}

// ---------------------------------------------------------------------------
int uvlm_batch_init(uvlm_ctx* c, const uvlm_batch_case* cases, size_t n,
                    const double* wing_nodes, size_t max_steps, int scheme,
                    size_t ignore_steps) {
  if (!c || !cases || !wing_nodes || n == 0) return UVLM_ERR_ARG;
  if (n > (1u << 20) || max_steps > (1u << 20) || ignore_steps > (1u << 20))
    return fail(c, UVLM_ERR_ARG, "batch size or step count out of range");
  if (!default_cut(c)) return fail(c, UVLM_ERR_ARG, "batches run with the default cut-off only");
  TRY(set_dev(c));
  delete c->batch;
  c->batch = new (std::nothrow) BatchEngine();
  if (!c->batch) return fail(c, UVLM_ERR_OOM, "batch engine");
  c->batch_pairs = 0.0;
  static_assert(sizeof(BatchCaseSpec) == sizeof(uvlm_batch_case), "batch case layout");
  std::string e;
  const int rc = c->batch->init(c->stream, reinterpret_cast<const BatchCaseSpec*>(cases),
                                wing_nodes, n, max_steps, scheme, (int)ignore_steps, &e);
  if (rc != UVLM_OK) {
    delete c->batch;
    c->batch = nullptr;
    return fail(c, rc, "%s", e.c_str());
  }
  return UVLM_OK;
}

int uvlm_batch_check_plan(const uvlm_batch_case* cases, size_t n, size_t max_steps, char* msg,
                          size_t msg_len) {
  if (!cases || n == 0 || n > (1u << 20) || max_steps > (1u << 20)) return UVLM_ERR_ARG;
  BatchPlan plan;
  std::string e;
  int rc = batch_plan(reinterpret_cast<const BatchCaseSpec*>(cases), n, max_steps, &plan, &e);
  if (rc == UVLM_OK) rc = batch_check_extents(plan, max_steps, &e);
  if (msg && msg_len) {
    std::snprintf(msg, msg_len, "%s", rc == UVLM_OK ? "" : e.c_str());
  }
  return rc;
}

namespace {
int need_batch(uvlm_ctx* c) {
  if (!c->batch) return fail(c, UVLM_ERR_STATE, "uvlm_batch_init not called");
  return set_dev(c);
}
int need_case(uvlm_ctx* c, size_t k) {
  TRY(need_batch(c));
  if (k >= (size_t)c->batch->n_cases())
    return fail(c, UVLM_ERR_ARG, "case %zu out of range (%d cases)", k, c->batch->n_cases());
  return UVLM_OK;
}
}  // namespace

int uvlm_batch_run(uvlm_ctx* c, size_t steps, int use_graph) {
  if (!c) return UVLM_ERR_ARG;
  TRY(need_batch(c));
  const int s0 = c->batch->steps_done();
  std::string e;
  const int rc = c->batch->run(steps, use_graph != 0, &e);
  for (int s = s0; s < c->batch->steps_done(); ++s) c->batch_pairs += c->batch->pairs_per_step_at(s);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  return UVLM_OK;
}

int uvlm_batch_autotune(uvlm_ctx* c, int* chosen, double* times_ms, int max_variants) {
  if (!c) return UVLM_ERR_ARG;
  TRY(need_batch(c));
  std::string e;
  const int rc = c->batch->autotune(&e, chosen, times_ms, max_variants);
  if (rc != UVLM_OK) return fail(c, rc, "%s", e.c_str());
  return UVLM_OK;
}

int uvlm_batch_steps(uvlm_ctx* c, size_t* steps) {
  if (!c || !steps) return UVLM_ERR_ARG;
  TRY(need_batch(c));
  *steps = (size_t)c->batch->steps_done();
  return UVLM_OK;
}

int uvlm_batch_get_gamma(uvlm_ctx* c, size_t k, double* gamma) {
  if (!c || !gamma) return UVLM_ERR_ARG;
  TRY(need_case(c, k));
  std::string e;
  const int rc = c->batch->get_gamma((int)k, gamma, &e);
  return rc == UVLM_OK ? rc : fail(c, rc, "%s", e.c_str());
}

int uvlm_batch_get_wing(uvlm_ctx* c, size_t k, double* nodes) {
  if (!c || !nodes) return UVLM_ERR_ARG;
  TRY(need_case(c, k));
  std::string e;
  const int rc = c->batch->get_wing((int)k, nodes, &e);
  return rc == UVLM_OK ? rc : fail(c, rc, "%s", e.c_str());
}

int uvlm_batch_get_wake(uvlm_ctx* c, size_t k, double* nodes, double* gamma) {
  if (!c) return UVLM_ERR_ARG;
  TRY(need_case(c, k));
  std::string e;
  const int rc = c->batch->get_wake((int)k, nodes, gamma, &e);
  return rc == UVLM_OK ? rc : fail(c, rc, "%s", e.c_str());
}

int uvlm_batch_get_loads(uvlm_ctx* c, size_t k, double* out, size_t* count_inout) {
  if (!c || !count_inout) return UVLM_ERR_ARG;
  TRY(need_case(c, k));
  std::string e;
  const int rc = c->batch->get_loads((int)k, out, count_inout, &e);
  return rc == UVLM_OK ? rc : fail(c, rc, "%s", e.c_str());
}

int uvlm_batch_pairs(uvlm_ctx* c, double* pairs) {
  if (!c || !pairs) return UVLM_ERR_ARG;
  TRY(need_batch(c));
  *pairs = c->batch_pairs;
  return UVLM_OK;
}

int uvlm_batch_set_option(uvlm_ctx* c, int option, int value) {
  if (!c) return UVLM_ERR_ARG;
  TRY(need_batch(c));
  if (option != UVLM_BATCH_FORCE_PIVOTED) return fail(c, UVLM_ERR_ARG, "unknown option %d", option);
  c->batch->set_force_pivoted(value != 0);
  return UVLM_OK;
}

int uvlm_batch_free(uvlm_ctx* c) {
  if (!c) return UVLM_ERR_ARG;
  TRY(set_dev(c));
  delete c->batch;
  c->batch = nullptr;
  return UVLM_OK;
}

}  // extern "C"
