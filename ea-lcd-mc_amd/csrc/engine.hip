// engine.hip -- the C ABI of include/eatube.h over the gfx950 kernels.
//
// An engine is one tube: its device state, a one-entry device tube table and
// the block maps of that table.  A batch is a table of several engines'
// tubes whose step (sweeps, PT, Houdayer, measures) is captured once in a
// HIP graph and replayed; the kernels are the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"

using namespace eatube;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(EA_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define RC_TRY(call)        \
  do {                      \
    const int rc_ = (call); \
    if (rc_) return rc_;    \
  } while (0)

template <typename T>
int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (!count) return EA_OK;
  if (hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return fail(EA_E_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
  }
  return EA_OK;
}

// Shape of one step of ea_run (main.cu:149-228): the loops whose launches a
// step holds.
struct StepShape {
  int mesfr = 0, houfr = 0, ptfr = 0;
  bool cluster = false;   // some tube has cluster moves
  bool operator==(const StepShape &o) const {
    return mesfr == o.mesfr && houfr == o.houfr && ptfr == o.ptfr && cluster == o.cluster;
  }
};

}  // namespace

struct ea_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Tube t{};             // host copy of the descriptor
  Tube *dt = nullptr;   // device copy
  DevPlan plan;
  Shape shape;          // launch shape of the engine's own calls
  uint32_t job = 0;
  std::vector<double> betas;
  // Call block: the descriptor, the step counter and the inputs of the calls
  // in flight, on the device (`dt` points at its start) and in a pinned host
  // mirror from which one asynchronous copy sends them (stage)
  uint8_t *call_dev = nullptr, *call_host = nullptr;
  size_t call_cap = 0;
  hipEvent_t staged = nullptr;  // recorded after the last staging copy
  ea_counter *noalloc_ctr = nullptr;  // this device's `scratch` address (not owned)
  int hist_rows = 0;
  int keep_on = 0;
  uint32_t *keep_buf = nullptr;
  // energy bookkeeping: the last sweep left partials by beta that no kernel
  // has summed yet (part_pending), or ener holds the current energies
  bool part_pending = false, ener_valid = false;
  // kernel timing: pending (kind, start, stop) event triples
  bool timing = false;
  struct Mark {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> spare;
  double ms[EA_K_KINDS] = {};
  int64_t launches[EA_K_KINDS] = {};

  const Lattice &g() const { return t.g; }

  hipEvent_t event() {
    if (!spare.empty()) {
      hipEvent_t ev = spare.back();
      spare.pop_back();
      return ev;
    }
    hipEvent_t ev = nullptr;
    (void)hipEventCreate(&ev);
    return ev;
  }

  ~ea_engine() {
    for (auto &m : marks) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    for (auto ev : spare) (void)hipEventDestroy(ev);
    free_plan(&plan);
    const void *ptrs[] = {t.spin, t.J,    t.wctb, t.lut,  t.ener, t.part,   t.pt,
                          t.acc,  t.att,  t.raw,  t.nd,   t.tab,  t.toff,   t.tlen,
                          t.esum, t.hist, t.hpar, t.hl16, t.hrow, t.hflag,  t.ovl,
                          t.thr,  t.spin2, keep_buf, call_dev};
    for (const void *p : ptrs)
      if (p) (void)hipFree(const_cast<void *>(p));
    if (call_host) (void)hipHostFree(call_host);
    if (staged) (void)hipEventDestroy(staged);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct ea_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  // the captured step's second branch (the whole-tube Houdayer kernels,
  // beside the slab chain) and the events that fork and join it
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::vector<ea_engine *> eng;
  Tube *dt = nullptr;
  DevPlan plan;
  std::vector<int> plan_cluster;  // cluster flags the plan was built for
  Shape launch;                   // launch shape asked for by the caller
  StepShape shape;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~ea_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    free_plan(&plan);
    if (dt) (void)hipFree(dt);
    if (join) (void)hipEventDestroy(join);
    if (fork) (void)hipEventDestroy(fork);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Counter blocks of ea_sweep_noalloc: one array per GPU, shared by every
// engine there and held until the process ends.  Host code uses only the
// per-device address hipGetSymbolAddress returns (scratch_address).
static __device__ __align__(256) ea_counter scratch[EA_NOALLOC_COUNTERS];

// Device address of `scratch` on `dev`, looked up once per device with that
// device current (ea_create calls it for the device it binds).
static int scratch_address(int dev, ea_counter **out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static ea_counter *addr[kMaxDevices] = {};
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return fail(EA_E_NODEV, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = static_cast<ea_counter *>(p);
  }
  *out = addr[dev];
  return EA_OK;
}

extern "C" const char *ea_last_error(void) { return g_err.c_str(); }

extern "C" int ea_device_count(int *count) {
  if (!count) return fail(EA_E_ARG, "count is NULL");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return fail(EA_E_NODEV, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return EA_OK;
}

extern "C" int ea_select_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail(EA_E_NODEV, "device=%d count_devices=%d", device, n);
  HIP_TRY(hipSetDevice(device));
  return EA_OK;
}

static int check_config(const ea_config *c) {
  if (!c) return fail(EA_E_ARG, "config is NULL");
  if (!valid_lattice(c->L, c->Lz))
    return fail(EA_E_ARG, "L=%d Lz=%d: both must be multiples of 8", c->L, c->Lz);
  if (const char *d = inexact_divisor(c->L, c->Lz, c->clone_threshold > 0))
    return fail(EA_E_ARG,
                "L=%d Lz=%d: the kernels' index quotients by %s are not exact over this "
                "tube's index range",
                c->L, c->Lz, d);
  if (c->nr < 2 || c->nr > EA_MAX_REPLICAS || (c->nr & (c->nr - 1)))
    return fail(EA_E_ARG, "nr=%d: a power of two in 2..%d", c->nr, EA_MAX_REPLICAS);
  if (c->nt < 2 || c->nt > EA_MAX_BETAS)
    return fail(EA_E_ARG, "nt=%d out of 2..%d", c->nt, EA_MAX_BETAS);
  if (c->ns < 1 || c->ns > EA_MAX_SAMPLES)
    return fail(EA_E_ARG, "ns=%d out of 1..%d", c->ns, EA_MAX_SAMPLES);
  if (c->bc != EA_PBC && c->bc != EA_OBC) return fail(EA_E_ARG, "bc=%d", c->bc);
  if (c->clone_threshold < 0 || c->clone_threshold > c->nt)
    return fail(EA_E_ARG, "clone_threshold=%d out of 0..nt", c->clone_threshold);
  if (c->job_id >= (1u << 14))
    return fail(EA_E_ARG, "job_id=%u exceeds the 14-bit counter field", c->job_id);
  if (!c->betas || !c->lut) return fail(EA_E_ARG, "betas and lut are required");
  for (int i = 0; i + 1 < c->nt; ++i)
    if (!(c->betas[i] >= c->betas[i + 1]))
      return fail(EA_E_ARG, "betas must be sorted from largest to smallest");
  return EA_OK;
}

// Layout of an engine's call block: the descriptor, the step counter, then
// the counter blocks and the replica pairs of the calls in flight.
constexpr size_t kOffStep = (sizeof(Tube) + 15) / 16 * 16;
constexpr size_t kOffCtr = kOffStep + 16;
constexpr size_t kCallInitial = 64 * 1024;  // bytes allocated at creation

inline size_t align16(size_t x) { return (x + 15) / 16 * 16; }

// Grows both copies of the call block to at least `bytes`; every call
// restages what it needs, so nothing is carried over.
static int grow_call_block(ea_engine *e, size_t bytes) {
  if (bytes <= e->call_cap) return EA_OK;
  bytes = std::max(bytes, 2 * e->call_cap);
  HIP_TRY(hipStreamSynchronize(e->stream));
  uint8_t *d = nullptr, *h = nullptr;
  RC_TRY(dalloc(&d, bytes));
  if (hipHostMalloc(reinterpret_cast<void **>(&h), bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(d);
    return fail(EA_E_NOMEM, "hipHostMalloc of %zu bytes failed", bytes);
  }
  if (e->call_dev) HIP_TRY(hipFree(e->call_dev));
  if (e->call_host) HIP_TRY(hipHostFree(e->call_host));
  e->call_dev = d;
  e->call_host = h;
  e->call_cap = bytes;
  e->dt = reinterpret_cast<Tube *>(d);
  e->t.step = reinterpret_cast<int *>(d + kOffStep);
  return EA_OK;
}

// Sends the host descriptor to the device -- with `reset` also the step
// counter (0) and the inputs of the calls in flight -- in one asynchronous
// copy on the engine's stream: it lands after the work queued before it, and
// the calls' launches follow without the host waiting.  The host mirror is
// reused once the previous copy has been made.
static int stage(ea_engine *e, bool reset, const ea_counter *ctr = nullptr, size_t nctr = 0,
                 const int32_t *pairs = nullptr, size_t npairs = 0) {
  const size_t off_pairs = align16(kOffCtr + sizeof(ea_counter) * nctr);
  const size_t end = reset ? off_pairs + sizeof(int32_t) * npairs : sizeof(Tube);
  RC_TRY(grow_call_block(e, end));
  HIP_TRY(hipEventSynchronize(e->staged));
  uint8_t *h = e->call_host;
  if (reset) {
    if (nctr) e->t.ctr = reinterpret_cast<const ea_counter *>(e->call_dev + kOffCtr);
    if (npairs) e->t.pairs = reinterpret_cast<const int32_t *>(e->call_dev + off_pairs);
    const int zero = 0;
    std::memcpy(h + kOffStep, &zero, sizeof zero);
    if (nctr) std::memcpy(h + kOffCtr, ctr, sizeof(ea_counter) * nctr);
    if (npairs) std::memcpy(h + off_pairs, pairs, sizeof(int32_t) * npairs);
  }
  std::memcpy(h, &e->t, sizeof(Tube));
  HIP_TRY(hipMemcpyAsync(e->call_dev, h, end, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipEventRecord(e->staged, e->stream));
  return EA_OK;
}

// The descriptor alone (the step counter keeps its value).
static int upload_tube(ea_engine *e) { return stage(e, false); }

// Sets up the calls in flight: nsteps steps of `halves` half-sweeps and
// `calls` Houdayer calls, their counter blocks and replica pairs, step
// counter at 0.
static int begin_calls(ea_engine *e, int nsteps, int halves, int calls, int it0, int frec,
                       const ea_counter *ctr = nullptr, size_t nctr = 0,
                       const int32_t *pairs = nullptr, size_t npairs = 0) {
  e->t.nsteps = nsteps;
  e->t.halves = halves;
  e->t.calls = calls;
  e->t.it0 = it0;
  e->t.frec = std::max(frec, 1);
  return stage(e, true, ctr, nctr, pairs, npairs);
}

// Grows the TRW history to hold `rows` rows; the rows written so far are
// kept (a caller may advance it0 across ea_run calls of one bin), new rows
// are 0.  It grows to at least 512 rows (the reference's MAXIT_TRW rows per
// bin) and at least doubles, so that a run's calls rarely allocate.  On
// failure the old buffer stays in place.
static int ensure_hist(ea_engine *e, int rows) {
  const Lattice &g = e->g();
  if (rows <= e->hist_rows) return EA_OK;
  rows = std::max({rows, 2 * e->hist_rows, 512});
  const size_t per = (size_t)g.ns * g.nr * g.nt;
  uint8_t *h = nullptr;
  RC_TRY(dalloc(&h, (size_t)rows * per));
  // on the engine's stream: the clear and the copy land after the work that
  // wrote the old rows and before the calls that write new ones
  hipError_t err = hipMemsetAsync(h, 0, (size_t)rows * per, e->stream);
  if (err == hipSuccess && e->t.hist)
    err = hipMemcpyAsync(h, e->t.hist, (size_t)e->hist_rows * per, hipMemcpyDeviceToDevice,
                         e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) {
    (void)hipFree(h);
    return fail(EA_E_HIP, "growing the history: %s", hipGetErrorString(err));
  }
  if (e->t.hist) HIP_TRY(hipFree(e->t.hist));
  e->t.hist = h;
  e->hist_rows = rows;
  return EA_OK;
}

// EA_PT_EXACT_WALK=1: every PT walk applies the exact rule pair by pair (the
// path a walk takes when a pair's swap threshold does not verify).  Same
// decisions either way; it lets that path be tested.
static int pt_exact_from_env() {
  const char *v = std::getenv("EA_PT_EXACT_WALK");
  return v && v[0] == '1' && v[1] == '\0';
}

// EA_HOU_SITE_FOREST=1: the whole-tube cluster labelling always uses its
// forest over sites, the path otherwise taken only by an instance with more
// runs than the run forest holds.  Same labels either way; it lets that
// path be tested.
static bool site_forest_from_env() {
  const char *v = std::getenv("EA_HOU_SITE_FOREST");
  return v && v[0] == '1' && v[1] == '\0';
}

extern "C" int ea_create(const ea_config *cfg, ea_engine **out) {
  if (!out) return fail(EA_E_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(EA_E_NODEV, "no HIP device is visible");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(EA_E_NODEV, "device %d is %s; this engine is built for gfx950", dev,
                prop.gcnArchName);

  auto e = std::make_unique<ea_engine>();
  e->device = dev;
  RC_TRY(scratch_address(dev, &e->noalloc_ctr));
  Tube &t = e->t;
  Lattice &g = t.g;
  g.L = cfg->L;
  g.Lz = cfg->Lz;
  g.ML = cfg->L / 4;
  g.MLz = cfg->Lz / 4;
  g.MS = g.ML * g.ML;
  g.V = cfg->L * cfg->L * cfg->Lz;
  g.n = g.V / 128;
  g.hl = g.ML / 2;
  g.hs = g.MS / 2;
  g.nr = cfg->nr;
  g.nt = cfg->nt;
  g.ns = cfg->ns;
  g.nr_shift = __builtin_ctz(cfg->nr);
  g.bc = cfg->bc;
  g.ctm = cfg->clone_threshold;
  g.half = (int64_t)g.n * g.nr * g.nt;
  set_divisors(g);
  t.nbx = sweep_partials_x(g, e->shape);
  t.sweep_planes = sweep_slab_planes(g);
  t.frec = 1;
  t.cluster = g.ctm > 0 && g.nr >= 4;
  t.pt_exact = pt_exact_from_env();
  t.hou_cap = site_forest_from_env() ? 0 : hou_run_cap(g);
  e->job = cfg->job_id;
  e->betas.assign(cfg->betas, cfg->betas + cfg->nt);

  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&e->staged, hipEventDisableTiming));
  const size_t chains = (size_t)g.ns * g.nr;
  uint64_t *J = nullptr;
  uint8_t *lut = nullptr;
  if ((rc = dalloc(&t.spin, (size_t)g.ns * 2 * g.half)) ||
      (rc = dalloc(&J, (size_t)g.ns * 12 * g.n)) || (rc = dalloc(&t.wctb, chains * g.nt)) ||
      (rc = dalloc(&lut, (size_t)g.nt * EA_LUT_BYTES)) ||
      (rc = dalloc(&t.ener, chains * g.nt)) || (rc = dalloc(&t.pt, chains)) ||
      (rc = dalloc(&t.part, energy_partials(g))) ||
      (rc = dalloc(&t.acc, chains * (g.nt - 1))) ||
      (rc = dalloc(&t.att, chains * (g.nt - 1))) ||
      (rc = dalloc(&t.raw, chains * (g.nt - 1))) || (rc = dalloc(&t.esum, chains * g.nt)) ||
      (sweep_full_thr_words(g) && (rc = dalloc(&t.thr, sweep_full_thr_words(g)))) ||
      (rc = grow_call_block(e.get(), kCallInitial))) {
    t.J = J;
    t.lut = lut;
    return rc;
  }
  t.J = J;
  t.lut = lut;
  HIP_TRY(hipMemsetAsync(t.spin, 0, sizeof(uint64_t) * g.ns * 2 * g.half, e->stream));
  HIP_TRY(hipMemsetAsync(J, 0, sizeof(uint64_t) * g.ns * 12 * g.n, e->stream));
  HIP_TRY(hipMemcpy(lut, cfg->lut, (size_t)g.nt * EA_LUT_BYTES, hipMemcpyHostToDevice));
  std::vector<uint8_t> ident(chains * g.nt);
  for (size_t c = 0; c < chains; ++c)
    for (int b = 0; b < g.nt; ++b) ident[c * g.nt + b] = (uint8_t)b;
  HIP_TRY(hipMemcpy(t.wctb, ident.data(), ident.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemsetAsync(t.ener, 0, sizeof(int32_t) * chains * g.nt, e->stream));
  HIP_TRY(hipMemsetAsync(t.pt, 0, sizeof(ea_hq_state) * chains, e->stream));
  HIP_TRY(hipMemsetAsync(t.acc, 0, sizeof(double) * chains * (g.nt - 1), e->stream));
  HIP_TRY(hipMemsetAsync(t.att, 0, sizeof(double) * chains * (g.nt - 1), e->stream));
  HIP_TRY(hipMemsetAsync(t.raw, 0, sizeof(double) * chains * (g.nt - 1), e->stream));
  HIP_TRY(hipMemsetAsync(t.esum, 0, sizeof(double) * chains * g.nt, e->stream));

  // exact exp(pt) tables for the device parallel tempering
  const ExpTables x = make_exp_tables(cfg->betas, g.nt, g.V);
  double *nd = nullptr, *tab = nullptr;
  int32_t *toff = nullptr, *tlen = nullptr;
  rc = dalloc(&nd, x.nd.size());
  if (!rc) rc = dalloc(&toff, x.off.size());
  if (!rc) rc = dalloc(&tlen, x.len.size());
  if (!rc) rc = dalloc(&tab, std::max<size_t>(x.val.size(), 1));
  t.nd = nd;
  t.tab = tab;
  t.toff = toff;
  t.tlen = tlen;
  if (rc) return rc;
  HIP_TRY(hipMemcpy(nd, x.nd.data(), sizeof(double) * x.nd.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(toff, x.off.data(), sizeof(int32_t) * x.off.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(tlen, x.len.data(), sizeof(int32_t) * x.len.size(), hipMemcpyHostToDevice));
  if (!x.val.empty())
    HIP_TRY(hipMemcpy(tab, x.val.data(), sizeof(double) * x.val.size(), hipMemcpyHostToDevice));

  // Houdayer: a tube is labelled in one workgroup per instance, or slab by
  // slab with a merge in HBM (hou_use_slabs)
  if (g.ctm > 0) {
    if (g.L > kHouMaxL)
      return fail(EA_E_ARG, "the Houdayer move supports L <= %d (L=%d)", kHouMaxL, g.L);
    int lds_block = 0, lds_optin = 0, lds_cu = 0, cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIP_TRY(hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
    HIP_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int lds_max = std::max(lds_block, lds_optin);
    if (hou_use_slabs(g, lds_max, lds_cu, cus)) {
      t.planes = hou_slab_planes(g, hou_slab_budget(lds_max, lds_cu));
      const size_t inst = (size_t)2 * g.ns * g.ctm;
      const size_t slabs = (g.Lz + t.planes - 1) / t.planes;
      if ((rc = dalloc(&t.hpar, inst * g.V)) || (rc = dalloc(&t.hl16, inst * g.V)) ||
          (rc = dalloc(&t.hrow, inst * (g.V / g.L))) ||
          (rc = dalloc(&t.hflag, inst * slabs * 2 * hou_slab_words(g, t.planes))))
        return rc;
    }
  }
  HIP_TRY(upload_plan(make_plan({t}, e->shape), &e->plan));
  RC_TRY(begin_calls(e.get(), 0, 0, 0, 0, 1));
  *out = e.release();
  return EA_OK;
}

// Device bytes ea_create allocates for cfg, with lds_max bytes of LDS per
// workgroup, lds_cu per CU and cus CUs deciding the Houdayer path.
static uint64_t engine_bytes(const ea_config *cfg, int lds_max, int lds_cu, int cus) {
  Lattice g{};
  g.L = cfg->L;
  g.Lz = cfg->Lz;
  g.V = cfg->L * cfg->L * cfg->Lz;
  g.n = g.V / 128;
  g.nr = cfg->nr;
  g.nt = cfg->nt;
  g.ns = cfg->ns;
  g.ctm = cfg->clone_threshold;
  g.half = (int64_t)g.n * g.nr * g.nt;
  const uint64_t chains = (uint64_t)g.ns * g.nr;
  uint64_t b = 8ull * g.ns * 2 * g.half + 8ull * g.ns * 12 * g.n + chains * g.nt +
               (uint64_t)g.nt * EA_LUT_BYTES + 4ull * chains * g.nt + sizeof(ea_hq_state) * chains +
               4ull * energy_partials(g) + 24ull * chains * (g.nt - 1) + 8ull * chains * g.nt +
               8ull * sweep_full_thr_words(g) + kCallInitial;
  const ExpTables x = make_exp_tables(cfg->betas, g.nt, g.V);
  b += 8ull * x.nd.size() + 8ull * x.off.size() + 8ull * std::max<size_t>(x.val.size(), 1);
  if (g.ctm > 0 && hou_use_slabs(g, lds_max, lds_cu, cus))
    b += hou_slab_bytes(g, hou_slab_planes(g, hou_slab_budget(lds_max, lds_cu)));
  return b;
}

extern "C" int ea_footprint(const ea_config *cfg, uint64_t *bytes) {
  if (!bytes) return fail(EA_E_ARG, "bytes is NULL");
  RC_TRY(check_config(cfg));
  *bytes = engine_bytes(cfg, 160 * 1024, 160 * 1024, 256);  // MI355X: 160 KB LDS, 256 CUs
  return EA_OK;
}

extern "C" int ea_destroy(ea_engine *e) {
  if (!e) return EA_OK;
  (void)hipStreamSynchronize(e->stream);
  delete e;
  return EA_OK;
}

extern "C" int ea_sync(ea_engine *e) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  HIP_TRY(hipStreamSynchronize(e->stream));
  return EA_OK;
}

extern "C" int ea_set_keys(ea_engine *e, const uint32_t *keys) {
  if (!e || !keys) return fail(EA_E_ARG, "NULL argument");
  std::memcpy(e->t.keys, keys, sizeof(uint32_t) * e->g().nr);
  return upload_tube(e);
}

// One parity array of one sample: the reference's [clone][h][replica] words
// to the device's [replica][clone][h] rows and back.
static void spins_to_device(const Lattice &g, const uint64_t *host, uint64_t *dev) {
  for (int c = 0; c < g.nt; ++c)
    for (int h = 0; h < g.n; ++h)
      for (int r = 0; r < g.nr; ++r)
        dev[spin_row(g, r, c) + h] = host[((size_t)c * g.n + h) * g.nr + r];
}
static void spins_from_device(const Lattice &g, const uint64_t *dev, uint64_t *host) {
  for (int c = 0; c < g.nt; ++c)
    for (int h = 0; h < g.n; ++h)
      for (int r = 0; r < g.nr; ++r)
        host[((size_t)c * g.n + h) * g.nr + r] = dev[spin_row(g, r, c) + h];
}

extern "C" int ea_upload_sample(ea_engine *e, int s, const uint64_t *even,
                                const uint64_t *odd, const uint64_t *J) {
  if (!e || !even || !odd || !J) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g();
  if (s < 0 || s >= g.ns) return fail(EA_E_ARG, "sample %d out of range", s);
  HIP_TRY(hipStreamSynchronize(e->stream));
  std::vector<uint64_t> w((size_t)2 * g.half);
  spins_to_device(g, even, w.data());
  spins_to_device(g, odd, w.data() + g.half);
  HIP_TRY(hipMemcpy(e->t.spin + (size_t)s * 2 * g.half, w.data(), sizeof(uint64_t) * w.size(),
                    hipMemcpyHostToDevice));
  // [parity][direction][h] -> [parity][direction pair][h][2]
  std::vector<uint64_t> t((size_t)12 * g.n);
  for (int p = 0; p < 2; ++p)
    for (int d = 0; d < 6; ++d)
      for (int h = 0; h < g.n; ++h)
        t[coupling_index(g, 0, p, h, d)] = J[((size_t)p * 6 + d) * g.n + h];
  HIP_TRY(hipMemcpy(const_cast<uint64_t *>(e->t.J) + (size_t)s * 12 * g.n, t.data(),
                    sizeof(uint64_t) * t.size(), hipMemcpyHostToDevice));
  e->part_pending = e->ener_valid = false;
  return EA_OK;
}

extern "C" int ea_download_spins(ea_engine *e, int s, uint64_t *even, uint64_t *odd) {
  if (!e || !even || !odd) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g();
  if (s < 0 || s >= g.ns) return fail(EA_E_ARG, "sample %d out of range", s);
  HIP_TRY(hipStreamSynchronize(e->stream));
  std::vector<uint64_t> w((size_t)2 * g.half);
  HIP_TRY(hipMemcpy(w.data(), e->t.spin + (size_t)s * 2 * g.half, sizeof(uint64_t) * w.size(),
                    hipMemcpyDeviceToHost));
  spins_from_device(g, w.data(), even);
  spins_from_device(g, w.data() + g.half, odd);
  return EA_OK;
}

extern "C" int ea_upload_perms(ea_engine *e, const uint8_t *wctb) {
  if (!e || !wctb) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g();
  for (int c = 0; c < g.ns * g.nr; ++c) {
    unsigned seen[EA_MAX_BETAS + 1] = {};
    for (int b = 0; b < g.nt; ++b) {
      const int v = wctb[(size_t)c * g.nt + b];
      if (v >= g.nt || seen[v]++)
        return fail(EA_E_ARG, "chain %d: not a permutation of 0..nt-1", c);
    }
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(e->t.wctb, wctb, (size_t)g.ns * g.nr * g.nt, hipMemcpyHostToDevice));
  e->part_pending = e->ener_valid = false;  // partials by beta map through the old perms
  return EA_OK;
}

extern "C" int ea_download_perms(ea_engine *e, uint8_t *wctb) {
  if (!e || !wctb) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g();
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(wctb, e->t.wctb, (size_t)g.ns * g.nr * g.nt, hipMemcpyDeviceToHost));
  return EA_OK;
}

extern "C" int ea_upload_pt_state(ea_engine *e, const ea_hq_state *st, const double *acc,
                                  const double *att, const double *raw) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const size_t chains = (size_t)e->g().ns * e->g().nr, np = chains * (e->g().nt - 1);
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (st) HIP_TRY(hipMemcpy(e->t.pt, st, sizeof(ea_hq_state) * chains, hipMemcpyHostToDevice));
  if (acc) HIP_TRY(hipMemcpy(e->t.acc, acc, sizeof(double) * np, hipMemcpyHostToDevice));
  if (att) HIP_TRY(hipMemcpy(e->t.att, att, sizeof(double) * np, hipMemcpyHostToDevice));
  if (raw) HIP_TRY(hipMemcpy(e->t.raw, raw, sizeof(double) * np, hipMemcpyHostToDevice));
  return EA_OK;
}

extern "C" int ea_download_pt_state(ea_engine *e, ea_hq_state *st, double *acc,
                                    double *att, double *raw) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const size_t chains = (size_t)e->g().ns * e->g().nr, np = chains * (e->g().nt - 1);
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (st) HIP_TRY(hipMemcpy(st, e->t.pt, sizeof(ea_hq_state) * chains, hipMemcpyDeviceToHost));
  if (acc) HIP_TRY(hipMemcpy(acc, e->t.acc, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (att) HIP_TRY(hipMemcpy(att, e->t.att, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (raw) HIP_TRY(hipMemcpy(raw, e->t.raw, sizeof(double) * np, hipMemcpyDeviceToHost));
  return EA_OK;
}

// ------------------------------------------------------------ the step ops

// Brackets the launches of one kernel kind with events when timing is on.
struct Timed {
  ea_engine *e;
  int kind;
  hipEvent_t a = nullptr;
  Timed(ea_engine *e_, int kind_) : e(e_), kind(kind_) {
    if (e && e->timing) {
      a = e->event();
      (void)hipEventRecord(a, e->stream);
    }
  }
  ~Timed() {
    if (a) {
      hipEvent_t b = e->event();
      (void)hipEventRecord(b, e->stream);
      e->marks.push_back({kind, a, b});
    }
  }
};

// Where the launches of one step go: a tube table, its block maps and a
// stream; `timer` is the engine whose timing brackets them (NULL in a batch).
// With a side stream (a batch), a cluster move that has both whole-tube and
// slab tubes runs the whole-tube kernels there, forked from and joined back
// to the stream by the two events.
struct Target {
  const Tube *dt;
  const DevPlan *plan;
  int ntubes;
  hipStream_t stream;
  ea_engine *timer;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// One step of main()'s `it` loop (main.cu:149-228) for every tube of the
// table: mesfr x { houfr x { ptfr sweeps, PT }, Houdayer }.  Every PT
// follows a sweep, so PT takes its energies from the sweep's partials; the
// measuring PT adds them to the energy sums and history, and a cluster move
// after it corrects the sums of the chains it changed.  With step >= 0
// (one tube launched step by step) the launches carry the step index;
// with step < 0 (a captured step, replayed) they read the tubes' step
// counters, which the step's last launch advances.
static int enqueue_step(const Target &x, const StepShape &sh, int step = -1) {
  int half = 0;
  for (int im = 0; im < sh.mesfr; ++im) {
    const bool last_m = im == sh.mesfr - 1;
    for (int ih = 0; ih < sh.houfr; ++ih) {
      for (int sw = 0; sw < sh.ptfr; ++sw) {
        if (sweep_is_whole(x.plan->shape)) {  // both halves in one launch
          Phase ph{half, sw == sh.ptfr - 1, 0, 0, step};
          half += 2;
          Timed tm(x.timer, EA_K_SWEEP);
          HIP_TRY(launch_sweep(x.dt, *x.plan, ph, x.stream));
          continue;
        }
        for (int p = 0; p < 2; ++p) {
          Phase ph{half++, sw == sh.ptfr - 1 && p == 1, 0, 0, step};
          Timed tm(x.timer, EA_K_SWEEP);
          HIP_TRY(launch_sweep(x.dt, *x.plan, ph, x.stream));
        }
      }
      Phase ph{0, 1, 0, last_m && ih == sh.houfr - 1, step};
      Timed tm(x.timer, EA_K_PT);
      HIP_TRY(launch_pt(x.dt, *x.plan, ph, x.stream));
    }
    if (sh.cluster) {
      Phase ph{0, 0, im, last_m, step};
      Timed tm(x.timer, EA_K_HOUDAYER);
      if (x.side && hou_has(*x.plan, kHouWhole) && hou_has(*x.plan, kHouSlabs)) {
        // the two paths touch disjoint tubes: the whole-tube kernels fill
        // the CUs that the slab chain's last workgroups and its small seam
        // and energy launches leave idle
        HIP_TRY(hipEventRecord(x.fork, x.stream));
        HIP_TRY(hipStreamWaitEvent(x.side, x.fork, 0));
        HIP_TRY(launch_houdayer(x.dt, *x.plan, ph, x.stream, kHouSlabs));
        HIP_TRY(launch_houdayer(x.dt, *x.plan, ph, x.side, kHouWhole));
        HIP_TRY(hipEventRecord(x.join, x.side));
        HIP_TRY(hipStreamWaitEvent(x.stream, x.join, 0));
      } else {
        HIP_TRY(launch_houdayer(x.dt, *x.plan, ph, x.stream));
      }
    }
  }
  if (step < 0) HIP_TRY(launch_step_end(x.dt, x.ntubes, x.stream));
  return EA_OK;
}

// Sweep kind 4 writes every sweep into a second spin array; it is allocated
// when the kind is first set and kept until the engine is destroyed.
static int ensure_spin2(ea_engine *e) {
  if (e->t.spin2) return EA_OK;
  const Lattice &g = e->g();
  return dalloc(&e->t.spin2, (size_t)g.ns * 2 * g.half);
}

// The host copy of the descriptor follows the swaps of spin and spin2 that
// `sweeps` kind-4 sweeps make on the device.
static void note_sweeps(ea_engine *e, const Shape &launch, long long sweeps) {
  if (sweep_swaps(launch) && (sweeps & 1)) std::swap(e->t.spin, e->t.spin2);
}

static Target engine_target(ea_engine *e) { return {e->dt, &e->plan, 1, e->stream, e}; }

// Makes ener hold the current energies by clone: sums the last sweep's
// partials, or recomputes them from the spins.  Needs a call in flight.
static int ensure_energy(ea_engine *e) {
  if (e->part_pending) {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_resolve(e->dt, e->plan, 0, e->stream));
  } else if (!e->ener_valid) {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_energy(e->dt, e->plan, 0, e->stream));
  }
  e->part_pending = false;
  e->ener_valid = true;
  return EA_OK;
}

// Enqueues the 2*nsweeps half-sweeps of the call in flight (timed by
// ea_set_timing when `timed`).
static int enqueue_sweeps(ea_engine *e, int nsweeps, bool timed) {
  const int per = sweep_is_whole(e->plan.shape) ? 2 : 1;  // halves per launch
  for (int half = 0; half < 2 * nsweeps; half += per) {
    Phase ph{half, half + per == 2 * nsweeps, 0, 0, 0};
    Timed t(timed ? e : nullptr, EA_K_SWEEP);
    HIP_TRY(launch_sweep(e->dt, e->plan, ph, e->stream));
  }
  note_sweeps(e, e->plan.shape, nsweeps);
  e->part_pending = true;
  e->ener_valid = false;
  return EA_OK;
}

extern "C" int ea_sweep(ea_engine *e, int nsweeps, const ea_counter *blocks) {
  if (!e || (nsweeps > 0 && !blocks)) return fail(EA_E_ARG, "NULL argument");
  if (nsweeps < 0) return fail(EA_E_ARG, "nsweeps < 0");
  if (nsweeps == 0) return EA_OK;
  RC_TRY(begin_calls(e, 1, 2 * nsweeps, 0, 0, 1, blocks, (size_t)nsweeps * 2 * e->g().nr));
  return enqueue_sweeps(e, nsweeps, true);
}

extern "C" int ea_sweep_noalloc(ea_engine *e, int nsweeps, const ea_counter *blocks) {
  if (!e || (nsweeps > 0 && !blocks)) return fail(EA_E_ARG, "NULL argument");
  if (nsweeps < 0) return fail(EA_E_ARG, "nsweeps < 0");
  const size_t nctr = (size_t)nsweeps * 2 * e->g().nr;
  if (nctr > EA_NOALLOC_COUNTERS)
    return fail(EA_E_ARG,
                "ea_sweep_noalloc: %d sweeps of %d replicas need %zu counter blocks, more than "
                "the %d of the shared buffer",
                nsweeps, e->g().nr, nctr, EA_NOALLOC_COUNTERS);
  if (nsweeps == 0) return EA_OK;
  if (!e->noalloc_ctr) return fail(EA_E_STATE, "no counter buffer on this engine's device");
  // the engine's previous kernels may still read the buffer
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(e->noalloc_ctr, blocks, sizeof(ea_counter) * nctr, hipMemcpyHostToDevice));
  // the descriptor points the kernels at the shared buffer; staging it
  // without inputs stays within the call block allocated at creation
  const ea_counter *own = e->t.ctr;
  e->t.ctr = e->noalloc_ctr;
  const int rc = begin_calls(e, 1, 2 * nsweeps, 0, 0, 1);
  e->t.ctr = own;
  RC_TRY(rc);
  return enqueue_sweeps(e, nsweeps, false);
}

extern "C" int ea_energy(ea_engine *e, int32_t *out) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const Lattice &g = e->g();
  RC_TRY(begin_calls(e, 1, 0, 0, 0, 1));
  {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_energy(e->dt, e->plan, 0, e->stream));
  }
  e->part_pending = false;
  e->ener_valid = true;
  if (out) {
    HIP_TRY(hipMemcpyAsync(out, e->t.ener, sizeof(int32_t) * g.ns * g.nr * g.nt,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  }
  return EA_OK;
}

extern "C" int ea_overlap(ea_engine *e, int32_t *out) {
  if (!e || !out) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g();
  const size_t cnt = (size_t)g.ns * (g.nr * (g.nr - 1) / 2) * g.nt;
  if (!e->t.ovl) RC_TRY(dalloc(&e->t.ovl, cnt));
  RC_TRY(begin_calls(e, 1, 0, 0, 0, 1));  // uploads the descriptor with the buffer
  {
    Timed t(e, EA_K_MEASURE);
    HIP_TRY(launch_overlap(e->dt, e->plan, 0, e->stream));
  }
  HIP_TRY(hipMemcpyAsync(out, e->t.ovl, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return EA_OK;
}

extern "C" int ea_pt(ea_engine *e) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  RC_TRY(begin_calls(e, 1, 0, 0, 0, 1));
  Phase ph{0, e->part_pending ? 1 : 0, 0, 0, 0};
  if (!e->part_pending) RC_TRY(ensure_energy(e));
  {
    Timed t(e, EA_K_PT);
    HIP_TRY(launch_pt(e->dt, e->plan, ph, e->stream));
  }
  e->part_pending = false;
  e->ener_valid = true;
  return EA_OK;
}

static int check_pairs(const ea_engine *e, const int32_t *pairs) {
  const int half = e->g().nr / 2;
  for (int k = 0; k < 4; ++k) {
    const int lo = (k / 2) * half;
    if (pairs[k] < lo || pairs[k] >= lo + half)
      return fail(EA_E_ARG, "replica pair entry %d = %d outside group %d", k, pairs[k], k / 2);
  }
  if (pairs[0] == pairs[1] || pairs[2] == pairs[3])
    return fail(EA_E_ARG, "a Houdayer pair needs two distinct replicas");
  return EA_OK;
}

static uint64_t coin_per_call(const Lattice &g) { return 2ull * g.V * g.nt * g.ns; }

// The block maps follow the tube's cluster flag.
static int set_cluster(ea_engine *e, int cluster) {
  if (e->t.cluster == cluster) return EA_OK;
  e->t.cluster = cluster;
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(upload_plan(make_plan({e->t}, e->shape), &e->plan));
  return EA_OK;
}

static int check_shape(const ea_shape *sh) {
  if (!sh) return fail(EA_E_ARG, "NULL argument");
  if (sh->sweep_positions != 32 && sh->sweep_positions != 64)
    return fail(EA_E_ARG, "sweep_positions must be 32 or 64 (got %d)", sh->sweep_positions);
  if (sh->cluster_threads < 256 || sh->cluster_threads > 1024 || sh->cluster_threads % 64)
    return fail(EA_E_ARG, "cluster_threads must be a multiple of 64 in 256..1024 (got %d)",
                sh->cluster_threads);
  if (sh->sweep_kind < 0 || sh->sweep_kind > 4)
    return fail(EA_E_ARG, "sweep_kind must be 0, 1, 2, 3 or 4 (got %d)", sh->sweep_kind);
  if (sh->sweep_kind != 0 && sh->sweep_positions != 64)
    return fail(EA_E_ARG, "sweep_kind %d takes sweep_positions 64", sh->sweep_kind);
  if (sh->sweep_xcd != 0 && sh->sweep_xcd != 1)
    return fail(EA_E_ARG, "sweep_xcd must be 0 or 1 (got %d)", sh->sweep_xcd);
  return EA_OK;
}

// Why a tube does not fit a sweep kind (shape_fits).
static const char *shape_misfit(int kind) {
  if (kind == 4)
    return "no slab thickness T divides Lz/4 with Lz/4 >= T + 4, (T + 2) L*L/32 <= 1024 "
           "and (2T + 6) L*L/32 <= 3072";
  return kind == 3 ? "the clones of a one-launch sweep need n = L*L*Lz/128 <= 512"
                   : "the tile (256 + L*L/16 half sites) exceeds 768";
}


static Shape to_shape(const ea_shape *sh) {
  Shape x;
  x.sweep_positions = sh->sweep_positions;
  x.cluster_threads = sh->cluster_threads;
  x.sweep_kind = sh->sweep_kind;
  x.sweep_xcd = sh->sweep_xcd;
  return x;
}

// Partials by sweep block change layout with the shape: pending ones are
// summed first, under the old shape.
extern "C" int ea_set_shape(ea_engine *e, const ea_shape *shape) {
  if (!e) return fail(EA_E_ARG, "NULL engine");
  RC_TRY(check_shape(shape));
  if (!shape_fits(e->g(), to_shape(shape)))
    return fail(EA_E_ARG, "sweep_kind %d: %s", shape->sweep_kind, shape_misfit(shape->sweep_kind));
  if (sweep_swaps(to_shape(shape))) RC_TRY(ensure_spin2(e));
  if (e->part_pending) {
    // the resolve launch exits unless a call is in flight (a previous
    // ea_run with n_it = 0 leaves none)
    RC_TRY(begin_calls(e, 1, 0, 0, 0, 1));
    RC_TRY(ensure_energy(e));
  }
  e->shape = to_shape(shape);
  e->t.nbx = sweep_partials_x(e->g(), e->shape);
  RC_TRY(upload_tube(e));
  HIP_TRY(upload_plan(make_plan({e->t}, e->shape), &e->plan));
  return EA_OK;
}

extern "C" int ea_houdayer(ea_engine *e, const int32_t pairs[4], uint64_t coin_seed,
                           uint64_t *rng_offset) {
  if (!e || !pairs || !rng_offset) return fail(EA_E_ARG, "NULL argument");
  if (e->g().nr < 4 || e->g().ctm <= 0)
    return fail(EA_E_STATE, "the Houdayer move needs nr >= 4 and clone_threshold > 0");
  RC_TRY(check_pairs(e, pairs));
  RC_TRY(set_cluster(e, 1));
  e->t.coin_seed = coin_seed;
  e->t.coin0 = *rng_offset;
  e->t.coin_per_call = coin_per_call(e->g());
  RC_TRY(begin_calls(e, 1, 0, 1, 0, 1, nullptr, 0, pairs, 4));
  RC_TRY(ensure_energy(e));  // the move updates the two changed chains' energies
  {
    Timed t(e, EA_K_HOUDAYER);
    HIP_TRY(launch_houdayer(e->dt, e->plan, Phase{0, 0, 0, 0, 0}, e->stream));
  }
  *rng_offset += e->t.coin_per_call;
  return EA_OK;
}

extern "C" int ea_keep_labels(ea_engine *e, int on) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  if (on && !e->keep_buf) {
    const size_t n = (size_t)2 * e->g().ns * e->g().ctm * e->g().V;
    RC_TRY(dalloc(&e->keep_buf, std::max<size_t>(n, 1)));
  }
  e->keep_on = on != 0;
  e->t.keep = e->keep_on ? e->keep_buf : nullptr;
  return upload_tube(e);
}

extern "C" int ea_download_labels(ea_engine *e, uint32_t *labels) {
  if (!e || !labels) return fail(EA_E_ARG, "NULL argument");
  if (!e->keep_buf) return fail(EA_E_STATE, "labels are kept only after ea_keep_labels(e, 1)");
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(labels, e->keep_buf,
                    sizeof(uint32_t) * 2 * e->g().ns * e->g().ctm * e->g().V,
                    hipMemcpyDeviceToHost));
  return EA_OK;
}

static int check_schedule(const ea_engine *e, const ea_schedule *sch) {
  if (sch->n_it < 0 || sch->mesfr < 1 || sch->houfr < 1 || sch->ptfr < 1 ||
      sch->frec_trw < 1 || sch->it0 < 0 || (sch->n_it > 0 && !sch->blocks))
    return fail(EA_E_ARG, "invalid schedule");
  if (sch->cluster) {
    if (!sch->pairs || e->g().nr < 4 || e->g().ctm <= 0)
      return fail(EA_E_ARG, "the Houdayer move needs pairs, nr >= 4 and clone_threshold > 0");
    for (int k = 0; k < sch->n_it * sch->mesfr; ++k) RC_TRY(check_pairs(e, sch->pairs + 4 * k));
  }
  return EA_OK;
}

// Loads a schedule's inputs into the engine's calls in flight.
static int load_schedule(ea_engine *e, const ea_schedule *sch) {
  const Lattice &g = e->g();
  const int halves = 2 * sch->ptfr * sch->houfr * sch->mesfr;
  RC_TRY(ensure_hist(e, (sch->it0 + sch->n_it) / sch->frec_trw));
  RC_TRY(set_cluster(e, sch->cluster ? 1 : 0));
  e->t.coin_seed = sch->coin_seed;
  e->t.coin0 = sch->rng_offset;
  e->t.coin_per_call = coin_per_call(g);
  return begin_calls(e, sch->n_it, halves, sch->mesfr, sch->it0, sch->frec_trw, sch->blocks,
                     (size_t)sch->n_it * halves * g.nr, sch->cluster ? sch->pairs : nullptr,
                     sch->cluster ? (size_t)sch->n_it * sch->mesfr * 4 : 0);
}

static void finish_schedule(ea_engine *e, const ea_schedule *sch, const Shape &launch,
                            uint64_t *rng_offset_out) {
  note_sweeps(e, launch, (long long)sch->n_it * sch->mesfr * sch->houfr * sch->ptfr);
  if (sch->n_it > 0) {
    e->part_pending = false;
    e->ener_valid = true;
  }
  if (rng_offset_out)
    *rng_offset_out = sch->rng_offset + (sch->cluster ? (uint64_t)sch->n_it * sch->mesfr *
                                                            e->t.coin_per_call
                                                      : 0);
}

extern "C" int ea_run(ea_engine *e, const ea_schedule *sch, uint64_t *rng_offset_out) {
  if (!e || !sch) return fail(EA_E_ARG, "NULL argument");
  RC_TRY(check_schedule(e, sch));
  RC_TRY(load_schedule(e, sch));
  StepShape sh;
  sh.mesfr = sch->mesfr;
  sh.houfr = sch->houfr;
  sh.ptfr = sch->ptfr;
  sh.cluster = sch->cluster != 0;
  const Target x = engine_target(e);
  // the launches carry their step index; the device step counter (read by
  // replayed batch steps only) is staged afresh by every call
  for (int it = 0; it < sch->n_it; ++it) RC_TRY(enqueue_step(x, sh, it));
  finish_schedule(e, sch, e->plan.shape, rng_offset_out);
  return EA_OK;
}

extern "C" int ea_read_measures(ea_engine *e, double *esum, uint8_t *history, int max_rows) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const Lattice &g = e->g();
  const size_t per = (size_t)g.ns * g.nr * g.nt;
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (esum) HIP_TRY(hipMemcpy(esum, e->t.esum, sizeof(double) * per, hipMemcpyDeviceToHost));
  if (history && e->t.hist) {
    const int rows = std::min(max_rows, e->hist_rows);
    if (rows > 0) HIP_TRY(hipMemcpy(history, e->t.hist, per * rows, hipMemcpyDeviceToHost));
  }
  HIP_TRY(launch_fill_f64(e->t.esum, 0.0, per, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return EA_OK;
}

extern "C" int ea_set_timing(ea_engine *e, int on) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  e->timing = on != 0;
  return EA_OK;
}

extern "C" int ea_read_timing(ea_engine *e, double *ms, int64_t *launches) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (auto &m : e->marks) {
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, m.a, m.b));
    e->ms[m.kind] += t;
    e->launches[m.kind] += 1;
    e->spare.push_back(m.a);
    e->spare.push_back(m.b);
  }
  e->marks.clear();
  for (int k = 0; k < EA_K_KINDS; ++k) {
    if (ms) ms[k] = e->ms[k];
    if (launches) launches[k] = e->launches[k];
    e->ms[k] = 0;
    e->launches[k] = 0;
  }
  return EA_OK;
}

// ------------------------------------------------------------------ batch

extern "C" int ea_batch_create(ea_engine *const *engines, int n, ea_batch **out) {
  if (!out || !engines || n < 1) return fail(EA_E_ARG, "invalid argument");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!engines[i]) return fail(EA_E_ARG, "engine %d is NULL", i);
    for (int j = 0; j < i; ++j)
      if (engines[j] == engines[i]) return fail(EA_E_ARG, "engine %d appears twice", i);
    if (engines[i]->device != engines[0]->device)
      return fail(EA_E_ARG, "engines of one batch must share a device");
  }
  auto b = std::make_unique<ea_batch>();
  b->eng.assign(engines, engines + n);
  b->device = engines[0]->device;
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
  RC_TRY(dalloc(&b->dt, (size_t)n));
  *out = b.release();
  return EA_OK;
}

extern "C" int ea_batch_destroy(ea_batch *b) {
  if (!b) return EA_OK;
  (void)hipStreamSynchronize(b->stream);
  delete b;
  return EA_OK;
}

static int drop_graph(ea_batch *b) {
  if (b->exec) HIP_TRY(hipGraphExecDestroy(b->exec));
  if (b->graph) HIP_TRY(hipGraphDestroy(b->graph));
  b->exec = nullptr;
  b->graph = nullptr;
  return EA_OK;
}

// Takes effect at the next ea_batch_run (the plan and graph are rebuilt).
extern "C" int ea_batch_set_shape(ea_batch *b, const ea_shape *shape) {
  if (!b) return fail(EA_E_ARG, "NULL batch");
  RC_TRY(check_shape(shape));
  for (size_t i = 0; i < b->eng.size(); ++i)
    if (!shape_fits(b->eng[i]->g(), to_shape(shape)))
      return fail(EA_E_ARG, "sweep_kind %d: tube %zu: %s", shape->sweep_kind, i,
                  shape_misfit(shape->sweep_kind));
  if (sweep_swaps(to_shape(shape)))
    for (ea_engine *e : b->eng) RC_TRY(ensure_spin2(e));
  b->launch = to_shape(shape);
  return EA_OK;
}

extern "C" int ea_batch_run(ea_batch *b, const ea_schedule *sch, uint64_t *rng_offsets_out) {
  if (!b || !sch) return fail(EA_E_ARG, "NULL argument");
  const int n = (int)b->eng.size();
  StepShape sh;
  sh.mesfr = sch[0].mesfr;
  sh.houfr = sch[0].houfr;
  sh.ptfr = sch[0].ptfr;
  int steps = 0;
  for (int i = 0; i < n; ++i) {
    const ea_schedule &s = sch[i];
    if (s.mesfr != sh.mesfr || s.houfr != sh.houfr || s.ptfr != sh.ptfr)
      return fail(EA_E_ARG, "tube %d: the tubes of a batch share mesfr, houfr and ptfr", i);
    RC_TRY(check_schedule(b->eng[i], &s));
    steps = std::max(steps, s.n_it);
  }
  std::vector<Tube> tubes(n);
  std::vector<int> cluster(n);
  for (int i = 0; i < n; ++i) {
    ea_engine *e = b->eng[i];
    RC_TRY(load_schedule(e, &sch[i]));
    tubes[i] = e->t;
    tubes[i].nbx = sweep_partials_x(e->g(), b->launch);
    cluster[i] = e->t.cluster;
    sh.cluster = sh.cluster || e->t.cluster;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpy(b->dt, tubes.data(), sizeof(Tube) * n, hipMemcpyHostToDevice));
  // the engines' step counters and inputs reach their call blocks on the
  // engines' own streams
  for (int i = 0; i < n; ++i) HIP_TRY(hipStreamWaitEvent(b->stream, b->eng[i]->staged, 0));
  if (cluster != b->plan_cluster || !b->plan.ready() || !(b->plan.shape == b->launch)) {
    RC_TRY(drop_graph(b));
    HIP_TRY(upload_plan(make_plan(tubes, b->launch), &b->plan));
    b->plan_cluster = cluster;
  }
  if (!b->exec || !(sh == b->shape)) {
    RC_TRY(drop_graph(b));
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const Target x{b->dt, &b->plan, n, b->stream, nullptr, b->side, b->fork, b->join};
    const int rc = enqueue_step(x, sh);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(ce);
    b->graph = graph;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->shape = sh;
  }
  for (int k = 0; k < steps; ++k) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i)
    finish_schedule(b->eng[i], &sch[i], b->launch, rng_offsets_out ? rng_offsets_out + i : nullptr);
  return EA_OK;
}
