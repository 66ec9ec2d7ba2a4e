// engine.hip -- the C ABI of include/eatube.h over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
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

}  // namespace

struct ea_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Lattice g{};
  uint32_t job = 0;
  std::vector<double> betas;
  uint32_t keys[EA_MAX_REPLICAS] = {};

  uint64_t *spin = nullptr;   // [ns][2][nt][n][nr]
  uint64_t *J = nullptr;      // [ns][2][n][6]
  uint8_t *wctb = nullptr;    // [ns][nr][nt]
  uint8_t *lut = nullptr;     // [nt][EA_LUT_BYTES]
  int32_t *ener = nullptr;    // [ns][nr][nt]
  int32_t *part = nullptr;    // partial energies (see energy_partials)
  // energy bookkeeping: the last sweep left partials by beta that no kernel
  // has summed yet (part_pending), or ener holds the current energies
  bool part_pending = false, ener_valid = false;
  ea_hq_state *pt = nullptr;  // [ns][nr]
  double *acc = nullptr, *att = nullptr, *raw = nullptr;
  double *nd = nullptr, *tab = nullptr;
  int32_t *toff = nullptr, *tlen = nullptr;
  uint32_t *hlab = nullptr, *hcb = nullptr;  // tiled Houdayer labels / coins
  int planes = 0;
  uint32_t *keep = nullptr;
  int keep_on = 0;
  double *esum = nullptr;
  uint8_t *hist = nullptr;
  int hist_rows = 0;
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
    void *ptrs[] = {spin, J,   wctb, lut,  ener, part, pt,   acc,  att, raw,
                    nd,   tab, toff, tlen, hlab, hcb,  keep, esum, hist};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

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

static int check_config(const ea_config *c) {
  if (!c) return fail(EA_E_ARG, "config is NULL");
  if (!valid_lattice(c->L, c->Lz))
    return fail(EA_E_ARG, "L=%d Lz=%d: both must be multiples of 8", c->L, c->Lz);
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
  Lattice &g = e->g;
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
  e->job = cfg->job_id;
  e->betas.assign(cfg->betas, cfg->betas + cfg->nt);

  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  const size_t chains = (size_t)g.ns * g.nr;
  if ((rc = dalloc(&e->spin, (size_t)g.ns * 2 * g.half)) ||
      (rc = dalloc(&e->J, (size_t)g.ns * 12 * g.n)) ||
      (rc = dalloc(&e->wctb, chains * g.nt)) ||
      (rc = dalloc(&e->lut, (size_t)g.nt * EA_LUT_BYTES)) ||
      (rc = dalloc(&e->ener, chains * g.nt)) || (rc = dalloc(&e->pt, chains)) ||
      (rc = dalloc(&e->part, energy_partials(g))) ||
      (rc = dalloc(&e->acc, chains * (g.nt - 1))) ||
      (rc = dalloc(&e->att, chains * (g.nt - 1))) ||
      (rc = dalloc(&e->raw, chains * (g.nt - 1))) ||
      (rc = dalloc(&e->esum, chains * g.nt)))
    return rc;
  HIP_TRY(hipMemset(e->spin, 0, sizeof(uint64_t) * g.ns * 2 * g.half));
  HIP_TRY(hipMemset(e->J, 0, sizeof(uint64_t) * g.ns * 12 * g.n));
  HIP_TRY(hipMemcpy(e->lut, cfg->lut, (size_t)g.nt * EA_LUT_BYTES, hipMemcpyHostToDevice));
  std::vector<uint8_t> ident(chains * g.nt);
  for (size_t c = 0; c < chains; ++c)
    for (int b = 0; b < g.nt; ++b) ident[c * g.nt + b] = (uint8_t)b;
  HIP_TRY(hipMemcpy(e->wctb, ident.data(), ident.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(e->ener, 0, sizeof(int32_t) * chains * g.nt));
  HIP_TRY(hipMemset(e->pt, 0, sizeof(ea_hq_state) * chains));
  HIP_TRY(hipMemset(e->acc, 0, sizeof(double) * chains * (g.nt - 1)));
  HIP_TRY(hipMemset(e->att, 0, sizeof(double) * chains * (g.nt - 1)));
  HIP_TRY(hipMemset(e->raw, 0, sizeof(double) * chains * (g.nt - 1)));
  HIP_TRY(hipMemset(e->esum, 0, sizeof(double) * chains * g.nt));

  // exact exp(pt) tables for the device parallel tempering
  const ExpTables t = make_exp_tables(cfg->betas, g.nt, g.V);
  if ((rc = dalloc(&e->nd, t.nd.size())) || (rc = dalloc(&e->toff, t.off.size())) ||
      (rc = dalloc(&e->tlen, t.len.size())) || (rc = dalloc(&e->tab, std::max<size_t>(t.val.size(), 1))))
    return rc;
  HIP_TRY(hipMemcpy(e->nd, t.nd.data(), sizeof(double) * t.nd.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->toff, t.off.data(), sizeof(int32_t) * t.off.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->tlen, t.len.data(), sizeof(int32_t) * t.len.size(), hipMemcpyHostToDevice));
  if (!t.val.empty())
    HIP_TRY(hipMemcpy(e->tab, t.val.data(), sizeof(double) * t.val.size(), hipMemcpyHostToDevice));

  // Houdayer: a tube whose labels fit one workgroup's LDS is labelled there,
  // larger tubes slab by slab with a merge in HBM
  if (g.ctm > 0) {
    if (g.L > kHouMaxL)
      return fail(EA_E_ARG, "the Houdayer move supports L <= %d (L=%d)", kHouMaxL, g.L);
    int lds_block = 0, lds_optin = 0;
    HIP_TRY(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIP_TRY(hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, dev));
    const int lds_max = std::max(lds_block, lds_optin);
    if (hou_whole_lds(g) > (size_t)lds_max) {
      int planes = std::max(1, std::min(g.Lz, kHouSlabSites / (g.L * g.L)));
      while (planes > 1 && hou_slab_lds(g, planes) > (size_t)lds_max) --planes;
      e->planes = planes;
      const size_t inst = (size_t)2 * g.ns * g.ctm;
      if ((rc = dalloc(&e->hlab, inst * g.V)) || (rc = dalloc(&e->hcb, inst * (g.V / 32))))
        return rc;
    }
  }
  *out = e.release();
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
  std::memcpy(e->keys, keys, sizeof(uint32_t) * e->g.nr);
  return EA_OK;
}

extern "C" int ea_upload_sample(ea_engine *e, int s, const uint64_t *even,
                                const uint64_t *odd, const uint64_t *J) {
  if (!e || !even || !odd || !J) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g;
  if (s < 0 || s >= g.ns) return fail(EA_E_ARG, "sample %d out of range", s);
  HIP_TRY(hipStreamSynchronize(e->stream));
  uint64_t *base = e->spin + (size_t)s * 2 * g.half;
  HIP_TRY(hipMemcpy(base, even, sizeof(uint64_t) * g.half, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + g.half, odd, sizeof(uint64_t) * g.half, hipMemcpyHostToDevice));
  // [parity][direction][h] -> [parity][h][direction]
  std::vector<uint64_t> t((size_t)12 * g.n);
  for (int p = 0; p < 2; ++p)
    for (int d = 0; d < 6; ++d)
      for (int h = 0; h < g.n; ++h)
        t[((size_t)p * g.n + h) * 6 + d] = J[((size_t)p * 6 + d) * g.n + h];
  HIP_TRY(hipMemcpy(e->J + (size_t)s * 12 * g.n, t.data(), sizeof(uint64_t) * t.size(),
                    hipMemcpyHostToDevice));
  e->part_pending = e->ener_valid = false;
  return EA_OK;
}

extern "C" int ea_download_spins(ea_engine *e, int s, uint64_t *even, uint64_t *odd) {
  if (!e || !even || !odd) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g;
  if (s < 0 || s >= g.ns) return fail(EA_E_ARG, "sample %d out of range", s);
  HIP_TRY(hipStreamSynchronize(e->stream));
  const uint64_t *base = e->spin + (size_t)s * 2 * g.half;
  HIP_TRY(hipMemcpy(even, base, sizeof(uint64_t) * g.half, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(odd, base + g.half, sizeof(uint64_t) * g.half, hipMemcpyDeviceToHost));
  return EA_OK;
}

extern "C" int ea_upload_perms(ea_engine *e, const uint8_t *wctb) {
  if (!e || !wctb) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g;
  for (int c = 0; c < g.ns * g.nr; ++c) {
    unsigned seen[EA_MAX_BETAS + 1] = {};
    for (int b = 0; b < g.nt; ++b) {
      const int v = wctb[(size_t)c * g.nt + b];
      if (v >= g.nt || seen[v]++)
        return fail(EA_E_ARG, "chain %d: not a permutation of 0..nt-1", c);
    }
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(e->wctb, wctb, (size_t)g.ns * g.nr * g.nt, hipMemcpyHostToDevice));
  e->part_pending = e->ener_valid = false;  // partials by beta map through the old perms
  return EA_OK;
}

extern "C" int ea_download_perms(ea_engine *e, uint8_t *wctb) {
  if (!e || !wctb) return fail(EA_E_ARG, "NULL argument");
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(wctb, e->wctb, (size_t)e->g.ns * e->g.nr * e->g.nt,
                    hipMemcpyDeviceToHost));
  return EA_OK;
}

extern "C" int ea_upload_pt_state(ea_engine *e, const ea_hq_state *st, const double *acc,
                                  const double *att, const double *raw) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const size_t chains = (size_t)e->g.ns * e->g.nr, np = chains * (e->g.nt - 1);
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (st) HIP_TRY(hipMemcpy(e->pt, st, sizeof(ea_hq_state) * chains, hipMemcpyHostToDevice));
  if (acc) HIP_TRY(hipMemcpy(e->acc, acc, sizeof(double) * np, hipMemcpyHostToDevice));
  if (att) HIP_TRY(hipMemcpy(e->att, att, sizeof(double) * np, hipMemcpyHostToDevice));
  if (raw) HIP_TRY(hipMemcpy(e->raw, raw, sizeof(double) * np, hipMemcpyHostToDevice));
  return EA_OK;
}

extern "C" int ea_download_pt_state(ea_engine *e, ea_hq_state *st, double *acc,
                                    double *att, double *raw) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const size_t chains = (size_t)e->g.ns * e->g.nr, np = chains * (e->g.nt - 1);
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (st) HIP_TRY(hipMemcpy(st, e->pt, sizeof(ea_hq_state) * chains, hipMemcpyDeviceToHost));
  if (acc) HIP_TRY(hipMemcpy(acc, e->acc, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (att) HIP_TRY(hipMemcpy(att, e->att, sizeof(double) * np, hipMemcpyDeviceToHost));
  if (raw) HIP_TRY(hipMemcpy(raw, e->raw, sizeof(double) * np, hipMemcpyDeviceToHost));
  return EA_OK;
}

// ------------------------------------------------------------ the step ops

// Brackets the launches of one kernel kind with events when timing is on.
struct Timed {
  ea_engine *e;
  int kind;
  hipEvent_t a = nullptr;
  Timed(ea_engine *e_, int kind_) : e(e_), kind(kind_) {
    if (e->timing) {
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

static int enqueue_sweeps(ea_engine *e, int nsweeps, const ea_counter *blocks,
                          bool energy) {
  SweepArgs a{};
  a.spin = e->spin;
  a.J = e->J;
  a.wctb = e->wctb;
  a.lut = e->lut;
  a.part = e->part;
  std::memcpy(a.keys, e->keys, sizeof a.keys);
  const int nr = e->g.nr;
  for (int sw = 0; sw < nsweeps; ++sw) {
    const bool last = energy && sw == nsweeps - 1;
    for (int p = 0; p < 2; ++p) {
      a.parity = p;
      a.energy = last && p == 1;
      std::memcpy(a.ctr, blocks + ((size_t)sw * 2 + p) * nr, sizeof(ea_counter) * nr);
      Timed t(e, EA_K_SWEEP);
      HIP_TRY(launch_sweep(e->g, a, e->stream));
    }
  }
  if (nsweeps > 0) {
    e->part_pending = energy;
    e->ener_valid = false;
  }
  return EA_OK;
}

// Makes ener hold the current energies by clone: sums the last sweep's
// partials, or recomputes them from the spins.
static int ensure_energy(ea_engine *e) {
  if (e->part_pending) {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_resolve(e->g, e->part, e->wctb, e->ener, e->stream));
  } else if (!e->ener_valid) {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_energy(e->g, e->spin, e->J, e->ener, e->part, e->stream));
  }
  e->part_pending = false;
  e->ener_valid = true;
  return EA_OK;
}

// measure: add the per-beta energies after the swaps to esum and, for
// row >= 0, write that history row.
static int enqueue_pt(ea_engine *e, bool measure = false, int row = -1) {
  PtArgs a{};
  if (e->part_pending) {
    a.part = e->part;
    a.nbx = sweep_blocks_x(e->g);
  } else {
    const int rc = ensure_energy(e);
    if (rc) return rc;
  }
  a.st = e->pt;
  a.ener = e->ener;
  a.wctb = e->wctb;
  a.acc = e->acc;
  a.att = e->att;
  a.raw = e->raw;
  a.nd = e->nd;
  a.tab = e->tab;
  a.off = e->toff;
  a.len = e->tlen;
  a.esum = e->esum;
  a.hist = e->hist;
  a.measure = measure;
  a.row = row;
  Timed t(e, EA_K_PT);
  HIP_TRY(launch_pt(e->g, a, e->stream));
  e->part_pending = false;
  e->ener_valid = true;
  return EA_OK;
}

// measure (whole-tube path only): correct the per-beta sums a measuring PT
// added just before for the configurations the move changes.
static int enqueue_houdayer(ea_engine *e, const int32_t pairs[4], uint64_t seed,
                            uint64_t offset, bool measure = false) {
  const int half = e->g.nr / 2;
  for (int k = 0; k < 4; ++k) {
    const int lo = (k / 2) * half;
    if (pairs[k] < lo || pairs[k] >= lo + half)
      return fail(EA_E_ARG, "replica pair entry %d = %d outside group %d", k, pairs[k], k / 2);
  }
  if (pairs[0] == pairs[1] || pairs[2] == pairs[3])
    return fail(EA_E_ARG, "a Houdayer pair needs two distinct replicas");
  if (e->planes == 0) {  // the whole-tube kernel updates the energies it moves
    const int rc = ensure_energy(e);
    if (rc) return rc;
  }
  HouArgs a{};
  a.spin = e->spin;
  a.J = e->J;
  a.ener = e->ener;
  a.wctb = e->wctb;
  std::memcpy(a.pairs, pairs, sizeof a.pairs);
  a.seed = seed;
  a.offset0 = offset;
  a.lab = e->hlab;
  a.cbits = e->hcb;
  a.labels_out = e->keep_on ? e->keep : nullptr;
  a.planes = e->planes;
  a.esum = e->esum;
  a.measure = measure && e->planes == 0;
  Timed t(e, EA_K_HOUDAYER);
  HIP_TRY(launch_houdayer(e->g, a, e->stream));
  if (e->planes > 0) e->part_pending = e->ener_valid = false;
  return EA_OK;
}

extern "C" int ea_sweep(ea_engine *e, int nsweeps, const ea_counter *blocks) {
  if (!e || (nsweeps > 0 && !blocks)) return fail(EA_E_ARG, "NULL argument");
  if (nsweeps < 0) return fail(EA_E_ARG, "nsweeps < 0");
  return enqueue_sweeps(e, nsweeps, blocks, true);
}

extern "C" int ea_energy(ea_engine *e, int32_t *out) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const Lattice &g = e->g;
  {
    Timed t(e, EA_K_ENERGY);
    HIP_TRY(launch_energy(g, e->spin, e->J, e->ener, e->part, e->stream));
  }
  e->part_pending = false;
  e->ener_valid = true;
  if (out) {
    HIP_TRY(hipMemcpyAsync(out, e->ener, sizeof(int32_t) * g.ns * g.nr * g.nt,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  }
  return EA_OK;
}

extern "C" int ea_pt(ea_engine *e) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  return enqueue_pt(e);
}

extern "C" int ea_houdayer(ea_engine *e, const int32_t pairs[4], uint64_t coin_seed,
                           uint64_t *rng_offset) {
  if (!e || !pairs || !rng_offset) return fail(EA_E_ARG, "NULL argument");
  if (e->g.nr < 4) return fail(EA_E_STATE, "the Houdayer move needs nr >= 4");
  const int rc = enqueue_houdayer(e, pairs, coin_seed, *rng_offset);
  if (rc) return rc;
  *rng_offset += 2ull * e->g.V * e->g.nt * e->g.ns;
  return EA_OK;
}

extern "C" int ea_keep_labels(ea_engine *e, int on) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  if (on && !e->keep) {
    const size_t n = (size_t)2 * e->g.ns * e->g.ctm * e->g.V;
    const int rc = dalloc(&e->keep, std::max<size_t>(n, 1));
    if (rc) return rc;
  }
  e->keep_on = on != 0;
  return EA_OK;
}

extern "C" int ea_download_labels(ea_engine *e, uint32_t *labels) {
  if (!e || !labels) return fail(EA_E_ARG, "NULL argument");
  if (!e->keep) return fail(EA_E_STATE, "labels are kept only after ea_keep_labels(e, 1)");
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(labels, e->keep, sizeof(uint32_t) * 2 * e->g.ns * e->g.ctm * e->g.V,
                    hipMemcpyDeviceToHost));
  return EA_OK;
}

extern "C" int ea_run(ea_engine *e, const ea_schedule *sch, uint64_t *rng_offset_out) {
  if (!e || !sch) return fail(EA_E_ARG, "NULL argument");
  const Lattice &g = e->g;
  if (sch->n_it < 0 || sch->mesfr < 1 || sch->houfr < 1 || sch->ptfr < 1 ||
      sch->frec_trw < 1 || sch->it0 < 0 || !sch->blocks)
    return fail(EA_E_ARG, "invalid schedule");
  if (sch->cluster && (!sch->pairs || g.nr < 4))
    return fail(EA_E_ARG, "the Houdayer move needs pairs and nr >= 4");
  const int rows = (sch->it0 + sch->n_it) / sch->frec_trw;
  if (rows > e->hist_rows) {
    if (e->hist) HIP_TRY(hipFree(e->hist));
    e->hist = nullptr;
    e->hist_rows = 0;
    const int rc = dalloc(&e->hist, (size_t)rows * g.ns * g.nr * g.nt);
    if (rc) return rc;
    HIP_TRY(hipMemset(e->hist, 0, (size_t)rows * g.ns * g.nr * g.nt));
    e->hist_rows = rows;
  }
  uint64_t offset = sch->rng_offset;
  const uint64_t per_call = 2ull * g.V * g.nt * g.ns;
  const size_t blocks_per_sweep = 2 * (size_t)g.nr;
  size_t bi = 0;
  // The step's energy measure (main.cu:208-228) rides on its last PT, which
  // adds the energies after the swaps; the whole-tube cluster move then
  // corrects the sums for what it changed.  The slab path recomputes the
  // energies and measures separately.
  const bool fused = !sch->cluster || e->planes == 0;
  for (int it = 0; it < sch->n_it; ++it) {
    const int itb = sch->it0 + it;
    const int row = (itb % sch->frec_trw == sch->frec_trw - 1) ? itb / sch->frec_trw : -1;
    for (int im = 0; im < sch->mesfr; ++im) {
      const bool last_m = im == sch->mesfr - 1;
      for (int ih = 0; ih < sch->houfr; ++ih) {
        int rc = enqueue_sweeps(e, sch->ptfr, sch->blocks + bi * blocks_per_sweep, true);
        if (rc) return rc;
        bi += sch->ptfr;
        const bool measure = fused && last_m && ih == sch->houfr - 1;
        if ((rc = enqueue_pt(e, measure, measure ? row : -1))) return rc;
      }
      if (sch->cluster) {
        const int rc = enqueue_houdayer(
            e, sch->pairs + ((size_t)it * sch->mesfr + im) * 4, sch->coin_seed, offset,
            fused && last_m);
        if (rc) return rc;
        offset += per_call;
      }
    }
    if (!fused) {
      int rc = ensure_energy(e);
      if (rc) return rc;
      MeasureArgs m{};
      m.ener = e->ener;
      m.wctb = e->wctb;
      m.esum = e->esum;
      m.hist = e->hist;
      m.row = row;
      Timed t(e, EA_K_MEASURE);
      HIP_TRY(launch_measure(g, m, e->stream));
    }
  }
  if (rng_offset_out) *rng_offset_out = offset;
  return EA_OK;
}

extern "C" int ea_read_measures(ea_engine *e, double *esum, uint8_t *history, int max_rows) {
  if (!e) return fail(EA_E_ARG, "engine is NULL");
  const Lattice &g = e->g;
  const size_t per = (size_t)g.ns * g.nr * g.nt;
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (esum) HIP_TRY(hipMemcpy(esum, e->esum, sizeof(double) * per, hipMemcpyDeviceToHost));
  if (history && e->hist) {
    const int rows = std::min(max_rows, e->hist_rows);
    if (rows > 0) HIP_TRY(hipMemcpy(history, e->hist, per * rows, hipMemcpyDeviceToHost));
  }
  HIP_TRY(launch_fill_f64(e->esum, 0.0, per, e->stream));
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
