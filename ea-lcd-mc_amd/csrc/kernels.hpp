// kernels.hpp -- launch interface of the gfx950 kernels (internal).
//
// Every kernel works on a table of tubes: a block reads a 64-bit entry of a
// block map (tube number << 32 | block index within the tube), then that
// tube's descriptor.  A single engine is a table of one tube; a batch
// is a table of many tubes of different sizes advancing in the same launches.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "eatube.h"

namespace eatube {

// Geometry and sizes of one tube.
struct Lattice {
  int L, Lz, ML, MLz, MS;  // MSC sub-lattice sides (L/4, Lz/4), MS = ML*ML
  int n, V;                // half MSC volume V/128, sites
  int hl, hs;              // ML/2, MS/2: strides of the half index
  int nr, nt, ns, nr_shift, bc, ctm;
  int64_t half;            // words of one parity array of one sample
  // ceil(2^32 / d) for d = L, ML, MLz, hl, hs (0 for d = 1): index
  // quotients by a multiply (set_divisors)
  uint32_t mL, mML, mMLz, mhl, mhs;
};

// Device layouts (host transfers convert to and from the reference's
// [parity][clone][h][replica] spin words and [parity][direction][h]
// couplings):
//   spins of sample s: [2 parities][nr replicas][nt clones][n] words, so the
//     words a replica's clone at one beta needs are contiguous;
//   couplings of sample s: [2 parities][3 direction pairs][n][2] words, the
//     two directions 2k, 2k+1 of a half site side by side.
inline __host__ __device__ size_t spin_row(const Lattice &g, int r, int clone) {
  return ((size_t)r * g.nt + clone) * g.n;
}
inline __host__ __device__ size_t coupling_index(const Lattice &g, int s, int p, int h, int d) {
  return ((((size_t)s * 2 + p) * 3 + (d >> 1)) * g.n + h) * 2 + (d & 1);
}

inline uint32_t divisor_magic(uint32_t d) {
  return d <= 1 ? 0u : (uint32_t)(((1ULL << 32) + d - 1) / d);
}
inline void set_divisors(Lattice &g) {
  g.mL = divisor_magic(g.L);
  g.mML = divisor_magic(g.ML);
  g.mMLz = divisor_magic(g.MLz);
  g.mhl = divisor_magic(g.hl);
  g.mhs = divisor_magic(g.hs);
}

// Whether the kernels' quotient by multiply, hi32(i * m) with
// m = ceil(2^32 / d) = (2^32 + e) / d, equals i / d for every i <= imax:
// i m / 2^32 = i / d + i e / (d 2^32), and the fraction of i / d is at most
// (d - 1) / d, so the quotient is exact when imax * e < 2^32.
inline bool qdiv_exact(uint32_t d, uint64_t imax) {
  if (d <= 1) return true;  // m = 0: the index itself
  const uint64_t e = (uint64_t)divisor_magic(d) * d - (1ULL << 32);
  return imax * e < (1ULL << 32);
}

// The first divisor of an L x L x Lz tube whose quotients the kernels would
// get wrong over that divisor's index range, or NULL when all are exact.
// Every kernel divides by L/8 the remainder within a plane of half sites
// (< L*L/32) and by L*L/32 the half-site index (< V/128); the Houdayer
// kernels (`cluster`) also divide site and row indices (< V) by L, the y
// coordinate (< L) by L/4 and the z coordinate (< Lz) by Lz/4.
inline const char *inexact_divisor(int L, int Lz, bool cluster) {
  const uint64_t V = (uint64_t)L * L * Lz;
  const uint32_t ML = L / 4, hs = ML * ML / 2;
  if (!qdiv_exact(ML / 2, hs - 1)) return "L/8";
  if (!qdiv_exact(hs, V / 128 - 1)) return "L*L/32";
  if (cluster) {
    if (!qdiv_exact(L, V - 1)) return "L";
    if (!qdiv_exact(ML, L - 1)) return "L/4";
    if (!qdiv_exact(Lz / 4, Lz - 1)) return "Lz/4";
  }
  return nullptr;
}

// Everything the kernels need about one tube: geometry, device state and the
// inputs of the calls in flight.  Per-call inputs are indexed by the tube's
// device step counter, so one captured step can be replayed: step k reads
// counter blocks ctr[k][half][nr], replica pairs pairs[k][call][4] and coin
// offset coin0 + (k * calls + call) * coin_per_call.
struct Tube {
  Lattice g;
  uint64_t *spin;           // [ns][2][nr][nt][n] (spin_row)
  const uint64_t *J;        // [ns][2][3][n][2] (coupling_index)
  uint8_t *wctb;            // [ns][nr][nt] clone at beta
  const uint8_t *lut;       // [nt][EA_LUT_BYTES]
  int32_t *ener;            // [ns][nr][nt] by clone
  int32_t *part;            // partial energies, energy_partials(g) slots
  ea_hq_state *pt;          // [ns][nr]
  double *acc, *att, *raw;  // [ns][nr][nt-1]
  const double *nd;         // [nt-1]  -(beta[i+1]-beta[i])
  const double *tab;        // exp tables (host libm values)
  const int32_t *toff, *tlen;
  double *esum;             // [ns][nr][nt] by beta
  uint8_t *hist;            // [rows][ns][nr][nt] beta of clone
  int32_t *ovl;             // [ns][replica pair r1 < r2][nt] site overlaps, or NULL
  // slab Houdayer, per instance: parents of the forest of the roots that
  // reach a seam [V], slab-local root of every site [V], q bits of every row
  // [V/L], and per slab the flip bits of the roots whose cluster stays inside
  // it followed by the bits of the roots that reach a seam [slabs][2][words]
  uint32_t *hpar;
  uint16_t *hl16;
  uint64_t *hrow, *hflag;
  uint32_t *keep;           // copy of the last labels, or NULL
  // sweep kind 3: the next sweep's thresholds (id1, id2) of every position,
  // [2 buffers][nr][nt][2 parities][n][2], made a slice per sample group by
  // the sweep before (NULL when n > kFullMaxN)
  uint64_t *thr;
  // sweep kind 4: the array the new words of a sweep go to, swapped with
  // `spin` after every such sweep (NULL until the kind is first set), and the
  // MSC planes of one sweep slab (sweep_slab_planes)
  uint64_t *spin2;
  int sweep_planes;
  int planes;               // slab Houdayer planes; 0: whole tube in one block
  int hou_cap;              // run nodes of the whole-tube labelling (hou_run_cap)
  int nbx;                  // sweep blocks per beta
  int cluster;              // Houdayer moves are part of this tube's steps
  int pt_exact;             // every PT walk applies the exact rule pair by pair
  // calls in flight
  int *step;                // device step counter
  int nsteps;               // blocks of a tube with *step >= nsteps do nothing
  int halves, calls;        // half-sweeps and Houdayer calls per step
  int it0, frec;            // history row of step k: (it0 + k) / frec when
                            // (it0 + k) % frec == frec - 1
  const ea_counter *ctr;
  const int32_t *pairs;
  uint64_t coin_seed, coin0, coin_per_call;
  uint32_t keys[EA_MAX_REPLICAS];
};

// Position of a launch inside a step, fixed when the step is enqueued.
struct Phase {
  int half;        // half-sweep index within the step (sweep), parity = half & 1
  int energy;      // sweep: write the energy partials
  int call;        // Houdayer call index within the step
  int measure;     // PT / Houdayer: accumulate esum and history
  int step = -1;   // the step, when the host knows it (one tube, launched
                   // step by step); -1: read the tube's step counter
};

// Block maps of every kernel kind over one table of tubes (kMapHouEnergy:
// one block per cluster instance of the tubes labelled in slabs;
// kMapOverlap: one per sample, replica pair and beta).
enum MapKind {
  kMapSweep, kMapEnergy, kMapChain, kMapWhole16, kMapWhole32, kMapWhole64, kMapSlab,
  kMapSeam, kMapApply, kMapHouEnergy, kMapOverlap, kMapKinds
};

// Launch shape of a table's kernels (chosen by measurement, ea_set_shape /
// ea_batch_set_shape).  Sweep: kind 0 shares each position's random words
// among the samples through LDS (1024-thread blocks of sweep_positions = 32
// or 64 half sites of one replica), kind 1 gives each thread one position
// and all samples (256-thread blocks, one energy partial per wave);
// kind 2 is kind 1 with the neighbour words staged in LDS tiles; kind 3
// makes a whole sweep (both halves) in one launch, each block holding whole
// clones in LDS (tubes with n <= kFullMaxN); kind 4 makes a whole sweep in
// one launch over z-slabs of the clones, into a second spin array;
// sweep_xcd = 1 hands each XCD a contiguous range of sweep blocks.  cluster_threads: threads of
// a whole-tube Houdayer workgroup (256, 512 or 1024).
struct Shape {
  int sweep_positions = 64;
  int cluster_threads = 1024;
  int sweep_kind = 0;
  int sweep_xcd = 0;
  bool operator==(const Shape &o) const {
    return sweep_positions == o.sweep_positions && cluster_threads == o.cluster_threads &&
           sweep_kind == o.sweep_kind && sweep_xcd == o.sweep_xcd;
  }
};
constexpr int kSweepPosMin = 32;

// Kind 3 (one launch per sweep): a block of kFullThreads threads holds the
// clones at one beta of one replica of sweep_full_samples(g) samples, both
// parities, one thread per half site of each; tubes with at most kFullMaxN
// half sites per parity take it.
constexpr int kFullThreads = 512;
constexpr int kFullMaxN = 512;
__host__ __device__ inline int sweep_full_samples(const Lattice &g) {
  const int sg = kFullThreads / g.n;
  return sg < 1 ? 1 : (sg < g.ns ? sg : g.ns);
}
// its dynamic LDS: clones [SG][2][n] u64 | thresholds [2][n][2] u64 |
// random words: the sweep's [2][n][6] u32 and a slice of the next sweep's
// (at most [n][6] with two or more sample groups)
__host__ __device__ inline size_t sweep_full_lds(const Lattice &g) {
  return (size_t)sweep_full_samples(g) * 16 * g.n + 32 * (size_t)g.n + 72 * (size_t)g.n;
}
// words of Tube::thr: two buffers of [nr][nt][2n][2] thresholds
inline size_t sweep_full_thr_words(const Lattice &g) {
  return g.n <= kFullMaxN ? (size_t)8 * g.nr * g.nt * g.n : 0;
}

// Kind 4 (one launch per sweep, for tubes that kind 3 does not take): a
// block of kSlabThreads threads owns a z-slab of T = sweep_slab_planes(g) MSC
// planes of one clone and takes the samples in turn, staging in LDS the odd
// words of planes z0-2 .. z0+T+1 and the even words of planes z0-1 .. z0+T
// (two buffers, one sample ahead), and the thresholds of the 2T+2 planes'
// positions it updates; it reads Tube::spin and writes Tube::spin2.
// T divides MLz, leaves at most kSlabPos positions of each parity per thread,
// keeps the halo planes distinct (MLz >= T + 4) and the LDS within
// kSlabLdsMax; 0 when no T does (the kind does not fit the tube).
constexpr int kSlabThreads = 512;
constexpr int kSlabPos = 2;
constexpr int kSlabChunks = 3;             // 16-byte chunks per thread and sample
constexpr size_t kSlabLdsMax = 69 * 1024;  // dynamic LDS: two blocks per CU
__host__ __device__ inline int sweep_slab_words(const Lattice &g, int T) {
  return (2 * T + 6) * g.hs;  // one sample's buffer: odd then even planes
}
__host__ __device__ inline size_t sweep_slab_lds(const Lattice &g, int T) {
  // two sample buffers, then (id1, id2) of the (2T + 2) hs positions
  return 16 * (size_t)sweep_slab_words(g, T) + 16 * (size_t)(2 * T + 2) * g.hs;
}
inline int sweep_slab_planes(const Lattice &g) {
  if (16 * g.half >= (1LL << 32)) return 0;  // 32-bit byte offsets within a sample
  // one energy partial per slab: within the slots every chain has
  // (energy_partials)
  const long long slots = std::max<long long>((g.n + 31) / 32, 4LL * ((g.n * g.nr + 255) / 256));
  for (int T = g.MLz; T >= 2; --T) {
    if (g.MLz % T || g.MLz < T + 4 || g.MLz / T > slots) continue;
    if ((T + 2) * g.hs > kSlabPos * kSlabThreads) continue;
    if (sweep_slab_words(g, T) > 2 * kSlabChunks * kSlabThreads) continue;
    if (sweep_slab_lds(g, T) > kSlabLdsMax) continue;
    return T;
  }
  return 0;
}

struct Plan {
  std::vector<uint64_t> map[kMapKinds];
  Shape shape;
  size_t lds16 = 0, lds32 = 0, lds64 = 0, lds_slab = 0, lds_apply = 0, lds_seam = 0;
  size_t lds_full = 0;  // dynamic LDS of a kind-3 or kind-4 sweep (the largest tube's)
  int tile_words = 0;  // 16-byte tile chunks per thread of a tiled sweep (the widest tube's)
  int ntubes = 0;
};

// Device copy of a plan: one buffer holding all maps.  The maps of a
// one-tube table are the identity (entry i = i) and are not stored: map()
// returns NULL, the tube is 0 and a block's local index is its block index.
struct DevPlan {
  uint64_t *buf = nullptr;
  size_t off[kMapKinds] = {};
  size_t cnt[kMapKinds] = {};
  size_t lds16 = 0, lds32 = 0, lds64 = 0, lds_slab = 0, lds_apply = 0, lds_seam = 0;
  size_t lds_full = 0;
  int tile_words = 0;  // 16-byte tile chunks per thread of a tiled sweep (the widest tube's)
  int ntubes = 0;
  Shape shape;
  bool identity = false;
  const uint64_t *map(int k) const { return identity ? nullptr : buf + off[k]; }
  bool ready() const { return identity || buf; }
};

// Energy partials per (sample, replica, beta) a sweep of this shape writes
// (a tube's nbx), and the partial-energy slots a tube needs under any shape.
int sweep_partials_x(const Lattice &g, const Shape &shape);
// Whether a tube can take a launch shape (the tiled sweep needs its tile,
// 256 + 2 hs half sites, within 768; the one-launch sweep n <= kFullMaxN).
bool shape_fits(const Lattice &g, const Shape &shape);
size_t energy_partials(const Lattice &g);

// Block maps of the tubes (host copies of the descriptors give geometry,
// Houdayer path and whether the tube has cluster moves).
// The tubes' nbx must be sweep_partials_x(g, shape).
Plan make_plan(const std::vector<Tube> &tubes, Shape shape);
hipError_t upload_plan(const Plan &p, DevPlan *d);
void free_plan(DevPlan *d);

// Launches over a device tube table.  Each returns hipGetLastError().
// One half-sweep (ph.half), or with kind 3 (sweep_is_whole) the whole sweep
// of halves ph.half (even) and ph.half + 1.
hipError_t launch_sweep(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s);
inline bool sweep_is_whole(const Shape &sh) { return sh.sweep_kind >= 3; }
// Kind 4 writes the new words to Tube::spin2; after each of its launches the
// descriptors of the tubes that ran swap spin and spin2.
inline bool sweep_swaps(const Shape &sh) { return sh.sweep_kind == 4; }
hipError_t launch_energy(const Tube *tubes, const DevPlan &p, int step, hipStream_t s);
hipError_t launch_overlap(const Tube *tubes, const DevPlan &p, int step, hipStream_t s);
hipError_t launch_resolve(const Tube *tubes, const DevPlan &p, int step, hipStream_t s);
hipError_t launch_pt(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s);
// parts: kHouWhole (the whole-tube kernels), kHouSlabs (the slab chain) or
// both; the two touch disjoint tubes, so they may run concurrently
constexpr int kHouWhole = 1, kHouSlabs = 2;
hipError_t launch_houdayer(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s,
                           int parts = kHouWhole | kHouSlabs);
// whether the plan has blocks for part kHouWhole or kHouSlabs
bool hou_has(const DevPlan &p, int part);
hipError_t launch_step_end(const Tube *tubes, int ntubes, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_f64(double *p, double v, size_t n, hipStream_t s);

// LDS bytes of the whole-tube and slab Houdayer kernels; a tube is labelled
// in one workgroup or in slabs as hou_use_slabs decides.  Both need
// L <= 64.
constexpr int kHouQueue = 192;  // union queue entries per wave (whole-tube kernel)
// Run nodes the whole-tube labelling holds (Tube::hou_cap): every run of a
// small tube, about 9/16 of the sites otherwise (a tube of uncorrelated q
// has about V/2 runs); an instance with more runs is labelled over sites.
int hou_run_cap(const Lattice &g);
// Byte offsets of the whole-tube kernel's dynamic LDS with `threads`
// threads, run capacity cap, rows of rowb bytes and site-forest parents of
// idxb bytes: q rows | q words | coin bits | first run of each row (u16) |
// the run forest (u32) or the site forest | the waves' union queues, then
// the runs' move bits (at most V/8 bytes, 16 n).
struct HouWholeLds {
  uint32_t qw, cb, rbase, forest, region, total;
};
__host__ __device__ inline uint32_t hou_up16(uint32_t x) { return (x + 15u) & ~15u; }
__host__ __device__ inline HouWholeLds hou_whole_layout(const Lattice &g, int cap, int threads,
                                                         int rowb, int idxb) {
  const uint32_t nrows = (uint32_t)(g.V / g.L);
  const uint32_t forest = (uint32_t)(4 * cap) > (uint32_t)(idxb * g.V) ? (uint32_t)(4 * cap)
                                                                        : (uint32_t)(idxb * g.V);
  const uint32_t queues = (uint32_t)(threads / 64) * 4u * (uint32_t)kHouQueue;
  HouWholeLds l;
  l.qw = hou_up16(nrows * (uint32_t)rowb);
  l.cb = l.qw + 16u * (uint32_t)g.n;
  l.rbase = l.cb + (uint32_t)g.V / 8u;
  l.forest = l.rbase + hou_up16(2u * nrows);
  l.region = l.forest + hou_up16(forest);
  l.total = l.region + (16u * (uint32_t)g.n > queues ? 16u * (uint32_t)g.n : queues);
  return l;
}
size_t hou_whole_lds(const Lattice &g, int threads, int cap);
size_t hou_slab_lds(const Lattice &g, int planes);
size_t hou_apply_lds(const Lattice &g, int planes);
constexpr int kHouSlabSites = 65536;  // slab-local labels are 16-bit
// Whether a tube's cluster moves take the slab path: always when its
// whole-tube labelling does not fit a workgroup's LDS (lds_max); also when
// it takes more than half a CU's LDS (lds_cu) and the tube has fewer
// instances than the device has CUs (cus), where the longest instance, not
// their number, sets the kernel's time and half-size slabs finish sooner.
bool hou_use_slabs(const Lattice &g, int lds_max, int lds_cu, int cus);
// LDS bytes of one slab's labelling: as much as lets two labelling
// workgroups share a CU (two 1024-thread workgroups fill its wave slots).
size_t hou_slab_budget(int lds_max, int lds_cu);
// Planes per slab: the thickest slab whose labelling fits lds_slab bytes.
int hou_slab_planes(const Lattice &g, size_t lds_slab);
// 64-bit words of one slab's flip (or seam-root) bits.
inline __host__ __device__ int hou_slab_words(const Lattice &g, int planes) {
  return (planes * g.L * g.L + 63) / 64;
}
// Device bytes of the slab Houdayer buffers of a tube.
size_t hou_slab_bytes(const Lattice &g, int planes);
constexpr int kHouMaxL = 64;

}  // namespace eatube
