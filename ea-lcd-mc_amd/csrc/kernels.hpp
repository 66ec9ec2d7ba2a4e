// kernels.hpp -- launch interface of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "eatube.h"

namespace eatube {

// Geometry and sizes, passed by value to every kernel.
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

// One half-sweep.  Counters and keys travel by value, as the reference's
// s_time / s_keys kernel arguments do (main.cu:299-305).
struct SweepArgs {
  uint64_t *spin;        // [ns][2][nt][n][nr]
  const uint64_t *J;     // [ns][2][n][6]
  const uint8_t *wctb;   // [ns][nr][nt] clone at beta
  const uint8_t *lut;    // [nt][EA_LUT_BYTES]
  int32_t *part;         // [ns][nr][nt][gridDim.x] partial energies by beta
  int parity;
  int energy;            // odd half-sweep: write the energy partials
  ea_counter ctr[EA_MAX_REPLICAS];
  uint32_t keys[EA_MAX_REPLICAS];
};

struct PtArgs {
  ea_hq_state *st;       // [ns][nr]
  int32_t *ener;         // [ns][nr][nt] by clone
  const int32_t *part;   // sweep partials [ns][nr][nt][nbx], or NULL: use ener
  int nbx;
  uint8_t *wctb;         // [ns][nr][nt]
  double *acc, *att, *raw;  // [ns][nr][nt-1]
  const double *nd;      // [nt-1]  -(beta[i+1]-beta[i])
  const double *tab;     // exp tables (host glibc values)
  const int32_t *off, *len;
  double *esum;          // per-beta energy sums, added after the swaps
  uint8_t *hist;         // beta-of-clone history row, when row >= 0
  int measure, row;
};

struct HouArgs {
  uint64_t *spin;
  const uint64_t *J;     // [ns][2][n][6]
  int32_t *ener;         // [ns][nr][nt] by clone; updated by the whole-tube path
  const uint8_t *wctb;
  int32_t pairs[4];
  uint64_t seed, offset0;  // coin index of (group 0, sample 0, beta 0, site 0)
  uint32_t *lab;         // [2][ns][ctm][V] global labels (tiled path)
  uint32_t *cbits;       // [2][ns][ctm][V/32] root coins (tiled path)
  uint32_t *labels_out;  // optional copy of the final labels
  int planes;            // z-planes per slab; 0: the whole tube in one block
  double *esum;          // whole-tube path: per-beta sums corrected for the
  int measure;           // moved configurations when measure != 0
};

struct MeasureArgs {
  const int32_t *ener;
  const uint8_t *wctb;
  double *esum;          // [ns][nr][nt] by beta
  uint8_t *hist;         // [rows][ns][nr][nt] beta of clone
  int row;               // history row to write, < 0 for none
};

hipError_t launch_sweep(const Lattice &g, const SweepArgs &a, hipStream_t s);
hipError_t launch_energy(const Lattice &g, const uint64_t *spin,
                         const uint64_t *J, int32_t *ener, int32_t *part,
                         hipStream_t s);
// ener by clone from the partials the last energy-writing sweep left by beta
// (wctb as it was during that sweep)
hipError_t launch_resolve(const Lattice &g, const int32_t *part, const uint8_t *wctb,
                          int32_t *ener, hipStream_t s);
// int32 slots of the partial-energy buffer both kinds of partials need
size_t energy_partials(const Lattice &g);
int sweep_blocks_x(const Lattice &g);
hipError_t launch_pt(const Lattice &g, const PtArgs &a, hipStream_t s);
hipError_t launch_houdayer(const Lattice &g, const HouArgs &a, hipStream_t s);
hipError_t launch_measure(const Lattice &g, const MeasureArgs &a,
                          hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_fill_f64(double *p, double v, size_t n, hipStream_t s);

// LDS bytes of the whole-tube and slab Houdayer kernels; a tube is labelled
// in one workgroup when hou_whole_lds() fits the device's per-block LDS,
// otherwise in slabs of about kHouSlabSites sites.  Both need L <= 64.
size_t hou_whole_lds(const Lattice &g);
size_t hou_slab_lds(const Lattice &g, int planes);
constexpr int kHouSlabSites = 16384;
constexpr int kHouMaxL = 64;

}  // namespace eatube
