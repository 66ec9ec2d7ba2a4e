// kernels.hip -- gfx950 kernels of the EA tube engine.
//
// Layouts (per sample s): spins [2 parities][nt clones][n half sites][nr
// replicas] of 64-bit words, bit = bx + 4 by + 16 bz of the 4x4x4 copies of
// an MSC site (ini.c:339-383); couplings [2][n][6] words, 48 contiguous
// bytes per half site so the nr replicas of a site read them as one
// broadcast.  All kernels are wave64.
#include "kernels.hpp"

#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace eatube {

namespace {

constexpr uint64_t kLcg = 3202034522624059733ULL;  // random_generator.h:133
constexpr double kFnorm = 5.4210108624275218e-20;  // random_generator.h:92

// ---------------------------------------------------------------- helpers

// 16-byte vectors for wide loads and stores
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// Global-memory view of the tube's arrays: the table holds generic
// pointers, and loads through them would compile to flat loads, whose waits
// also drain LDS traffic and serialise the samples' loads.
typedef __attribute__((address_space(1))) uint64_t gword;
typedef __attribute__((address_space(1))) uint8_t gbyte;
__device__ __forceinline__ gword *gmem(uint64_t *p) { return (gword *)p; }
__device__ __forceinline__ const gword *gmem(const uint64_t *p) { return (const gword *)p; }
// The same for any element type.  Stores through a flat pointer also count
// against the LDS wait counter, so an LDS read after them waits for the
// store to reach memory; through a global view it does not.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ global_t<T> *gv(T *p) {
  return (global_t<T> *)p;
}


// Philox4x32-10, Salmon et al. SC'11 (the round and key schedule of
// main.cu:996-1043).
__device__ __forceinline__ void philox10(uint32_t &v0, uint32_t &v1,
                                         uint32_t &v2, uint32_t &v3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const uint32_t lo0 = 0xD2511F53u * v0, hi0 = __umulhi(0xD2511F53u, v0);
    const uint32_t lo1 = 0xCD9E8D57u * v2, hi1 = __umulhi(0xCD9E8D57u, v2);
    v0 = hi1 ^ v1 ^ k0;
    v1 = lo1;
    v2 = hi0 ^ v3 ^ k1;
    v3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// R = (m:l) < umbral[i]; the LUT sits in LDS as 256 (hi, lo) threshold
// pairs followed by 4096 16-bit (min | max << 8) prefix windows.
__device__ __forceinline__ bool below(const uint32_t *lut, uint32_t i,
                                      uint32_t m, uint32_t l) {
  const uint32_t hx = lut[2 * i], lx = lut[2 * i + 1];
  return m == hx ? l < lx : m < hx;
}

// Byte whose bits are Bernoulli(exp(-4 beta)): the interval of the 64-bit
// uniform among the 256 cumulative thresholds, narrowed by the 12-bit prefix
// window and then bisected (main.cu:1063-1078).
__device__ __forceinline__ uint32_t lut_byte(const uint32_t *lut, uint32_t m,
                                             uint32_t l) {
  const uint16_t *pre = reinterpret_cast<const uint16_t *>(lut + 512);
  uint32_t lo = 255, hi = 255;
  if (below(lut, 254, m, l)) {
    const uint32_t w = pre[m >> 20];
    lo = w & 255;
    hi = w >> 8;
  }
  while (hi - lo > 1) {
    const uint32_t mid = (hi + lo) >> 1;
    if (below(lut, mid, m, l))
      hi = mid;
    else
      lo = mid;
  }
  return below(lut, lo, m, l) ? lo : hi;
}

// One 32-bit random word: Philox, the 64-bit uniforms (v[j] : lsb[j]), four
// LUT bytes and a data-dependent rotation (main.cu:1045-1092).  A rotation
// by 0 leaves the word unchanged.
__device__ __forceinline__ uint32_t rnd_word(uint32_t x, uint32_t y, uint32_t z,
                                             uint32_t w, uint32_t k0,
                                             uint32_t k1, const uint32_t *lut) {
  uint32_t v0 = x, v1 = y, v2 = z, v3 = w;
  philox10(v0, v1, v2, v3, k0, k1);
  const uint32_t hi = v0 ^ v2, lo = v1 ^ v3;
  const uint32_t l0 = ((v1 << 17) | (v0 >> 15)) ^ hi ^ ((hi << 21) | (lo >> 11));
  const uint32_t l1 = ((v0 << 17) | (v1 >> 15)) ^ lo ^ (lo << 21);
  const uint32_t l2 = (hi << 28) | (lo >> 4);
  const uint32_t l3 = (lo << 28) | (hi >> 4);
  const uint32_t rot = (l0 + v1 + l2 + v3) >> 27;
  const uint32_t bb = lut_byte(lut, v0, l0) | (lut_byte(lut, v1, l1) << 8) |
                      (lut_byte(lut, v2, l2) << 16) |
                      (lut_byte(lut, v3, l3) << 24);
  return (bb << rot) | (bb >> ((32 - rot) & 31));
}

// Quotient i / d from m = ceil(2^32 / d) (m = 0 for d = 1), one multiply
// instead of a division sequence: exact while i * (m d - 2^32) < 2^32, which
// holds for every divisor and index range of the tubes the engine accepts
// (inexact_divisor refuses the others at creation).
__device__ __forceinline__ uint32_t qdiv(uint32_t i, uint32_t m) {
  return m ? __umulhi(i, m) : i;
}

// The tube of this block and the block's index within the tube (block map
// entry tube << 32 | local; no map: tube 0, local = block index), and the
// step: `step` when the launch carries it, else the tube's step counter.
struct Where {
  const Tube *t;
  int local, step;
};
__device__ __forceinline__ Where locate_block(const Tube *tubes, const uint64_t *map, int step,
                                             uint32_t block) {
  const uint64_t e = map ? map[block] : (uint64_t)block;
  Where w;
  w.t = tubes + (e >> 32);
  w.local = (int)(uint32_t)e;
  w.step = step >= 0 ? step : *w.t->step;
  return w;
}
__device__ __forceinline__ Where locate(const Tube *tubes, const uint64_t *map, int step) {
  return locate_block(tubes, map, step, blockIdx.x);
}

// Workgroups are dealt to the 8 XCDs in turn (block b runs on XCD b % 8).
// This bijection of [0, nb) gives XCD x a contiguous range of logical
// blocks, so neighbouring blocks share an L2.
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Bit permutations bringing a neighbour word whose MSC site wrapped around
// the sub-lattice to the copy layout of the centre word.
constexpr uint64_t kMaskE = 0x1111111111111111ULL;
constexpr uint64_t kMaskN = 0x000F000F000F000FULL;
constexpr uint64_t kMaskU = 0x000000000000FFFFULL;
constexpr uint64_t kMaskO = 0x8888888888888888ULL;
constexpr uint64_t kMaskS = 0xF000F000F000F000ULL;
constexpr uint64_t kMaskD = 0xFFFF000000000000ULL;
__device__ __forceinline__ uint64_t wrap_xp(uint64_t o) {
  return (o & kMaskE) << 3 | (o & ~kMaskE) >> 1;
}
__device__ __forceinline__ uint64_t wrap_yp(uint64_t o) {
  return (o & kMaskN) << 12 | (o & ~kMaskN) >> 4;
}
__device__ __forceinline__ uint64_t wrap_zp(uint64_t o) {
  return (o & kMaskU) << 48 | (o & ~kMaskU) >> 16;
}
__device__ __forceinline__ uint64_t wrap_xm(uint64_t o) {
  return (o & kMaskO) >> 3 | (o & ~kMaskO) << 1;
}
__device__ __forceinline__ uint64_t wrap_ym(uint64_t o) {
  return (o & kMaskS) >> 12 | (o & ~kMaskS) << 4;
}
__device__ __forceinline__ uint64_t wrap_zm(uint64_t o) {
  return (o & kMaskD) >> 48 | (o & ~kMaskD) << 16;
}

// Neighbours of the parity-p site at half index h, in closed form: half
// index of each of the six opposite-parity neighbours (+x,+y,+z,-x,-y,-z)
// and a 6-bit mask of the directions that wrapped around the sub-lattice.
struct Nbrs {
  int h[6];
  int wrap;
  int gz;
};

__device__ __forceinline__ Nbrs neighbours(const Lattice &g, int p, int h) {
  const int gz = qdiv(h, g.mhs), rem = h - gz * g.hs, by = qdiv(rem, g.mhl);
  const int ax = 2 * (rem - by * g.hl) + ((p + by + gz) & 1);
  Nbrs nb;
  nb.gz = gz;
  int w = 0;
  const int row = g.ML * by + g.MS * gz;  // site index of (0, by, gz)
  const int xp = ax + 1 == g.ML ? (w |= 1, 0) : ax + 1;
  const int xm = ax == 0 ? (w |= 8, g.ML - 1) : ax - 1;
  nb.h[0] = (xp + row) >> 1;
  nb.h[3] = (xm + row) >> 1;
  const int yp = by + 1 == g.ML ? (w |= 2, 0) : by + 1;
  const int ym = by == 0 ? (w |= 16, g.ML - 1) : by - 1;
  nb.h[1] = (ax + g.ML * yp + g.MS * gz) >> 1;
  nb.h[4] = (ax + g.ML * ym + g.MS * gz) >> 1;
  const int zp = gz + 1 == g.MLz ? (w |= 4, 0) : gz + 1;
  const int zm = gz == 0 ? (w |= 32, g.MLz - 1) : gz - 1;
  nb.h[2] = (ax + g.ML * by + g.MS * zp) >> 1;
  nb.h[5] = (ax + g.ML * by + g.MS * zm) >> 1;
  nb.wrap = w;
  return nb;
}

// The six neighbour words of (clone row `old`, replica r), brought to the
// centre's copy layout; with open z boundaries the z bond of the top
// (bottom) copy block is replaced by the y+1 (even) or y-1 (odd) neighbour
// (main.cu:1460-1472).
template <typename P>
__device__ __forceinline__ void load_neighbours(const Lattice &g, P old, int r,
                                                const Nbrs &nb, int parity,
                                                uint64_t sc[6], int stride) {
#pragma unroll
  for (int d = 0; d < 6; ++d) sc[d] = old[(size_t)nb.h[d] * stride + r];
  if (nb.wrap) {
    if (nb.wrap & 1) sc[0] = wrap_xp(sc[0]);
    if (nb.wrap & 2) sc[1] = wrap_yp(sc[1]);
    if (nb.wrap & 4) sc[2] = wrap_zp(sc[2]);
    if (nb.wrap & 8) sc[3] = wrap_xm(sc[3]);
    if (nb.wrap & 16) sc[4] = wrap_ym(sc[4]);
    if (nb.wrap & 32) sc[5] = wrap_zm(sc[5]);
  }
  if (g.bc == EA_OBC) {
    const uint64_t aux = parity == 0 ? sc[1] : sc[4];
    if (nb.gz == g.MLz - 1) sc[2] = (sc[2] & ~kMaskD) | (aux & kMaskD);
    if (nb.gz == 0) sc[5] = (sc[5] & ~kMaskU) | (aux & kMaskU);
  }
}

// The six couplings of half site h of parity p of sample s.
__device__ __forceinline__ void load_couplings(const Lattice &g, const uint64_t *J, int s, int p,
                                               int h, uint64_t out[6]) {
  const auto j = gmem(J) + coupling_index(g, s, p, h, 0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    out[2 * k] = j[(size_t)k * 2 * g.n];
    out[2 * k + 1] = j[(size_t)k * 2 * g.n + 1];
  }
}

// Sum of e over the lanes of each `width`-lane group of this wave holding
// the same replica (lane & (nr-1)); result valid in every lane.
__device__ __forceinline__ int replica_sum(int e, int nr, int width = 64) {
  for (int off = nr; off < width; off <<= 1) e += __shfl_xor(e, off, 64);
  return e;
}

// ------------------------------------------------------------------ sweep

constexpr int kSwThreads = 1024;

// The beta's LUT, the clone at beta of every (sample, replica) (with ROWS,
// the word offset of that clone's row, spin_row), and the replicas' counter
// blocks and keys of this half-sweep, into LDS.
template <bool ROWS, typename C>
__device__ __forceinline__ void sweep_tables(const Tube &a, int beta, int step, int half,
                                             uint32_t *lut, C *clone_at, uint32_t (*ctr)[5]) {
  const Lattice &g = a.g;
  const int tx = threadIdx.x, T = blockDim.x;
  const auto src = gv(reinterpret_cast<const uint4 *>(a.lut + (size_t)beta * EA_LUT_BYTES));
  u32x4 *dst = reinterpret_cast<u32x4 *>(lut);
  const auto vsrc = reinterpret_cast<const global_t<u32x4> *>(src);
  for (int i = tx; i < EA_LUT_BYTES / 16; i += T) dst[i] = vsrc[i];
  for (int i = tx; i < g.ns * g.nr; i += T) {
    const int c = gv(a.wctb)[(size_t)i * g.nt + beta];
    clone_at[i] = ROWS ? (C)spin_row(g, i & (g.nr - 1), c) : (C)c;
  }
  if (tx < g.nr) {
    const auto c = gv(a.ctr) + ((size_t)step * a.halves + half) * g.nr + tx;
    ctr[tx][0] = c->x;
    ctr[tx][1] = c->y;
    ctr[tx][2] = c->z;
    ctr[tx][3] = c->w;
    ctr[tx][4] = a.keys[tx];
  }
}


// Sample-independent addressing of one (h, r) position of parity p: word
// offsets of its own word and of its six neighbours within a clone row of a
// sample's spin block, of its couplings within a sample's coupling block,
// and the boundary fix-ups its neighbours need (main.cu:1440-1472).
struct LaneSite {
  uint32_t own, nb[6], jo;
  int wrap;
  bool top, bottom;
};

__device__ __forceinline__ LaneSite lane_site(const Lattice &g, int p, int h) {
  const Nbrs nb = neighbours(g, p, h);
  LaneSite ls;
  ls.own = (uint32_t)(p * g.half) + (uint32_t)h;
#pragma unroll
  for (int d = 0; d < 6; ++d) ls.nb[d] = (uint32_t)((1 - p) * g.half) + (uint32_t)nb.h[d];
  ls.jo = (uint32_t)coupling_index(g, 0, p, h, 0);
  ls.wrap = nb.wrap;
  ls.top = g.bc == EA_OBC && nb.gz == g.MLz - 1;
  ls.bottom = g.bc == EA_OBC && nb.gz == 0;
  return ls;
}

// One sample's inputs at a position: its word, the six neighbours and the
// six couplings.
struct SiteIn {
  uint64_t sv, sc[6], J[6];
};

__device__ __forceinline__ void load_lane(const gword *S, const gword *Jb, uint32_t c,
                                          uint32_t jpair, const LaneSite &ls, SiteIn &x) {
  x.sv = S[c + ls.own];
#pragma unroll
  for (int d = 0; d < 6; ++d) x.sc[d] = S[c + ls.nb[d]];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x.J[2 * k] = Jb[ls.jo + k * jpair];
    x.J[2 * k + 1] = Jb[ls.jo + k * jpair + 1];
  }
}

// v_bitop3_b32 (gfx950): bit i of the result is bit (a_i << 2 | b_i << 1 |
// c_i) of an 8-bit table.  Tables are written as the function applied to
// A = 0xF0, B = 0xCC, C = 0xAA.
constexpr uint32_t kT3A = 0xF0, kT3B = 0xCC, kT3C = 0xAA;
constexpr uint32_t kXor3 = kT3A ^ kT3B ^ kT3C;                        // a ^ b ^ c
constexpr uint32_t kMaj3 = (kT3A & kT3B) | (kT3C & (kT3A | kT3B));    // majority
constexpr uint32_t kAnd3 = kT3A & kT3B & kT3C;                        // a & b & c
constexpr uint32_t kOr3 = kT3A | kT3B | kT3C;                         // a | b | c
constexpr uint32_t kAndOr = (kT3A & kT3B) | kT3C;                     // (a & b) | c

template <uint32_t T>
__device__ __forceinline__ uint64_t op3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t lo =
      __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, (T & 0xFF));
  const uint32_t hi = __builtin_amdgcn_bitop3_b32(
      (uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32), (T & 0xFF));
  return (uint64_t)hi << 32 | lo;
}

// Neighbours brought to the centre's copy layout (load_neighbours) without
// selects: each x or y permutation is (a & m) | (b & ~m) of two shifts of
// the word, and a lane that does not wrap shifts by 0, which leaves the word
// as it is (both forms, with the shift amounts of each direction, 1/3 for x
// and 4/12 for y, equal the wrap_* functions bit for bit).  The rare z wraps
// (first and last MSC plane) and the open-boundary substitutes sit behind
// branches that almost every wave skips; `obc` may be uniform (the tube's
// boundary) or per lane (top or bottom).  No branch splits two samples'
// loads: every caller issues them before the first fix-up.
constexpr uint32_t kSelAB = (kT3A & kT3C) | (kT3B & ~kT3C & 0xFF);  // (a & c) | (b & ~c)
__device__ __forceinline__ uint64_t wrap_up(uint64_t o, uint32_t f, uint32_t a, uint32_t b,
                                            uint64_t m) {
  return op3<kSelAB>(o << (f * a), o >> (f * b), m);
}
__device__ __forceinline__ uint64_t wrap_down(uint64_t o, uint32_t f, uint32_t a, uint32_t b,
                                              uint64_t m) {
  return op3<kSelAB>(o >> (f * a), o << (f * b), m);
}
__device__ __forceinline__ void fix_flags_sel(uint32_t w, bool top, bool bottom, int parity,
                                              bool obc, SiteIn &x) {
  x.sc[0] = wrap_up(x.sc[0], w & 1, 3, 1, 0x8888888888888888ULL);
  x.sc[3] = wrap_down(x.sc[3], (w >> 3) & 1, 3, 1, 0x1111111111111111ULL);
  x.sc[1] = wrap_up(x.sc[1], (w >> 1) & 1, 12, 4, 0xF000F000F000F000ULL);
  x.sc[4] = wrap_down(x.sc[4], (w >> 4) & 1, 12, 4, 0x000F000F000F000FULL);
  if (w & 4) x.sc[2] = wrap_zp(x.sc[2]);
  if (w & 32) x.sc[5] = wrap_zm(x.sc[5]);
  if (obc) {  // uniform: periodic tubes skip the substitutes
    const uint64_t aux = parity == 0 ? x.sc[1] : x.sc[4];
    x.sc[2] = top ? (x.sc[2] & ~kMaskD) | (aux & kMaskD) : x.sc[2];
    x.sc[5] = bottom ? (x.sc[5] & ~kMaskU) | (aux & kMaskU) : x.sc[5];
  }
}

__device__ __forceinline__ void fix_flags(int w, bool top, bool bottom, int parity, SiteIn &x) {
  fix_flags_sel((uint32_t)w, top, bottom, parity, top || bottom, x);
}
__device__ __forceinline__ void fix_lane(const LaneSite &ls, int parity, SiteIn &x) {
  fix_flags(ls.wrap, ls.top, ls.bottom, parity, x);
}

// The Metropolis flips of one word: (j2 j1) = f0+f1+f2 and (j4 j3) =
// f3+f4+f5 unsatisfied bonds, id = id1 + 2 id2; flip iff j + id >= 3, i.e.
// j1 j3 id1 all set, or a weight-2 term together with a weight-1 term, or
// two weight-2 terms (majority of j2, j4, id2).  With `energy` also the
// word's bond energy after the move.
__device__ __forceinline__ uint64_t site_flip(const SiteIn &x, uint64_t id1, uint64_t id2,
                                              bool energy, int &e) {
  uint64_t f[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) f[d] = op3<kXor3>(x.sv, x.J[d], x.sc[d]);  // 1 = unsatisfied
  const uint64_t j1 = op3<kXor3>(f[0], f[1], f[2]), j2 = op3<kMaj3>(f[0], f[1], f[2]);
  const uint64_t j3 = op3<kXor3>(f[3], f[4], f[5]), j4 = op3<kMaj3>(f[3], f[4], f[5]);
  const uint64_t one = op3<kOr3>(j1, j3, id1), two = op3<kOr3>(j2, j4, id2);
  const uint64_t flip =
      op3<kAndOr>(one, two, op3<kMaj3>(j2, j4, id2)) | op3<kAnd3>(j1, j3, id1);
  if (energy) {
    int u = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) u += __popcll(f[d] ^ flip);
    e = 6 * 64 - 2 * u;
  }
  return flip;
}

// Sweep block order: a block holds consecutive half sites of one replica
// (a contiguous run of each clone row it touches), and the launch order is
// site range, then replica, then beta, so that the blocks reading the same
// couplings run together and share them in L2.
struct SweepBlock {
  int hb, r, beta;
};
__device__ __forceinline__ SweepBlock sweep_block(const Lattice &g, int local) {
  SweepBlock b;
  const int rest = local / g.nt;
  b.beta = local - rest * g.nt;
  b.hb = rest >> g.nr_shift;
  b.r = rest & (g.nr - 1);
  return b;
}

// The six random words of position (h, r) at beta, as the reference thread
// tid = beta*n*nr + h*nr + r makes them (d_gen_rndbits_cuda,
// main.cu:1095-1150), folded into the Metropolis thresholds id = id1 + 2 id2
// with P(id >= k) = exp(-4 k beta).
__device__ __forceinline__ uint32_t sweep_word(const Lattice &g, const uint32_t (*ctr)[5],
                                               const uint32_t *lut, int beta, int h, int r,
                                               int k) {
  const uint32_t tid = ((uint32_t)beta * (uint32_t)g.n + (uint32_t)h) * (uint32_t)g.nr + r;
  return rnd_word(ctr[r][0] + k, ctr[r][1], ctr[r][2] + tid, ctr[r][3], tid, ctr[r][4], lut);
}
__device__ __forceinline__ void thresholds(const uint32_t w[6], uint64_t &id1, uint64_t &id2) {
  const uint64_t b0 = ((uint64_t)w[0] << 32) | w[1];
  const uint64_t b1 = (((uint64_t)w[2] << 32) | w[3]) & b0;
  const uint64_t b2 = (((uint64_t)w[4] << 32) | w[5]) & b1;
  id1 = (b0 ^ b1) | b2;
  id2 = b1;
}

// One sample's update at a position: loads, boundary fix-ups, flips and
// the store, two samples at a time so that both samples' loads are in
// flight together.  Returns the words' energies after the move in e0, e1
// when `energy`.
__device__ __forceinline__ void update_two(uint64_t *spin, const uint64_t *J, size_t sstride,
                                           size_t jstride, uint32_t jpair, int s0, int s1,
                                           uint32_t c0, uint32_t c1, const LaneSite &ls, int p,
                                           uint64_t id1, uint64_t id2, bool energy, int &e0,
                                           int &e1) {
  gword *S0 = gmem(spin + s0 * sstride), *S1 = gmem(spin + s1 * sstride);
  SiteIn x0, x1;
  load_lane(S0, gmem(J + s0 * jstride), c0, jpair, ls, x0);
  load_lane(S1, gmem(J + s1 * jstride), c1, jpair, ls, x1);
  fix_lane(ls, p, x0);
  fix_lane(ls, p, x1);
  const uint64_t f0 = site_flip(x0, id1, id2, energy, e0);
  const uint64_t f1 = site_flip(x1, id1, id2, energy, e1);
  S0[c0 + ls.own] = x0.sv ^ f0;
  if (s1 != s0) S1[c1 + ls.own] = x1.sv ^ f1;
}

// Workgroup barrier that waits for this wave's LDS writes only: global loads
// issued before it stay in flight (__syncthreads waits for all memory
// traffic).  The signal fences keep the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt, expcnt untouched
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Within one wave: wait until this wave's LDS accesses issued so far are
// done (another lane of the wave reads what they wrote next).
__device__ __forceinline__ void lds_wave_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Inclusive prefix sum of e within each W-lane group of the wave (W = 32
// or 64), so that the group's last lane holds the group's sum: a scan within
// each 16-lane row by DPP row shifts (a lane whose source lies before its
// row adds 0), then each row's total carried into the next by the row
// broadcasts (row 0's into row 1 and row 2's into row 3; for W = 64 also
// lane 31's into rows 2 and 3).  Six VALU instructions and no LDS traffic.
template <int W>
__device__ __forceinline__ int group_scan(int e) {
  static_assert(W == 32 || W == 64, "groups of 32 or 64 lanes");
  e += __builtin_amdgcn_update_dpp(0, e, 0x111, 0xf, 0xf, true);  // row_shr:1
  e += __builtin_amdgcn_update_dpp(0, e, 0x112, 0xf, 0xf, true);  // row_shr:2
  e += __builtin_amdgcn_update_dpp(0, e, 0x114, 0xf, 0xf, true);  // row_shr:4
  e += __builtin_amdgcn_update_dpp(0, e, 0x118, 0xf, 0xf, true);  // row_shr:8
  e += __builtin_amdgcn_update_dpp(0, e, 0x142, 0xa, 0xf, false);  // row_bcast:15
  if (W == 64) e += __builtin_amdgcn_update_dpp(0, e, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return e;
}

// One half-sweep (d_oneMCstepBN_multisample, main.cu:1385-1510), kind 0: a
// block of 1024 threads owns POS consecutive half sites of one replica at
// one beta.  First 6*POS threads each make one of the six random words of a
// position with the beta's LUT in LDS; then each group of POS lanes
// updates one sample's clone at beta, 1024/POS samples at a time.  The
// random words are shared by all samples, are made once and never reach
// HBM.  One energy partial per (sample, replica, beta, block).
template <int POS, bool XCD>
__global__ void __launch_bounds__(kSwThreads) k_sweep(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map,
                                                      Phase ph) {
  static_assert(POS == 32 || POS == 64, "a sample's positions stay within one wave");
  constexpr int kSlots = kSwThreads / POS;
  __shared__ __align__(16) uint32_t lut[EA_LUT_BYTES / 4];
  __shared__ uint32_t clone_off[EA_MAX_SAMPLES * EA_MAX_REPLICAS];
  __shared__ uint32_t words[POS][6];
  __shared__ uint64_t ids[POS][2];
  __shared__ uint32_t ctr[EA_MAX_REPLICAS][5];
  const Where wh = locate_block(tubes, map, ph.step,
                                XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const SweepBlock b = sweep_block(g, wh.local);
  const int tx = threadIdx.x, h0 = b.hb * POS, r = b.r;
  sweep_tables<true>(a, b.beta, wh.step, ph.half, lut, clone_off, ctr);
  __syncthreads();

  const int pos = tx & (POS - 1), slot = tx / POS;
  const bool valid = h0 + pos < g.n;
  const int h = valid ? h0 + pos : 0;
  const int p = ph.half & 1;
  const LaneSite ls = lane_site(g, p, h);
  uint64_t *const spin = a.spin;
  const uint64_t *const J = a.J;
  int32_t *const part = a.part;
  const int nbx = a.nbx;
  const size_t sstride = 2 * (size_t)g.half, jstride = 12 * (size_t)g.n;
  const uint32_t jpair = 2 * (uint32_t)g.n;
  const uint32_t pstride = (uint32_t)(g.nr * g.nt * nbx);  // partials of one sample
  const uint32_t pbase = (uint32_t)((r * g.nt + b.beta) * nbx + b.hb);
  // the first pair of samples' loads go out before the random words are
  // made, so that their latency overlaps that work (the barriers until the
  // update wait for LDS traffic only)
  SiteIn x0, x1;
  auto load_pair = [&](int s0, int s1) {
    load_lane(gmem(spin + s0 * sstride), gmem(J + s0 * jstride), clone_off[s0 * g.nr + r], jpair,
              ls, x0);
    load_lane(gmem(spin + s1 * sstride), gmem(J + s1 * jstride), clone_off[s1 * g.nr + r], jpair,
              ls, x1);
  };
  if (valid && slot < g.ns) load_pair(slot, slot + kSlots < g.ns ? slot + kSlots : slot);

  if (tx < 6 * POS) {
    const int wp = tx / 6, k = tx - 6 * wp;
    if (h0 + wp < g.n) words[wp][k] = sweep_word(g, ctr, lut, b.beta, h0 + wp, r, k);
  }
  lds_barrier();
  if (tx < POS) thresholds(words[tx], ids[tx][0], ids[tx][1]);
  lds_barrier();
  const uint64_t id1 = ids[pos][0], id2 = ids[pos][1];

  for (int s0 = slot; s0 < g.ns; s0 += 2 * kSlots) {
    const int s1 = s0 + kSlots < g.ns ? s0 + kSlots : s0;
    int e0 = 0, e1 = 0;
    if (valid) {
      if (s0 != slot) load_pair(s0, s1);
      fix_lane(ls, p, x0);
      fix_lane(ls, p, x1);
      const uint64_t f0 = site_flip(x0, id1, id2, ph.energy, e0);
      const uint64_t f1 = site_flip(x1, id1, id2, ph.energy, e1);
      gmem(spin + s0 * sstride)[clone_off[s0 * g.nr + r] + ls.own] = x0.sv ^ f0;
      if (s1 != s0) gmem(spin + s1 * sstride)[clone_off[s1 * g.nr + r] + ls.own] = x1.sv ^ f1;
    }
    if (ph.energy) {
      // every bond has exactly one odd end: the odd half-sweep sees them
      // all; one partial per (sample, replica, beta, block), no atomics
      e0 = group_scan<POS>(e0);
      e1 = group_scan<POS>(e1);
      if (pos == POS - 1) {
        part[(uint32_t)s0 * pstride + pbase] = e0;
        if (s1 != s0) part[(uint32_t)s1 * pstride + pbase] = e1;
      }
    }
  }
}

// One half-sweep, kind 1: one thread per half site of one replica at one
// beta makes the position's six random words itself and updates every
// sample's clone at beta, two samples' loads in flight at a time.  Blocks of
// 256 threads (256 consecutive half sites), one energy partial per wave.
constexpr int kLaneThreads = 256;

template <bool XCD>
__global__ void __launch_bounds__(kLaneThreads) k_sweep_lane(const Tube *__restrict__ tubes,
                                                             const uint64_t *__restrict__ map, Phase ph) {
  __shared__ __align__(16) uint32_t lut[EA_LUT_BYTES / 4];
  __shared__ uint32_t clone_off[EA_MAX_SAMPLES * EA_MAX_REPLICAS];
  __shared__ uint32_t ctr[EA_MAX_REPLICAS][5];
  const Where wh = locate_block(tubes, map, ph.step,
                                XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const SweepBlock b = sweep_block(g, wh.local);
  const int tx = threadIdx.x, r = b.r;
  sweep_tables<true>(a, b.beta, wh.step, ph.half, lut, clone_off, ctr);
  __syncthreads();

  const int hh = b.hb * kLaneThreads + tx;
  const bool valid = hh < g.n;
  const int h = valid ? hh : 0;
  uint64_t id1 = 0, id2 = 0;
  if (valid) {
    uint32_t w[6];
    // not unrolled: the words' Philox states would otherwise be live
    // together and set the kernel's register count
#pragma unroll 1
    for (int k = 0; k < 6; ++k) w[k] = sweep_word(g, ctr, lut, b.beta, h, r, k);
    thresholds(w, id1, id2);
  }
  const int p = ph.half & 1;
  const LaneSite ls = lane_site(g, p, h);
  // the tube's pointers stay in registers: stores through them could
  // otherwise alias the descriptor and force reloads
  uint64_t *const spin = a.spin;
  const uint64_t *const J = a.J;
  int32_t *const part = a.part;
  const int nbx = a.nbx, bx = b.hb * (kLaneThreads / 64) + (tx >> 6);
  const uint32_t pstride = (uint32_t)(g.nr * g.nt * nbx);  // partials of one sample
  const uint32_t pbase = (uint32_t)((r * g.nt + b.beta) * nbx + bx);
  const bool pwrite = (tx & 63) == 63 && bx < nbx;  // a wave past the last site has none
  const size_t sstride = 2 * (size_t)g.half, jstride = 12 * (size_t)g.n;
  const uint32_t jpair = 2 * (uint32_t)g.n;  // words between direction pairs
  for (int s0 = 0; s0 < g.ns; s0 += 2) {
    const int s1 = s0 + 1 < g.ns ? s0 + 1 : s0;
    int e0 = 0, e1 = 0;
    if (valid)
      update_two(spin, J, sstride, jstride, jpair, s0, s1, clone_off[s0 * g.nr + r],
                 clone_off[s1 * g.nr + r], ls, p, id1, id2, ph.energy, e0, e1);
    if (ph.energy) {
      // every bond has exactly one odd end: one partial per (sample,
      // replica, beta, 64 sites), no atomics
      e0 = group_scan<64>(e0);
      e1 = group_scan<64>(e1);
      if (pwrite) {
        part[(uint32_t)s0 * pstride + pbase] = e0;
        if (s1 != s0) part[(uint32_t)s1 * pstride + pbase] = e1;
      }
    }
  }
}


// One half-sweep, kind 2 (tiled): blocks as kind 1 (256 consecutive half
// sites of one replica at one beta), but a sample's six neighbour words come
// from an LDS tile of the opposite parity's clone row over the block's run
// widened by hs on each side (every x, y and unwrapped z neighbour), read
// as 16-byte chunks from the even half site at or below its start: each
// thread loads KC <= 2 chunks per sample instead of six neighbours.
// Samples go one at a time; the next sample's tile is loaded with this
// sample's words and staged after the update (double buffer, barriers that
// leave global traffic in flight).  All loads are unconditional (clamped
// addresses), so they stay ahead of the one wait of the update instead of
// sinking towards their use.  Only lanes whose z neighbour wraps around the
// tube read it from global memory.  One sample in flight per thread (not
// two) and the six random words made in a loop that is not unrolled keep
// the kernel at 72 VGPRs, 7 waves per SIMD instead of 4: 230 -> 210 us per
// half-sweep at L=48, Lz=384.
constexpr int kTileWordsMax = 3;  // tile of at most 768 half sites

__host__ __device__ inline int tile_width(const Lattice &g) { return kLaneThreads + 2 * g.hs; }
inline int tile_words(const Lattice &g) {
  return (tile_width(g) + kLaneThreads - 1) / kLaneThreads;
}
// 16-byte chunks per thread: the tile from the even half site at or below
// its start covers at most tile_width + 1 half sites.
inline int tile_chunks(const Lattice &g) {
  return (tile_width(g) + 1 + 2 * kLaneThreads - 1) / (2 * kLaneThreads);
}

template <int KC, bool XCD>
__global__ void __launch_bounds__(kLaneThreads) k_sweep_tile(const Tube *__restrict__ tubes,
                                                             const uint64_t *__restrict__ map,
                                                             Phase ph) {
  __shared__ __align__(16) uint32_t lut[EA_LUT_BYTES / 4];
  __shared__ uint32_t clone_off[EA_MAX_SAMPLES * EA_MAX_REPLICAS];
  __shared__ uint32_t ctr[EA_MAX_REPLICAS][5];
  __shared__ __align__(16) uint64_t tile[2][2 * KC * kLaneThreads];  // [buffer][x]
  const Where wh = locate_block(tubes, map, ph.step,
                                XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const SweepBlock b = sweep_block(g, wh.local);
  const int tx = threadIdx.x, r = b.r;
  sweep_tables<true>(a, b.beta, wh.step, ph.half, lut, clone_off, ctr);
  __syncthreads();

  const int h0 = b.hb * kLaneThreads, hh = h0 + tx;
  const bool valid = hh < g.n;
  const int h = valid ? hh : 0;
  const int p = ph.half & 1;
  const LaneSite ls = lane_site(g, p, h);
  const uint32_t opp = (uint32_t)((1 - p) * g.half);
  // the neighbours lie in half sites [lo, lo + W); slot x of the tile holds
  // half site lo2 + x, lo2 even, so that slot pairs are 16-byte chunks
  const int W = tile_width(g), lo = h0 - g.hs, lo2 = lo & ~1;
  const int nchunk = (lo + W - lo2 + 1) >> 1;
  int ti[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int x = (int)(ls.nb[d] - opp) - lo;
    ti[d] = x >= 0 && x < W ? x + (lo - lo2) : -1;  // -1: a z neighbour across the wrap
  }
  // this thread's chunks (clamped: a chunk past the tile or the tube repeats
  // words nobody reads; n is even, so no chunk straddles the row's ends)
  int tc_[KC];
  uint32_t th_[KC];
#pragma unroll
  for (int i = 0; i < KC; ++i) {
    tc_[i] = min(tx + i * kLaneThreads, nchunk - 1);
    th_[i] = (uint32_t)min(max(lo2 + 2 * tc_[i], 0), g.n - 2);
  }
  uint64_t *const spin = a.spin;
  const uint64_t *const J = a.J;
  int32_t *const part = a.part;
  const int nbx = a.nbx, bx = b.hb * (kLaneThreads / 64) + (tx >> 6);
  const uint32_t pstride = (uint32_t)(g.nr * g.nt * nbx);
  const uint32_t pbase = (uint32_t)((r * g.nt + b.beta) * nbx + bx);
  const bool pwrite = (tx & 63) == 63 && bx < nbx;
  const size_t sstride = 2 * (size_t)g.half, jstride = 12 * (size_t)g.n;
  const uint32_t jpair = 2 * (uint32_t)g.n;

  // row offsets are even (n is), so a chunk is 16-byte aligned
  u64x2 tv[KC];
  auto fetch = [&](int s) {
    const gword *S = gmem(spin + s * sstride) + clone_off[s * g.nr + r] + opp;
    const auto S2 = reinterpret_cast<const global_t<u64x2> *>(S);
#pragma unroll
    for (int i = 0; i < KC; ++i) tv[i] = S2[th_[i] >> 1];
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < KC; ++i) *reinterpret_cast<u64x2 *>(&tile[buf][2 * tc_[i]]) = tv[i];
  };
  // the first sample's tile is loaded before the random words are made, so
  // that its latency overlaps that work
  const int last = g.ns - 1;
  fetch(0);
  uint64_t id1 = 0, id2 = 0;
  if (valid) {
    uint32_t w[6];
    // not unrolled: the words' Philox states would otherwise be live
    // together and set the kernel's register count
#pragma unroll 1
    for (int k = 0; k < 6; ++k) w[k] = sweep_word(g, ctr, lut, b.beta, h, r, k);
    thresholds(w, id1, id2);
  }
  stage(0);
  lds_barrier();
  for (int s0 = 0, buf = 0; s0 < g.ns; ++s0, buf ^= 1) {
    // the sample's loads issue at a raised wave priority, ahead of other
    // waves' arithmetic, so more of them are in flight (206 instead of 210 us
    // per half-sweep at L=48, Lz=384)
    __builtin_amdgcn_s_setprio(1);
    const uint32_t c0 = clone_off[s0 * g.nr + r];
    gword *S0 = gmem(spin + s0 * sstride);
    fetch(min(s0 + 1, last));  // the next sample's tile (after the last: a repeat)
    SiteIn x0;
    const gword *J0 = gmem(J + s0 * jstride);
    x0.sv = S0[c0 + ls.own];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      x0.J[2 * k] = J0[ls.jo + k * jpair];
      x0.J[2 * k + 1] = J0[ls.jo + k * jpair + 1];
    }
    uint64_t z0[2] = {0, 0};
    if (ti[2] < 0) z0[0] = S0[c0 + ls.nb[2]];
    if (ti[5] < 0) z0[1] = S0[c0 + ls.nb[5]];
    __builtin_amdgcn_s_setprio(0);
    int e0 = 0;
    uint64_t n0 = 0;
    if (valid) {
#pragma unroll
      for (int d = 0; d < 6; ++d) x0.sc[d] = ti[d] >= 0 ? tile[buf][ti[d]] : 0;
      if (ti[2] < 0) x0.sc[2] = z0[0];
      if (ti[5] < 0) x0.sc[5] = z0[1];
      fix_lane(ls, p, x0);
      n0 = x0.sv ^ site_flip(x0, id1, id2, ph.energy, e0);
    }
    stage(buf ^ 1);  // after the last sample: into a buffer nobody reads
    if (valid) S0[c0 + ls.own] = n0;
    if (ph.energy) {
      e0 = group_scan<64>(e0);
      if (pwrite) part[(uint32_t)s0 * pstride + pbase] = e0;
    }
    lds_barrier();
  }
}

// Sample-independent addressing of half site h of parity p within one
// sample's clone staged in LDS as [2 parities][n]: lane_site with the clone
// row's own and neighbour words counted from the row's start.
__device__ __forceinline__ LaneSite lane_site_row(const Lattice &g, int p, int h) {
  const Nbrs nb = neighbours(g, p, h);
  LaneSite ls;
  ls.own = (uint32_t)(p * g.n + h);
#pragma unroll
  for (int d = 0; d < 6; ++d) ls.nb[d] = (uint32_t)((1 - p) * g.n + nb.h[d]);
  ls.jo = (uint32_t)coupling_index(g, 0, p, h, 0);
  ls.wrap = nb.wrap;
  ls.top = g.bc == EA_OBC && nb.gz == g.MLz - 1;
  ls.bottom = g.bc == EA_OBC && nb.gz == 0;
  return ls;
}

// One sweep, kind 3: both half-sweeps (ph.half, even, and ph.half + 1) in
// one launch.  A block of kFullThreads threads holds in LDS the clones at beta of
// replica r of SG = sweep_full_samples(g) consecutive samples, both parities
// (16 n bytes each), one thread per half site of each clone.  The even
// half-sweep updates the clones in LDS and, after a barrier, the odd one
// reads its neighbours there: each word is loaded and stored once per sweep,
// and a sweep is one launch instead of two.
// Random words: the thresholds (id1, id2) of the 2 n positions of (r, beta)
// for both halves are the same for every sample group.  With G > 1 groups,
// each launch's group gi makes slice gi of the NEXT sweep's thresholds into
// a.thr (buffer (q + 1) & 1, q the sweep's index in the call), and takes its
// own from buffer q & 1, where the previous launch left them; the first
// sweep of a call (and every sweep of a one-group tube) makes its own, six
// lanes per word as kind 0 does, with the half's counter block.  The words
// depend on the counters and beta only, so either way they equal the
// reference's.  Each thread's couplings of both halves, its thresholds and
// its clone row are loaded first, then its 16-byte chunk of the clones,
// whose latency overlaps the random words.  Blocks (sample group, replica,
// beta), beta fastest; one energy partial per (sample, replica, beta)
// (nbx = 1) when ph.energy.
template <bool XCD>
__global__ void __launch_bounds__(kFullThreads) k_sweep_full(const Tube *__restrict__ tubes,
                                                           const uint64_t *__restrict__ map,
                                                           Phase ph) {
  __shared__ __align__(16) uint32_t lut[EA_LUT_BYTES / 4];
  __shared__ uint32_t ctr[4][5];  // this sweep's two halves, then the next sweep's
  __shared__ int esum[EA_MAX_SAMPLES];
  extern __shared__ uint64_t dyn[];
  const Where wh = locate_block(tubes, map, ph.step,
                                XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int n = g.n, tx = threadIdx.x, T = blockDim.x;
  const int rest = wh.local / g.nt, beta = wh.local - rest * g.nt;
  const int r = rest & (g.nr - 1), gi = rest >> g.nr_shift, SG = sweep_full_samples(g);
  const int G = (g.ns + SG - 1) / SG, s0 = gi * SG, sgn = min(SG, g.ns - s0);
  const int hidx = wh.step * a.halves + ph.half, q = hidx >> 1;  // the call's half, sweep
  const bool own = G == 1 || q == 0;
  const bool next = G > 1 && hidx + 2 < a.nsteps * a.halves;
  const int nlo = gi * 2 * n / G, nhi = (gi + 1) * 2 * n / G;  // slice of the next sweep
  const int town = own ? 12 * n : 0, tall = town + (next ? 6 * (nhi - nlo) : 0);
  uint64_t *sp = dyn;                                           // [SG][2][n]
  uint64_t *ids = sp + (size_t)2 * SG * n;                      // [2][n][2]
  uint32_t *words = reinterpret_cast<uint32_t *>(ids + 4 * n);  // own [2n][6], next slice
  const size_t tbuf = (size_t)g.nr * g.nt * 4 * n, trow = ((size_t)r * g.nt + beta) * 4 * n;
  // Thread tx: half site h of sample s0 + j in both halves, and the clones'
  // 16-byte chunk tx (words 2 tx, 2 tx + 1 of sp: half sites ch, ch + 1 of
  // parity cp of the same sample).  Idle threads read sample s0's first
  // chunk, couplings and thresholds, and write nothing.
  const bool act = tx < sgn * n;
  const int j = act ? tx / n : 0, h = act ? tx - j * n : 0;
  const int cp = 2 * h >= n, ch = 2 * h - cp * n;
  // Every load is issued first, into registers (the clone row's chunk after
  // its row offset): the block waits for one memory round trip, then two
  // dependent ones, before its LDS stores, instead of one per array.
  constexpr int kLutChunks = (EA_LUT_BYTES / 16 + kFullThreads - 1) / kFullThreads;
  const bool makes = own || next;  // random words in this launch
  u32x4 lv[kLutChunks];
  if (makes) {
    const auto src = gv(reinterpret_cast<const uint4 *>(a.lut + (size_t)beta * EA_LUT_BYTES));
    const auto vsrc = reinterpret_cast<const global_t<u32x4> *>(src);
#pragma unroll
    for (int i = 0; i < kLutChunks; ++i) lv[i] = vsrc[min(tx + i * T, EA_LUT_BYTES / 16 - 1)];
  }
  uint32_t cw[4] = {0, 0, 0, 0};
  const bool cnt = tx < 4 && (tx < 2 ? own : next);
  if (cnt) {
    const auto c = gv(a.ctr) + ((size_t)hidx + tx) * g.nr + r;
    cw[0] = c->x;
    cw[1] = c->y;
    cw[2] = c->z;
    cw[3] = c->w;
  }
  const int clone = gv(a.wctb)[((size_t)(s0 + j) * g.nr + r) * g.nt + beta];
  uint64_t Jv[2][6];
  {
    const gword *Jb = gmem(a.J + (size_t)(s0 + j) * 12 * n);
    const uint32_t jpair = 2 * (uint32_t)n;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const uint32_t jo = (uint32_t)coupling_index(g, 0, p, h, 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        Jv[p][2 * k] = Jb[jo + k * jpair];
        Jv[p][2 * k + 1] = Jb[jo + k * jpair + 1];
      }
    }
  }
  u64x2 idv[2] = {u64x2{0, 0}, u64x2{0, 0}};
  if (!own) {  // made by the previous launch
    const auto cur = reinterpret_cast<const global_t<u64x2> *>(gmem(a.thr + (q & 1) * tbuf + trow));
    idv[0] = cur[h];
    idv[1] = cur[n + h];
  }
  // the chunk's words (row offsets and n are even: 16-byte aligned)
  const auto chunk = reinterpret_cast<global_t<u64x2> *>(
      gmem(a.spin + (size_t)(s0 + j) * 2 * g.half + (size_t)cp * g.half) +
      spin_row(g, r, clone) + ch);
  const u64x2 cv = *chunk;
  if (makes) {
    u32x4 *dst = reinterpret_cast<u32x4 *>(lut);
#pragma unroll
    for (int i = 0; i < kLutChunks; ++i)
      if (tx + i * T < EA_LUT_BYTES / 16) dst[tx + i * T] = lv[i];
  }
  if (cnt) {
#pragma unroll
    for (int k = 0; k < 4; ++k) ctr[tx][k] = cw[k];
    ctr[tx][4] = a.keys[r];
  }
  if (tx < sgn) esum[tx] = 0;
  // Words k of positions pos = p n + h' of this sweep (own: task t < town)
  // and of the next sweep's slice [nlo, nhi) (t >= town), then their
  // thresholds: own ones into ids, the next sweep's into a.thr.
  auto make_words = [&]() {
    lds_barrier();  // LUT and counters staged
    for (int t = tx; t < tall; t += T) {
      const bool mine = t < town;
      const int tt = mine ? t : t - town, pq = tt / 6, k = tt - 6 * pq;
      const int pos = mine ? pq : nlo + pq, p = pos >= n, hp = pos - p * n;
      const uint32_t(&c)[5] = ctr[(mine ? 0 : 2) + p];
      const uint32_t tid = ((uint32_t)beta * (uint32_t)n + (uint32_t)hp) * (uint32_t)g.nr + r;
      words[t] = rnd_word(c[0] + k, c[1], c[2] + tid, c[3], tid, c[4], lut);
    }
    lds_barrier();
    const auto nxt = reinterpret_cast<global_t<u64x2> *>(gmem(a.thr + ((q + 1) & 1) * tbuf + trow));
    for (int t = tx; 6 * t < tall; t += T) {
      uint64_t i1, i2;
      thresholds(words + 6 * t, i1, i2);
      if (6 * t < town) {
        ids[2 * t] = i1;
        ids[2 * t + 1] = i2;
      } else {
        nxt[nlo + t - 2 * n * own] = u64x2{i1, i2};
      }
    }
  };
  // a launch that makes its own words makes them before the update; one
  // that only makes the next sweep's slice makes it after its stores are
  // issued, so that the slice overlaps their latency
  if (own) make_words();
  if (act) *reinterpret_cast<u64x2 *>(sp + 2 * tx) = cv;
  lds_barrier();
  if (own) {
    idv[0] = u64x2{ids[2 * h], ids[2 * h + 1]};
    idv[1] = u64x2{ids[2 * (n + h)], ids[2 * (n + h) + 1]};
  }
  uint64_t *S = sp + (size_t)j * 2 * n;  // this thread's sample's clone
  int e = 0;
  // the two halves: parity p reads the other parity's words in LDS
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    if (act) {
      const LaneSite ls = lane_site_row(g, p, h);
      SiteIn x;
      x.sv = S[ls.own];
#pragma unroll
      for (int d = 0; d < 6; ++d) x.sc[d] = S[ls.nb[d]];
#pragma unroll
      for (int d = 0; d < 6; ++d) x.J[d] = Jv[p][d];
      fix_lane(ls, p, x);
      const uint64_t f = site_flip(x, idv[p][0], idv[p][1], p == 1 && ph.energy, e);
      S[ls.own] = x.sv ^ f;
    }
    lds_barrier();
  }
  if (ph.energy) {
    // every bond has exactly one odd end: the odd half's words hold them all
    if ((n & 63) == 0) {  // each wave's lanes belong to one sample
      e = group_scan<64>(e);
      if ((tx & 63) == 63 && act) atomicAdd(&esum[j], e);
    } else if (act) {
      atomicAdd(&esum[j], e);
    }
    lds_barrier();
    if (tx < sgn)
      gv(a.part)[(uint32_t)(s0 + tx) * (uint32_t)(g.nr * g.nt) + (uint32_t)(r * g.nt + beta)] =
          esum[tx];
  }
  if (act) *chunk = *reinterpret_cast<const u64x2 *>(sp + 2 * tx);
  if (!own && next) make_words();
}


// ------------------------------------------------------------ sweep kind 4

typedef __attribute__((address_space(3))) uint8_t lbyte;

// LDS byte offset, within a sample's buffer, of the word of parity `par` at
// half index hh of a slab whose first plane is z0: odd words of planes
// z0-2 .. z0+T+1 first, then even words of planes z0-1 .. z0+T, each range
// taken mod MLz (so a slab at either end of the tube reads its halo planes
// from the other end, as the periodic z neighbours and the open-boundary
// substitutes of main.cu:1440-1472 need).
__device__ __forceinline__ uint32_t slab_offset(const Lattice &g, int T, int z0, int par,
                                                int hh) {
  const int pl = hh / g.hs, rr = hh - pl * g.hs;
  int lp = pl - z0 + (par ? 2 : 1);
  lp += lp < 0 ? g.MLz : 0;
  lp -= lp >= g.MLz ? g.MLz : 0;
  return 8u * (uint32_t)((par ? 0 : (T + 4) * g.hs) + lp * g.hs + rr);
}

// One position a thread updates in every sample, packed into two words:
// `own` holds the LDS byte offset of its word in a sample buffer (bits
// 0-15), the directions whose neighbour wrapped around the sub-lattice (bits
// 16-21) and the open-boundary top and bottom flags (bits 22, 23); `dxy`
// holds, as signed bytes, the in-plane word offsets of the +x, -x, +y and -y
// neighbours from the position's own half index within its plane.  The z
// neighbours sit at the same in-plane index one plane up or down, and the
// other parity's copy of the position's plane lies at a fixed distance from
// its own word, so every neighbour's offset is own + a per-phase constant
// (+ 8 dxy for x and y).
struct SlabSite {
  uint32_t own, dxy;
  int h;
};

__device__ __forceinline__ SlabSite slab_site(const Lattice &g, int T, int z0, int p, int gz,
                                              int rem) {
  SlabSite ss;
  ss.h = gz * g.hs + rem;
  const Nbrs nb = neighbours(g, p, ss.h);
  const uint32_t own = slab_offset(g, T, z0, p, ss.h);
  const bool top = g.bc == EA_OBC && nb.gz == g.MLz - 1;
  const bool bottom = g.bc == EA_OBC && nb.gz == 0;
  ss.own = own | (uint32_t)nb.wrap << 16 | (uint32_t)top << 22 | (uint32_t)bottom << 23;
  const int dirs[4] = {0, 3, 1, 4};  // +x, -x, +y, -y
  ss.dxy = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int hn = nb.h[dirs[k]];
    const int d = (hn - (hn / g.hs) * g.hs) - rem;
    ss.dxy |= (uint32_t)(d & 0xFF) << (8 * k);
  }
  return ss;
}

// The seven words of a packed position from a sample buffer at LDS byte
// offset buf: cxy, czp, czm are the phase's distances from the own word to
// the other parity's copy of the same plane and of the planes above and
// below.
template <typename LW>
__device__ __forceinline__ void slab_words(LW lword, uint32_t buf, const SlabSite &ss,
                                           int cxy, int czp, int czm, SiteIn &x) {
  const uint32_t own = (ss.own & 0xFFFFu) + buf;
  x.sv = lword(own);
  const uint32_t pl = own + (uint32_t)cxy;
  const int dx = (int)(ss.dxy << 24) >> 24, dxm = (int)(ss.dxy << 16) >> 24;
  const int dy = (int)(ss.dxy << 8) >> 24, dym = (int)ss.dxy >> 24;
  x.sc[0] = lword(pl + 8 * dx);
  x.sc[3] = lword(pl + 8 * dxm);
  x.sc[1] = lword(pl + 8 * dy);
  x.sc[4] = lword(pl + 8 * dym);
  x.sc[2] = lword(own + (uint32_t)czp);
  x.sc[5] = lword(own + (uint32_t)czm);
}

// A copy the compiler cannot see through: per-lane byte offsets made
// inside a loop stay 32-bit there (a uniform base plus a 32-bit offset is
// one load with a scalar base), instead of being hoisted as 64-bit
// addresses that hold twice the registers for the whole loop.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "v"(v));
  return v;
}

// One sweep, kind 4: both half-sweeps (ph.half, even, and ph.half + 1) in
// one launch, for tubes whose clones do not fit one block's LDS.  Block
// (slab, replica r, beta), beta fastest: T = a.sweep_planes MSC planes
// [z0, z0 + T) of the clone at beta of replica r, every sample in turn.
// Per sample the block stages in LDS the odd words of planes z0-2 .. z0+T+1
// and the even words of planes z0-1 .. z0+T (from a.spin), updates the even
// words of planes z0-1 .. z0+T in LDS -- the two outer planes belong to the
// neighbouring slabs, which make the same new words from the same inputs --
// then the odd words of its T planes from the new even ones, and writes the
// new words of its T planes to a.spin2: no block reads a word that another
// has already rewritten, and every word of a.spin2 is written once.  The
// next sample's words load while this one is updated, into the other of
// two LDS buffers.  The thresholds of the block's (2T + 2) hs positions are
// made once, from the two halves' counter blocks, and serve every sample
// (the reference's one set of random words per half-sweep for all samples,
// main.cu:299,308).  One energy partial per (sample, replica, beta, slab)
// when ph.energy, from the odd half (every bond has exactly one odd end).
template <bool XCD>
__global__ void __launch_bounds__(kSlabThreads, 4) k_sweep_slab(const Tube *__restrict__ tubes,
                                                            const uint64_t *__restrict__ map,
                                                            Phase ph) {
  __shared__ __align__(16) uint32_t lut[EA_LUT_BYTES / 4];
  __shared__ uint32_t crow[EA_MAX_SAMPLES];  // clone row of (s, r) at beta, in words
  __shared__ uint32_t ctr[2][5];
  __shared__ int esum[2];
  extern __shared__ __align__(16) uint64_t dyn[];  // buffers [2][kSlabBuf], thresholds
  const Where wh = locate_block(tubes, map, ph.step,
                                XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int T = a.sweep_planes, tx = threadIdx.x;
  const int rest = wh.local / g.nt, beta = wh.local - rest * g.nt;
  const int r = rest & (g.nr - 1), slab = rest >> g.nr_shift, z0 = slab * T;
  const int hidx = wh.step * a.halves + ph.half;
  const int ne = (T + 2) * g.hs, no = T * g.hs;  // positions of each parity
  const int nch = sweep_slab_words(g, T) / 2;     // 16-byte chunks per sample

  // tables: every load first, into registers
  constexpr int kLutChunks = (EA_LUT_BYTES / 16 + kSlabThreads - 1) / kSlabThreads;
  u32x4 lv[kLutChunks];
  {
    const auto src =
        reinterpret_cast<const global_t<u32x4> *>(gv(a.lut + (size_t)beta * EA_LUT_BYTES));
#pragma unroll
    for (int i = 0; i < kLutChunks; ++i) lv[i] = src[min(tx + i * kSlabThreads, EA_LUT_BYTES / 16 - 1)];
  }
  int clone = 0;
  if (tx < g.ns) clone = gv(a.wctb)[((size_t)tx * g.nr + r) * g.nt + beta];
  uint32_t cw[4] = {0, 0, 0, 0};
  if (tx < 2) {
    const auto c = gv(a.ctr) + ((size_t)hidx + tx) * g.nr + r;
    cw[0] = c->x;
    cw[1] = c->y;
    cw[2] = c->z;
    cw[3] = c->w;
  }
  // the thread's chunks of a sample buffer (chunk c fills LDS bytes 16c ..
  // 16c + 15): word offset within a sample's spins (parity array, then the
  // clone row's plane)
  uint32_t cg[kSlabChunks];
#pragma unroll
  for (int k = 0; k < kSlabChunks; ++k) {
    const int c = min(tx + k * kSlabThreads, nch - 1), w = 2 * c;
    const int odd = w < (T + 4) * g.hs;
    const int w2 = odd ? w : w - (T + 4) * g.hs;
    const int lp = w2 / g.hs, off = w2 - lp * g.hs;
    int gz = z0 + lp - (odd ? 2 : 1);
    gz += gz < 0 ? g.MLz : 0;
    gz -= gz >= g.MLz ? g.MLz : 0;
    cg[k] = (uint32_t)(odd ? g.half : 0) + (uint32_t)(gz * g.hs + off);
  }
  {
    u32x4 *dst = reinterpret_cast<u32x4 *>(lut);
#pragma unroll
    for (int i = 0; i < kLutChunks; ++i)
      if (tx + i * kSlabThreads < EA_LUT_BYTES / 16) dst[tx + i * kSlabThreads] = lv[i];
  }
  if (tx < g.ns) crow[tx] = (uint32_t)spin_row(g, r, clone);
  if (tx < 2) {
#pragma unroll
    for (int k = 0; k < 4; ++k) ctr[tx][k] = cw[k];
    ctr[tx][4] = a.keys[r];
    esum[tx] = 0;
  }
  lds_barrier();

  const size_t sstride = 2 * (size_t)g.half;
  const gword *src = gmem(a.spin);
  gword *dst = gmem(a.spin2);
  lbyte *const lds = (lbyte *)(dyn);
  typedef __attribute__((address_space(3))) uint64_t lu64;
  auto lword = [&](uint32_t off) -> lu64 & { return *(lu64 *)(lds + off); };
  u64x2 tv[kSlabChunks];
  auto fetch = [&](int s) {
    const gbyte *S2 =
        (const gbyte *)(src + (size_t)s * sstride + __builtin_amdgcn_readfirstlane(crow[s]));
#pragma unroll
    for (int k = 0; k < kSlabChunks; ++k)
      tv[k] = *(const global_t<u64x2> *)(S2 + opaque(8u * cg[k]));
  };
  auto stage = [&](uint32_t buf) {
#pragma unroll
    for (int k = 0; k < kSlabChunks; ++k) {
      const uint32_t c = (uint32_t)min(tx + k * kSlabThreads, nch - 1);  // chunk k's LDS slot
      *(__attribute__((address_space(3))) u64x2 *)(lds + buf + 16u * c) = tv[k];
    }
  };
  fetch(0);  // in flight while the thresholds are made
  // thresholds of positions pos = tx, tx + kSlabThreads, ...: even ones
  // (pos < ne) with the even half's counter block, odd ones with the odd's
  const uint32_t bstride = 8u * (uint32_t)sweep_slab_words(g, T);  // one sample buffer
  const uint32_t thr0 = 2 * bstride;
  for (int pos = tx; pos < ne + no; pos += kSlabThreads) {
    const int odd = pos >= ne, q = odd ? pos - ne : pos;
    const int lp = q / g.hs, rem = q - lp * g.hs;
    int gz = z0 + lp - (odd ? 0 : 1);
    gz += gz < 0 ? g.MLz : 0;
    gz -= gz >= g.MLz ? g.MLz : 0;
    const uint32_t hh = (uint32_t)(gz * g.hs + rem);
    const uint32_t tid = ((uint32_t)beta * (uint32_t)g.n + hh) * (uint32_t)g.nr + r;
    const uint32_t(&c)[5] = ctr[odd];
    uint32_t w[6];
#pragma unroll 1
    for (int k = 0; k < 6; ++k) w[k] = rnd_word(c[0] + k, c[1], c[2] + tid, c[3], tid, c[4], lut);
    uint64_t i1, i2;
    thresholds(w, i1, i2);
    *(__attribute__((address_space(3))) u64x2 *)(lds + thr0 + 16u * pos) = u64x2{i1, i2};
  }
  // (made after the thresholds, so that fewer values are live there)
  // the thread's positions: even ones tx, tx + kSlabThreads of planes
  // z0-1 .. z0+T, odd ones of planes z0 .. z0+T-1; a thread past the end
  // takes position 0's addressing and stores nothing
  SlabSite es[kSlabPos], os[kSlabPos];
  bool ev[kSlabPos], ov[kSlabPos], est[kSlabPos];
#pragma unroll
  for (int i = 0; i < kSlabPos; ++i) {
    const int pe = tx + i * kSlabThreads, qe = pe < ne ? pe : 0;
    ev[i] = pe < ne;
    const int le = qe / g.hs, re = qe - le * g.hs;
    int gz = z0 - 1 + le;
    gz += gz < 0 ? g.MLz : 0;
    gz -= gz >= g.MLz ? g.MLz : 0;
    es[i] = slab_site(g, T, z0, 0, gz, re);
    est[i] = ev[i] && le >= 1 && le <= T;
    const int po = tx + i * kSlabThreads, qo = po < no ? po : 0;
    ov[i] = po < no;
    const int lo = qo / g.hs;
    os[i] = slab_site(g, T, z0, 1, z0 + lo, qo - lo * g.hs);
  }
  stage(0);
  lds_barrier();

  const uint32_t jpb = 16 * (uint32_t)g.n, jodd = 3 * jpb;  // direction pair, odd parity (bytes)
  // distances (bytes) from an even word of local plane le (LDS offset
  // 8 (E0 + le hs + rem)) to the odd words of the same plane and of the planes
  // above and below (odd local planes le + 1, le + 2, le); an odd word's to
  // the even ones are their negatives
  const int E0 = (T + 4) * g.hs;
  const int exy = 8 * (g.hs - E0), ezp = 8 * (2 * g.hs - E0), ezm = -8 * E0;
  const uint32_t pstride = (uint32_t)(g.nr * g.nt * a.nbx);
  const uint32_t pbase = (uint32_t)((r * g.nt + beta) * a.nbx + slab);
  const int last = g.ns - 1;
  // the passes in which this wave has a valid position (uniform per wave)
  const int w0 = __builtin_amdgcn_readfirstlane(tx) & ~63;
  bool wev[kSlabPos], wov[kSlabPos];
#pragma unroll
  for (int i = 0; i < kSlabPos; ++i) {
    wev[i] = w0 + i * kSlabThreads < ne;
    wov[i] = w0 + i * kSlabThreads < no;
  }
  // couplings of sample s (bytes from its start); the even half's of the
  // next sample are loaded a sample ahead (a wave with no valid position in
  // a pass skips its loads and its update)
  auto jbase = [&](int s) { return (const gbyte *)(gmem(a.J) + (size_t)s * 12 * g.n); };
  u64x2 Je[kSlabPos][3];
  auto load_even = [&](int s) {
    const gbyte *Jb = jbase(s);
#pragma unroll
    for (int i = 0; i < kSlabPos; ++i) {
      const uint32_t je = opaque(16u * (uint32_t)es[i].h);
      if (wev[i]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) Je[i][k] = *(const global_t<u64x2> *)(Jb + (je + k * jpb));
      }
    }
  };
  load_even(0);
  // one sample, its words in the buffer at LDS byte offset BUF
  auto sample = [&](int s, uint32_t BUF) {
    fetch(min(s + 1, last));  // the next sample's words (after the last: a repeat)
    // uniform bases and 32-bit byte offsets (within a sample's spins or
    // couplings), so that the loads and stores take a scalar base
    gbyte *Db = (gbyte *)(dst + (size_t)s * sstride + __builtin_amdgcn_readfirstlane(crow[s]));
    // the odd half's couplings arrive during the even half; the even half's
    // came in during the previous sample's odd half (load_even)
    u64x2 Jd[kSlabPos][3];
    {
      const gbyte *Jb = jbase(s);
#pragma unroll
      for (int i = 0; i < kSlabPos; ++i) {
        const uint32_t jd = opaque(16u * (uint32_t)os[i].h + jodd);
        if (wov[i]) {
#pragma unroll
          for (int k = 0; k < 3; ++k) Jd[i][k] = *(const global_t<u64x2> *)(Jb + (jd + k * jpb));
        }
      }
    }
    // even half: planes z0-1 .. z0+T, in LDS; planes z0 .. z0+T-1 also to spin2
#pragma unroll
    for (int i = 0; i < kSlabPos; ++i) {
      if (!wev[i]) break;  // the passes' valid waves only shrink
      // one pass at a time: both in flight together would need more
      // registers than four waves per SIMD leave
      __builtin_amdgcn_sched_barrier(0);
      const SlabSite &ss = es[i];
      SiteIn x;
#pragma unroll
      for (int d = 0; d < 6; ++d) x.J[d] = Je[i][d >> 1][d & 1];
      slab_words(lword, BUF, ss, exy, ezp, ezm, x);
      const uint32_t ti = ev[i] ? (uint32_t)(tx + i * kSlabThreads) : 0u;
      const u64x2 id = *(__attribute__((address_space(3))) u64x2 *)(lds + thr0 + 16u * ti);
      const uint32_t fl = opaque(ss.own) >> 16;  // derived per sample, not held
      fix_flags_sel(fl & 63, (fl >> 6) & 1, (fl >> 7) & 1, 0, g.bc == EA_OBC, x);
      int e = 0;
      const uint64_t nw = x.sv ^ site_flip(x, id[0], id[1], false, e);
      if (ev[i]) lword(BUF + (ss.own & 0xFFFFu)) = nw;
      if (est[i]) *(global_t<uint64_t> *)(Db + opaque(8u * (uint32_t)ss.h)) = nw;
    }
    lds_barrier();
    load_even(min(s + 1, last));  // the next sample's even couplings
    // odd half: planes z0 .. z0+T-1, from the new even words
    int e = 0;
#pragma unroll
    for (int i = 0; i < kSlabPos; ++i) {
      if (!wov[i]) break;
      __builtin_amdgcn_sched_barrier(0);
      const SlabSite &ss = os[i];
      const bool vo = ov[i];
      SiteIn x;
#pragma unroll
      for (int d = 0; d < 6; ++d) x.J[d] = Jd[i][d >> 1][d & 1];
      slab_words(lword, BUF, ss, -exy, -ezm, -ezp, x);
      const uint32_t ti = (uint32_t)ne + (vo ? (uint32_t)(tx + i * kSlabThreads) : 0u);
      const u64x2 id = *(__attribute__((address_space(3))) u64x2 *)(lds + thr0 + 16u * ti);
      const uint32_t fl = opaque(ss.own) >> 16;
      fix_flags_sel(fl & 63, (fl >> 6) & 1, (fl >> 7) & 1, 1, g.bc == EA_OBC, x);
      int ei = 0;
      const uint64_t nw = x.sv ^ site_flip(x, id[0], id[1], ph.energy, ei);
      if (vo) {
        *(global_t<uint64_t> *)(Db + opaque(8u * (uint32_t)(g.half + ss.h))) = nw;
        e += ei;
      }
    }
    if (ph.energy) {
      e = group_scan<64>(e);
      if ((tx & 63) == 63) atomicAdd(&esum[s & 1], e);
    }
    stage(bstride - BUF);  // the next sample's words, into the other buffer
    lds_barrier();
    if (ph.energy && tx == 0) {
      gv(a.part)[(uint32_t)s * pstride + pbase] = esum[s & 1];
      esum[s & 1] = 0;
    }
  };
  for (int s = 0; s < g.ns; ++s) sample(s, (uint32_t)(s & 1) * bstride);
}

// After a kind-4 sweep: the tubes that ran it swap their two spin arrays.
__global__ void k_swap_spins(Tube *tubes, int ntubes, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntubes) return;
  Tube &t = tubes[i];
  if ((step >= 0 ? step : *t.step) >= t.nsteps) return;
  uint64_t *const a = t.spin;
  t.spin = t.spin2;
  t.spin2 = a;
}

// ----------------------------------------------------------------- energy

// Energy of every (sample, replica, clone) from the even words
// (d_computeEne, main.cu:1209-1381): sum of 384 - 2 popcount(unsatisfied).
// Block (bx, clone, s) of a tube; each wave writes one partial per replica,
// k_energy_sum adds them up.
__device__ __forceinline__ int energy_blocks_x(const Lattice &g) {
  return (g.n * g.nr + 255) / 256;
}

__global__ void __launch_bounds__(256) k_energy(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map,
                                                int step) {
  const Where wh = locate(tubes, map, step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int nbx = energy_blocks_x(g);
  const int bxi = wh.local % nbx, cs = wh.local / nbx;
  const int clone = cs % g.nt, s = cs / g.nt;
  const int nwords = g.n * g.nr;
  const int t = bxi * blockDim.x + threadIdx.x;
  const bool valid = t < nwords;
  const int h = valid ? t >> g.nr_shift : 0, r = t & (g.nr - 1);
  int e = 0;
  if (valid) {
    const auto base = gmem(a.spin) + (size_t)s * 2 * g.half + spin_row(g, r, clone);
    const uint64_t sv = base[h];
    const Nbrs nb = neighbours(g, 0, h);
    uint64_t sc[6], Jw[6];
    load_neighbours(g, base + g.half, 0, nb, 0, sc, 1);
    load_couplings(g, a.J, s, 0, h, Jw);
    int u = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) u += __popcll(sv ^ Jw[d] ^ sc[d]);
    e = 6 * 64 - 2 * u;
  }
  e = replica_sum(e, g.nr);
  const int lane = threadIdx.x & 63, nw = 4 * nbx;
  const int wv = bxi * 4 + (threadIdx.x >> 6);
  if (lane < g.nr) gv(a.part)[(((size_t)s * g.nr + r) * g.nt + clone) * nw + wv] = e;
}

// x of another lane of the same quad, by DPP quad_perm (SEL holds the
// source lane of each quad position, two bits each).
template <int SEL>
__device__ __forceinline__ uint64_t quad_swap(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)x, SEL, 0xf, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(x >> 32), SEL, 0xf, 0xf, false);
  return (uint64_t)hi << 32 | lo;
}

// Sum of v over the wave, in every lane.
__device__ __forceinline__ int wave_sum(int v) {
  return __builtin_amdgcn_readlane(group_scan<64>(v), 63);
}

// Site overlap Q = sum_i s_i(r1) s_i(r2) of replicas r1 < r2, each at the
// clone it holds at beta, for every sample, pair and beta: one workgroup per
// (sample, pair, beta), 64 - 2 popcount(w1 ^ w2) per word of 64 sites,
// summed by wave reductions.  Output [ns][pair][nt] (pairs in the order
// (0,1), (0,2), ..., (nr-2, nr-1)).  The reference's program has no such
// measure; its overlaps are taken from the configuration files offline.
__global__ void __launch_bounds__(256) k_overlap(const Tube *__restrict__ tubes,
                                                 const uint64_t *__restrict__ map, int step) {
  __shared__ int wsum[256 / 64];
  const Where wh = locate(tubes, map, step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int np = g.nr * (g.nr - 1) / 2;
  const int beta = wh.local % g.nt, sp = wh.local / g.nt, s = sp / np;
  int r1 = 0, k = sp - s * np;
  while (k >= g.nr - 1 - r1) {
    k -= g.nr - 1 - r1;
    ++r1;
  }
  const int r2 = r1 + 1 + k;
  const int c1 = gv(a.wctb)[((size_t)s * g.nr + r1) * g.nt + beta];
  const int c2 = gv(a.wctb)[((size_t)s * g.nr + r2) * g.nt + beta];
  const auto base = gmem(a.spin) + (size_t)s * 2 * g.half;
  const size_t o1 = spin_row(g, r1, c1), o2 = spin_row(g, r2, c2);
  int d = 0;
  for (int w = threadIdx.x; w < 2 * g.n; w += blockDim.x) {
    const int p = w >= g.n, h = w - p * g.n;
    const size_t off = (size_t)p * g.half + h;
    d += __popcll(base[o1 + off] ^ base[o2 + off]);
  }
  d = wave_sum(d);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += wsum[i];
    gv(a.ovl)[wh.local] = 128 * g.n - 2 * t;  // V = 128 n sites
  }
}

// Adds the nw partials of each of a chain's nt slots (slot-major) to
// out[slot], which the caller has zeroed: lane l takes the l-th of 64
// contiguous ranges of the nt*nw partials, loads them 8 at a time with all 8
// loads in flight, and adds each run of one slot to out[slot] with one LDS
// atomic.  Integer sums, so the order of the atomics does not matter.  The
// caller synchronises before reading out[].
template <typename P>
__device__ __forceinline__ void sum_partials(P part, int nt, int nw, int lane, int32_t *out) {
  const int total = nt * nw, per = (total + 63) / 64;
  const int e0 = min(lane * per, total), e1 = min(e0 + per, total);
  if (e0 >= e1) return;
  int b = e0 / nw, next = (b + 1) * nw, sum = 0;
  for (int e = e0; e < e1; e += 8) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = e + u < e1 ? (int)part[e + u] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (e + u < e1 && e + u == next) {  // a new slot starts here
        atomicAdd(out + b, sum);
        sum = 0;
        ++b;
        next += nw;
      }
      sum += v[u];
    }
  }
  atomicAdd(out + b, sum);
}

// ener[chain][c] from the chain's partials: the sweep's partials by beta
// (BY_BETA, clone c = wctb[chain][b]) or k_energy's partials by clone.
template <bool BY_BETA>
__global__ void __launch_bounds__(64) k_energy_sum(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map,
                                                   int step) {
  __shared__ int32_t E[EA_MAX_BETAS];
  const Where wh = locate(tubes, map, step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const int chain = wh.local, nt = a.g.nt, lane = threadIdx.x;
  const int nw = BY_BETA ? a.nbx : 4 * energy_blocks_x(a.g);
  for (int b = lane; b < nt; b += 64) E[b] = 0;
  __syncthreads();
  sum_partials(a.part + (size_t)chain * nt * nw, nt, nw, lane, E);
  __syncthreads();
  for (int b = lane; b < nt; b += 64) {
    const int c = BY_BETA ? a.wctb[(size_t)chain * nt + b] : b;
    gv(a.ener)[(size_t)chain * nt + c] = E[b];
  }
}

// ------------------------------------------------------ parallel tempering

// exp(pt) of the host's libm for pair i and energy difference d < 0: the
// tabulated value, 0 where exp underflows.
__device__ __forceinline__ double exact_exp(const Tube &a, int i, int d) {
  // energies are sums of 384 - 2*popcount terms, so d is even and k >= 1
  const int k = -d / 2;
  if (k < 1) return exp(a.nd[i] * (double)d);
  return k <= a.tlen[i] ? a.tab[a.toff[i] + k - 1] : 0.0;
}

// f^m(z) for the congruential step f(z) = kLcg z + 1 (mod 2^64) of the
// Parisi-Rapuano generator (random_generator.h:133): the affine map is
// squared (A, C) -> (A A, A C + C) and applied for each set bit of m.
__device__ __forceinline__ uint64_t lcg_jump(uint64_t z, uint32_t m) {
  uint64_t A = kLcg, C = 1;
  for (; m; m >>= 1) {
    if (m & 1u) z = A * z + C;
    C = A * C + C;
    A = A * A;
  }
  return z;
}

// Parallel_Tempering (upd.c:3-48), one wave per (sample, replica) chain.
// The chain's Parisi-Rapuano state is staged in LDS.  Its draws do not
// depend on the swap decisions, so the wave makes them first: with lags 24,
// 55 and 61, draws k .. k+23 only read earlier values, so 24 lanes draw at a
// time.  Then each pair's uniform u = FNORM*final and log(u) are formed in
// parallel, and lane 0 walks the pairs in order: swap when pt >= 0, else when
// log(u) < pt -- unless the two are within 1e-12 (or u = 0), where the exact
// comparison u < exp(pt) is made with the host libm's exp from the table.
// Both sides' errors are far below 1e-12, so every decision equals the
// reference's.  The statistics are then added by one lane per pair with the
// tabulated exp values.  The block is one wave and its barriers wait for
// LDS traffic only (lds_barrier): __syncthreads would also wait for every
// global store in flight, such as the generator state's, in mid-walk.
__global__ void __launch_bounds__(64) k_pt(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map, Phase ph) {
  __shared__ uint64_t ira[256];
  __shared__ uint64_t zs[EA_MAX_BETAS], prs[EA_MAX_BETAS];
  __shared__ double uu[EA_MAX_BETAS], lu[EA_MAX_BETAS];
  __shared__ int32_t E[EA_MAX_BETAS];   // by clone
  __shared__ int32_t Ep[EA_MAX_BETAS];  // by position (beta) before the swaps
  __shared__ uint8_t perm[EA_MAX_BETAS + 1];
  __shared__ int32_t dE[EA_MAX_BETAS];
  __shared__ uint8_t took[EA_MAX_BETAS];
  __shared__ double ndv[EA_MAX_BETAS];
  __shared__ double sacc[EA_MAX_BETAS], satt[EA_MAX_BETAS], sraw[EA_MAX_BETAS],
      sesum[EA_MAX_BETAS];
  __shared__ int32_t stlen[EA_MAX_BETAS], stoff[EA_MAX_BETAS];
  __shared__ int32_t dmin[EA_MAX_BETAS];  // pair i swaps iff its energy step >= dmin[i]
  __shared__ int exact_walk;
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int chain = wh.local, nt = g.nt, np = nt - 1, lane = threadIdx.x;
  // global views of the chain's arrays (see gv)
  const auto st = gv(a.pt + chain);
  const auto wctb = gv(a.wctb) + (size_t)chain * nt;
  const auto ener = gv(a.ener) + (size_t)chain * nt;
  const auto esum = gv(a.esum) + (size_t)chain * nt;
  // everything the serial pass reads is staged before it: no global load
  // sits on the chain of dependent swap decisions.  The loads of every array
  // are issued into registers before the first of them is stored to LDS, so
  // the staging costs one memory round trip rather than one per array.
  constexpr int kPer = (EA_MAX_BETAS + 63) / 64;  // slots per lane
  const bool measure = ph.measure, energy = ph.energy;
  const size_t k0 = (size_t)chain * np;
  const auto acc = gv(a.acc) + k0, att = gv(a.att) + k0, raw = gv(a.raw) + k0;
  const uint64_t zseed = st->zseed;
  const uint8_t ip = st->ip, ip1 = st->ip1, ip2 = st->ip2, ip3 = st->ip3;
  uint64_t r_ira[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) r_ira[j] = st->ira[lane + 64 * j];
  uint32_t r_perm[kPer];
  int32_t r_en[kPer], r_tlen[kPer], r_toff[kPer];
  double r_es[kPer], r_nd[kPer], r_acc[kPer], r_att[kPer], r_raw[kPer];
  // (no branches around the loads: indexes past an array's end read its
  // last element, nt >= 2 so every array has one, and are not stored)
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = lane + 64 * j, it = min(i, nt - 1), ip_ = min(i, np - 1);
    r_perm[j] = wctb[it];
    r_es[j] = esum[it];
    r_en[j] = ener[it];
    r_nd[j] = gv(a.nd)[ip_];
    r_tlen[j] = gv(a.tlen)[ip_];
    r_toff[j] = gv(a.toff)[ip_];
    r_acc[j] = acc[ip_];
    r_att[j] = att[ip_];
    r_raw[j] = raw[ip_];
  }
  if (energy) {
    // the sweep left partial energies by beta: dE sums them by beta for a
    // moment, E receives them by clone
    for (int b = lane; b < nt; b += 64) dE[b] = 0;
    lds_barrier();  // the staging loads stay in flight
    sum_partials(gv(a.part) + (size_t)chain * nt * a.nbx, nt, a.nbx, lane, dE);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) ira[lane + 64 * j] = r_ira[j];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = lane + 64 * j;
    if (i < nt) {
      perm[i] = (uint8_t)r_perm[j];
      if (measure) sesum[i] = r_es[j];
      if (!energy) E[i] = r_en[j];
    }
    if (i < np) {
      ndv[i] = r_nd[j];
      stlen[i] = r_tlen[j];
      stoff[i] = r_toff[j];
      sacc[i] = r_acc[j];
      satt[i] = r_att[j];
      sraw[i] = r_raw[j];
    }
  }
  lds_barrier();
  if (energy) {
    for (int b = lane; b < nt; b += 64) {
      E[perm[b]] = dE[b];
      ener[perm[b]] = dE[b];
    }
    lds_barrier();
  }
  for (int p = lane; p < nt; p += 64) Ep[p] = E[perm[p]];
  // the congruential half: zs[k] = f^(k+1)(zseed) with f(z) = kLcg z + 1,
  // every k at once by squaring the affine map (lcg_jump)
  for (int k = lane; k < np; k += 64) zs[k] = lcg_jump(zseed, (uint32_t)k + 1);
  lds_barrier();
  // a ring slot is rewritten 256 draws after it was written, and every
  // value a draw reads was made 24 to 61 draws before it: each round reads
  // only slots of earlier rounds that no later draw has touched yet
  for (int k0 = 0; k0 < np; k0 += 24) {
    const int k = k0 + lane;
    const bool act = lane < 24 && k < np;
    uint64_t nv = 0, pr = 0;
    if (act) {
      nv = ira[(uint8_t)(ip + k - 24)] + ira[(uint8_t)(ip + k - 55)];
      pr = nv ^ ira[(uint8_t)(ip + k - 61)];
    }
    lds_barrier();
    if (act) {
      ira[(uint8_t)(ip + k)] = nv;
      prs[k] = pr;
    }
    lds_barrier();
  }
  if (lane == 0) exact_walk = a.pt_exact;
  for (int k = lane; k < np; k += 64) {
    const uint64_t fin = zs[k] + prs[k];
    const double u = kFnorm * (double)fin;
    uu[k] = u;
    lu[k] = log(u);
  }
  lds_barrier();
  // Each pair's decision is monotone in its (even) energy step d: pt grows
  // with d, and both the log comparison and the exact one against the
  // table of exp values accept more as pt grows.  So pair i swaps iff
  // d >= dmin[i], the smallest even d its own decision rule accepts, found
  // by evaluating that rule around log(u)/nd; a threshold that does not
  // verify (rule true at dmin, false at dmin - 2) sends the walk back to
  // the rule itself.
  for (int k = lane; k < np; k += 64) {
    auto take = [&](int d) {
      const double pt = ndv[k] * (double)d;
      if (pt >= 0) return true;
      const double u = uu[k], l = lu[k];
      if (u == 0.0 || fabs(l - pt) <= 1e-12) return u < exact_exp(a, k, d);
      return l < pt;
    };
    int dm;
    if (!(ndv[k] > 0.0)) {
      dm = INT_MIN;  // pt = 0: always swap
    } else if (uu[k] == 0.0) {
      dm = -2 * stlen[k];  // u = 0 < exp(pt) while the table has a value
    } else {
      // the rule turns near d = lu / nd < 0: the search starts at the even
      // d at or below an estimate of it from the approximate reciprocal and
      // steps up until the rule accepts.  The check below catches a start
      // above the threshold or too far below it (an estimate off by more
      // than a few steps, possible only for |d| far beyond the tube's
      // energy steps) and sends the walk to the exact rule.
      const double x = lu[k] * __builtin_amdgcn_rcp(ndv[k]);
      dm = x < -1e9 ? INT_MIN : 2 * (int)floor(x * 0.5);
      for (int t = 0; t < 8 && dm != INT_MIN && !take(dm); ++t) dm += 2;
    }
    if (dm > 0) dm = 0;
    if (dm != INT_MIN && (!take(dm) || take(dm - 2))) exact_walk = 1;
    dmin[k] = dm;
  }
  lds_barrier();
  if (!exact_walk) {
    // every decision is d >= dmin[i]: the walk runs in scalar registers, 64
    // pairs at a time.  Lane k holds pair base+k's clone and energy above it
    // and its threshold, which the walk reads with readlane, and keeps pair
    // base+k's outcome, stored once the chunk's steps are made.  The clone
    // moving up the ladder and its energy stay in registers.
    int pc = __builtin_amdgcn_readfirstlane(perm[0]);
    int32_t ec = __builtin_amdgcn_readfirstlane(Ep[0]);
    for (int base = 0; base < np; base += 64) {
      const int i = base + lane, cnt = min(64, np - base);
      const int vpb = i < np ? perm[i + 1] : 0;
      const int32_t veb = i < np ? Ep[i + 1] : 0, vdm = i < np ? dmin[i] : 0;
      // the serial pass keeps only the energy carried up the ladder and the
      // take bits, in scalar registers; the readlanes do not depend on it,
      // so eight pairs' operands are read first and then stepped through
      // (a loop step per pair, readlane to compare to branch, took about
      // 110 cycles)
      uint64_t takes = 0;
      int32_t e = ec;
      for (int k0 = 0; k0 < cnt; k0 += 8) {  // k0 + 7 <= 63
        int32_t eb[8], dm[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          eb[u] = __builtin_amdgcn_readlane(veb, k0 + u);
          dm[u] = __builtin_amdgcn_readlane(vdm, k0 + u);
        }
        // (steps past the chunk's last pair read lanes holding zeros: their
        // bits are masked off below, and the energy they carry is not used)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool t = eb[u] - e >= dm[u];
          const uint64_t bit = 1ull << (k0 + u);
          takes = t ? takes | bit : takes;  // selects: the step stays scalar
          e = t ? e : eb[u];
        }
      }
      // each pair's outcome from the bits: the clone that meets position
      // i + 1 is the one carried from the last refusal j < i of the chunk
      // (position j + 1's clone and energy, lane j's), or, with none, the one
      // carried into the chunk (pc, ec)
      const uint64_t refused = ~takes & (cnt == 64 ? ~0ull : (1ull << cnt) - 1);
      const uint64_t prior = refused & ((1ull << lane) - 1);
      const int j = prior ? 63 - __clzll((long long)prior) : 0;
      const int cj = __shfl(vpb, j), ej = __shfl(veb, j);
      const int pcl = prior ? cj : pc;
      const int32_t ecl = prior ? ej : ec;
      const bool take = (takes >> lane) & 1ull;
      const int32_t myd = veb - ecl;
      const int mykeep = take ? vpb : pcl;  // the carried clone moves on when taken
      if (refused) {
        const int jl = 63 - __clzll((long long)refused);
        pc = __builtin_amdgcn_readlane(vpb, jl);
        ec = __builtin_amdgcn_readlane(veb, jl);
      }
      lds_barrier();  // the chunk's reads of perm are done
      if (i < np) {
        dE[i] = myd;
        took[i] = (uint8_t)take;
        perm[i] = (uint8_t)mykeep;
      }
      lds_barrier();
    }
    if (lane == 0) perm[np] = (uint8_t)pc;
  } else if (lane == 0) {
    // the exact rule, pair by pair: position i+1 still holds its clone from
    // before the walk, so no LDS access depends on an earlier decision
    int pc = perm[0];
    int32_t ec = Ep[0];
    for (int i = 0; i < np; ++i) {
      const int pb = perm[i + 1];
      const int32_t eb = Ep[i + 1];
      const int d = eb - ec;
      const double pt = ndv[i] * (double)d;
      bool take = true;
      if (!(pt >= 0)) {
        const double u = uu[i], l = lu[i];
        if (u == 0.0 || fabs(l - pt) <= 1e-12)
          take = u < exact_exp(a, i, d);
        else
          take = l < pt;
      }
      dE[i] = d;
      took[i] = take;
      if (take) {  // pc moves to position i+1
        perm[i] = (uint8_t)pb;
      } else {
        perm[i] = (uint8_t)pc;
        pc = pb;
        ec = eb;
      }
    }
    perm[np] = (uint8_t)pc;
  }
  if (lane == 0) {  // the generator's state after the call's last draw
    st->final_ = zs[np - 1] + prs[np - 1];
    st->PR = prs[np - 1];
    st->zseed = zs[np - 1];
    st->ip = (uint8_t)(ip + np);
    st->ip1 = (uint8_t)(ip1 + np);
    st->ip2 = (uint8_t)(ip2 + np);
    st->ip3 = (uint8_t)(ip3 + np);
  }
  lds_barrier();
  for (int i = lane; i < 256; i += 64) st->ira[i] = ira[i];
  for (int i = lane; i < nt; i += 64) wctb[i] = perm[i];
  if (measure) {  // energy sums, history (main.cu:208-228)
    const int itb = a.it0 + wh.step;
    const int row = itb % a.frec == a.frec - 1 ? itb / a.frec : -1;
    for (int b = lane; b < nt; b += 64) {
      const int c = perm[b];
      esum[b] = sesum[b] + (double)E[c];
      if (row >= 0) gv(a.hist)[((size_t)row * g.ns * g.nr + chain) * nt + c] = (uint8_t)b;
    }
  }
  // statistics (upd.c:20-44): the exp of a rejected-side pair from the
  // table of host exp values (exact_exp), with the table's bounds staged
  for (int i = lane; i < np; i += 64) {
    const int d = dE[i];
    const double pt = ndv[i] * (double)d;
    double x = 1.0;
    if (!(pt >= 0)) {
      const int kk = -d / 2;
      x = kk < 1 ? exp(pt) : kk <= stlen[i] ? gv(a.tab)[stoff[i] + kk - 1] : 0.0;
    }
    att[i] = satt[i] + 1.0;
    acc[i] = sacc[i] + x;
    if (took[i]) raw[i] = sraw[i] + 1.0;
  }
}

// ------------------------------------------------------------ step counter

// Advances the step counter of every tube that has steps left.
__global__ void k_step_end(const Tube *__restrict__ tubes, int ntubes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntubes && *tubes[i].step < tubes[i].nsteps) *tubes[i].step += 1;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f64(double *p, double v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// --------------------------------------------------------------- Houdayer

// Cluster coins.  The reference draws cuRAND XORWOW doubles
// (HoudayerXEA.cu:174-195,373): coin[site] = U(idx) < 1/2 with
// idx = offset + s*V*NT + beta*V + site.  Here idx keeps that scheme and the
// coin is bit (idx mod 128) of Philox4x32-10 at counter
// (b_lo, b_hi, 0x486f7564, 0), b = idx / 128, key (seed_lo, seed_hi): one
// Philox call serves 128 consecutive sites.  coin_word returns the coins of
// the 32 indices idx0 .. idx0+31 (idx0 a multiple of 32).
__device__ __forceinline__ uint32_t coin_word(uint64_t seed, uint64_t idx0) {
  const uint64_t b = idx0 >> 7;
  uint32_t v0 = (uint32_t)b, v1 = (uint32_t)(b >> 32), v2 = 0x486f7564u, v3 = 0;
  philox10(v0, v1, v2, v3, (uint32_t)seed, (uint32_t)(seed >> 32));
  switch ((idx0 >> 5) & 3) {
    case 0: return v0;
    case 1: return v1;
    case 2: return v2;
    default: return v3;
  }
}

// A Houdayer instance (group, sample, beta < ctm) of a tube: local index
// beta + ctm * (2 * s + grp), and the inputs of the call in flight.
struct Inst {
  int beta, s, grp;
  const int32_t *pairs;  // r1, r2 of group 0, then of group 1
  uint64_t coin;         // coin index of (this group, sample 0, beta 0, site 0)
};
__device__ __forceinline__ Inst instance_of(const Tube &a, int local, int step, int call) {
  Inst in;
  in.beta = local % a.g.ctm;
  const int sg = local / a.g.ctm;
  in.s = sg >> 1;
  in.grp = sg & 1;
  const uint64_t k = (uint64_t)step * a.calls + call;
  in.pairs = a.pairs + k * 4;
  in.coin = a.coin0 + k * a.coin_per_call + (uint64_t)in.grp * a.g.V * a.g.nt * a.g.ns;
  return in;
}

// Word offsets of replicas r1, r2 of the instance (group, sample, beta).
struct Pair {
  size_t o1, o2;
};
__device__ __forceinline__ Pair pair_words(const Lattice &g, const Tube &a, const int32_t *pairs,
                                           int s, int grp, int beta) {
  const int r1 = gv(pairs)[2 * grp], r2 = gv(pairs)[2 * grp + 1];
  const int c1 = gv(a.wctb)[((size_t)s * g.nr + r1) * g.nt + beta];
  const int c2 = gv(a.wctb)[((size_t)s * g.nr + r2) * g.nt + beta];
  const size_t sb = (size_t)s * 2 * g.half;
  Pair q;
  q.o1 = sb + spin_row(g, r1, c1);
  q.o2 = sb + spin_row(g, r2, c2);
  return q;
}

__device__ __forceinline__ uint32_t find_root(const uint32_t *lab, uint32_t a) {
  const volatile uint32_t *v = lab;
  uint32_t p = v[a];
  while (p != a) {
    a = p;
    p = v[a];
  }
  return a;
}

// Root of a with path splitting: every node passed is re-pointed at its
// grandparent.  The re-pointing is an atomicMin to an ancestor, so it never
// cuts a node from its tree, and a link it displaces was made by a union
// that does not rely on it (see unite).
__device__ __forceinline__ uint32_t find_split(uint32_t *lab, uint32_t a) {
  const volatile uint32_t *v = lab;
  uint32_t p = v[a];
  while (p != a) {
    const uint32_t gp = v[p];
    if (gp != p) atomicMin(lab + a, gp);
    a = p;
    p = gp;
  }
  return a;
}

// Link the trees of a and b, the larger root under the smaller one; a link
// displaced by a concurrent union is re-linked (Komura, CPC 194 (2015) 54).
// A union only relies on its own link when it found a root there; nodes
// that are not roots are only ever re-pointed at ancestors.
__device__ __forceinline__ void unite(uint32_t *lab, uint32_t a, uint32_t b) {
  a = find_split(lab, a);
  b = find_split(lab, b);
  while (a != b) {
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicMin(lab + a, b);
    if (old == a) return;
    a = old;
  }
}

// Union-find forest in LDS, parents as 32-bit node indices.  Loads and stores
// are relaxed workgroup-scope atomics: every step re-reads the parent from
// LDS, and the accesses stay LDS instructions.  Every parent index is below
// its child's, so the forest has no cycles; a node that is not a root only
// ever gets re-pointed at an ancestor.
template <typename I = uint32_t, bool Swz = false>
struct Forest {
  I *p;
  // LDS slot of entry i.  With Swz, bits 8-13 of i are folded into the bits
  // that pick the LDS bank (bits 1-6 for 16-bit entries, 0-5 for 32-bit
  // ones), so entries a multiple of 256 apart sit in different banks.  The
  // map is its own inverse, keeps every 128-entry block in place (the forest
  // size is a multiple of 128) and keeps the two 16-bit entries of a 32-bit
  // word together.
  static __device__ __forceinline__ uint32_t ix(uint32_t i) {
    if constexpr (Swz)
      return i ^ (((i >> 8) & 63u) << (sizeof(I) == 2 ? 1 : 0));
    else
      return i;
  }
  __device__ __forceinline__ uint32_t get(uint32_t i) const {
    return __hip_atomic_load(p + ix(i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ void put(uint32_t i, uint32_t v) const {
    __hip_atomic_store(p + ix(i), (I)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // p[a] = min(p[a], b), returning the previous p[a].  16-bit parents sit
  // two to a 32-bit LDS word and are lowered by compare-and-swap of the
  // word (LDS has no 16-bit atomic minimum).
  __device__ __forceinline__ uint32_t lower(uint32_t a, uint32_t b) const {
    if constexpr (sizeof(I) == 4) {
      return atomicMin(p + ix(a), b);
    } else {
      uint32_t *w = reinterpret_cast<uint32_t *>(p) + (ix(a) >> 1);
      const int sh = (a & 1) * 16;
      uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      for (;;) {
        const uint32_t cur = (old >> sh) & 0xFFFFu;
        if (cur <= b) return cur;
        const uint32_t nw = (old & ~(0xFFFFu << sh)) | (b << sh);
        if (__hip_atomic_compare_exchange_strong(w, &old, nw, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
          return cur;
      }
    }
  }
  __device__ __forceinline__ uint32_t root(uint32_t a) const {
    uint32_t q = get(a);
    while (q != a) {
      a = q;
      q = get(a);
    }
    return a;
  }
  // root with path splitting (each node passed is pointed at its grandparent)
  __device__ __forceinline__ uint32_t find(uint32_t a) const {
    uint32_t q = get(a);
    while (q != a) {
      const uint32_t gq = get(q);
      if (gq != q) put(a, gq);
      a = q;
      q = gq;
    }
    return a;
  }
  // Komura's union: the larger root goes under the smaller one; a link
  // displaced by a concurrent union is made again from the displaced parent.
  __device__ __forceinline__ void unite(uint32_t a, uint32_t b) const {
    a = find(a);
    b = find(b);
    while (a != b) {
      if (a < b) {
        const uint32_t t = a;
        a = b;
        b = t;
      }
      const uint32_t old = lower(a, b);
      if (old == a) return;
      a = old;
    }
  }
};

// Rows of L <= 64 sites as bit masks (bit x = site x); RowT is 32 bits wide
// when L <= 32.
template <typename RowT>
struct Rows {
  static __device__ __forceinline__ int hi(RowT v) {
    if constexpr (sizeof(RowT) == 4) return 31 - __clz(v);
    else return 63 - __clzll(v);
  }
  static __device__ __forceinline__ int lo(RowT v) {
    if constexpr (sizeof(RowT) == 4) return __ffs(v) - 1;
    else return __ffsll((unsigned long long)v) - 1;
  }
  // run boundaries of the periodic row: bit x where q[x] != q[x-1 mod L]
  static __device__ __forceinline__ RowT bounds(RowT q, int L, RowT full) {
    return (q ^ ((q << 1) | (q >> (L - 1)))) & full;
  }
  // first site of the run holding x, 0 for the run that wraps past L-1 (and
  // for a constant row): the smallest x of the run
  static __device__ __forceinline__ int rep(RowT b, int x) {
    const RowT m = b & ((RowT(2) << x) - 1);
    if (m == 0) return 0;
    const int s = hi(m);
    return (!(b & 1) && s == hi(b)) ? 0 : s;
  }
  // rep(b, x) without a branch, from two constants of the row: b1 = b | 1
  // and hbe = wrap_start(b, L), where the run that wraps past L-1 starts (L
  // when none does); sites from hbe on belong to the row start, and bit 0 of
  // b1 gives the sites before the first boundary rep 0
  static __device__ __forceinline__ int wrap_start(RowT b, int L) {
    return (b && !(b & 1)) ? hi(b) : L;
  }
  static __device__ __forceinline__ int rep_at(RowT b1, int hbe, int x) {
    return x >= hbe ? 0 : hi(b1 & ((RowT(2) << x) - 1));
  }
  static __device__ __forceinline__ int popc(RowT v) {
    if constexpr (sizeof(RowT) == 4) return __popc(v);
    else return __popcll((unsigned long long)v);
  }
  // runs of a periodic row with boundaries b (one for a constant row)
  static __device__ __forceinline__ int runs(RowT b) { return b ? popc(b) : 1; }
  // rank of the run holding x among the row's runs ordered by their first
  // sites, the run that wraps past L-1 (first site 0) being rank 0; b1 and
  // hbe as for rep_at
  static __device__ __forceinline__ int rank(RowT b1, int hbe, int x) {
    return x >= hbe ? 0 : popc(b1 & ((RowT(2) << x) - 1)) - 1;
  }
  // first site of the run of rank k
  static __device__ __forceinline__ int run_first(RowT b, int k) {
    RowT b1 = b | 1;
    for (int i = 0; i < k; ++i) b1 &= b1 - 1;
    return lo(b1);
  }
  // fn(k, first, seg) for every run of the row, seg being its sites
  template <typename F>
  static __device__ __forceinline__ void for_runs(RowT b, int L, RowT full, F fn) {
    if (b == 0) {
      fn(0, 0, full);
      return;
    }
    const int first = lo(b);
    int k = first == 0 ? 0 : 1;
    RowT bb = b;
    while (bb) {
      const int st = lo(bb);
      bb &= bb - 1;
      RowT seg = (bb ? (RowT(1) << lo(bb)) : RowT(0)) - (RowT(1) << st);
      if (!bb) seg &= full;
      if (!bb && first != 0) {  // the last run wraps into the row start
        fn(0, 0, seg | ((RowT(1) << first) - 1));
      } else {
        fn(k, st, seg);
        ++k;
      }
    }
  }
};

// Cluster labelling of a slab of `planes` full z-planes in LDS, all threads
// of the block taking part.  rows[r] holds the q bits of row r = y + L*z
// (bit x; L <= 64, a multiple of 8); lab[i] is the parent of local site
// i = x + L*r.  Label values are local site indices; after link_rows, the
// first site of each run of equal q along x is a node of a forest whose
// roots are the smallest index of each equal-q cluster, and every other
// site points at its run's first site:
//  1. every site points at the first site of its run along the periodic x
//     row (a run wrapping past x = L-1 belongs to the row start), 8 sites
//     per thread: the first one's run start is found from the run
//     boundaries, the others step along them;
//  2. y and z bonds between rows, from bit masks: a bond at x is united
//     only where a run overlap starts (when the bond at x-1 joins the same
//     two runs the union is implied).
// z bonds leave the slab only when wrap_z (the whole periodic tube).
// Row order of the linking loops: the lanes of a wave take rows about
// nrows/64 apart, so that concurrent unions mostly meet different clusters
// instead of contending for the same roots: lane l = rr mod 64 of loop row
// rr takes row l * m + j of its range of m = nrows/64 rows, with
// j = (rr / 64 + l / 2) mod m.  The rotation by l / 2 puts the lanes' rows
// (and their run bases) in different LDS banks; without it rows m words
// apart share a few banks.  Rows from 64 * m on (m = nrows / 64 rounded
// down) keep their place.
__device__ __forceinline__ int spread_row(int rr, int nrows) {
  const int m = nrows >> 6;
  if (rr >= m << 6) return rr;
  const int lane = rr & 63;
  int j = (rr >> 6) + (lane >> 1) % m;
  if (j >= m) j -= m;
  return lane * m + j;
}

template <typename I>
__device__ void link_rows(const Lattice &g, I *lab, const uint64_t *rows, int planes,
                          bool wrap_z) {
  static_assert(sizeof(I) == 2, "8 sites per 16-byte store");
  using R = Rows<uint64_t>;
  const Forest<I> f{lab};
  const int L = g.L, nrows = planes * L, T = blockDim.x, tid = threadIdx.x;
  const uint64_t full = L == 64 ? ~0ULL : ((1ULL << L) - 1);
  for (int c = tid; 8 * c < nrows * L; c += T) {
    const int r = qdiv(8 * c, g.mL), x0 = 8 * c - r * L;
    const uint64_t b = R::bounds(rows[r], L, full);
    const int hb = (b && !(b & 1)) ? R::hi(b) : L;  // start of a run that wraps
    uint32_t rep = (uint32_t)R::rep(b, x0), v[4];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int x = x0 + k;
      if (k > 0 && ((b >> x) & 1ULL)) rep = x == hb ? 0u : (uint32_t)x;
      const uint32_t l = (uint32_t)(r * L) + rep;
      if (k & 1)
        v[k >> 1] |= l << 16;
      else
        v[k >> 1] = l;
    }
    *reinterpret_cast<u32x4 *>(lab + 8 * c) = u32x4{v[0], v[1], v[2], v[3]};
  }
  __syncthreads();
  {
    // Links: id = r + nrows * dir joins row r to row r+1 in y (dir 0) or to
    // the next plane (dir 1), taken id = tid, tid + T, ... in one loop.
    uint64_t todo = 0;
    int base = 0, base2 = 0;
    for (int id = tid;;) {
      while (!todo && id < 2 * nrows) {
        const int dir = id >= nrows, r = spread_row(id - dir * nrows, nrows);
        id += T;
        const int zl = qdiv(r, g.mL), y = r - zl * L;
        int r2;
        if (dir == 0)
          r2 = y + 1 == L ? r - (L - 1) : r + 1;
        else if (zl + 1 < planes)
          r2 = r + L;
        else if (wrap_z)
          r2 = r - L * (planes - 1);
        else
          continue;
        const uint64_t q = rows[r], q2 = rows[r2];
        const uint64_t lq = ~(q ^ (q << 1)) & full & ~1ULL;  // q[x] == q[x-1]
        const uint64_t l2 = ~(q2 ^ (q2 << 1)) & full & ~1ULL;
        todo = ~(q ^ q2) & full & ~(lq & l2);
        base = r * L;
        base2 = r2 * L;
      }
      if (!todo) break;
      const int x = R::lo(todo);
      todo &= todo - 1;
      f.unite(base + x, base2 + x);
    }
  }
  __syncthreads();
}

// q bits of row r = y + L z from the two replicas' words: the L sites of a
// row use ML words (parity alternating with x), each for its 4 x copies (bx),
// which are the 4 consecutive bits 4 by + 16 bz ... +3 of the word.  RowT is
// the row's width (32 bits when L <= 32).
template <typename RowT = uint64_t, typename Load>
__device__ __forceinline__ RowT row_bits(const Lattice &g, int r, Load load) {
  const int z = qdiv(r, g.mL), y = r - z * g.L;
  const int by = qdiv(y, g.mML), yb = y - by * g.ML;
  const int bz = qdiv(z, g.mMLz), zg = z - bz * g.MLz;
  const int hrow = g.ML * yb + g.MS * zg, bitrow = 4 * by + 16 * bz;
  const int prow = (y ^ z) & 1;
  RowT bits = 0;
  for (int ax = 0; ax < g.ML; ++ax) {
    const uint32_t nib = (uint32_t)(load((ax + prow) & 1, (ax + hrow) >> 1) >> bitrow);
#pragma unroll
    for (int bx = 0; bx < 4; ++bx) bits |= (RowT)((nib >> bx) & 1u) << (bx * g.ML + ax);
  }
  return bits;
}

// Move bits of word (p, h) gathered from the per-row move masks mrow; 4
// threads per word, thread `quarter` gathering the 16 copies of z block
// bz = quarter.
template <typename RowT>
__device__ __forceinline__ uint64_t word_move_bits(const Lattice &g, const RowT *mrow,
                                                   int p, int h, int quarter) {
  const int gz = qdiv(h, g.mhs), rem = h - gz * g.hs, yb = qdiv(rem, g.mhl);
  const int ax = 2 * (rem - yb * g.hl) + ((p + yb + gz) & 1);
  const int z = quarter * g.MLz + gz;
  uint64_t m = 0;
#pragma unroll
  for (int by = 0; by < 4; ++by) {
    const uint64_t row = mrow[(by * g.ML + yb) + g.L * z] >> ax;
#pragma unroll
    for (int bx = 0; bx < 4; ++bx)
      m |= ((row >> (bx * g.ML)) & 1ULL) << (bx + 4 * by + 16 * quarter);
  }
  return m;
}

// The run forest: entry i holds the parent of run node i in its high 16
// bits and the node's own first site (its smallest site index) in the low
// 16 bits, which every write keeps, so that an atomic minimum on an entry
// compares parents.  Loads and stores are relaxed workgroup-scope atomics,
// so every step re-reads LDS.
__device__ __forceinline__ uint32_t run_entry(const uint32_t *p, uint32_t i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of run node a with path splitting (each node passed is pointed at
// its grandparent, an ancestor, so the trees keep their members).  Returns
// the root's entry: its number in the high bits, its first site in the low.
__device__ __forceinline__ uint32_t run_root(uint32_t *p, uint32_t a) {
  uint32_t ea = run_entry(p, a);
  while ((ea >> 16) != a) {
    const uint32_t q = ea >> 16, eq = run_entry(p, q);
    if ((eq >> 16) != q)
      __hip_atomic_store(p + a, (eq & 0xFFFF0000u) | (ea & 0xFFFFu), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
    a = q;
    ea = eq;
  }
  return ea;
}

// Union of the trees of run nodes a and b: Rem's algorithm with splicing
// (Patwary, Blair and Manne, 2010).  Every parent is at most its node, and
// parents only ever decrease, so a tree's root is its smallest node.  Of
// the two current nodes, the one with the larger parent takes the other's
// parent by an atomic minimum: on a root that links the two trees; on any
// other node it moves that node's subtree into the other tree (a splice),
// and the walk goes on from the node's old parent.  It ends when both nodes
// have one parent, or at a link; concurrent unions only lower parents
// further, so every node stays connected to what it was joined with.
// One step of the walk; true when the union is complete.  The pair (a, b)
// is the walk's whole state, so a union left unfinished can be resumed
// later from the pair, after any number of other unions.
__device__ __forceinline__ bool rem_step(uint32_t *p, uint32_t &a, uint32_t &b) {
  const uint32_t ea = run_entry(p, a), eb = run_entry(p, b);
  if ((ea ^ eb) < 0x10000u) return true;  // one parent
  // x: the node with the larger parent, y the other (selects, no branch)
  const bool sw = ea < eb;
  const uint32_t x = sw ? b : a, ex = sw ? eb : ea, ey = sw ? ea : eb;
  b = sw ? a : b;
  const uint32_t old = __hip_atomic_fetch_min(p + x, (ey & 0xFFFF0000u) | (ex & 0xFFFFu),
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  a = old >> 16;
  return a == x;  // x was a root: linked
}

// The unions of a whole tube's labelling: for every row r, the y and z bonds
// where a run overlap with the next row starts (when the bond at x-1 joins
// the same two runs the union is implied), between the nodes of the runs
// they join.  A lane finds the bonds of its rows and pushes the node pairs
// onto its wave's queue (kHouQueue entries of two 16-bit nodes); once 64
// are queued the wave takes them, one per lane, so every lane stays busy
// however unevenly the bonds fall on the rows.  unite(u, v) works on the
// pair and returns whether the union is complete; a pair it leaves
// unfinished (its walk state, which it updates) goes back onto the queue,
// so that a drain takes a fixed number of steps instead of as many as its
// slowest union needs.  BY_RUNS: nodes are run numbers (the first run
// number of row r, rbase[r], plus the run's rank); otherwise sites (the
// run's first site, Rows::rep_at).  Lanes take rows nrows/64 apart
// (spread_row), so concurrent unions mostly meet different clusters.
template <bool BY_RUNS, typename RowT, typename Unite>
__device__ __forceinline__ void queue_links(const Lattice &g, const RowT *rows, RowT full,
                                            const uint16_t *rbase, uint32_t *queue,
                                            Unite unite) {
  using R = Rows<RowT>;
  const int L = g.L, nrows = g.V / L, T = blockDim.x, tid = threadIdx.x, lane = tid & 63;
  uint32_t *wq = queue + (tid >> 6) * kHouQueue;
  int qn = 0;  // pairs in the wave's queue, the same in every lane
  // takes the top cnt <= 64 pairs; with qn < kHouQueue before, qn stays
  // below kHouQueue (at most 64 come back)
  auto drain = [&](int cnt) {
    // the drain is a chain of dependent LDS round trips: at a raised wave
    // priority each step issues as soon as its data is back, ahead of the
    // other waves' bond pushes (the CU's arbiter otherwise favours the
    // oldest waves)
    __builtin_amdgcn_s_setprio(1);
    lds_wave_sync();  // the queue's writes are done
    uint32_t u = 0, v = 0;
    const bool got = lane < cnt;
    if (got) {
      const uint32_t e = wq[qn - cnt + lane];
      u = e & 0xFFFFu;
      v = e >> 16;
    }
    qn -= cnt;
    lds_wave_sync();  // and read before the slots are written again
    const bool left = got && !unite(u, v);
    const uint64_t m = __ballot(left);
    if (m) {  // the unfinished pairs, back onto the queue
      if (left)
        wq[qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
            u | v << 16;
      qn += __popcll(m);
    }
    __builtin_amdgcn_s_setprio(0);
  };
  auto node_base = [&](int r) { return BY_RUNS ? (uint32_t)rbase[r] : (uint32_t)(r * L); };
  auto node = [&](uint32_t nb, RowT b1, int hbe, int x) {
    return nb + (uint32_t)(BY_RUNS ? R::rank(b1, hbe, x) : R::rep_at(b1, hbe, x));
  };
  for (int base = tid - lane; base < nrows; base += T) {
    const int rr = base + lane;
    RowT ty = 0, tz = 0, b1 = 0, by1 = 0, bz1 = 0;
    int hbe = L, hby = L, hbz = L;
    uint32_t n0 = 0, ny = 0, nz = 0;
    if (rr < nrows) {
      const int r = spread_row(rr, nrows);
      const RowT q = rows[r], b = R::bounds(q, L, full);
      const RowT lq = ~(q ^ (q << 1)) & full & ~RowT(1);  // q[x] == q[x-1]
      b1 = b | 1;
      hbe = R::wrap_start(b, L);
      n0 = node_base(r);
      const int zl = qdiv(r, g.mL), y = r - zl * L;
      // where a run overlap with row r2 starts, and r2's constants
      auto todo = [&](int r2, RowT &b21, int &hbe2) {
        const RowT q2 = rows[r2], b2 = R::bounds(q2, L, full);
        const RowT l2 = ~(q2 ^ (q2 << 1)) & full & ~RowT(1);
        b21 = b2 | 1;
        hbe2 = R::wrap_start(b2, L);
        return (RowT)(~(q ^ q2) & full & ~(lq & l2));
      };
      const int ry = y + 1 == L ? r - (L - 1) : r + 1;
      ty = todo(ry, by1, hby);
      ny = node_base(ry);
      int rz = -1;
      if (zl + 1 < g.Lz)
        rz = r + L;
      else if (g.bc == EA_PBC)
        rz = r - L * (g.Lz - 1);
      if (rz >= 0) {
        tz = todo(rz, bz1, hbz);
        nz = node_base(rz);
      }
    }
    // the y bonds of every lane's row, then the z bonds: one loop per
    // direction needs a few more steps than one over both, but each step
    // is shorter without choosing the direction's constants.  A lane
    // pushes up to two bonds per step (its two lowest), so the loop takes
    // half as many steps as the lane with the most bonds has; a step starts
    // with fewer than 64 pairs queued and adds at most 128.
    auto bonds = [&](RowT t, uint32_t n2, RowT b21, int hbe2) {
      auto below = [](uint64_t m) {
        return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      };
      for (;;) {
        const bool has1 = t != 0;
        const uint64_t m1 = __ballot(has1);
        if (!m1) break;
        const RowT t2 = t & (t - 1);
        const bool has2 = t2 != 0;
        const uint64_t m2 = __ballot(has2);
        const int slot = qn + below(m1) + below(m2);
        if (has1) {
          const int x = R::lo(t);
          wq[slot] = node(n0, b1, hbe, x) | node(n2, b21, hbe2, x) << 16;
        }
        if (has2) {
          const int x = R::lo(t2);
          wq[slot + 1] = node(n0, b1, hbe, x) | node(n2, b21, hbe2, x) << 16;
        }
        t = t2 & (t2 - 1);
        qn += __popcll(m1) + __popcll(m2);
        while (qn >= 64) drain(64);
      }
    };
    bonds(ty, ny, by1, hby);
    bonds(tz, nz, bz1, hbz);
  }
  while (qn > 0) drain(qn < 64 ? qn : 64);
}

// Whole-tube Houdayer move of one instance (group, sample, beta < ctm) in a
// single workgroup: overlap q = s1 ^ s2 of replicas r1, r2 at beta
// (d_computeoverlap), clusters of equal q labelled in LDS
// (init_YK_custom/ReduceLabels/analysis, HoudayerXEA.cu:77-121,233-354),
// label = minimum linear site index, one coin per cluster root
// (spin_select, :358-374), then q' = q ^ coin(label) and s1 ^= q', s2 ^= q'
// (spin_flip_YK_custom + d_pack + d_useoverlap).
//
// The nodes of the labelling are the runs of equal q along x.  They are
// numbered in site order -- row by row (row r = y + L z), and within a row
// in the order of their first sites, the run that wraps past x = L-1
// belonging to the row start -- so the smallest run number of a cluster is
// the run holding its smallest site.  A forest of 32-bit parents over the
// run numbers is linked with Rem's union with splicing (parents only ever
// decrease, so a root is its tree's smallest node; the atomic minimum on a
// root links it, on any other node it splices that node's subtree into the
// other tree, and the walk goes on from its old parent until both sides
// meet); the unions come from a per-wave queue (queue_links).  Then every
// root run draws the coin of its first site, every run takes its root's
// coin as its move bit, and each row's move mask is built site by site
// from its runs' bits -- passes whose lanes all take the same number of
// steps, unlike loops over the runs of a row.  An instance with more runs
// than the run forest holds (hou_cap, near the largest possible count of V
// runs) is labelled with a forest over sites instead (16-bit parents,
// Komura's union), as a whole tube whose runs fit always is.
// The kernel also recomputes the energies of the two configurations it
// changed (the Measure_Energy that follows a cluster move, main.cu:209).
template <typename RowT, typename IdxT>
__global__ void __launch_bounds__(1024) k_hou_whole(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map,
                                                    Phase ph) {
  using R = Rows<RowT>;
  extern __shared__ uint64_t lds[];
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const auto spin = gmem(a.spin);
  const Inst in = instance_of(a, wh.local, wh.step, ph.call);
  const int V = g.V, L = g.L, nrows = V / L, T = blockDim.x, tid = threadIdx.x;
  const int cap = a.hou_cap;
  const HouWholeLds lay = hou_whole_layout(g, cap, T, sizeof(RowT), sizeof(IdxT));
  uint8_t *base = reinterpret_cast<uint8_t *>(lds);
  RowT *rows = reinterpret_cast<RowT *>(base);  // q rows, then move rows
  uint64_t *qw = reinterpret_cast<uint64_t *>(base + lay.qw);  // q words, then r1's new words
  uint32_t *cb = reinterpret_cast<uint32_t *>(base + lay.cb);  // coin of every site
  uint16_t *rbase = reinterpret_cast<uint16_t *>(base + lay.rbase);  // first run of a row
  uint32_t *rp = reinterpret_cast<uint32_t *>(base + lay.forest);    // run forest
  const Forest<IdxT, true> f{reinterpret_cast<IdxT *>(base + lay.forest)};  // site forest
  uint32_t *queue = reinterpret_cast<uint32_t *>(base + lay.region);  // while linking
  __shared__ int esum[1];
  __shared__ int wtot[16];
  __shared__ int nruns;
  const int beta = in.beta, s = in.s, grp = in.grp;
  // Wave priority 2 until the links: the set-up is a few dependent memory
  // round trips, and a CU runs up to three of these workgroups, the later
  // ones arriving while the first link.  Raised, a newcomer's set-up is not
  // held behind the older workgroups' link instructions, so the three finish
  // closer together and the last one's end runs less alone.
  __builtin_amdgcn_s_setprio(2);
  // The instance's two chains: replicas r1, r2 of the call's pairs and their
  // clones at beta (the clones of all replicas are loaded at once, one per
  // lane, beside the pairs), then each chain's energy before the move and
  // its running sum at beta, read here so that the block's end only stores.
  const int lane = tid & 63;
  const int clone = gv(a.wctb)[((size_t)s * g.nr + min(lane, g.nr - 1)) * g.nt + beta];
  const int r1 = gv(in.pairs)[2 * grp], r2 = gv(in.pairs)[2 * grp + 1];
  const int c1 = __builtin_amdgcn_readlane(clone, r1), c2 = __builtin_amdgcn_readlane(clone, r2);
  const size_t ch1 = (size_t)s * g.nr + r1, ch2 = (size_t)s * g.nr + r2;
  const size_t sb = (size_t)s * 2 * g.half;
  const size_t o1 = sb + spin_row(g, r1, c1), o2 = sb + spin_row(g, r2, c2);
  int e_old1 = 0, e_old2 = 0;
  double e_sum = 0.0;
  if (tid < 2) {
    e_old1 = gv(a.ener)[ch1 * g.nt + c1];
    e_old2 = gv(a.ener)[ch2 * g.nt + c2];
    if (ph.measure) e_sum = gv(a.esum)[(tid ? ch2 : ch1) * g.nt + beta];
    if (tid == 0) esum[0] = 0;
  }
  // the q words; the first of every thread is loaded (at a clamped index)
  // before the coins are drawn, so that its latency overlaps them
  const int w0 = min(tid, 2 * g.n - 1);
  const size_t oo = (size_t)(w0 >= g.n) * g.half + (w0 - (w0 >= g.n) * g.n);
  const uint64_t x1 = spin[o1 + oo], x2 = spin[o2 + oo];
  // coins of 128 consecutive sites per Philox call (they do not depend on
  // the clusters)
  const uint64_t cidx = in.coin + (uint64_t)s * V * g.nt + (uint64_t)beta * V;
  for (int c = tid; c < V / 128; c += T) {
    const uint64_t blk = (cidx >> 7) + c;
    uint32_t v0 = (uint32_t)blk, v1 = (uint32_t)(blk >> 32), v2 = 0x486f7564u, v3 = 0;
    philox10(v0, v1, v2, v3, (uint32_t)a.coin_seed, (uint32_t)(a.coin_seed >> 32));
    cb[4 * c] = v0;
    cb[4 * c + 1] = v1;
    cb[4 * c + 2] = v2;
    cb[4 * c + 3] = v3;
  }
  if (tid < 2 * g.n) qw[tid] = x1 ^ x2;
  for (int w = tid + T; w < 2 * g.n; w += T) {
    const int p = w >= g.n, h = w - p * g.n;
    const size_t o = (size_t)p * g.half + h;
    qw[w] = spin[o1 + o] ^ spin[o2 + o];
  }
  __syncthreads();
  const RowT full = (RowT)(L == 64 ? ~0ULL : ((1ULL << L) - 1));
  // the q rows, each thread a contiguous range of them, and the run numbers:
  // rbase[r] = runs of the rows before r (a block-wide scan)
  {
    const int per = (nrows + T - 1) / T, r0 = min(tid * per, nrows), r1 = min(r0 + per, nrows);
    int own = 0;
    for (int r = r0; r < r1; ++r) {
      const RowT q = row_bits<RowT>(g, r, [&](int p, int h) { return qw[p * g.n + h]; });
      rows[r] = q;
      own += R::runs(R::bounds(q, L, full));
    }
    const int lane = tid & 63;
    int x = own;
    x = group_scan<64>(x);
    if (lane == 63) wtot[tid >> 6] = x;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < T / 64; ++w) {
      before += w < (tid >> 6) ? wtot[w] : 0;
      total += wtot[w];
    }
    int run = before + x - own;  // runs of the rows before r0
    for (int r = r0; r < r1; ++r) {
      rbase[r] = (uint16_t)run;
      const RowT b = R::bounds(rows[r], L, full);
      if (total <= cap) {
        // every run its own root; its first sites: bit 0 (rank 0, the run
        // that wraps included) and the boundaries but the wrapping one
        RowT st = b | 1;
        const int hbe = R::wrap_start(b, L);
        if (hbe < L) st &= ~(RowT(1) << hbe);
        for (uint32_t id = run; st; ++id) {
          rp[id] = id << 16 | (uint32_t)(r * L + R::lo(st));
          st &= st - 1;
        }
      }
      run += R::runs(b);
    }
    if (tid == T - 1) nruns = run;
  }
  __syncthreads();
  const int nrun = nruns;
  const bool by_runs = nrun <= cap;

  // The apply's first spin words (of thread tid's word, quarter 0) and the
  // energy's first couplings (half site tid), loaded once the links are made
  // so that their latency overlaps the move bits.  No other workgroup of the
  // launch writes the two replicas' words at this beta, and this one writes
  // them only in the apply.
  uint64_t pre1 = 0, pre2 = 0, Jpre[6] = {0, 0, 0, 0, 0, 0};
  auto prefetch = [&]() {
    if (tid < 8 * g.n && (tid & 3) == 0) {
      const int w = tid >> 2, p = w >= g.n, h = w - p * g.n;
      const size_t o = (size_t)p * g.half + h;
      pre1 = spin[o1 + o];
      pre2 = spin[o2 + o];
    }
    if (tid < g.n) load_couplings(g, a.J, s, 0, tid, Jpre);
  };
  if (by_runs) {
    // two steps of Rem's walk per drain: most unions end within them (1.5
    // to 1.8 steps on average on uncorrelated q), the rest resume later
    __builtin_amdgcn_s_setprio(0);
    queue_links<true>(g, rows, full, rbase, queue, [&](uint32_t &u, uint32_t &v) {
      return rem_step(rp, u, v) || rem_step(rp, u, v);
    });
    __syncthreads();
    prefetch();
    if (a.keep) {
      const auto out = gv(a.keep) + (((size_t)grp * g.ns + s) * g.ctm + beta) * V;
      for (int i = tid; i < V; i += T) {
        const int r = qdiv(i, g.mL), x = i - r * L;
        const RowT b = R::bounds(rows[r], L, full);
        out[i] = run_root(rp, rbase[r] + R::rank(b | 1, R::wrap_start(b, L), x)) & 0xFFFFu;
      }
      __syncthreads();  // the q rows are read before they become move rows
    }
    // the move bit of every run is the coin of its root's first site, 64
    // runs per wave step (the queues are done with, so the bits take their
    // place)
    uint64_t *mv = reinterpret_cast<uint64_t *>(queue);
    {
      const int lane = tid & 63;
      for (int b0 = tid - lane; b0 < nrun; b0 += T) {
        const int id = b0 + lane;
        bool bit = false;
        if (id < nrun) {
          const uint32_t site = run_root(rp, id) & 0xFFFFu;
          bit = (cb[site >> 5] >> (site & 31)) & 1u;
        }
        const uint64_t m = __ballot(bit);
        if (lane == 0) mv[b0 >> 6] = m;
      }
    }
    __syncthreads();
    // the move mask of each row: every site takes its run's move bit.  The
    // mask toggles where a run starts whose bit differs from the run before
    // (and at the start of the run that wraps, which is run 0 again); a
    // prefix XOR of the toggles then fills the runs
    for (int r = tid; r < nrows; r += T) {
      const RowT b = R::bounds(rows[r], L, full);
      const int hbe = R::wrap_start(b, L);
      const uint32_t id0 = rbase[r], sh = id0 & 63;
      const int nr_ = R::runs(b);
      const uint64_t w0 = mv[id0 >> 6];
      const uint64_t w1 = sh + nr_ > 64 ? mv[(id0 >> 6) + 1] : 0;
      const uint64_t rm = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;  // bit k: run k moves
      const uint64_t dm = rm ^ (rm >> 1);  // bit k: runs k and k+1 differ
      RowT st = b & ~RowT(1);  // the starts of runs 1 .. nr_-1, in order
      if (hbe < L) st &= ~(RowT(1) << hbe);
      RowT t = hbe < L ? (RowT)(((rm >> (nr_ - 1)) ^ rm) & 1u) << hbe : RowT(0);
      for (int k = 0; st; ++k) {
        t |= (RowT)((dm >> k) & 1u) << R::lo(st);
        st &= st - 1;
      }
#pragma unroll
      for (int sft = 1; sft < (int)(8 * sizeof(RowT)); sft <<= 1) t ^= t << sft;
      rows[r] = (t ^ ((rm & 1u) ? full : RowT(0))) & full;  // only this thread reads row r here
    }
  } else {
    for (int i = tid; i < V; i += T) f.p[i] = f.ix(i);  // every site its own root
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    queue_links<false>(g, rows, full, rbase, queue, [&](uint32_t &u, uint32_t &v) {
      f.unite(u, v);
      return true;
    });
    __syncthreads();
    prefetch();
    if (a.keep) {
      const auto out = gv(a.keep) + (((size_t)grp * g.ns + s) * g.ctm + beta) * V;
      for (int i = tid; i < V; i += T) {
        const int r = qdiv(i, g.mL), x = i - r * L;
        out[i] = f.root(r * L + R::rep(R::bounds(rows[r], L, full), x));
      }
      __syncthreads();  // the q rows are read before they become move rows
    }
    for (int r = tid; r < nrows; r += T) {  // move mask of row r, run by run
      RowT m = 0;
      R::for_runs(R::bounds(rows[r], L, full), L, full, [&](int, int first, RowT seg) {
        const uint32_t root = f.root(r * L + first);
        if ((cb[root >> 5] >> (root & 31)) & 1u) m |= seg;
      });
      rows[r] = m;  // only this thread reads row r in this loop
    }
  }
  __syncthreads();
  for (int t0 = 0; t0 < 8 * g.n; t0 += T) {  // every lane reaches the shuffles
    const int t = t0 + tid;
    const bool live = t < 8 * g.n;
    const int w = live ? t >> 2 : 0, quarter = t & 3;
    const int p = w >= g.n, h = w - p * g.n;
    uint64_t m = live ? word_move_bits(g, rows, p, h, quarter) : 0;
    m |= quad_swap<0xB1>(m);  // the quad's lanes 1, 0, 3, 2
    m |= quad_swap<0x4E>(m);  // and 2, 3, 0, 1
    if (live && quarter == 0) {
      const size_t o = (size_t)p * g.half + h;
      const uint64_t qn = qw[w] ^ m;  // q' = q ^ coin(label)
      const uint64_t y1 = (t0 ? spin[o1 + o] : pre1) ^ qn, y2 = (t0 ? spin[o2 + o] : pre2) ^ qn;
      spin[o1 + o] = y1;
      spin[o2 + o] = y2;
      qw[w] = y1;  // r1's new words, for its energy
    }
  }
  __syncthreads();
  // The energy of r1's new configuration, from its even words.  r2's follows
  // from E(r1) + E(r2), which the move conserves: every bond joins either
  // two sites of one cluster (both flipped or neither) or two sites of
  // unequal q, whose two replicas' terms cancel before and after (the
  // open-boundary substitute bonds join y neighbours, bonds of the
  // labelling too).
  const int lane_e = tid & 63;
  for (int t0 = 0; t0 < g.n; t0 += T) {
    if (t0 + tid - lane_e >= g.n) break;  // the whole wave is past the words
    const int h = t0 + tid;
    int e0 = 0;
    if (h < g.n) {
      const Nbrs nb = neighbours(g, 0, h);
      uint64_t sc[6];
      load_neighbours(g, qw + g.n, 0, nb, 0, sc, 1);
      const uint64_t sv = qw[h];
      uint64_t Jw[6];
      if (t0) {
        load_couplings(g, a.J, s, 0, h, Jw);
      } else {
#pragma unroll
        for (int d = 0; d < 6; ++d) Jw[d] = Jpre[d];
      }
      int u = 0;
#pragma unroll
      for (int d = 0; d < 6; ++d) u += __popcll(sv ^ Jw[d] ^ sc[d]);
      e0 = 6 * 64 - 2 * u;
    }
    e0 = group_scan<64>(e0);
    if (lane_e == 63 && e0) atomicAdd(esum, e0);
  }
  __syncthreads();
  if (tid < 2) {
    const size_t ch = tid ? ch2 : ch1;
    const int e1 = esum[0], e = tid ? e_old1 + e_old2 - e1 : e1;
    gv(a.ener)[ch * g.nt + (tid ? c2 : c1)] = e;
    // the sums already hold the pre-move energy; all values are integers
    // below 2^53, so adding the difference is exact
    if (ph.measure) gv(a.esum)[ch * g.nt + beta] = e_sum + (double)(e - (tid ? e_old2 : e_old1));
  }
}

// Houdayer for tubes larger than one workgroup's LDS, in slabs of `planes`
// z-planes; slab block local index: slab + slabs * instance.
//  1. k_hou_slab labels a slab in LDS (link_rows, then one pass of roots)
//     and stores every site's slab-local root (16 bits) and the q bits of
//     every row.  A cluster that reaches neither of the slab's seam planes
//     ends there: its label is gbase + its local root, and its flip bit is
//     drawn at once.  The roots of the clusters that reach a seam plane
//     become nodes of a forest in HBM (parent = itself).
//  2. k_hou_seam unites, across each seam, the nodes of the two slabs' runs
//     that touch in z.
//  3. k_hou_apply gives each seam root the coin of its root in that forest,
//     every site the flip bit of its root, and applies the move to the
//     slab's 16-bit pieces of the two replicas' spin words (piece bz of a
//     word holds its sites in z block bz, and each slab owns the pieces of
//     its planes, so blocks never write the same bits).
//  4. k_hou_energy recomputes r1's changed energy; r2's follows from the
//     conserved sum E(r1) + E(r2).
__device__ __forceinline__ size_t instance_index(const Lattice &g, const Inst &in) {
  return ((size_t)in.grp * g.ns + in.s) * g.ctm + in.beta;
}

__device__ __forceinline__ int slab_count(const Tube &a) {
  return (a.g.Lz + a.planes - 1) / a.planes;
}

// coin of site index idx of the call: bit idx mod 128 of the Philox block
// idx / 128 (as the whole-tube kernel's coin words)
__device__ __forceinline__ uint32_t coin_bit(uint64_t seed, uint64_t idx) {
  return (coin_word(seed, idx & ~31ULL) >> (idx & 31)) & 1u;
}

// coins of the 32 indices c .. c+31 (bit k: index c+k), at any alignment
__device__ __forceinline__ uint32_t coin_run(uint64_t seed, uint64_t c) {
  const uint32_t sh = (uint32_t)(c & 31);
  const uint32_t lo = coin_word(seed, c - sh);
  if (!sh) return lo;
  return (lo >> sh) | (coin_word(seed, c - sh + 32) << (32 - sh));
}

// Coin index of site 0 of the instance.
__device__ __forceinline__ uint64_t coin_base(const Lattice &g, const Inst &in) {
  return in.coin + (uint64_t)in.s * g.V * g.nt + (uint64_t)in.beta * g.V;
}

// Dynamic LDS: rows[planes*L] u64 (then the seam-root bitmap) |
// lab[planes*L*L] u16.
__global__ void __launch_bounds__(1024) k_hou_slab(const Tube *__restrict__ tubes, const uint64_t *__restrict__ map,
                                                   Phase ph) {
  extern __shared__ uint64_t lds[];
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int ns = slab_count(a), slab = wh.local % ns;
  const Inst in = instance_of(a, wh.local / ns, wh.step, ph.call);
  const int L = g.L, S = L * L, z0 = slab * a.planes;
  const int planes = min(a.planes, g.Lz - z0), nsite = planes * S, nrows = planes * L;
  const int T = blockDim.x, tid = threadIdx.x;
  uint64_t *rows = lds;
  uint16_t *lab = reinterpret_cast<uint16_t *>(rows + a.planes * L);  // < 65536 sites
  const Pair pw = pair_words(g, a, in.pairs, in.s, in.grp, in.beta);
  const size_t ib = instance_index(g, in);
  const int gbase = z0 * S;
  const auto grow = gv(a.hrow) + ib * (g.V / L) + z0 * L;
  for (int r = tid; r < nrows; r += T) {
    const uint64_t q = row_bits(g, z0 * L + r, [&](int p, int h) {
      const size_t o = (size_t)p * g.half + h;
      return gmem(a.spin)[pw.o1 + o] ^ gmem(a.spin)[pw.o2 + o];
    });
    rows[r] = q;
    grow[r] = q;
  }
  __syncthreads();
  link_rows(g, lab, rows, planes, false);  // ends with a barrier; rows are not read again
  uint32_t *seam = reinterpret_cast<uint32_t *>(rows);  // nsite bits fit the rows' words
  for (int w = tid; w < (nsite + 31) / 32; w += T) seam[w] = 0;
  __syncthreads();
  // Every site gets its cluster's root, 8 sites per thread and each run's
  // root found once, into LDS and HBM.  Root lookups only follow parents,
  // which are runs' first sites, so overwriting other sites is safe while
  // other threads still look up roots.  The roots of the sites on a seam
  // plane are marked: plane 0 unless it starts an open tube, plane
  // planes-1 unless it ends one.
  const bool pbc = g.bc == EA_PBC;
  const bool below = z0 > 0 || pbc, above = z0 + planes < g.Lz || pbc;
  const auto gl16 = reinterpret_cast<global_t<u32x4> *>(gv(a.hl16) + ib * g.V + gbase);
  {
    const Forest<uint16_t> f{lab};
    for (int c = tid; 8 * c < nsite; c += T) {
      const u32x4 v = *reinterpret_cast<const u32x4 *>(lab + 8 * c);
      const bool mark = (below && 8 * c < S) || (above && 8 * c >= nsite - S);
      uint32_t o[4], prev = ~0u, root = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t l = (v[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
        if (l != prev) {
          prev = l;
          root = f.root(l);
          if (mark) atomicOr(seam + (root >> 5), 1u << (root & 31));
        }
        if (k & 1)
          o[k >> 1] |= root << 16;
        else
          o[k >> 1] = root;
      }
      const u32x4 w = u32x4{o[0], o[1], o[2], o[3]};
      *reinterpret_cast<u32x4 *>(lab + 8 * c) = w;
      gl16[c] = w;
    }
  }
  __syncthreads();
  // roots, 32 sites (one coin word) per thread: flip bits of the roots of
  // clusters inside the slab, seam roots as forest nodes
  const auto gpar = gv(a.hpar) + ib * g.V + gbase;
  const int wps = hou_slab_words(g, a.planes);
  const auto gfl = gv(reinterpret_cast<uint32_t *>(a.hflag)) + (ib * ns + slab) * 4 * wps;
  const uint64_t cidx = coin_base(g, in) + gbase;
  for (int w = tid; w < 2 * wps; w += T) {
    uint32_t fb = 0, nb = 0;
    if (32 * w < nsite) {
      const uint32_t coins = coin_run(a.coin_seed, cidx + 32 * w), sb = seam[w];
      // the group's labels in four 16-byte reads (lanes 64 bytes apart)
      const u32x4 *lv = reinterpret_cast<const u32x4 *>(lab + 32 * w);
      u32x4 v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = lv[q];
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        const int i = 32 * w + k;
        if (((v[k >> 3][(k >> 1) & 3] >> (16 * (k & 1))) & 0xFFFFu) != (uint32_t)i) continue;
        if ((sb >> k) & 1u) {
          nb |= 1u << k;
          gpar[i] = gbase + i;
        } else {
          fb |= coins & (1u << k);
        }
      }
    }
    gfl[w] = fb;
    gfl[2 * wps + w] = nb;
  }
}

// z bonds between plane z0-1 (or Lz-1 for the periodic wrap) and plane z0,
// the first plane of the slab: the two planes' labels are staged in LDS
// (16-byte loads), then one lane per row y takes the two rows' q bits, which
// give the overlaps of equal-q runs, and unites each overlap once, at its
// start, between the two runs' slab-local roots (every site of a run has
// its run's root).  A pair of roots is united once per seam: repeats in a
// row are skipped, and a set in LDS lets only the first lane to meet a pair
// unite it (the same clusters touch across a seam in many rows).
__global__ void __launch_bounds__(64) k_hou_seam(const Tube *__restrict__ tubes,
                                                 const uint64_t *__restrict__ map, Phase ph) {
  using R = Rows<uint64_t>;
  // dynamic LDS: the two planes' 16-bit labels, L*L each (sized for the
  // widest tube of the launch, not for the largest L, so that more seams
  // share a CU)
  extern __shared__ __align__(16) uint16_t seam_lab[];
  constexpr int kSetBits = 8, kSet = 1 << kSetBits;  // pairs already united by this seam
  __shared__ uint64_t seen[kSet];
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  uint16_t *const lu = seam_lab, *const ld = seam_lab + g.L * g.L;
  const int ns = slab_count(a), slab = wh.local % ns;
  const Inst in = instance_of(a, wh.local / ns, wh.step, ph.call);
  const int L = g.L, S = L * L, y = threadIdx.x, z0 = slab * a.planes;
  int zl = z0 - 1;
  if (z0 == 0) {
    if (g.bc != EA_PBC) return;  // the whole block: no seam below the tube
    zl = g.Lz - 1;
  }
  const size_t ib = instance_index(g, in);
  const auto l16 = gv(a.hl16) + ib * g.V;
  {
    const auto su = reinterpret_cast<const global_t<u32x4> *>(l16 + (size_t)z0 * S);
    const auto sd = reinterpret_cast<const global_t<u32x4> *>(l16 + (size_t)zl * S);
    for (int c = y; 8 * c < S; c += blockDim.x) {
      reinterpret_cast<u32x4 *>(lu)[c] = su[c];
      reinterpret_cast<u32x4 *>(ld)[c] = sd[c];
    }
    for (int i = y; i < kSet; i += blockDim.x) seen[i] = ~0ULL;
  }
  __syncthreads();
  if (y >= L) return;  // no barrier follows
  const auto grow = gv(a.hrow) + ib * (g.V / L);
  uint32_t *par = a.hpar + ib * g.V;  // global unions, no LDS
  const uint64_t full = L == 64 ? ~0ULL : ((1ULL << L) - 1);
  const uint64_t qu = grow[z0 * L + y], qd = grow[zl * L + y];
  const uint64_t lqu = ~(qu ^ (qu << 1)) & full & ~1ULL;  // q[x] == q[x-1]
  const uint64_t lqd = ~(qd ^ (qd << 1)) & full & ~1ULL;
  uint64_t todo = ~(qu ^ qd) & full & ~(lqu & lqd);
  const uint32_t bu = (uint32_t)z0 * S, bd = (uint32_t)(zl / a.planes) * a.planes * S;
  uint32_t pu = ~0u, pd = ~0u;
  while (todo) {
    const int x = R::lo(todo);
    todo &= todo - 1;
    const uint32_t nu = bu + lu[y * L + x], nd = bd + ld[y * L + x];
    if (nu == pu && nd == pd) continue;
    pu = nu;
    pd = nd;
    // the first lane to meet a pair unites it; a full set just unites
    const uint64_t key = (uint64_t)nu << 32 | nd;
    uint32_t hsh = (uint32_t)((key * 0x9E3779B97F4A7C15ULL) >> (64 - kSetBits));
    bool fresh = true;
    for (int t = 0; t < kSet; ++t, hsh = (hsh + 1) & (kSet - 1)) {
      const unsigned long long old =
          atomicCAS(reinterpret_cast<unsigned long long *>(seen + hsh), ~0ULL, key);
      if (old == ~0ULL) break;
      if (old == key) {
        fresh = false;
        break;
      }
    }
    if (fresh) unite(par, nu, nd);
  }
}

// Dynamic LDS: flip[words] u64 | move[words] u64 | mrow[planes*L] u64 |
// seam roots[2*L*L] u16.
__global__ void __launch_bounds__(256) k_hou_apply(const Tube *__restrict__ tubes,
                                                   const uint64_t *__restrict__ map, Phase ph) {
  extern __shared__ uint64_t lds[];
  __shared__ int nseam;
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const int ns = slab_count(a), slab = wh.local % ns;
  const Inst in = instance_of(a, wh.local / ns, wh.step, ph.call);
  const int L = g.L, S = L * L, z0 = slab * a.planes;
  const int planes = min(a.planes, g.Lz - z0), nsite = planes * S, nrows = planes * L;
  const int T = blockDim.x, tid = threadIdx.x;
  const int wps = hou_slab_words(g, a.planes);
  uint64_t *flip = lds, *move = lds + wps, *mrow = move + wps;
  uint16_t *seam = reinterpret_cast<uint16_t *>(mrow + a.planes * L);
  const size_t ib = instance_index(g, in);
  const uint32_t gbase = (uint32_t)z0 * S;
  const auto gfl = gv(a.hflag) + (ib * ns + slab) * 2 * wps;
  const auto gl16 = gv(a.hl16) + ib * g.V + gbase;
  const uint32_t *par = a.hpar + ib * g.V;
  const uint64_t cidx = coin_base(g, in);
  if (tid == 0) nseam = 0;
  for (int w = tid; w < wps; w += T) flip[w] = gfl[w];
  __syncthreads();
  for (int w = tid; w < wps; w += T) {  // list the seam roots (at most the 2 S seam sites)
    uint64_t nb = gfl[wps + w];
    if (!nb) continue;
    int j = atomicAdd(&nseam, __popcll(nb));
    while (nb) {
      seam[j++] = (uint16_t)(64 * w + __ffsll((unsigned long long)nb) - 1);
      nb &= nb - 1;
    }
  }
  __syncthreads();
  // each seam root takes the coin of its root in the forest
  for (int j = tid; j < nseam; j += T) {
    const uint32_t l = seam[j];
    const uint32_t root = find_root(par, gbase + l);
    if (coin_bit(a.coin_seed, cidx + root))
      atomicOr(reinterpret_cast<unsigned long long *>(flip + (l >> 6)), 1ULL << (l & 63));
  }
  __syncthreads();
  // move bit of every site from its root's flip bit, 8 sites per lane
  uint8_t *mb = reinterpret_cast<uint8_t *>(move);
  const auto lab8 = reinterpret_cast<const global_t<u32x4> *>(gl16);
  for (int c = tid; 8 * c < nsite; c += T) {
    const u32x4 v = lab8[c];
    uint32_t byte = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t l = (v[k >> 1] >> (16 * (k & 1))) & 0xFFFFu;
      byte |= (uint32_t)((flip[l >> 6] >> (l & 63)) & 1ULL) << k;
    }
    mb[c] = (uint8_t)byte;
  }
  __syncthreads();
  const uint64_t full = L == 64 ? ~0ULL : ((1ULL << L) - 1);
  for (int r = tid; r < nrows; r += T) {  // row r: move bits r L .. r L + L - 1
    const int b0 = r * L, w0 = b0 >> 6, sh = b0 & 63;
    uint64_t m = move[w0] >> sh;
    if (sh + L > 64) m |= move[w0 + 1] << (64 - sh);
    mrow[r] = m & full;
  }
  __syncthreads();
  // pieces (p, h, bz) of the slab's planes: plane z holds the words of
  // gz = z mod MLz, piece bz = z / MLz, bits bx + 4 by
  const auto s16 = gv(reinterpret_cast<uint16_t *>(a.spin));
  const Pair pw = pair_words(g, a, in.pairs, in.s, in.grp, in.beta);
  const int pp = 2 * g.hs;  // pieces per plane
  for (int t = tid; t < planes * pp; t += T) {
    const int zl = t / pp, rem = t - zl * pp, p = rem >= g.hs, hh = rem - p * g.hs;
    const int z = z0 + zl, bz = qdiv(z, g.mMLz), gz = z - bz * g.MLz;
    const int yb = qdiv(hh, g.mhl), ax = 2 * (hh - yb * g.hl) + ((p + yb + gz) & 1);
    uint32_t m = 0;
#pragma unroll
    for (int by = 0; by < 4; ++by) {
      const uint64_t row = mrow[by * g.ML + yb + L * zl] >> ax;
#pragma unroll
      for (int bx = 0; bx < 4; ++bx) m |= (uint32_t)((row >> (bx * g.ML)) & 1ULL) << (bx + 4 * by);
    }
    const size_t o = (size_t)p * g.half + (size_t)gz * g.hs + hh;
    const size_t k1 = (pw.o1 + o) * 4 + bz, k2 = (pw.o2 + o) * 4 + bz;
    const uint32_t x1 = s16[k1], x2 = s16[k2];
    const uint32_t qn = x1 ^ x2 ^ m;  // q' = q ^ coin(label)
    s16[k1] = (uint16_t)(x1 ^ qn);
    s16[k2] = (uint16_t)(x2 ^ qn);
  }
  if (a.keep) {  // the label of every site: its cluster's smallest site index
    const auto out = gv(a.keep) + ib * g.V + gbase;
    for (int i = tid; i < nsite; i += T) {
      const uint32_t l = gl16[i];
      const bool node = (gfl[wps + (l >> 6)] >> (l & 63)) & 1ULL;
      out[i] = node ? find_root(par, gbase + l) : gbase + l;
    }
  }
}

// Energies of the two configurations a slab move changed, as the
// whole-tube kernel recomputes them (the Measure_Energy after a cluster
// move, main.cu:209): one workgroup per instance over the even words of the
// two replicas; ener of the two chains is replaced and, when the step
// measures, their energy sums move by the difference.
__global__ void __launch_bounds__(512) k_hou_energy(const Tube *__restrict__ tubes,
                                                    const uint64_t *__restrict__ map, Phase ph) {
  __shared__ int esum[1];
  const Where wh = locate(tubes, map, ph.step);
  const Tube &a = *wh.t;
  if (wh.step >= a.nsteps) return;
  const Lattice g = a.g;
  const Inst in = instance_of(a, wh.local, wh.step, ph.call);
  const Pair pw = pair_words(g, a, in.pairs, in.s, in.grp, in.beta);
  const int T = blockDim.x, tid = threadIdx.x;
  // the two chains' energies before the move, read here so that the end only
  // stores
  const int r1 = gv(in.pairs)[2 * in.grp], r2 = gv(in.pairs)[2 * in.grp + 1];
  const size_t ch1 = (size_t)in.s * g.nr + r1, ch2 = (size_t)in.s * g.nr + r2;
  const int c1 = gv(a.wctb)[ch1 * g.nt + in.beta], c2 = gv(a.wctb)[ch2 * g.nt + in.beta];
  int e_old1 = 0, e_old2 = 0;
  if (tid < 2) {
    e_old1 = gv(a.ener)[ch1 * g.nt + c1];
    e_old2 = gv(a.ener)[ch2 * g.nt + c2];
  }
  if (tid == 0) esum[0] = 0;
  __syncthreads();
  // r1's new energy from its even words; r2's follows from the sum, which
  // the move conserves (as in k_hou_whole)
  int e0 = 0;
  for (int h = tid; h < g.n; h += T) {
    const auto base = gmem(a.spin) + pw.o1;
    const Nbrs nb = neighbours(g, 0, h);
    uint64_t sc[6], Jw[6];
    load_neighbours(g, base + g.half, 0, nb, 0, sc, 1);
    load_couplings(g, a.J, in.s, 0, h, Jw);
    const uint64_t sv = base[h];
    int u = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) u += __popcll(sv ^ Jw[d] ^ sc[d]);
    e0 += 6 * 64 - 2 * u;
  }
  e0 = wave_sum(e0);
  if ((tid & 63) == 0 && e0) atomicAdd(esum, e0);
  __syncthreads();
  if (tid < 2) {
    const size_t chain = tid ? ch2 : ch1;
    const int e1 = esum[0], e = tid ? e_old1 + e_old2 - e1 : e1;
    gv(a.ener)[chain * g.nt + (tid ? c2 : c1)] = e;
    // the sums already hold the pre-move energy; all values are integers
    // below 2^53, so adding the difference is exact
    if (ph.measure) gv(a.esum)[chain * g.nt + in.beta] += (double)(e - (tid ? e_old2 : e_old1));
  }
}

inline unsigned blocks_for(long long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ------------------------------------------------------------- launchers

int sweep_partials_x(const Lattice &g, const Shape &shape) {
  if (shape.sweep_kind == 3) return 1;
  if (shape.sweep_kind == 4) {
    const int T = sweep_slab_planes(g);
    return T ? g.MLz / T : 1;
  }
  return (int)blocks_for(g.n, shape.sweep_kind ? 64 : shape.sweep_positions);
}

bool shape_fits(const Lattice &g, const Shape &shape) {
  if (shape.sweep_kind == 3) return g.n <= kFullMaxN;
  if (shape.sweep_kind == 4) return sweep_slab_planes(g) > 0;
  return shape.sweep_kind != 2 || tile_words(g) <= kTileWordsMax;
}

static long long sweep_grid(const Tube &t, const Shape &shape) {
  const Lattice &g = t.g;
  if (shape.sweep_kind == 3)
    return blocks_for(g.ns, sweep_full_samples(g)) * g.nr * g.nt;
  if (shape.sweep_kind == 4) return (long long)(g.MLz / t.sweep_planes) * g.nr * g.nt;
  const long long per = shape.sweep_kind ? blocks_for(g.n, kLaneThreads) : t.nbx;
  return per * g.nr * g.nt;
}

static int energy_blocks_host(const Lattice &g) { return (int)blocks_for((long long)g.n * g.nr, 256); }

size_t energy_partials(const Lattice &g) {
  const size_t per = (size_t)g.ns * g.nr * g.nt;
  return per * std::max<size_t>(blocks_for(g.n, kSweepPosMin),
                                4 * (size_t)energy_blocks_host(g));
}

// 16-bit parents when every site index fits (and the rows are 32 bits).
static bool hou_idx16(const Lattice &g) { return g.L <= 32 && g.V <= 65536; }

int hou_run_cap(const Lattice &g) {
  if (g.V <= 8192) return g.V;
  return std::min((g.V * 9 / 16 + 63) / 64 * 64, 65535);
}

size_t hou_whole_lds(const Lattice &g, int threads, int cap) {
  return hou_whole_layout(g, cap, threads, g.L <= 32 ? 4 : 8, hou_idx16(g) ? 2 : 4).total;
}

size_t hou_slab_lds(const Lattice &g, int planes) {
  return 8 * (size_t)planes * g.L + 2 * (size_t)planes * g.L * g.L;
}

size_t hou_apply_lds(const Lattice &g, int planes) {
  return 8 * (2 * (size_t)hou_slab_words(g, planes) + (size_t)planes * g.L) +
         2 * 2 * (size_t)g.L * g.L;
}

bool hou_use_slabs(const Lattice &g, int lds_max, int lds_cu, int cus) {
  if (g.V > 65536) return true;  // the whole-tube queues hold 16-bit node pairs
  const size_t whole = hou_whole_lds(g, 1024, hou_run_cap(g));
  if (whole > (size_t)lds_max) return true;
  return whole > (size_t)lds_cu / 2 && 2LL * g.ns * g.ctm < cus;
}

size_t hou_slab_budget(int lds_max, int lds_cu) {
  return std::min<size_t>(lds_max, std::max(lds_cu / 2, 1 << 15));
}

int hou_slab_planes(const Lattice &g, size_t lds_slab) {
  int planes = std::max(1, std::min(g.Lz, kHouSlabSites / (g.L * g.L)));
  while (planes > 1 && hou_slab_lds(g, planes) > lds_slab) --planes;
  // as many slabs as the thickest that fits needs, made as even as they go:
  // a slab's workgroup takes time in proportion to its planes, and the
  // thickest one sets the kernel's (Lz = 192 at 60 planes: 48+48+48+48
  // instead of 60+60+60+12)
  const int slabs = (g.Lz + planes - 1) / planes;
  return (g.Lz + slabs - 1) / slabs;
}

size_t hou_slab_bytes(const Lattice &g, int planes) {
  const size_t inst = (size_t)2 * g.ns * g.ctm;
  const size_t slabs = (g.Lz + planes - 1) / planes;
  return inst * (4 * (size_t)g.V + 2 * (size_t)g.V + 8 * (size_t)(g.V / g.L) +
                 16 * slabs * hou_slab_words(g, planes));
}

Plan make_plan(const std::vector<Tube> &tubes, Shape shape) {
  Plan p;
  p.ntubes = (int)tubes.size();
  p.shape = shape;
  for (size_t ti = 0; ti < tubes.size(); ++ti) {
    const Tube &t = tubes[ti];
    const Lattice &g = t.g;
    const uint64_t tag = (uint64_t)ti << 32;
    auto add = [&](int kind, long long count) {
      for (long long i = 0; i < count; ++i) p.map[kind].push_back(tag | (uint64_t)i);
    };
    add(kMapSweep, sweep_grid(t, shape));
    if (shape.sweep_kind == 2) p.tile_words = std::max(p.tile_words, tile_chunks(g));
    if (shape.sweep_kind == 3) p.lds_full = std::max(p.lds_full, sweep_full_lds(g));
    if (shape.sweep_kind == 4)
      p.lds_full = std::max(p.lds_full, sweep_slab_lds(g, t.sweep_planes));
    add(kMapEnergy, (long long)energy_blocks_host(g) * g.nt * g.ns);
    add(kMapChain, (long long)g.ns * g.nr);
    add(kMapOverlap, (long long)g.ns * (g.nr * (g.nr - 1) / 2) * g.nt);
    if (!t.cluster || g.ctm <= 0) continue;
    const long long inst = 2LL * g.ns * g.ctm;
    if (t.planes == 0) {
      if (hou_idx16(g)) {
        add(kMapWhole16, inst);
        p.lds16 = std::max(p.lds16, hou_whole_lds(g, shape.cluster_threads, t.hou_cap));
      } else if (g.L <= 32) {
        add(kMapWhole32, inst);
        p.lds32 = std::max(p.lds32, hou_whole_lds(g, shape.cluster_threads, t.hou_cap));
      } else {
        add(kMapWhole64, inst);
        p.lds64 = std::max(p.lds64, hou_whole_lds(g, shape.cluster_threads, t.hou_cap));
      }
    } else {
      const long long slabs = (g.Lz + t.planes - 1) / t.planes;
      add(kMapSlab, slabs * inst);
      add(kMapSeam, slabs * inst);
      add(kMapApply, slabs * inst);
      add(kMapHouEnergy, inst);
      p.lds_slab = std::max(p.lds_slab, hou_slab_lds(g, t.planes));
      p.lds_apply = std::max(p.lds_apply, hou_apply_lds(g, t.planes));
      p.lds_seam = std::max(p.lds_seam, (size_t)4 * g.L * g.L);  // two planes of 16-bit labels
    }
  }
  return p;
}

hipError_t upload_plan(const Plan &p, DevPlan *d) {
  free_plan(d);
  size_t total = 0;
  for (int k = 0; k < kMapKinds; ++k) {
    d->off[k] = total;
    d->cnt[k] = p.map[k].size();
    total += p.map[k].size();
  }
  d->lds16 = p.lds16;
  d->lds32 = p.lds32;
  d->lds64 = p.lds64;
  d->lds_slab = p.lds_slab;
  d->lds_apply = p.lds_apply;
  d->lds_seam = p.lds_seam;
  d->lds_full = p.lds_full;
  d->ntubes = p.ntubes;
  d->tile_words = p.tile_words;
  d->shape = p.shape;
  d->identity = p.ntubes == 1;
  if (!total || d->identity) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->buf), total * sizeof(uint64_t));
  if (e != hipSuccess) return e;
  for (int k = 0; k < kMapKinds; ++k)
    if (d->cnt[k]) {
      e = hipMemcpy(d->buf + d->off[k], p.map[k].data(), d->cnt[k] * sizeof(uint64_t),
                    hipMemcpyHostToDevice);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

void free_plan(DevPlan *d) {
  if (d->buf) (void)hipFree(d->buf);
  *d = DevPlan{};
}

static void allow_lds(const void *kernel, size_t bytes) {
  // raise the dynamic LDS cap of a kernel once per process, to the largest
  // size requested so far
  static std::mutex mu;
  static std::vector<std::pair<const void *, size_t>> done;
  std::lock_guard<std::mutex> lock(mu);
  for (auto &d : done)
    if (d.first == kernel) {
      if (d.second >= bytes) return;
      d.second = bytes;
      (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes);
      return;
    }
  done.push_back({kernel, bytes});
  (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_sweep(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s) {
  if (!p.cnt[kMapSweep]) return hipGetLastError();
  const dim3 grid((unsigned)p.cnt[kMapSweep]);
  const uint64_t *m = p.map(kMapSweep);
  const Shape &sh = p.shape;
  if (sh.sweep_kind == 4) {
    auto k = sh.sweep_xcd ? k_sweep_slab<true> : k_sweep_slab<false>;
    allow_lds(reinterpret_cast<const void *>(k), p.lds_full);
    hipLaunchKernelGGL(k, grid, dim3(kSlabThreads), p.lds_full, s, tubes, m, ph);
    // the descriptors are not written by any kernel but this one, which runs
    // after the sweep has finished reading them
    hipLaunchKernelGGL(k_swap_spins, dim3(blocks_for(p.ntubes, 64)), dim3(64), 0, s,
                       const_cast<Tube *>(tubes), p.ntubes, ph.step);
  } else if (sh.sweep_kind == 3) {
    auto k = sh.sweep_xcd ? k_sweep_full<true> : k_sweep_full<false>;
    allow_lds(reinterpret_cast<const void *>(k), p.lds_full);
    hipLaunchKernelGGL(k, grid, dim3(kFullThreads), p.lds_full, s, tubes, m, ph);
  } else if (sh.sweep_kind == 2) {
    const dim3 blk(kLaneThreads);
    if (p.tile_words <= 1)
      hipLaunchKernelGGL((sh.sweep_xcd ? k_sweep_tile<1, true> : k_sweep_tile<1, false>), grid,
                         blk, 0, s, tubes, m, ph);
    else
      hipLaunchKernelGGL((sh.sweep_xcd ? k_sweep_tile<2, true> : k_sweep_tile<2, false>), grid,
                         blk, 0, s, tubes, m, ph);
  } else if (sh.sweep_kind == 1) {
    if (sh.sweep_xcd)
      hipLaunchKernelGGL(k_sweep_lane<true>, grid, dim3(kLaneThreads), 0, s, tubes, m, ph);
    else
      hipLaunchKernelGGL(k_sweep_lane<false>, grid, dim3(kLaneThreads), 0, s, tubes, m, ph);
  } else if (sh.sweep_positions == 32) {
    if (sh.sweep_xcd)
      hipLaunchKernelGGL((k_sweep<32, true>), grid, dim3(kSwThreads), 0, s, tubes, m, ph);
    else
      hipLaunchKernelGGL((k_sweep<32, false>), grid, dim3(kSwThreads), 0, s, tubes, m, ph);
  } else {
    if (sh.sweep_xcd)
      hipLaunchKernelGGL((k_sweep<64, true>), grid, dim3(kSwThreads), 0, s, tubes, m, ph);
    else
      hipLaunchKernelGGL((k_sweep<64, false>), grid, dim3(kSwThreads), 0, s, tubes, m, ph);
  }
  return hipGetLastError();
}

hipError_t launch_energy(const Tube *tubes, const DevPlan &p, int step, hipStream_t s) {
  if (p.cnt[kMapEnergy])
    hipLaunchKernelGGL(k_energy, dim3((unsigned)p.cnt[kMapEnergy]), dim3(256), 0, s, tubes,
                       p.map(kMapEnergy), step);
  if (p.cnt[kMapChain])
    hipLaunchKernelGGL(k_energy_sum<false>, dim3((unsigned)p.cnt[kMapChain]), dim3(64), 0, s,
                       tubes, p.map(kMapChain), step);
  return hipGetLastError();
}

hipError_t launch_overlap(const Tube *tubes, const DevPlan &p, int step, hipStream_t s) {
  if (p.cnt[kMapOverlap])
    hipLaunchKernelGGL(k_overlap, dim3((unsigned)p.cnt[kMapOverlap]), dim3(256), 0, s, tubes,
                       p.map(kMapOverlap), step);
  return hipGetLastError();
}

hipError_t launch_resolve(const Tube *tubes, const DevPlan &p, int step, hipStream_t s) {
  if (p.cnt[kMapChain])
    hipLaunchKernelGGL(k_energy_sum<true>, dim3((unsigned)p.cnt[kMapChain]), dim3(64), 0, s,
                       tubes, p.map(kMapChain), step);
  return hipGetLastError();
}

hipError_t launch_pt(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s) {
  if (p.cnt[kMapChain])
    hipLaunchKernelGGL(k_pt, dim3((unsigned)p.cnt[kMapChain]), dim3(64), 0, s, tubes,
                       p.map(kMapChain), ph);
  return hipGetLastError();
}


hipError_t launch_step_end(const Tube *tubes, int ntubes, hipStream_t s) {
  if (ntubes > 0)
    hipLaunchKernelGGL(k_step_end, dim3(blocks_for(ntubes, 64)), dim3(64), 0, s, tubes, ntubes);
  return hipGetLastError();
}


bool hou_has(const DevPlan &p, int part) {
  return (part == kHouWhole) ? p.cnt[kMapWhole16] || p.cnt[kMapWhole32] || p.cnt[kMapWhole64]
                             : p.cnt[kMapSlab] != 0;
}

hipError_t launch_houdayer(const Tube *tubes, const DevPlan &p, Phase ph, hipStream_t s,
                           int parts) {
  const dim3 wg(p.shape.cluster_threads);
  if (parts & kHouWhole) {
    if (p.cnt[kMapWhole16]) {
      auto k = k_hou_whole<uint32_t, uint16_t>;
      allow_lds(reinterpret_cast<const void *>(k), p.lds16);
      hipLaunchKernelGGL(k, dim3((unsigned)p.cnt[kMapWhole16]), wg, p.lds16, s, tubes,
                         p.map(kMapWhole16), ph);
    }
    if (p.cnt[kMapWhole32]) {
      auto k = k_hou_whole<uint32_t, uint32_t>;
      allow_lds(reinterpret_cast<const void *>(k), p.lds32);
      hipLaunchKernelGGL(k, dim3((unsigned)p.cnt[kMapWhole32]), wg, p.lds32, s, tubes,
                         p.map(kMapWhole32), ph);
    }
    if (p.cnt[kMapWhole64]) {
      auto k = k_hou_whole<uint64_t, uint32_t>;
      allow_lds(reinterpret_cast<const void *>(k), p.lds64);
      hipLaunchKernelGGL(k, dim3((unsigned)p.cnt[kMapWhole64]), wg, p.lds64, s, tubes,
                         p.map(kMapWhole64), ph);
    }
  }
  if ((parts & kHouSlabs) && p.cnt[kMapSlab]) {
    allow_lds(reinterpret_cast<const void *>(k_hou_slab), p.lds_slab);
    hipLaunchKernelGGL(k_hou_slab, dim3((unsigned)p.cnt[kMapSlab]), dim3(1024), p.lds_slab, s,
                       tubes, p.map(kMapSlab), ph);
    hipLaunchKernelGGL(k_hou_seam, dim3((unsigned)p.cnt[kMapSeam]), dim3(64), p.lds_seam, s, tubes,
                       p.map(kMapSeam), ph);
    allow_lds(reinterpret_cast<const void *>(k_hou_apply), p.lds_apply);
    hipLaunchKernelGGL(k_hou_apply, dim3((unsigned)p.cnt[kMapApply]), dim3(256), p.lds_apply, s,
                       tubes, p.map(kMapApply), ph);
    hipLaunchKernelGGL(k_hou_energy, dim3((unsigned)p.cnt[kMapHouEnergy]), dim3(512), 0, s,
                       tubes, p.map(kMapHouEnergy), ph);
  }
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  const unsigned b = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(b), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, double v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  const unsigned b = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f64, dim3(b), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace eatube
