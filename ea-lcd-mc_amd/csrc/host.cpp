// host.cpp -- host-side generators, seeding, sample initialisation, counter
// blocks and LUT generation of the EA tube engine (no GPU code).
#include "host.hpp"

#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace eatube {

// ------------------------------------------------------------- generators

Luscher::Luscher(uint64_t seed) {
  // Init_luxury (random_generator.cu:27-55): 24 lags seeded from the top
  // bytes of an LCG, one borrow bit, then 1111 luxury draws.
  std::memset(ring_, 0, sizeof ring_);
  ip_ = 128;
  ip1_ = 118;
  ip2_ = 104;
  for (int i = 0; i < 1111; ++i) seed = seed * kLcg + 1;
  for (int j = 104; j < 128; ++j) {
    uint32_t v = 0;
    for (int byte = 0; byte < 3; ++byte) {
      seed = seed * kLcg + 1;
      v |= uint32_t(seed >> 56) << (8 * byte);
    }
    ring_[j] = v;
  }
  seed = seed * kLcg + 1;
  borrow_ = int(seed >> 63);
  for (int i = 0; i < 1111; ++i) next();
}

uint32_t Luscher::next() {
  // luxury() (random_generator.cu:57-75): 389 subtract-with-borrow steps,
  // the last value is returned.
  for (int i = 0; i < 389; ++i) {
    const int d = int(ring_[ip1_]) - int(ring_[ip2_]) - borrow_;
    ring_[ip_] = uint32_t(d + (1 << 24)) & 0xFFFFFFu;
    borrow_ = d < 0;
    ++ip_;
    ++ip1_;
    ++ip2_;
  }
  return ring_[ip_];
}

uint64_t Luscher::next64() {
  uint64_t v = 0;
  for (int j = 0; j < 8; ++j) v |= uint64_t((next() >> 16) & 0xFF) << (8 * j);
  return v;
}

uint32_t Luscher::bits32() {
  uint32_t v = 0;
  for (int i = 0; i < 32; ++i) v |= ((next() >> 23) & 1u) << i;
  return v;
}

}  // namespace eatube

using namespace eatube;

extern "C" uint64_t ea_hq_next(ea_hq_state *g) {
  // Parisi-Rapuano lagged sum (lags 24, 55, 61) plus a 64-bit LCG
  // (random_generator.h:130-135)
  g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
  g->PR = g->ira[g->ip] ^ g->ira[g->ip3];
  g->zseed = g->zseed * kLcg + 1;
  g->final_ = g->zseed + g->PR;
  ++g->ip;
  ++g->ip1;
  ++g->ip2;
  ++g->ip3;
  return g->final_;
}

static void hq_set_pointers(ea_hq_state *g) {
  g->ip = 128;
  g->ip1 = 128 - 24;
  g->ip2 = 128 - 55;
  g->ip3 = 128 - 61;
}

extern "C" void ea_hq_init(ea_hq_state *g, uint64_t seed) {
  // Init_Rand_HQ_64bits (random_generator.cu:160-201)
  std::memset(g, 0, sizeof *g);
  uint64_t z = seed;
  for (int i = 0; i < 111; ++i) z = z * kLcg + 1;
  for (int k = 67; k < 128; ++k) {
    uint64_t v = 0;
    for (int b = 0; b < 8; ++b) {
      z = z * kLcg + 1;
      v |= (z >> 56) << (8 * b);
    }
    g->ira[k] = v;
  }
  g->zseed = z;
  hq_set_pointers(g);
  for (int i = 0; i < 1111; ++i) {  // warm the lagged sum alone
    g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
    ++g->ip;
    ++g->ip1;
    ++g->ip2;
    ++g->ip3;
  }
  for (int i = 0; i < 1111; ++i) ea_hq_next(g);
}

extern "C" void ea_seeds_from_one(uint64_t seed, uint64_t *out, int n) {
  // generate_seeds_from_one (random_generator.cu:203-241): between outputs
  // the Lüscher stream skips 709/109/1373/2099 draws chosen by the top two
  // bits of a draw; each output is a Knuth MMIX step of a 64-bit draw.
  static constexpr int kSkip[4] = {709, 109, 1373, 2099};
  Luscher lx(seed);
  for (int i = 0; i < 1109; ++i) lx.next();
  for (int k = 0; k < n; ++k) {
    const uint64_t sel = (lx.next64() + kLcg + 1) >> 62;
    for (int i = 0; i < kSkip[sel]; ++i) lx.next();
    out[k] = lx.next64() * 6364136223846793005ULL + 14426950408889634071ULL;
  }
}

extern "C" void ea_init_generators(uint64_t seed, ea_hq_state *prc,
                                   ea_xo_state *xo, uint32_t *key) {
  // Inicia_generadores_CPU (random_generator.cu:94-158)
  Luscher lx(seed);
  for (int i = 0; i < 1111; ++i) lx.next();
  ea_hq_state g;
  std::memset(&g, 0, sizeof g);
  g.zseed = lx.next64();
  for (int k = 67; k < 128; ++k) g.ira[k] = lx.next64();
  hq_set_pointers(&g);
  for (int i = 0; i < 2222; ++i) ea_hq_next(&g);
  *prc = g;
  std::memset(xo, 0, sizeof *xo);
  while (!(xo->s[0] | xo->s[1] | xo->s[2] | xo->s[3]))
    for (int j = 0; j < 4; ++j) {
      uint64_t v = 0;
      for (int i = 0; i < 64; ++i) v |= uint64_t((lx.next() >> 23) & 1u) << i;
      xo->s[j] = v;
    }
  *key = lx.bits32();
}

static inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

static uint64_t xo_next(ea_xo_state *u) {
  // xoshiro256++ (random_generator.h:118-127)
  u->final_ = rotl64(u->s[0] + u->s[3], 23) + u->s[0];
  u->t = u->s[1] << 17;
  u->s[2] ^= u->s[0];
  u->s[3] ^= u->s[1];
  u->s[1] ^= u->s[2];
  u->s[0] ^= u->s[3];
  u->s[2] ^= u->t;
  u->s[3] = rotl64(u->s[3], 45);
  return u->final_;
}

extern "C" void ea_fill_counters(ea_hq_state *prc, ea_xo_state *xo, int nr,
                                 int iter0, int nhalf, uint32_t job_id,
                                 ea_counter *out) {
  // _fill_bits (fill_bits.h:47-68).  The 128-bit Philox counter holds: in x
  // two internal-call bits, the parity bit and 15 low bits of the half-sweep
  // counter under 14 entropy bits; in y the counter's high 25 bits under 7
  // entropy bits; in z room for a 19-bit thread index under 13 entropy bits;
  // in w the 14-bit job id under 18 entropy bits.  The 52 entropy bits are
  // the top bits of PRC + xoshiro draws of the replica.
  for (int k = 0; k < nhalf; ++k) {
    const int iter = iter0 + k, parity = k & 1;
    for (int r = 0; r < nr; ++r) {
      xo_next(&xo[r]);
      ea_hq_next(&prc[r]);
      uint64_t e = (prc[r].final_ + xo[r].final_) >> 12;
      ea_counter c;
      c.x = uint32_t((uint64_t(iter) & 0x7FFFu) << 3) | uint32_t(parity << 2) |
            uint32_t((e & 0x3FFFu) << 18);
      c.y = uint32_t(iter >> 15) | uint32_t(((e >> 14) & 0x7Fu) << 25);
      c.z = uint32_t(((e >> 21) & 0x1FFFu) << 19);
      c.w = uint32_t((e >> 34) << 14) | job_id;
      out[size_t(k) * nr + r] = c;
    }
  }
}

extern "C" void ea_make_lut(double beta, uint8_t *out) {
  // prepara_LUT (create_LUT.c:265-388): a byte whose 8 bits are i.i.d. with
  // P(1) = exp(-4 beta) takes value v with probability q1^popc(v)
  // q0^(8-popc(v)); the 256 cumulative 64-bit thresholds of that law (in
  // quad precision, truncated) make umbrales, and prebusqueda stores for
  // every 12-bit prefix the first and last candidate byte.
  const __float128 q1 = expq(-4.0Q * (__float128)beta);
  const __float128 q0 = 1.0Q - q1;
  uint64_t weight[9];
  for (int k = 0; k <= 8; ++k)
    weight[k] =
        (uint64_t)(18446744073709551616.Q * (powq(q0, 8 - k) * powq(q1, k)));
  uint64_t cum[256];
  cum[0] = weight[0];
  for (int v = 1; v < 255; ++v) cum[v] = cum[v - 1] + weight[__builtin_popcount(v)];
  cum[255] = ~0ULL;
  std::memset(out, 0, EA_LUT_BYTES);
  uint32_t *umb = reinterpret_cast<uint32_t *>(out);
  for (int v = 0; v < 256; ++v) {
    umb[2 * v] = uint32_t(cum[v] >> 32);
    umb[2 * v + 1] = uint32_t(cum[v]);
  }
  uint16_t *pre = reinterpret_cast<uint16_t *>(out + 2048);
  const uint64_t tail = (1ULL << 52) - 1;
  for (uint64_t m = 0; m < 4096; ++m) {
    const uint64_t lo = m << 52, hi = lo | tail;
    const int a = int(std::upper_bound(cum, cum + 256, lo) - cum);
    const int b = int(std::upper_bound(cum, cum + 256, hi) - cum);
    pre[m] = uint16_t(std::min(a, 255) | (std::min(b, 255) << 8));
  }
}

// ------------------------------------------------------------- lattice

namespace eatube {

bool valid_lattice(int L, int Lz) {
  return L >= 8 && Lz >= 8 && L % 8 == 0 && Lz % 8 == 0 && L <= 1024 &&
         (int64_t)L * L * Lz < (1LL << 31);
}

Geometry::Geometry(int L_, int Lz_)
    : L(L_), Lz(Lz_), ML(L_ / 4), MLz(Lz_ / 4), MS((L_ / 4) * (L_ / 4)) {
  V = L * L * Lz;
  n = V / 128;
}

void init_spins(uint64_t seed_u, const Geometry &g, int nr, int nt,
                uint64_t *even, uint64_t *odd, std::vector<uint8_t> *first) {
  // Init_u (ini.c:316-337): one HQ stream per sample, replica-major then
  // clone then site order, spin = top bit; placed directly into the packed
  // words (packing_u, ini.c:339-383).
  const size_t half = size_t(g.n) * nr * nt;
  std::memset(even, 0, 8 * half);
  std::memset(odd, 0, 8 * half);
  uint64_t *arr[2] = {even, odd};
  std::vector<int> ph(g.V), pb(g.V), pp(g.V);
  for (int s = 0; s < g.V; ++s) g.locate(s, pp[s], ph[s], pb[s]);
  if (first) first->assign(g.V, 0);
  ea_hq_state hq;
  ea_hq_init(&hq, seed_u);
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nt; ++c)
      for (int s = 0; s < g.V; ++s) {
        const uint64_t bit = ea_hq_next(&hq) >> 63;
        if (first && r == 0 && c == 0) (*first)[s] = uint8_t(bit);
        arr[pp[s]][size_t(c) * g.n * nr + size_t(ph[s]) * nr + r] |= bit << pb[s];
      }
}

int init_couplings(uint64_t seed_J, const Geometry &g, int bc, int variant,
                   bool mattis, const uint8_t *u00, uint8_t *J6) {
  // Init_J (ini.c:465-670).  J6 = Jx, Jy, Jz, Jmx, Jmy, Jmz (1 = J of -1).
  const int L = g.L, Lz = g.Lz, S = L * L, V = g.V;
  uint8_t *Jx = J6, *Jy = J6 + V, *Jz = J6 + 2 * V;
  uint8_t *Jmx = J6 + 3 * V, *Jmy = J6 + 4 * V, *Jmz = J6 + 5 * V;
  auto nbr = [&](int s, int dx, int dy, int dz) {
    const int x = s % L, y = (s / L) % L, z = s / S;
    return (x + dx + L) % L + L * ((y + dy + L) % L) + S * ((z + dz + Lz) % Lz);
  };
  auto parity = [&](int s) { return ((s % L) ^ ((s / L) % L) ^ (s / S)) & 1; };
  std::vector<uint8_t> bottom(S, 0);  // J_falso_z0: the z=0 plane's extra bond
  int countJ = 0;
  if (mattis) {
    for (int s = 0; s < V; ++s) {
      const int z = s / S;
      Jx[s] = u00[s] ^ u00[nbr(s, 1, 0, 0)];
      Jy[s] = u00[s] ^ u00[nbr(s, 0, 1, 0)];
      if (bc == EA_PBC || z < Lz - 1)
        Jz[s] = u00[s] ^ u00[nbr(s, 0, 0, 1)];
      else
        Jz[s] = Jy[s];
      if (bc == EA_OBC && z == 0) bottom[s] = Jy[s];
    }
  } else {
    ea_hq_state hq;
    ea_hq_init(&hq, seed_J);
    for (int s = 0; s < V; ++s) {
      Jx[s] = uint8_t(ea_hq_next(&hq) >> 63);
      Jy[s] = uint8_t(ea_hq_next(&hq) >> 63);
      Jz[s] = uint8_t(ea_hq_next(&hq) >> 63);
      countJ += Jx[s] + Jy[s] + Jz[s];
    }
    if (bc == EA_OBC)
      for (int s = 0; s < S; ++s) bottom[s] = uint8_t(ea_hq_next(&hq) >> 63);
  }
  if (bc == EA_OBC) {
    const int top = (Lz - 1) * S;
    if (variant == 1 || variant == 2) {
      // weak: the substituted bond cancels the y bond; strong: doubles it
      const uint8_t inv = variant == 1 ? 1 : 0;
      for (int s = 0; s < S; ++s)
        if (!parity(s)) bottom[s] = Jy[s] ^ inv;
      for (int s = top; s < top + S; ++s)
        if (!parity(s)) Jz[s] = Jy[s] ^ inv;
    }
    // odd boundary sites share the coupling of their even y-1 neighbour
    for (int s = 0; s < S; ++s)
      if (parity(s)) bottom[s] = bottom[nbr(s, 0, -1, 0)];
    for (int s = top; s < top + S; ++s)
      if (parity(s)) Jz[s] = Jz[nbr(s, 0, -1, 0)];
  }
  for (int s = 0; s < V; ++s) {
    const int z = s / S;
    Jmx[s] = Jx[nbr(s, -1, 0, 0)];
    Jmy[s] = Jy[nbr(s, 0, -1, 0)];
    Jmz[s] = (bc == EA_PBC || z > 0) ? Jz[nbr(s, 0, 0, -1)] : bottom[s];
  }
  return countJ;
}

void pack_couplings(const Geometry &g, const uint8_t *J6, uint64_t *Jw) {
  // packing_J (ini.c:672-730): word [parity][direction][h]
  std::memset(Jw, 0, sizeof(uint64_t) * 12 * g.n);
  for (int s = 0; s < g.V; ++s) {
    int par, h, bit;
    g.locate(s, par, h, bit);
    for (int d = 0; d < 6; ++d)
      Jw[(size_t(par) * 6 + d) * g.n + h] |= uint64_t(J6[size_t(d) * g.V + s] & 1)
                                              << bit;
  }
}

ExpTables make_exp_tables(const double *betas, int nt, int V) {
  ExpTables t;
  t.nd.resize(nt > 1 ? nt - 1 : 0);
  t.off.resize(t.nd.size());
  t.len.resize(t.nd.size());
  for (int i = 0; i + 1 < nt; ++i) {
    const double nd = -(betas[i + 1] - betas[i]);
    t.nd[i] = nd;
    t.off[i] = int32_t(t.val.size());
    int64_t K = 0;
    if (nd > 0) K = std::min<int64_t>(int64_t(std::ceil(380.0 / nd)) + 2, 3LL * V);
    t.len[i] = int32_t(K);
    for (int64_t k = 1; k <= K; ++k) t.val.push_back(std::exp(nd * double(-2 * k)));
  }
  return t;
}

}  // namespace eatube

extern "C" int ea_init_sample(int L, int Lz, int nr, int nt, int bc,
                              int obc_variant, int mattis, uint64_t seed_J,
                              uint64_t seed_u, uint64_t *even, uint64_t *odd,
                              uint64_t *J, uint8_t *couplings) {
  if (!valid_lattice(L, Lz) || nr < 1 || nt < 1 || (bc != 0 && bc != 1) ||
      obc_variant < 0 || obc_variant > 2 || !even || !odd || !J)
    return EA_E_ARG;
  const Geometry g(L, Lz);
  std::vector<uint8_t> u00;
  init_spins(seed_u, g, nr, nt, even, odd, mattis ? &u00 : nullptr);
  std::vector<uint8_t> J6(size_t(6) * g.V);
  const int countJ =
      init_couplings(seed_J, g, bc, obc_variant, mattis != 0, u00.data(), J6.data());
  pack_couplings(g, J6.data(), J);
  if (couplings) std::memcpy(couplings, J6.data(), J6.size());
  return countJ;
}

extern "C" int ea_houdayer_pairs(ea_hq_state *cl, int nr, int32_t pairs[4]) {
  // main.cu:409-421: a partial Fisher-Yates shuffle of each half of the
  // replicas; the first two entries are the pair.
  const int half = nr / 2;
  if (half < 2 || half > 64) return EA_E_ARG;
  for (int grp = 0; grp < 2; ++grp) {
    int idx[64];
    for (int k = 0; k < half; ++k) idx[k] = k + grp * half;
    for (int k = 0; k < half - 1; ++k) {
      const uint32_t hi = uint32_t(ea_hq_next(cl) >> 32);
      const int j = k + int(hi % uint32_t(half - k));
      std::swap(idx[k], idx[j]);
    }
    pairs[2 * grp] = idx[0];
    pairs[2 * grp + 1] = idx[1];
  }
  return EA_OK;
}

extern "C" int ea_unpack_lattice(int L, int Lz, int nr, int nt,
                                 const uint64_t *even, const uint64_t *odd,
                                 int replica, int clone, uint8_t *sites) {
  if (!valid_lattice(L, Lz) || replica < 0 || replica >= nr || clone < 0 ||
      clone >= nt)
    return EA_E_ARG;
  const Geometry g(L, Lz);
  const uint64_t *arr[2] = {even, odd};
  for (int s = 0; s < g.V; ++s) {
    int par, h, bit;
    g.locate(s, par, h, bit);
    sites[s] = uint8_t(
        (arr[par][size_t(clone) * g.n * nr + size_t(h) * nr + replica] >> bit) & 1);
  }
  return EA_OK;
}
