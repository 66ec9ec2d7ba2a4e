// conf.cpp -- the reference's checkpoint record ("conf" file, one per sample
// and replica) and the MSC <-> one-byte-per-site conversions it needs.  Host
// code only.
#include <cstring>
#include <vector>

#include "eatube.h"
#include "host.hpp"

using namespace eatube;

namespace {

// Byte layout of one conf (write_conf, io.c:17-104), in file order:
//   s_data                        72 bytes (9 int32, 4 padding, 4 uint64)
//   betas[nt]                     double
//   random_PT, random_PRC, random_c   3 x 2080 bytes
//   random_xoshiro256pp           48 bytes
//   cuda_rng_offset               uint64
//   which_clon_this_beta[nt]      bytes
//   spins[nt][V/8]                bit (site & 7) of byte site >> 3, site =
//                                 x + L y + L^2 z, per clone (io.c:239-258)
//   energy[nt]                    int32, of the clone at each beta
struct Layout {
  size_t data = 0, betas, pt, prc, cl, xo, off, perm, spins, ener, total;
  Layout(int L, int Lz, int nt) {
    const size_t V8 = (size_t)L * L * Lz / 8;
    betas = data + sizeof(ea_run_data);
    pt = betas + 8 * (size_t)nt;
    prc = pt + sizeof(ea_hq_state);
    cl = prc + sizeof(ea_hq_state);
    xo = cl + sizeof(ea_hq_state);
    off = xo + sizeof(ea_xo_state);
    perm = off + sizeof(uint64_t);
    spins = perm + (size_t)nt;
    ener = spins + V8 * nt;
    total = ener + 4 * (size_t)nt;
  }
};

bool valid(int L, int Lz, int nr, int nt, int replica) {
  return valid_lattice(L, Lz) && nr >= 1 && nt >= 1 && nt <= EA_MAX_BETAS && replica >= 0 &&
         replica < nr;
}

size_t word(const Geometry &g, int nr, int replica, int clone, int h) {
  return (size_t)replica + (size_t)nr * h + (size_t)nr * g.n * clone;
}

}  // namespace

static_assert(sizeof(ea_run_data) == 72, "s_data is 72 bytes");

extern "C" size_t ea_conf_bytes(int L, int Lz, int nt) {
  if (!valid_lattice(L, Lz) || nt < 1) return 0;
  return Layout(L, Lz, nt).total;
}

extern "C" int ea_conf_encode(const ea_run_data *d, int L, int Lz, int nr, int nt,
                              const double *betas, const ea_hq_state *pt,
                              const ea_hq_state *prc, const ea_hq_state *cl,
                              const ea_xo_state *xo, uint64_t coin_offset,
                              const uint8_t *wctb, const uint64_t *even, const uint64_t *odd,
                              int replica, const int32_t *ener_by_clone, uint8_t *out) {
  if (!d || !betas || !pt || !prc || !cl || !xo || !wctb || !even || !odd || !ener_by_clone ||
      !out || !valid(L, Lz, nr, nt, replica))
    return EA_E_ARG;
  const Geometry g(L, Lz);
  const Layout lay(L, Lz, nt);
  std::memset(out, 0, lay.total);
  ea_run_data dd = *d;
  dd.pad_ = 0;
  std::memcpy(out + lay.data, &dd, sizeof dd);
  std::memcpy(out + lay.betas, betas, 8 * (size_t)nt);
  std::memcpy(out + lay.pt, pt, sizeof *pt);
  std::memcpy(out + lay.prc, prc, sizeof *prc);
  std::memcpy(out + lay.cl, cl, sizeof *cl);
  std::memcpy(out + lay.xo, xo, sizeof *xo);
  std::memcpy(out + lay.off, &coin_offset, sizeof coin_offset);
  std::memcpy(out + lay.perm, wctb, (size_t)nt);
  uint8_t *sp = out + lay.spins;
  const size_t V8 = (size_t)g.V / 8;
  for (int c = 0; c < nt; ++c)
    for (int site = 0; site < g.V; ++site) {
      int par, h, bit;
      g.locate(site, par, h, bit);
      const uint64_t w = (par ? odd : even)[word(g, nr, replica, c, h)];
      sp[c * V8 + (site >> 3)] |= (uint8_t)(((w >> bit) & 1u) << (site & 7));
    }
  for (int b = 0; b < nt; ++b) {
    if (wctb[b] >= nt) return EA_E_ARG;
    const int32_t e = ener_by_clone[wctb[b]];
    std::memcpy(out + lay.ener + 4 * (size_t)b, &e, 4);
  }
  return EA_OK;
}

extern "C" int ea_conf_decode(const uint8_t *buf, size_t len, int L, int Lz, int nr, int nt,
                              ea_run_data *d, double *betas, ea_hq_state *pt, ea_hq_state *prc,
                              ea_hq_state *cl, ea_xo_state *xo, uint64_t *coin_offset,
                              uint8_t *wctb, uint64_t *even, uint64_t *odd, int replica,
                              int32_t *ener_by_beta) {
  if (!buf || !valid(L, Lz, nr, nt, replica)) return EA_E_ARG;
  const Geometry g(L, Lz);
  const Layout lay(L, Lz, nt);
  if (len != lay.total) return EA_E_ARG;  // the size check of read_conf (io.c:128-134)
  // the permutation must be one (io.c:196-205)
  std::vector<int> seen(nt, 0);
  for (int b = 0; b < nt; ++b) {
    const int c = buf[lay.perm + b];
    if (c >= nt || seen[c]++) return EA_E_ARG;
  }
  if (d) std::memcpy(d, buf + lay.data, sizeof *d);
  if (betas) std::memcpy(betas, buf + lay.betas, 8 * (size_t)nt);
  if (pt) std::memcpy(pt, buf + lay.pt, sizeof *pt);
  if (prc) std::memcpy(prc, buf + lay.prc, sizeof *prc);
  if (cl) std::memcpy(cl, buf + lay.cl, sizeof *cl);
  if (xo) std::memcpy(xo, buf + lay.xo, sizeof *xo);
  if (coin_offset) std::memcpy(coin_offset, buf + lay.off, sizeof *coin_offset);
  if (wctb) std::memcpy(wctb, buf + lay.perm, (size_t)nt);
  if (even && odd) {
    const uint8_t *sp = buf + lay.spins;
    const size_t V8 = (size_t)g.V / 8;
    for (int c = 0; c < nt; ++c)
      for (int site = 0; site < g.V; ++site) {
        int par, h, bit;
        g.locate(site, par, h, bit);
        uint64_t &w = (par ? odd : even)[word(g, nr, replica, c, h)];
        const uint64_t v = (sp[c * V8 + (site >> 3)] >> (site & 7)) & 1u;
        w = (w & ~(1ULL << bit)) | (v << bit);
      }
  }
  if (ener_by_beta) std::memcpy(ener_by_beta, buf + lay.ener, 4 * (size_t)nt);
  return EA_OK;
}

extern "C" int ea_pack_lattice(int L, int Lz, int nr, int nt, const uint8_t *sites, int replica,
                               int clone, uint64_t *even, uint64_t *odd) {
  if (!sites || !even || !odd || !valid(L, Lz, nr, nt, replica) || clone < 0 || clone >= nt)
    return EA_E_ARG;
  const Geometry g(L, Lz);
  for (int site = 0; site < g.V; ++site) {
    int par, h, bit;
    g.locate(site, par, h, bit);
    uint64_t &w = (par ? odd : even)[word(g, nr, replica, clone, h)];
    w = (w & ~(1ULL << bit)) | ((uint64_t)(sites[site] & 1) << bit);
  }
  return EA_OK;
}
