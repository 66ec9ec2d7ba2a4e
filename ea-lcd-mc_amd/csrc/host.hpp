// host.hpp -- host-side pieces of the engine: the reference's generators and
// seeding chain, lattice geometry of the multispin-coded (MSC) layout, sample
// initialisation, counter blocks and the LUT generator.  No GPU code.
#pragma once

#include <cstdint>
#include <vector>

#include "eatube.h"

namespace eatube {

constexpr uint64_t kLcg = 3202034522624059733ULL;  // random_generator.h:133
constexpr double kFnorm = 5.4210108624275218e-20;  // random_generator.h:92

// Lüscher's 24-bit lagged subtract-with-borrow generator with luxury 389,
// used only to derive seeds (random_generator.cu:27-91).
class Luscher {
 public:
  explicit Luscher(uint64_t seed);
  uint32_t next();
  uint64_t next64();  // bits 16..23 of eight draws (Luescher_PRNG)
  uint32_t bits32();  // bit 23 of 32 draws, LSB first

 private:
  uint32_t ring_[256];
  uint8_t ip_, ip1_, ip2_;
  int borrow_;
};

// MSC geometry: a 64-bit word holds the 64 copies of one site of the
// (L/4) x (L/4) x (Lz/4) sub-lattice; bit = bx + 4 by + 16 bz.  Sites of
// parity p = (x^y^z)&1 live in array p at half index h = MSC_site/2
// (ini.c:339-383).
struct Geometry {
  int L = 0, Lz = 0, ML = 0, MLz = 0, MS = 0, n = 0, V = 0;
  Geometry() = default;
  Geometry(int L_, int Lz_);
  void locate(int site, int &par, int &h, int &bit) const {
    const int x = site % L, y = (site / L) % L, z = site / (L * L);
    par = (x ^ y ^ z) & 1;
    bit = x / ML + 4 * (y / ML) + 16 * (z / MLz);
    h = ((x % ML) + ML * (y % ML) + MS * (z % MLz)) >> 1;
  }
  int site_of(int par, int h, int bit) const {
    const int g = (2 * h) / MS, rem = 2 * h - g * MS, b = rem / ML;
    const int a = (rem - b * ML) + ((par + b + g) & 1);
    const int x = (bit & 3) * ML + a, y = ((bit >> 2) & 3) * ML + b,
              z = (bit >> 4) * MLz + g;
    return x + L * (y + L * z);
  }
};

bool valid_lattice(int L, int Lz);

// Sample initialisation (ini.c:316-730).  Spins are drawn straight into the
// packed layout; couplings are built as the six byte arrays of
// couplings.bin and then packed per parity and direction.
void init_spins(uint64_t seed_u, const Geometry &g, int nr, int nt,
                uint64_t *even, uint64_t *odd, std::vector<uint8_t> *first);
int init_couplings(uint64_t seed_J, const Geometry &g, int bc, int variant,
                   bool mattis, const uint8_t *u00, uint8_t *J6);
void pack_couplings(const Geometry &g, const uint8_t *J6, uint64_t *Jw);

// Exact per-pair tables of exp(pt) for the device parallel tempering:
// for pair i (betas i, i+1) and k >= 1, entry k-1 holds
// exp(-(beta[i+1]-beta[i]) * (-2k)), the value upd.c:29 computes for an
// energy difference of -2k.  Beyond the table exp() underflows to 0.
struct ExpTables {
  std::vector<double> nd;       // -(beta[i+1]-beta[i])
  std::vector<int32_t> off, len;
  std::vector<double> val;
};
ExpTables make_exp_tables(const double *betas, int nt, int V);

}  // namespace eatube
