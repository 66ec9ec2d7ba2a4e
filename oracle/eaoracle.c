/*
 * eaoracle.c -- CPU oracle for the EA tube Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see eaoracle.h for the pinning status of each
 * function).  A deliberately plain, scalar, table-driven restatement of the
 * reference algorithm: it keeps the reference's own neighbour/rotate tables,
 * its loop orders and its integer widths so that it can be compared line by
 * line with the cited reference code.  Speed is not a goal.
 */
#include "eaoracle.h"

#include <math.h>
#include <omp.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define LCG_A 3202034522624059733ULL /* random_generator.h:133 */
#define FNORM_D (5.4210108624275218e-20) /* random_generator.h:92 */

/* ------------------------------------------------------------------------ */
/* host generators                                                          */
/* ------------------------------------------------------------------------ */

/* _actualiza_aleatorio_HQ_escalar, random_generator.h:130-135 */
uint64_t orc_hq_next(orc_hq *g) {
  g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
  g->PR = g->ira[g->ip] ^ g->ira[g->ip3];
  g->zseed = g->zseed * LCG_A + 1;
  g->final_ = g->zseed + g->PR;
  g->ip++;
  g->ip1++;
  g->ip2++;
  g->ip3++;
  return g->final_;
}

/* Init_Rand_HQ_64bits, random_generator.cu:160-201 (seed must be non-zero:
 * the reference replaces 0 by /dev/urandom, comprueba_semilla :11-16) */
void orc_hq_init(orc_hq *g, uint64_t seed) {
  memset(g, 0, sizeof *g);
  g->zseed = seed;
  for (int i = 0; i < 111; i++) g->zseed = g->zseed * LCG_A + 1;
  for (int k = 128 - 61; k < 128; k++) {
    g->ira[k] = 0;
    for (int i = 0; i < 8; i++) {
      g->zseed = g->zseed * LCG_A + 1;
      g->ira[k] |= (g->zseed >> 56) << (8 * i);
    }
  }
  g->ip = 128;
  g->ip1 = (uint8_t)(128 - 24);
  g->ip2 = (uint8_t)(128 - 55);
  g->ip3 = (uint8_t)(128 - 61);
  for (int i = 0; i < 1111; i++) {
    g->ira[g->ip] = g->ira[g->ip1] + g->ira[g->ip2];
    g->ip++;
    g->ip1++;
    g->ip2++;
    g->ip3++;
  }
  for (int i = 0; i < 1111; i++) orc_hq_next(g);
}

static inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

/* _actualiza_xoshiro256pp, random_generator.h:118-127 */
uint64_t orc_xo_next(orc_xo *u) {
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

/* luxury(), random_generator.cu:57-75 */
uint32_t orc_luscher_next(orc_lux *l) {
  for (int i = 0; i < 389; i++) {
    int d = (int)l->lusch[l->ip1] - (int)l->lusch[l->ip2] - l->bit;
    l->lusch[l->ip] = (uint32_t)((d + 16777216) & 16777215);
    l->bit = d < 0 ? 1 : 0; /* (diferencia & (1UL<<31)) >> 31 */
    l->ip++;
    l->ip1++;
    l->ip2++;
  }
  return l->lusch[l->ip];
}

/* Init_luxury, random_generator.cu:27-55 */
void orc_luscher_init(orc_lux *l, uint64_t seed) {
  memset(l, 0, sizeof *l);
  l->ip = 128;
  l->ip1 = 128 - 10;
  l->ip2 = 128 - 24;
  for (int i = 0; i < 1111; i++) seed = seed * LCG_A + 1;
  for (int j = l->ip2; j < l->ip; j++) {
    uint64_t t = 0;
    for (int i = 0; i < 3; i++) {
      seed = seed * LCG_A + 1;
      t |= (seed >> 56) << (8 * i);
    }
    l->lusch[j] = (uint32_t)t;
  }
  seed = seed * LCG_A + 1;
  l->bit = (seed >> 63) ? 1 : 0;
  for (int i = 0; i < 1111; i++) orc_luscher_next(l);
}

/* Luescher_PRNG, random_generator.cu:77-91 */
uint64_t orc_luscher_u64(orc_lux *l) {
  uint64_t out = 0;
  for (int j = 0; j < 8; j++) {
    uint64_t b = orc_luscher_next(l);
    out |= (b >> 16) << (8 * j);
  }
  return out;
}

/* Inicia_generadores_CPU, random_generator.cu:94-158 */
void orc_init_generators(orc_hq *prc, orc_xo *xo, uint32_t *key,
                         uint64_t seed) {
  orc_lux l;
  orc_hq g;
  memset(&g, 0, sizeof g);
  orc_luscher_init(&l, seed);
  for (int i = 0; i < 1111; i++) orc_luscher_next(&l);
  g.zseed = orc_luscher_u64(&l);
  for (int k = 128 - 61; k < 128; k++) g.ira[k] = orc_luscher_u64(&l);
  g.ip = 128;
  g.ip1 = (uint8_t)(128 - 24);
  g.ip2 = (uint8_t)(128 - 55);
  g.ip3 = (uint8_t)(128 - 61);
  for (int i = 0; i < 2222; i++) orc_hq_next(&g);
  *prc = g;

  uint64_t st[4] = {0, 0, 0, 0};
  while (!(st[0] | st[1] | st[2] | st[3])) {
    for (int j = 0; j < 4; j++) {
      st[j] = 0;
      for (int i = 0; i < 64; i++) {
        uint64_t r = orc_luscher_next(&l);
        st[j] |= ((r >> 23) & 1ULL) << i;
      }
    }
  }
  memset(xo, 0, sizeof *xo);
  for (int j = 0; j < 4; j++) xo->s[j] = st[j];

  uint32_t k = 0;
  for (int i = 0; i < 32; i++) {
    uint32_t r = orc_luscher_next(&l);
    k |= ((r >> 23) & 1U) << i;
  }
  *key = k;
}

/* generate_seeds_from_one, random_generator.cu:203-241 */
void orc_seeds_from_one(uint64_t seed, uint64_t *out, int n) {
  static const int paranoia[4] = {709, 109, 1373, 2099};
  orc_lux l;
  orc_luscher_init(&l, seed);
  for (int i = 0; i < 1109; i++) orc_luscher_next(&l);
  for (int s = 0; s < n; s++) {
    uint64_t t = orc_luscher_u64(&l);
    t += LCG_A + 1;
    t >>= 62;
    for (int i = 0; i < paranoia[t]; i++) orc_luscher_next(&l);
    t = orc_luscher_u64(&l);
    out[s] = t * 6364136223846793005ULL + 14426950408889634071ULL;
  }
}

/* _fill_bits, fill_bits.h:47-68 (52 entropy bits: 14 in .x, 7 in .y,
 * 13 in .z, 18 in .w) */
void orc_fill_bits(orc_xo *xo, orc_hq *prc, int nr, int iter, int parity,
                   uint32_t job_id, uint32_t *out) {
  for (int r = 0; r < nr; r++) {
    orc_xo_next(&xo[r]);
    orc_hq_next(&prc[r]);
    uint64_t e = prc[r].final_ + xo[r].final_;
    e >>= 64 - 52;
    uint64_t t = (uint64_t)(iter >> 15);
    uint32_t y = (uint32_t)t;
    t = ((uint64_t)iter & 32767ULL) << 3;
    t |= ((uint64_t)parity & 1ULL) << 2;
    uint32_t x = (uint32_t)t;
    t = (e & ((1ULL << 14) - 1)) << (32 - 14);
    x |= (uint32_t)t;
    e >>= 14;
    t = (e & ((1ULL << 7) - 1)) << (32 - 7);
    y |= (uint32_t)t;
    e >>= 7;
    t = (e & ((1ULL << 13) - 1)) << (32 - 13);
    uint32_t z = (uint32_t)t;
    e >>= 13;
    t = e << 14;
    uint32_t w = ((uint32_t)t) | job_id;
    out[4 * r + 0] = x;
    out[4 * r + 1] = y;
    out[4 * r + 2] = z;
    out[4 * r + 3] = w;
  }
}

/* ------------------------------------------------------------------------ */
/* counter-based RNG and LUT (main.cu:995-1150)                             */
/* ------------------------------------------------------------------------ */

static inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * b) >> 32);
}

/* _philox_4x32_10 with _update_state/_update_key, main.cu:996-1043 */
void orc_philox(const uint32_t ctr[4], const uint32_t key_in[2],
                uint32_t out[4]) {
  uint32_t v0 = ctr[0], v1 = ctr[1], v2 = ctr[2], v3 = ctr[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; round++) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t lo0 = 0xD2511F53u * v0, hi0 = mulhi32(0xD2511F53u, v0);
    uint32_t lo1 = 0xCD9E8D57u * v2, hi1 = mulhi32(0xCD9E8D57u, v2);
    v0 = hi1 ^ v1 ^ k0;
    v1 = lo1;
    v2 = hi0 ^ v3 ^ k1;
    v3 = lo0;
  }
  out[0] = v0;
  out[1] = v1;
  out[2] = v2;
  out[3] = v3;
}

static inline uint32_t lut_umb_hi(const uint8_t *lut, int i) {
  uint32_t v;
  memcpy(&v, lut + 8 * i, 4);
  return v;
}
static inline uint32_t lut_umb_lo(const uint8_t *lut, int i) {
  uint32_t v;
  memcpy(&v, lut + 8 * i + 4, 4);
  return v;
}
static inline uint32_t lut_pre(const uint8_t *lut, uint32_t msb) {
  uint16_t v;
  memcpy(&v, lut + 2048 + 2 * msb, 2);
  return v;
}

/* R < umbral[i] as in the `decision` expression of _bisection */
static inline uint32_t below(const uint8_t *lut, int i, uint32_t m,
                             uint32_t l) {
  uint32_t hx = lut_umb_hi(lut, i), lx = lut_umb_lo(lut, i);
  int c = (m == hx) ? (l < lx) : (m < hx);
  return (uint32_t)(~(c - 1));
}

/* _bisection, main.cu:1063-1078 */
uint32_t orc_bisect(const uint8_t *lut, uint32_t Rmsb, uint32_t Rlsb) {
  const uint32_t msb = Rmsb >> (32 - 12);
  uint32_t mn, mx, decision, npr;
  decision = below(lut, 254, Rmsb, Rlsb);
  mn = lut_pre(lut, msb) & 255;
  mx = lut_pre(lut, msb) >> 8;
  mn = (decision & mn) | ((~decision) & 255);
  mx = (decision & mx) | ((~decision) & 255);
  while ((mx - mn) > 1) {
    npr = (mx + mn) >> 1;
    decision = below(lut, (int)npr, Rmsb, Rlsb);
    mx = (decision & npr) | ((~decision) & mx);
    mn = (decision & mn) | ((~decision) & npr);
  }
  decision = below(lut, (int)mn, Rmsb, Rlsb);
  npr = (decision & mn) | (~decision & mx);
  return npr;
}

/* _obten_aleatorio_contador + _get_rnd_word, main.cu:1045-1092 */
static uint32_t rnd_word(uint32_t *x, uint32_t y, uint32_t z, uint32_t w,
                         uint32_t tid, uint32_t useed, const uint8_t *lut) {
  uint32_t ctr[4] = {*x, y, z, w}, key[2] = {tid, useed}, v[4], lsb[4];
  orc_philox(ctr, key, v);
  (*x)++;
  uint32_t hi1 = v[0] ^ v[2];
  uint32_t lo1 = v[1] ^ v[3];
  lsb[0] = ((v[1] << 17) | (v[0] >> 15)) ^ hi1 ^ ((hi1 << 21) | lo1 >> 11);
  lsb[1] = ((v[0] << 17) | (v[1] >> 15)) ^ lo1 ^ (lo1 << 21);
  lsb[2] = (hi1 << 28) | (lo1 >> 4);
  lsb[3] = (lo1 << 28) | (hi1 >> 4);
  const uint32_t rot = (lsb[0] + v[1] + lsb[2] + v[3]) >> 27;
  uint32_t bb = orc_bisect(lut, v[0], lsb[0]);
  bb |= orc_bisect(lut, v[1], lsb[1]) << 8;
  bb |= orc_bisect(lut, v[2], lsb[2]) << 16;
  bb |= orc_bisect(lut, v[3], lsb[3]) << 24;
  /* (bb<<rot)|(bb>>(32-rot)); rot == 0 leaves bb unchanged on the hardware
   * (the shift by 32 is undefined in C, so it is special-cased here) */
  if (rot) bb = (bb << rot) | (bb >> (32 - rot));
  return bb;
}

/* d_gen_rndbits_cuda, main.cu:1095-1150 */
void orc_rand_words(int n, int nr, int nt, const uint32_t *ctr,
                    const uint32_t *keys, const uint8_t *lut, uint64_t *out) {
  const size_t tot = (size_t)nt * n * nr;
  for (int beta = 0; beta < nt; beta++)
    for (int h = 0; h < n; h++)
      for (int r = 0; r < nr; r++) {
        uint32_t tid = (uint32_t)(((size_t)beta * n + h) * nr + r);
        uint32_t x = ctr[4 * r], y = ctr[4 * r + 1], z = ctr[4 * r + 2],
                 w = ctr[4 * r + 3];
        z += tid;
        const uint8_t *lb = lut + (size_t)beta * ORC_LUT_BYTES;
        uint64_t wd[6];
        for (int k = 0; k < 6; k++)
          wd[k] = rnd_word(&x, y, z, w, tid, keys[r], lb);
        out[tid] = (wd[0] << 32) | wd[1];
        out[tot + tid] = (wd[2] << 32) | wd[3];
        out[2 * tot + tid] = (wd[4] << 32) | wd[5];
      }
}

/* prepara_LUT, create_LUT.c:265-388, NUMBITSPREBUSQUEDAS = 12 */
void orc_make_lut(double beta, uint8_t *out) {
  __float128 qbeta = (__float128)beta;
  __float128 q0 = 1.0Q - expq(-4.0Q * qbeta);
  __float128 q1 = expq(-4.0Q * qbeta);
  uint64_t caso[9], cum[256];
  for (int i = 0; i < 9; i++) {
    __float128 p = powq(q0, 8 - i) * powq(q1, i);
    __float128 t = 18446744073709551616.Q * p;
    caso[i] = (uint64_t)t;
  }
  cum[0] = caso[0];
  for (int i = 1; i < 255; i++)
    cum[i] = cum[i - 1] + caso[__builtin_popcount(i)];
  cum[255] = ~0ULL;
  memset(out, 0, ORC_LUT_BYTES);
  for (int i = 0; i < 256; i++) {
    uint32_t hi = (uint32_t)(cum[i] >> 32), lo = (uint32_t)cum[i];
    memcpy(out + 8 * i, &hi, 4);
    memcpy(out + 8 * i + 4, &lo, 4);
  }
  const uint64_t cola = (1ULL << (64 - 12)) - 1ULL;
  for (uint64_t m = 0; m < 4096; m++) {
    uint64_t t = m << (64 - 12);
    int i;
    for (i = 0; i < 256; i++)
      if (t < cum[i]) break;
    if (i > 255) i = 255;
    uint32_t tmin = (uint32_t)i;
    t |= cola;
    for (i = 0; i < 256; i++)
      if (t < cum[i]) break;
    if (i > 255) i = 255;
    uint32_t tmax = (uint32_t)i;
    uint16_t v = (uint16_t)(tmin | (tmax << 8));
    memcpy(out + 2048 + 2 * m, &v, 2);
  }
}

/* ------------------------------------------------------------------------ */
/* lattice tables (ini.c:103-299)                                           */
/* ------------------------------------------------------------------------ */

/* Init_MSC_neighbours + Init_MSC_index_and_rotate, ini.c:160-299 */
int orc_msc_tables(int L, int Lz, orc_tables *t) {
  const int ML = L / 4, MLz = Lz / 4, MS = ML * ML, MV = MS * MLz;
  const int n = MV / 2;
  int side[3] = {ML, ML, MLz}, off[4] = {1, ML, MS, MV};
  memset(t, 0, sizeof *t);
  t->L = L;
  t->Lz = Lz;
  t->n = n;
  t->V = L * L * Lz;
  unsigned *nb = malloc(sizeof(unsigned) * 6 * MV);
  int *eoro = malloc(sizeof(int) * MV);
  t->neig = malloc(sizeof(int) * 12 * n);
  t->rot = malloc(2 * n);
  t->white = malloc(sizeof(uint32_t) * n);
  t->black = malloc(sizeof(uint32_t) * n);
  for (int c = 0; c < MV; c++) {
    int ix[3];
    for (int d = 0; d < 3; d++) ix[d] = (c % off[d + 1]) / off[d];
    for (int d = 0; d < 3; d++) {
      int iy[3] = {ix[0], ix[1], ix[2]};
      iy[d] = (ix[d] + 1) % side[d];
      int p = iy[0] + iy[1] * off[1] + iy[2] * off[2];
      nb[6 * c + d] = (unsigned)p;
      nb[6 * p + d + 3] = (unsigned)c;
    }
  }
  for (int c = 0; c < MV; c++) eoro[c] = -1;
  int cw = 0, cb = 0;
  for (int c = 0; c < MV; c++) {
    int skip = 0;
    for (int d = 0; d < 6; d++)
      if (eoro[nb[6 * c + d]] != -1) {
        skip = 1;
        break;
      }
    if (!skip) {
      if (cw == n) return -1;
      t->white[cw++] = (uint32_t)c;
      eoro[c] = c;
    }
  }
  for (int i = 0; i < MV; i++)
    if (eoro[i] == -1) {
      if (cb == n) return -1;
      t->black[cb++] = (uint32_t)i;
    }
  if (cw != cb) return -1;
  for (int h = 0; h < n; h++)
    for (int d = 0; d < 6; d++) {
      for (int par = 0; par < 2; par++) {
        const uint32_t *own = par ? t->black : t->white;
        const uint32_t *oth = par ? t->white : t->black;
        unsigned value = nb[6 * own[h] + d];
        unsigned lo = 0, hi = (unsigned)n;
        while (lo < hi) {
          unsigned md = (hi + lo) / 2;
          if (oth[md] < value)
            lo = md + 1;
          else
            hi = md;
        }
        if (oth[lo] != value) return -1;
        t->neig[(par * 6 + d) * n + h] = (int)lo;
      }
    }
  for (int h = 0; h < n; h++)
    for (int par = 0; par < 2; par++) {
      const uint32_t *own = par ? t->black : t->white;
      const uint32_t *oth = par ? t->white : t->black;
      uint8_t r = 0;
      for (int d = 0; d < 6; d++) {
        int p = t->neig[(par * 6 + d) * n + h];
        unsigned diff = oth[p] - own[h];
        int want = d < 3 ? off[d] : -off[d - 3];
        if (diff != (unsigned)want) r |= (uint8_t)(1u << d);
      }
      t->rot[par * n + h] = r;
    }
  free(nb);
  free(eoro);
  return 0;
}

void orc_free_tables(orc_tables *t) {
  free(t->neig);
  free(t->rot);
  free(t->white);
  free(t->black);
  memset(t, 0, sizeof *t);
}

/* packing_u, ini.c:339-383 */
void orc_pack_spins(int L, int Lz, int nr, int nt, const uint8_t *uu,
                    uint64_t *even, uint64_t *odd) {
  const int ML = L / 4, MLz = Lz / 4, MS = ML * ML;
  const int V = L * L * Lz, n = V / 128;
  const size_t half = (size_t)n * nr * nt;
  uint64_t *ag[2] = {even, odd};
  memset(even, 0, 8 * half);
  memset(odd, 0, 8 * half);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nt; c++) {
      int site = 0;
      for (int z = 0; z < Lz; z++) {
        int bz = z / MLz, g = z - bz * MLz;
        for (int y = 0; y < L; y++) {
          int by = y / ML, b = y - by * ML;
          for (int x = 0; x < L; x++) {
            int bx = x / ML, a = x - bx * ML;
            int par = (x ^ y ^ z) & 1;
            int bit = bx + 4 * by + 16 * bz;
            int ms = (a + ML * b + g * MS) / 2;
            size_t tid = (size_t)r + (size_t)nr * ms + (size_t)n * nr * c;
            ag[par][tid] |=
                ((uint64_t)uu[((size_t)r * nt + c) * V + site]) << bit;
            site++;
          }
        }
      }
    }
}

/* unpacking_u, ini.c:385-433 */
void orc_unpack_spins(int L, int Lz, int nr, int nt, const uint64_t *even,
                      const uint64_t *odd, uint8_t *uu) {
  const int ML = L / 4, MLz = Lz / 4, MS = ML * ML;
  const int V = L * L * Lz, n = V / 128;
  const uint64_t *ag[2] = {even, odd};
  const size_t tot = (size_t)n * nr * nt;
  for (size_t tid = 0; tid < tot; tid++) {
    int c = (int)(tid / ((size_t)n * nr));
    size_t rest = tid - (size_t)c * n * nr;
    int ms = (int)(rest / nr);
    int r = (int)(rest - (size_t)ms * nr);
    ms *= 2;
    int g = ms / MS, rem = ms - g * MS, b = rem / ML, a = rem - b * ML;
    int bit = 0;
    for (int bz = 0; bz < 4; bz++) {
      int z = bz * MLz + g;
      for (int by = 0; by < 4; by++) {
        int y = by * ML + b;
        for (int bx = 0; bx < 4; bx++) {
          int aux = bx * ML + a;
          for (int par = 0; par < 2; par++) {
            int x = aux | (par ^ ((y ^ z) & 1));
            int site = x + y * L + z * L * L;
            uu[((size_t)r * nt + c) * V + site] =
                (uint8_t)((ag[par][tid] >> bit) & 1);
          }
          bit++;
        }
      }
    }
  }
}

/* packing_J, ini.c:672-730: J6 = {Jx,Jy,Jz,Jmx,Jmy,Jmz}[V] */
void orc_pack_J(const orc_tables *t, const uint8_t *J6, uint64_t *Jw) {
  const int L = t->L, Lz = t->Lz, ML = L / 4, MLz = Lz / 4, MS = ML * ML;
  const int V = t->V, n = t->n, MV = 2 * n;
  uint64_t *J = calloc((size_t)6 * MV, 8);
  int site = 0;
  for (int z = 0; z < Lz; z++) {
    int bz = z / MLz, g = z - bz * MLz;
    for (int y = 0; y < L; y++) {
      int by = y / ML, b = y - by * ML;
      for (int x = 0; x < L; x++) {
        int bx = x / ML, a = x - bx * ML;
        int bit = bx + 4 * by + 16 * bz;
        int ms = a + ML * b + g * MS;
        for (int d = 0; d < 3; d++) {
          J[6 * ms + d] |= ((uint64_t)(J6[d * V + site] & 1)) << bit;
          J[6 * ms + d + 3] |= ((uint64_t)(J6[(d + 3) * V + site] & 1)) << bit;
        }
        site++;
      }
    }
  }
  for (int h = 0; h < n; h++)
    for (int d = 0; d < 6; d++) {
      Jw[(0 * 6 + d) * n + h] = J[6 * t->white[h] + d];
      Jw[(1 * 6 + d) * n + h] = J[6 * t->black[h] + d];
    }
  free(J);
}

/* Init_u for one sample, ini.c:316-337 */
void orc_init_u(uint64_t seed_u, int V, int nr, int nt, uint8_t *uu) {
  orc_hq g;
  orc_hq_init(&g, seed_u);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nt; c++)
      for (int s = 0; s < V; s++)
        uu[((size_t)r * nt + c) * V + s] = (uint8_t)(orc_hq_next(&g) >> 63);
}

/* Init_J for one sample, ini.c:465-670 (neighbour offsets of
 * Init_neighbours, ini.c:103-132) */
int orc_init_J(uint64_t seed_J, int L, int Lz, int bc, int variant,
               int mattis, const uint8_t *u, uint8_t *J6) {
  const int S = L * L, V = S * Lz;
  uint8_t *Jx = J6, *Jy = J6 + V, *Jz = J6 + 2 * V;
  uint8_t *Jmx = J6 + 3 * V, *Jmy = J6 + 4 * V, *Jmz = J6 + 5 * V;
  uint8_t *f0 = calloc((size_t)V, 1); /* J_falso_z0 */
  int countJ = 0;
#define XP(x) ((x) == L - 1 ? -(L - 1) : 1)
#define XM(x) ((x) == 0 ? L - 1 : -1)
#define YP(y) ((y) == L - 1 ? -(L - 1) * L : L)
#define YM(y) ((y) == 0 ? (L - 1) * L : -L)
#define ZP(z) ((z) == Lz - 1 ? -(Lz - 1) * S : S)
#define ZM(z) ((z) == 0 ? (Lz - 1) * S : -S)
  if (mattis) {
    int site = 0;
    for (int z = 0; z < Lz; z++)
      for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
          Jx[site] = u[site] ^ u[site + XP(x)];
          Jy[site] = u[site] ^ u[site + YP(y)];
          if (bc == 0) {
            Jz[site] = u[site] ^ u[site + ZP(z)];
          } else {
            if (z < Lz - 1)
              Jz[site] = u[site] ^ u[site + ZP(z)];
            else
              Jz[site] = Jy[site];
            if (z == 0) f0[site] = Jy[site];
          }
          site++;
        }
  } else {
    orc_hq g;
    orc_hq_init(&g, seed_J);
    for (int site = 0; site < V; site++) {
      Jx[site] = (uint8_t)(orc_hq_next(&g) >> 63);
      countJ += Jx[site];
      Jy[site] = (uint8_t)(orc_hq_next(&g) >> 63);
      countJ += Jy[site];
      Jz[site] = (uint8_t)(orc_hq_next(&g) >> 63);
      countJ += Jz[site];
    }
    if (bc)
      for (int site = 0; site < S; site++)
        f0[site] = (uint8_t)(orc_hq_next(&g) >> 63);
  }
  if (bc) {
    if (variant == 1 || variant == 2) { /* WEAK / STRONG */
      int flipv = (variant == 1) ? 1 : 0;
      int z = 0, site = 0;
      for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
          if (((x ^ y ^ z) & 1) == 0) f0[site] = Jy[site] ^ flipv;
          site++;
        }
      z = Lz - 1;
      site = z * S;
      for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
          if (((x ^ y ^ z) & 1) == 0) Jz[site] = Jy[site] ^ flipv;
          site++;
        }
    }
    int z = 0, site = 0;
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        if ((x ^ y ^ z) & 1) f0[site] = f0[site + YM(y)];
        site++;
      }
    z = Lz - 1;
    site = z * S;
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        if ((x ^ y ^ z) & 1) Jz[site] = Jz[site + YM(y)];
        site++;
      }
  }
  int site = 0;
  for (int z = 0; z < Lz; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        Jmx[site] = Jx[site + XM(x)];
        Jmy[site] = Jy[site + YM(y)];
        if (bc == 0 || z > 0)
          Jmz[site] = Jz[site + ZM(z)];
        else
          Jmz[site] = f0[site];
        site++;
      }
  free(f0);
  return countJ;
}

/* ------------------------------------------------------------------------ */
/* Metropolis half-sweep and energies (main.cu:1154-1510)                   */
/* ------------------------------------------------------------------------ */

#define MASK_E 0x1111111111111111ULL
#define MASK_N 0x000F000F000F000FULL
#define MASK_U 0x000000000000FFFFULL
#define MASK_O 0x8888888888888888ULL
#define MASK_S 0xF000F000F000F000ULL
#define MASK_D 0xFFFF000000000000ULL

static uint64_t rotE(uint64_t o) { return (o & MASK_E) << 3 | (o & ~MASK_E) >> 1; }
static uint64_t rotN(uint64_t o) { return (o & MASK_N) << 12 | (o & ~MASK_N) >> 4; }
static uint64_t rotU(uint64_t o) { return (o & MASK_U) << 48 | (o & ~MASK_U) >> 16; }
static uint64_t rotO(uint64_t o) { return (o & MASK_O) >> 3 | (o & ~MASK_O) << 1; }
static uint64_t rotS(uint64_t o) { return (o & MASK_S) >> 12 | (o & ~MASK_S) << 4; }
static uint64_t rotD(uint64_t o) { return (o & MASK_D) >> 48 | (o & ~MASK_D) << 16; }

/* Metropolis algebra, main.cu:1474-1506 */
uint64_t orc_flip_mask(uint64_t f0, uint64_t f1, uint64_t f2, uint64_t f3,
                       uint64_t f4, uint64_t f5, uint64_t b0, uint64_t b1,
                       uint64_t b2) {
  uint64_t k1 = f0 ^ f1, k2 = f0 & f1, j1 = k1 ^ f2, k3 = k1 & f2;
  uint64_t j2 = k2 ^ k3;
  k1 = f3 ^ f4;
  k2 = f3 & f4;
  k3 = k1 & f5;
  uint64_t j3 = k1 ^ f5, j4 = k2 ^ k3;
  b1 = b0 & b1;
  b2 = b1 & b2;
  uint64_t id2 = b1, id1 = (b0 ^ b1) | b2;
  uint64_t j24 = j2 | j4;
  return (j1 & j3 & id1) | ((j24 | id2) & (j1 | j3 | id1)) |
         ((j2 & j4) | (j24 & id2));
}

/* d_oneMCstepBN_multisample, main.cu:1385-1510 (one sample) */
void orc_half_sweep(const orc_tables *t, int nr, int nt, int bc, int parity,
                    uint64_t *spin, const uint64_t *Jw,
                    const uint32_t *whichclone, const uint64_t *rw) {
  const int n = t->n, MS = (t->L / 4) * (t->L / 4), MLz = t->Lz / 4;
  const size_t half = (size_t)n * nr * nt, tot = half;
  uint64_t *nw = spin + (size_t)parity * half;
  const uint64_t *od = spin + (size_t)(1 - parity) * half;
  const uint64_t *J = Jw + (size_t)parity * 6 * n;
  const int *nb = t->neig + (size_t)parity * 6 * n;
  const uint8_t *rot = t->rot + (size_t)parity * n;
  for (int c = 0; c < nt; c++)
    for (int h = 0; h < n; h++)
      for (int r = 0; r < nr; r++) {
        const size_t coff = (size_t)c * n * nr;
        uint32_t beta = whichclone[c * nr + r];
        size_t ri = (size_t)beta * n * nr + (size_t)h * nr + r;
        uint64_t s = nw[coff + (size_t)h * nr + r];
        uint64_t sc[6];
        for (int d = 0; d < 6; d++)
          sc[d] = od[coff + (size_t)nb[d * n + h] * nr + r];
        uint8_t rf = rot[h];
        if (rf & 0x01) sc[0] = rotE(sc[0]);
        if (rf & 0x02) sc[1] = rotN(sc[1]);
        if (rf & 0x04) sc[2] = rotU(sc[2]);
        if (rf & 0x08) sc[3] = rotO(sc[3]);
        if (rf & 0x10) sc[4] = rotS(sc[4]);
        if (rf & 0x20) sc[5] = rotD(sc[5]);
        if (bc) {
          int g = (2 * h) / MS;
          uint64_t aux = parity == 0 ? sc[1] : sc[4];
          if (g == MLz - 1) sc[2] = (sc[2] & ~MASK_D) | (aux & MASK_D);
          if (g == 0) sc[5] = (sc[5] & ~MASK_U) | (aux & MASK_U);
        }
        uint64_t f[6];
        for (int d = 0; d < 6; d++) f[d] = s ^ J[d * n + h] ^ sc[d];
        uint64_t flip = orc_flip_mask(f[0], f[1], f[2], f[3], f[4], f[5],
                                      rw[ri], rw[tot + ri], rw[2 * tot + ri]);
        nw[coff + (size_t)h * nr + r] = s ^ flip;
      }
}

/* d_computeEne, main.cu:1209-1381 (one sample, even words, rotation decided
 * from the white/black index tables as the kernel does) */
void orc_energy_msc(const orc_tables *t, int nr, int nt, int bc,
                    const uint64_t *spin, const uint64_t *Jw, int32_t *ener) {
  const int n = t->n, ML = t->L / 4, MS = ML * ML, MLz = t->Lz / 4;
  const size_t half = (size_t)n * nr * nt;
  const uint64_t *nw = spin, *od = spin + half;
  const int delta[6] = {1, ML, MS, -1, -ML, -MS};
  const int order[6] = {0, 3, 1, 4, 2, 5};
  uint64_t (*const rf[6])(uint64_t) = {rotE, rotN, rotU, rotO, rotS, rotD};
  memset(ener, 0, sizeof(int32_t) * nr * nt);
  for (int c = 0; c < nt; c++)
    for (int h = 0; h < n; h++)
      for (int r = 0; r < nr; r++) {
        const size_t coff = (size_t)c * n * nr;
        int g = (2 * h) / MS, le = 0;
        uint64_t s = nw[coff + (size_t)h * nr + r], aux = 0;
        for (int k = 0; k < 6; k++) {
          int d = order[k];
          int p = t->neig[d * n + h];
          uint64_t sc = od[coff + (size_t)p * nr + r];
          if ((int)(t->black[p] - t->white[h]) != delta[d]) sc = rf[d](sc);
          if (bc) {
            if (d == 1) aux = sc;
            if (d == 2 && g == MLz - 1) sc = (sc & ~MASK_D) | (aux & MASK_D);
            if (d == 5 && g == 0) sc = (sc & ~MASK_U) | (aux & MASK_U);
          }
          le += __builtin_popcountll(s ^ (Jw[d * n + h] ^ sc));
        }
        ener[r * nt + c] += 6 * 64 - 2 * le;
      }
}

/* calculate_scalar_energy, med.c:3-44 */
int orc_energy_scalar(int L, int Lz, int bc, const uint8_t *u,
                      const uint8_t *J6) {
  const int S = L * L, V = S * Lz;
  const uint8_t *Jx = J6, *Jy = J6 + V, *Jz = J6 + 2 * V, *Jmz = J6 + 5 * V;
  int e = 0, site = 0;
  for (int z = 0; z < Lz; z++)
    for (int y = 0; y < L; y++)
      for (int x = 0; x < L; x++) {
        int par = (x ^ y ^ z) & 1;
        uint8_t s = u[site];
        e += s ^ Jx[site] ^ u[site + XP(x)];
        e += s ^ Jy[site] ^ u[site + YP(y)];
        if (bc == 0) {
          e += s ^ Jz[site] ^ u[site + ZP(z)];
        } else if (z == Lz - 1) {
          if (par == 0) e += s ^ Jz[site] ^ u[site + YP(y)];
        } else {
          e += s ^ Jz[site] ^ u[site + ZP(z)];
          if (z == 0 && par == 0) e += s ^ Jmz[site] ^ u[site + YP(y)];
        }
        site++;
      }
  return 3 * V - 2 * e;
}

/* ------------------------------------------------------------------------ */
/* parallel tempering (upd.c:3-48)                                          */
/* ------------------------------------------------------------------------ */
void orc_pt(orc_hq *g, int nt, const double *betas, uint8_t *wctb,
            uint8_t *wbtc, const int32_t *E, double *acc, double *att,
            double *raw) {
  for (int i = 0; i < nt - 1; i++) {
    att[i]++;
    double pt = -(betas[i + 1] - betas[i]) * (E[wctb[i + 1]] - E[wctb[i]]);
    orc_hq_next(g);
    int change = 0;
    if (pt >= 0) {
      acc[i]++;
      change = 1;
    } else {
      double e = exp(pt);
      acc[i] += e;
      if ((FNORM_D * g->final_) < e) change = 1;
    }
    if (change) {
      uint8_t tmp = wctb[i];
      wctb[i] = wctb[i + 1];
      wctb[i + 1] = tmp;
      raw[i]++;
    }
  }
  for (int i = 0; i < nt; i++) wbtc[wctb[i]] = (uint8_t)i;
}

/* ------------------------------------------------------------------------ */
/* Houdayer (main.cu:351-453, HoudayerXEA.cu:54-374)                        */
/* ------------------------------------------------------------------------ */

/* The reference labels clusters with Komura's label-equivalence kernels
 * (init_YK_custom, device_ReduceLabels, root_find; HoudayerXEA.cu:77-121,
 * 254-354); its final label is the minimum linear site index of the
 * connected set of equal overlap.  Restated here as a breadth-first search
 * in increasing site order. */
void orc_labels(int L, int Lz, int bc, const uint8_t *q, uint32_t *label) {
  const int S = L * L, V = S * Lz;
  uint32_t *queue = malloc(sizeof(uint32_t) * V);
  for (int i = 0; i < V; i++) label[i] = 0xFFFFFFFFu;
  for (int i = 0; i < V; i++) {
    if (label[i] != 0xFFFFFFFFu) continue;
    int head = 0, tail = 0;
    queue[tail++] = (uint32_t)i;
    label[i] = (uint32_t)i;
    while (head < tail) {
      int s = (int)queue[head++];
      int x = s % L, y = (s / L) % L, z = s / S;
      int nbs[6], k = 0;
      nbs[k++] = (x + 1) % L + y * L + z * S;
      nbs[k++] = (x + L - 1) % L + y * L + z * S;
      nbs[k++] = x + ((y + 1) % L) * L + z * S;
      nbs[k++] = x + ((y + L - 1) % L) * L + z * S;
      if (bc == 0 || z < Lz - 1) nbs[k++] = x + y * L + ((z + 1) % Lz) * S;
      if (bc == 0 || z > 0) nbs[k++] = x + y * L + ((z + Lz - 1) % Lz) * S;
      for (int j = 0; j < k; j++) {
        int t = nbs[j];
        if (label[t] == 0xFFFFFFFFu && q[t] == q[s]) {
          label[t] = (uint32_t)i;
          queue[tail++] = (uint32_t)t;
        }
      }
    }
  }
  free(queue);
}

/* Cluster coin.  The reference draws cuRAND XORWOW doubles
 * (HoudayerXEA.cu:174-178,189,373): coin[beta][site] = U(idx) < 0.5 with
 * idx = offset + s*V*NT + beta*V + site.  That generator is closed; this
 * build keeps the index scheme and takes the coin as bit (idx mod 128) of
 * Philox4x32-10 at counter (b_lo, b_hi, 0x486f7564, 0), b = idx / 128, key
 * (seed_lo, seed_hi) -- word (idx mod 128)/32, bit idx mod 32. */
int orc_coin(uint64_t seed, uint64_t idx) {
  const uint64_t b = idx >> 7;
  uint32_t ctr[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0x486f7564u, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
  orc_philox(ctr, key, o);
  const unsigned j = (unsigned)(idx & 127);
  return (int)((o[j >> 5] >> (j & 31)) & 1u);
}

/* site -> (parity, half index, bit) of the MSC layout (ini.c:355-375) */
static void site_pos(int L, int Lz, int site, int *par, int *h, int *bit) {
  const int ML = L / 4, MLz = Lz / 4, MS = ML * ML, S = L * L;
  int x = site % L, y = (site / L) % L, z = site / S;
  int bz = z / MLz, g = z - bz * MLz, by = y / ML, b = y - by * ML;
  int bx = x / ML, a = x - bx * ML;
  *par = (x ^ y ^ z) & 1;
  *bit = bx + 4 * by + 16 * bz;
  *h = (a + ML * b + g * MS) / 2;
}

/* ------------------------------------------------------------------------ */
/* simulation driver                                                        */
/* ------------------------------------------------------------------------ */

int orc_sim_init(orc_sim *s, const orc_cfg *c) {
  memset(s, 0, sizeof *s);
  s->c = *c;
  const int nr = c->nr, nt = c->nt, ns = c->ns;
  if (c->L % 8 || c->Lz % 8 || nr < 1 || nt < 2 || ns < 1 ||
      ns >= ORC_MAXNSAMPLES)
    return -1;
  if (orc_msc_tables(c->L, c->Lz, &s->t)) return -2;
  s->n = s->t.n;
  s->V = s->t.V;
  const int n = s->n, V = s->V;
  s->betas = malloc(sizeof(double) * nt);
  memcpy(s->betas, c->betas, sizeof(double) * nt);
  s->lut = malloc((size_t)nt * ORC_LUT_BYTES);
  for (int b = 0; b < nt; b++) {
    if (c->lut)
      memcpy(s->lut + (size_t)b * ORC_LUT_BYTES,
             c->lut + (size_t)b * ORC_LUT_BYTES, ORC_LUT_BYTES);
    else
      orc_make_lut(c->betas[b], s->lut + (size_t)b * ORC_LUT_BYTES);
  }
  /* generate_seed_vectors, ini.c:32-59 */
  uint64_t tmp[ORC_MAXNSAMPLES];
  orc_seeds_from_one(c->seed_J, tmp, ORC_MAXNSAMPLES);
  for (int i = 0; i < ns; i++) s->seeds_J[i] = tmp[c->list_samples[i]];
  orc_seeds_from_one(c->seed_u, tmp, ORC_MAXNSAMPLES);
  for (int i = 0; i < ns; i++) s->seeds_u[i] = tmp[c->list_samples[i]];
  s->seeds_MC = malloc(sizeof(uint64_t) * nr);
  orc_seeds_from_one(c->seed_MC, s->seeds_MC, nr);
  /* Init_Random, ini.c:61-101 */
  orc_hq_init(&s->cl, c->seed_cl);
  s->keys = malloc(sizeof(uint32_t) * nr);
  s->prc = malloc(sizeof(orc_hq) * nr);
  s->xo = malloc(sizeof(orc_xo) * nr);
  s->pt = malloc(sizeof(orc_hq) * ns * nr);
  for (int r = 0; r < nr; r++)
    orc_init_generators(&s->prc[r], &s->xo[r], &s->keys[r], s->seeds_MC[r]);
  for (int r = 0; r < nr; r++) {
    orc_seeds_from_one(s->seeds_MC[r], tmp, ORC_MAXNSAMPLES);
    for (int i = 0; i < ns; i++) {
      int seed_pt = (int)(uint32_t)tmp[c->list_samples[i]]; /* int seed_PT */
      orc_hq_init(&s->pt[i * nr + r], (uint64_t)(int64_t)seed_pt);
    }
  }
  s->rng_offset = 0;
  /* Init_u, packing_u, Init_J, packing_J */
  const size_t half = (size_t)n * nr * nt;
  s->spin = malloc(sizeof(uint64_t) * 2 * half * ns);
  s->J = malloc(sizeof(uint64_t) * 12 * n * ns);
  s->J6 = malloc((size_t)6 * V * ns);
  uint8_t *uu = malloc((size_t)nr * nt * V);
  for (int i = 0; i < ns; i++) {
    orc_init_u(s->seeds_u[i], V, nr, nt, uu);
    orc_pack_spins(c->L, c->Lz, nr, nt, uu, s->spin + 2 * half * i,
                   s->spin + 2 * half * i + half);
    s->countJ[i] = orc_init_J(s->seeds_J[i], c->L, c->Lz, c->bc, c->variant,
                              c->mattis, uu, s->J6 + (size_t)6 * V * i);
    orc_pack_J(&s->t, s->J6 + (size_t)6 * V * i, s->J + (size_t)12 * n * i);
  }
  free(uu);
  /* init_tempering, ini.c:763-781 */
  s->wctb = malloc((size_t)ns * nr * nt);
  s->wbtc = malloc((size_t)ns * nr * nt);
  for (int i = 0; i < ns * nr; i++)
    for (int b = 0; b < nt; b++) {
      s->wctb[i * nt + b] = (uint8_t)b;
      s->wbtc[i * nt + b] = (uint8_t)b;
    }
  s->ener = calloc((size_t)ns * nr * nt, sizeof(int32_t));
  s->acc = calloc((size_t)ns * nr * (nt - 1), sizeof(double));
  s->att = calloc((size_t)ns * nr * (nt - 1), sizeof(double));
  s->raw = calloc((size_t)ns * nr * (nt - 1), sizeof(double));
  s->esum = calloc((size_t)ns * nr * nt, sizeof(double));
  s->rw = malloc(sizeof(uint64_t) * 3 * half);
  s->iter = 0;
  return 0;
}

void orc_sim_free(orc_sim *s) {
  orc_free_tables(&s->t);
  free(s->betas);
  free(s->lut);
  free(s->seeds_MC);
  free(s->keys);
  free(s->prc);
  free(s->xo);
  free(s->pt);
  free(s->spin);
  free(s->J);
  free(s->J6);
  free(s->wctb);
  free(s->wbtc);
  free(s->ener);
  free(s->acc);
  free(s->att);
  free(s->raw);
  free(s->esum);
  free(s->rw);
  free(s->last_labels);
  memset(s, 0, sizeof *s);
}

/* Monte_Carlo, main.cu:285-349 */
/* Threads the simulation loops use (samples, chains and cluster instances
 * are independent, so any count gives the same results); 1 by default. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

void orc_sim_sweeps(orc_sim *s, int nsweeps) {
  const int nr = s->c.nr, nt = s->c.nt, n = s->n;
  const size_t half = (size_t)n * nr * nt;
  uint32_t *ctr = malloc(sizeof(uint32_t) * 4 * nr);
  for (int sw = 0; sw < nsweeps; sw++)
    for (int par = 0; par < 2; par++) {
      orc_fill_bits(s->xo, s->prc, nr, s->iter, par, s->c.job_id, ctr);
      orc_rand_words(n, nr, nt, ctr, s->keys, s->lut, s->rw);
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
      for (int i = 0; i < s->c.ns; i++) {
        uint32_t wc[256 * 16];
        /* d_whichclone[c*NR + r] = which_beta_this_clon (main.cu:733-738) */
        for (int r = 0; r < nr; r++)
          for (int c = 0; c < nt; c++)
            wc[c * nr + r] = s->wbtc[((size_t)i * nr + r) * nt + c];
        orc_half_sweep(&s->t, nr, nt, s->c.bc, par, s->spin + 2 * half * i,
                       s->J + (size_t)12 * n * i, wc, s->rw);
      }
      s->iter++;
    }
  free(ctr);
}

/* Measure_Energy, main.cu:531-554 */
void orc_sim_energy(orc_sim *s) {
  const int nr = s->c.nr, nt = s->c.nt, n = s->n;
  const size_t half = (size_t)n * nr * nt;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
  for (int i = 0; i < s->c.ns; i++)
    orc_energy_msc(&s->t, nr, nt, s->c.bc, s->spin + 2 * half * i,
                   s->J + (size_t)12 * n * i, s->ener + (size_t)i * nr * nt);
}

/* Site overlap Q = sum_i s_i(r1) s_i(r2) of replicas r1 < r2, each at the
 * clone it holds at beta, for every sample, pair and beta: per word of 64
 * sites, 64 - 2 popcount(w1 ^ w2).  The reference's program has no such
 * measure (overlaps are taken from its configuration files offline); this
 * is that sum over the MSC words of main.cu:1385-1510's layout. */
void orc_overlap(const orc_sim *s, int32_t *out) {
  const int nr = s->c.nr, nt = s->c.nt, n = s->n;
  const size_t half = (size_t)n * nr * nt;
  const int np = nr * (nr - 1) / 2;
  for (int i = 0; i < s->c.ns; i++) {
    int k = 0;
    for (int r1 = 0; r1 < nr; r1++)
      for (int r2 = r1 + 1; r2 < nr; r2++, k++)
        for (int b = 0; b < nt; b++) {
          const int c1 = s->wctb[((size_t)i * nr + r1) * nt + b];
          const int c2 = s->wctb[((size_t)i * nr + r2) * nt + b];
          long q = 0;
          for (int p = 0; p < 2; p++) {
            const uint64_t *w = s->spin + 2 * half * i + p * half;
            for (int h = 0; h < n; h++) {
              const uint64_t a = w[(size_t)nr * n * c1 + (size_t)nr * h + r1];
              const uint64_t c = w[(size_t)nr * n * c2 + (size_t)nr * h + r2];
              q += 64 - 2 * __builtin_popcountll(a ^ c);
            }
          }
          out[((size_t)i * np + k) * nt + b] = (int32_t)q;
        }
  }
}

/* PT loop of main.cu:170-178 */
void orc_sim_pt(orc_sim *s) {
  const int nr = s->c.nr, nt = s->c.nt;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(dynamic)
  for (int i = 0; i < s->c.ns; i++)
    for (int r = 0; r < nr; r++) {
      size_t k = (size_t)i * nr + r;
      orc_pt(&s->pt[k], nt, s->betas, s->wctb + k * nt, s->wbtc + k * nt,
             s->ener + k * nt, s->acc + k * (nt - 1), s->att + k * (nt - 1),
             s->raw + k * (nt - 1));
    }
}

/* Houdayer, main.cu:351-453, with ClusterStep HoudayerXEA.cu:183-211 */
void orc_sim_houdayer(orc_sim *s) {
  const int nr = s->c.nr, nt = s->c.nt, ns = s->c.ns, n = s->n, V = s->V;
  const int L = s->c.L, Lz = s->c.Lz, ctm = s->c.clonetbm;
  const size_t half = (size_t)n * nr * nt;
  if (s->keep_labels && !s->last_labels)
    s->last_labels = malloc(sizeof(uint32_t) * 2 * ns * ctm * (size_t)V);
  for (int grp = 0; grp < 2; grp++) {
    int index[64];
    for (int k = 0; k < nr / 2; k++) index[k] = k + grp * (nr / 2);
    for (int k = 0; k < nr / 2 - 1; k++) {
      orc_hq_next(&s->cl);
      int ind1 = k + (int)(((uint32_t)(s->cl.final_ >> 32)) %
                           (uint32_t)(nr / 2 - k));
      int tmp = index[k];
      index[k] = index[ind1];
      index[ind1] = tmp;
    }
    const int r1 = index[0], r2 = index[1];
    s->last_pairs[2 * grp] = r1;
    s->last_pairs[2 * grp + 1] = r2;
    /* the (sample, beta) instances of a group touch disjoint words */
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(dynamic)
    for (int i = 0; i < ns; i++) {
      for (int b = 0; b < ctm; b++) {
        uint64_t *sp = s->spin + 2 * half * i;
        uint8_t *q = malloc((size_t)V);
        uint32_t *lab = malloc(sizeof(uint32_t) * V);
        int c1 = s->wctb[((size_t)i * nr + r1) * nt + b];
        int c2 = s->wctb[((size_t)i * nr + r2) * nt + b];
        /* d_computeoverlap + d_unpack */
        for (int site = 0; site < V; site++) {
          int par, h, bit;
          site_pos(L, Lz, site, &par, &h, &bit);
          uint64_t w1 = sp[par * half + (size_t)c1 * n * nr + (size_t)h * nr + r1];
          uint64_t w2 = sp[par * half + (size_t)c2 * n * nr + (size_t)h * nr + r2];
          q[site] = (uint8_t)(((w1 ^ w2) >> bit) & 1);
        }
        orc_labels(L, Lz, s->c.bc, q, lab);
        if (s->last_labels)
          memcpy(s->last_labels + (((size_t)grp * ns + i) * ctm + b) * V, lab,
                 sizeof(uint32_t) * V);
        /* spin_select + spin_flip_YK_custom + d_pack + d_useoverlap */
        for (int site = 0; site < V; site++) {
          uint64_t idx = s->rng_offset + (uint64_t)i * V * nt +
                         (uint64_t)b * V + lab[site];
          uint8_t qn = q[site] ^ (uint8_t)orc_coin(s->c.seed_cl, idx);
          int par, h, bit;
          site_pos(L, Lz, site, &par, &h, &bit);
          uint64_t m = (uint64_t)qn << bit;
          sp[par * half + (size_t)c1 * n * nr + (size_t)h * nr + r1] ^= m;
          sp[par * half + (size_t)c2 * n * nr + (size_t)h * nr + r2] ^= m;
        }
        free(q);
        free(lab);
      }
    }
    s->rng_offset += (uint64_t)V * nt * ns; /* cuda_rng_offset, :195 */
  }
}

/* n_it iterations of main()'s `it` loop, main.cu:149-228: mesfr x { houfr x
 * { ptfr sweeps, Measure_Energy, PT }, Houdayer (NO_CLUSTER: none) }, then
 * Measure_Energy, the average_Energy sums (:209-218) and the TRW history
 * (:221-228): iteration it0 + k with (it0 + k) % frec == frec - 1 writes
 * which_beta_this_clon into row (it0 + k) / frec of hist[rows][ns][nr][nt]
 * (rows past hist_rows are dropped; hist may be NULL). */
void orc_sim_run(orc_sim *s, int n_it, int mesfr, int houfr, int ptfr,
                 int cluster, int it0, int frec, uint8_t *hist, int hist_rows) {
  const int nr = s->c.nr, nt = s->c.nt, ns = s->c.ns;
  for (int k = 0; k < n_it; k++) {
    for (int im = 0; im < mesfr; im++) {
      for (int ih = 0; ih < houfr; ih++) {
        orc_sim_sweeps(s, ptfr);
        orc_sim_energy(s);
        orc_sim_pt(s);
      }
      if (cluster) orc_sim_houdayer(s);
    }
    orc_sim_energy(s);
    for (int c = 0; c < ns * nr; c++)
      for (int b = 0; b < nt; b++)
        s->esum[(size_t)c * nt + b] +=
            (double)s->ener[(size_t)c * nt + s->wctb[(size_t)c * nt + b]];
    const int it = it0 + k;
    if (hist && it % frec == frec - 1 && it / frec < hist_rows)
      memcpy(hist + (size_t)(it / frec) * ns * nr * nt, s->wbtc, (size_t)ns * nr * nt);
  }
}

void orc_sim_steps(orc_sim *s, int nsteps, int ptfr, int houfr, int cluster) {
  orc_sim_run(s, nsteps, 1, houfr, ptfr, cluster, 0, 1, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* access for the Python tests                                              */
/* ------------------------------------------------------------------------ */
orc_sim *orc_sim_create(const orc_cfg *c) {
  orc_sim *s = malloc(sizeof *s);
  if (orc_sim_init(s, c)) {
    free(s);
    return NULL;
  }
  return s;
}

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  orc_sim_free(s);
  free(s);
}

void orc_sim_keep_labels(orc_sim *s, int on) { s->keep_labels = on; }

void *orc_sim_field(orc_sim *s, const char *name, size_t *nbytes) {
  const size_t nr = s->c.nr, nt = s->c.nt, ns = s->c.ns, n = s->n;
#define F(key, ptr, sz)                                                        \
  if (!strcmp(name, key)) {                                                    \
    *nbytes = (sz);                                                            \
    return (void *)(ptr);                                                      \
  }
  F("spin", s->spin, 8 * 2 * n * nr * nt * ns);
  F("J", s->J, 8 * 12 * n * ns);
  F("J6", s->J6, 6 * (size_t)s->V * ns);
  F("wctb", s->wctb, ns * nr * nt);
  F("wbtc", s->wbtc, ns * nr * nt);
  F("ener", s->ener, 4 * ns * nr * nt);
  F("acc", s->acc, 8 * ns * nr * (nt - 1));
  F("att", s->att, 8 * ns * nr * (nt - 1));
  F("raw", s->raw, 8 * ns * nr * (nt - 1));
  F("esum", s->esum, 8 * ns * nr * nt);
  F("pt", s->pt, sizeof(orc_hq) * ns * nr);
  F("prc", s->prc, sizeof(orc_hq) * nr);
  F("xo", s->xo, sizeof(orc_xo) * nr);
  F("keys", s->keys, 4 * nr);
  F("seeds_J", s->seeds_J, 8 * ns);
  F("seeds_u", s->seeds_u, 8 * ns);
  F("seeds_MC", s->seeds_MC, 8 * nr);
  F("cl", &s->cl, sizeof(orc_hq));
  F("iter", &s->iter, sizeof(int));
  F("rng_offset", &s->rng_offset, 8);
  F("lut", s->lut, ORC_LUT_BYTES * nt);
  F("countJ", s->countJ, 4 * ns);
  F("last_pairs", s->last_pairs, sizeof s->last_pairs);
  F("last_labels", s->last_labels,
    s->last_labels ? 4 * 2 * ns * (size_t)s->c.clonetbm * s->V : 0);
#undef F
  *nbytes = 0;
  return NULL;
}
