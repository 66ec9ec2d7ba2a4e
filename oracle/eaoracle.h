/*
 * eaoracle.h -- CPU oracle for the EA tube Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain, scalar C restatement of the
 * reference program's algorithm (IsidoroGlez/EA-LCD-MC, PROGRAM/).  It exists
 * to check the HIP engine: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the shipped engine
 * links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates.  Pinning status
 * (what checks each function against something other than itself) is listed
 * per function below; "parity unpinned" marks what only review and
 * statistics cover.
 *
 *   orc_hq_*, orc_xo_next, orc_luscher_*, orc_init_generators,
 *   orc_seeds_from_one ........ pinned: fixtures from the reference host C
 *                                (random_generator.cu) built in-container
 *   orc_fill_bits ............. pinned: fixtures from fill_bits.h itself
 *   orc_init_u / orc_init_J /
 *   orc_pack_* / orc_msc_tables pinned: fixtures from ini.c built in-container
 *   orc_philox ................ pinned: Random123 Philox4x32-10 known answers
 *   orc_bisect ................ pinned: brute-force search over the cumulative
 *                                thresholds + the LUT self-check
 *   orc_rand_words (post-mix,
 *   rotation, word packing) ... parity unpinned: the lsb mix/rotation of
 *                                main.cu:1055-1060,1083,1092 exists only in GPU
 *                                code; covered by review and Bernoulli
 *                                statistics of the produced bits
 *   orc_half_sweep ............ pinned: flip rule truth table (all 2^9 inputs),
 *                                zero-temperature quench vs brute force on the
 *                                unpacked lattice, beta=0 full flip
 *   orc_energy_msc ............ pinned: == reference scalar energy (med.c)
 *   orc_pt .................... pinned: fixtures from upd.c built in-container
 *   orc_labels ................ pinned: BFS minimum-index brute force
 *   orc_sim_houdayer's replica
 *   pairs ..................... pinned: the cluster stream by fixtures from
 *                                the reference host C; NR/2-1 draws per group
 *                                per call, the pair as the first two entries
 *                                of a forward Fisher-Yates shuffle in closed
 *                                form, and uniform ordered pairs (chi-square)
 *   orc_coin .................. substitute generator (reference uses closed
 *                                cuRAND): parity unpinned vs the reference
 *   orc_make_lut .............. pinned: byte-identical to the shipped
 *                                LUT_for_PRNG_nbits12_NB40.bin
 *   orc_overlap ............... pinned: == sum of s1*s2 over the unpacked
 *                                lattices (orc_unpack_spins, itself pinned)
 *   orc_sim_run ............... loop nesting and TRW rows restated from
 *                                main.cu:149-228, built from the pinned parts
 *                                (sweeps, energy, PT, Houdayer); the counts the
 *                                nesting fixes per `it` (half-sweeps, PT calls,
 *                                Houdayer calls and their cluster draws, one
 *                                energy sum, rows every frecTRW) are checked;
 *                                the files made from its sums and rows are
 *                                checked against the bytes of the reference's
 *                                write_measures / write_history
 */
#ifndef EAORACLE_H
#define EAORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXNSAMPLES 32
#define ORC_LUT_BYTES 10240 /* uint4 umbrales[128] + uint4 prebusqueda[512] */

/* s_aleatorio_HQ_64bits, random_generator.h:96-104 (2080 bytes) */
typedef struct {
  uint64_t ira[256];
  uint64_t PR;
  uint64_t zseed;
  uint64_t final_;
  uint8_t ip, ip1, ip2, ip3;
  uint8_t padding[4];
} orc_hq;

/* s_xoshiro256pp, random_generator.h:107-112 (48 bytes) */
typedef struct {
  uint64_t s[4];
  uint64_t t;
  uint64_t final_;
} orc_xo;

/* Lüscher 24-bit generator state (globals at random_generator.cu:6-8) */
typedef struct {
  uint32_t lusch[256];
  int bit;
  uint8_t ip, ip1, ip2;
} orc_lux;

/* host generators ------------------------------------------------------ */
void orc_hq_init(orc_hq *g, uint64_t seed);
uint64_t orc_hq_next(orc_hq *g);
uint64_t orc_xo_next(orc_xo *x);
void orc_luscher_init(orc_lux *l, uint64_t seed);
uint32_t orc_luscher_next(orc_lux *l);
uint64_t orc_luscher_u64(orc_lux *l);
void orc_init_generators(orc_hq *prc, orc_xo *xo, uint32_t *key, uint64_t seed);
void orc_seeds_from_one(uint64_t seed, uint64_t *out, int n);

/* counter block for one half-sweep: out[r*4 + {x,y,z,w}] */
void orc_fill_bits(orc_xo *xo, orc_hq *prc, int nr, int iter, int parity,
                   uint32_t job_id, uint32_t *out);

/* Philox4x32-10 (Salmon et al. SC'11) */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* LUT bisection: byte for 64-bit uniform (msb:lsb) under one beta's LUT */
uint32_t orc_bisect(const uint8_t *lut_beta, uint32_t msb, uint32_t lsb);

/* LUT generator (create_LUT.c:265-388), NUMBITSPREBUSQUEDAS = 12 */
void orc_make_lut(double beta, uint8_t *out /* 10240 bytes */);

/* random words b0,b1,b2 for every (beta, h, r) of one half-sweep.
 * out[k * nt*n*nr + (beta*n + h)*nr + r], k = 0..2 */
void orc_rand_words(int n, int nr, int nt, const uint32_t *ctr /*[nr][4]*/,
                    const uint32_t *keys /*[nr]*/, const uint8_t *lut,
                    uint64_t *out);

/* lattice helpers --------------------------------------------------------*/
typedef struct {
  int L, Lz, n, V;
  int *neig;        /* [2][6][n]: opposite-parity half index */
  uint8_t *rot;     /* [2][n] */
  uint32_t *white;  /* [n] bianchi_index */
  uint32_t *black;  /* [n] neri_index */
} orc_tables;

int orc_msc_tables(int L, int Lz, orc_tables *t);
void orc_free_tables(orc_tables *t);

/* spins uu[r][c][site] (0/1 bytes) -> packed [parity][c][h][r] */
void orc_pack_spins(int L, int Lz, int nr, int nt, const uint8_t *uu,
                    uint64_t *even, uint64_t *odd);
void orc_unpack_spins(int L, int Lz, int nr, int nt, const uint64_t *even,
                      const uint64_t *odd, uint8_t *uu);
/* couplings J6[6][V] (Jx,Jy,Jz,Jmx,Jmy,Jmz bytes) -> words [parity][6][h] */
void orc_pack_J(const orc_tables *t, const uint8_t *J6, uint64_t *Jw);

/* Init_u for one sample: uu[nr][nt][V] */
void orc_init_u(uint64_t seed_u, int V, int nr, int nt, uint8_t *uu);
/* Init_J for one sample.  bc: 0 PBC, 1 OBC; variant: 0 independent, 1 weak,
 * 2 strong; mattis: couplings from uu0 (replica 0, clone 0 spins).
 * Returns countJ. */
int orc_init_J(uint64_t seed_J, int L, int Lz, int bc, int variant,
               int mattis, const uint8_t *uu0, uint8_t *J6);

/* one half-sweep of one sample; spin = [parity][nt][n][nr] */
void orc_half_sweep(const orc_tables *t, int nr, int nt, int bc, int parity,
                    uint64_t *spin, const uint64_t *Jw,
                    const uint32_t *whichclone /*[nt*nr]: beta of clone*/,
                    const uint64_t *rw /*orc_rand_words output*/);
/* same, with explicit per-word id values (id in 0..3 broadcast): testing aid */
uint64_t orc_flip_mask(uint64_t f0, uint64_t f1, uint64_t f2, uint64_t f3,
                       uint64_t f4, uint64_t f5, uint64_t b0, uint64_t b1,
                       uint64_t b2);

/* MSC energy of one sample: ener[r*nt + c] */
void orc_energy_msc(const orc_tables *t, int nr, int nt, int bc,
                    const uint64_t *spin, const uint64_t *Jw, int32_t *ener);
/* scalar energy of one unpacked lattice (med.c:3-44) */
int orc_energy_scalar(int L, int Lz, int bc, const uint8_t *u,
                      const uint8_t *J6);

/* Parallel_Tempering (upd.c:3-48) for one (sample, replica) chain */
void orc_pt(orc_hq *g, int nt, const double *betas, uint8_t *wctb,
            uint8_t *wbtc, const int32_t *ener, double *acc, double *att,
            double *raw);

/* cluster labels: q[V] bytes -> label[V] = minimum linear site index of the
 * equal-q component (PBC in x,y; bc along z) */
void orc_labels(int L, int Lz, int bc, const uint8_t *q, uint32_t *label);
/* coin of the cluster generator for global index idx */
int orc_coin(uint64_t seed, uint64_t idx);

/* full simulation ---------------------------------------------------------*/
typedef struct {
  int L, Lz, nr, nt, ns, bc, variant, mattis, clonetbm;
  uint32_t job_id;
  uint64_t seed_J, seed_u, seed_MC, seed_cl;
  int list_samples[ORC_MAXNSAMPLES];
  const double *betas; /* nt, descending */
  const uint8_t *lut;  /* nt*10240, may be NULL -> generated */
} orc_cfg;

typedef struct {
  orc_cfg c;
  orc_tables t;
  int n, V;
  double *betas;
  uint8_t *lut;
  uint64_t seeds_J[ORC_MAXNSAMPLES], seeds_u[ORC_MAXNSAMPLES];
  uint64_t *seeds_MC;   /* [nr] */
  uint32_t *keys;       /* [nr] */
  orc_hq *prc;          /* [nr] */
  orc_xo *xo;           /* [nr] */
  orc_hq *pt;           /* [ns][nr] */
  orc_hq cl;
  uint64_t rng_offset;
  int countJ[ORC_MAXNSAMPLES];
  uint64_t *spin;       /* [ns][2][nt][n][nr] */
  uint64_t *J;          /* [ns][2][6][n] */
  uint8_t *J6;          /* [ns][6][V] */
  uint8_t *wctb;        /* [ns][nr][nt] which_clon_this_beta */
  uint8_t *wbtc;        /* [ns][nr][nt] which_beta_this_clon */
  int32_t *ener;        /* [ns][nr][nt] by clone */
  double *acc, *att, *raw; /* [ns][nr][nt-1] */
  double *esum;         /* [ns][nr][nt] by beta: per-step energy sums */
  uint64_t *rw;         /* scratch random words */
  int iter;
  int last_pairs[4];    /* r1,r2 of group 0 and 1 of the last Houdayer call */
  uint32_t *last_labels; /* optional [2][ns][clonetbm][V] of the last call */
  int keep_labels;
} orc_sim;

int orc_sim_init(orc_sim *s, const orc_cfg *c);
void orc_sim_free(orc_sim *s);
/* Monte_Carlo (main.cu:285-349): nsweeps full sweeps */
void orc_sim_sweeps(orc_sim *s, int nsweeps);
void orc_sim_energy(orc_sim *s);
void orc_sim_pt(orc_sim *s);
void orc_sim_houdayer(orc_sim *s);
/* Site overlaps [ns][pair r1 < r2][beta] of replicas at each beta. */
void orc_overlap(const orc_sim *s, int32_t *out);
/* n_it iterations of main()'s `it` loop (main.cu:149-228): mesfr x
 * { houfr x (ptfr sweeps, energy, PT), Houdayer (if cluster) }, energy, the
 * per-beta energy sums and the TRW history rows which_beta_this_clon
 * hist[(it0+k)/frec][ns][nr][nt] for (it0+k) % frec == frec-1 (rows beyond
 * hist_rows dropped; hist may be NULL). */
void orc_sim_run(orc_sim *s, int n_it, int mesfr, int houfr, int ptfr,
                 int cluster, int it0, int frec, uint8_t *hist, int hist_rows);
/* orc_sim_run with mesfr = 1 and no history */
void orc_sim_steps(orc_sim *s, int nsteps, int ptfr, int houfr, int cluster);
/* Threads of the simulation loops (default 1); results do not depend on it. */
void orc_set_threads(int n);
orc_sim *orc_sim_create(const orc_cfg *c);
void orc_sim_destroy(orc_sim *s);
void orc_sim_keep_labels(orc_sim *s, int on);
/* named array of the simulation state (see eaoracle.c for the names) */
void *orc_sim_field(orc_sim *s, const char *name, size_t *nbytes);

#ifdef __cplusplus
}
#endif
#endif
