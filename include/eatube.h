/*
 * eatube.h -- C ABI of the MI355X (gfx950) engine for Monte Carlo of the 3D
 * Edwards-Anderson +-J spin glass on L x L x Lz tubes (periodic in x and y,
 * periodic or open along z): multispin-coded checkerboard Metropolis with a
 * counter-based Philox+LUT random stream, parallel tempering and the
 * Houdayer cluster move.
 *
 * Drop-in boundary: these functions replace the per-step engine the
 * reference program (IsidoroGlez/EA-LCD-MC, PROGRAM/) drives from main();
 * each declaration names the reference function it replaces.  Data crossing
 * the boundary keeps the reference's host layouts, so a caller holding the
 * reference's host arrays passes them unchanged:
 *
 *   spins   even/odd arrays of n*NR*NT words, word index r + NR*h + NR*n*clone
 *           (h = MSC_site/2, bit = bx + 4*by + 16*bz; ini.c:339-383),
 *           n = L*L*Lz/128;
 *   J       12*n words [parity][direction][h], direction 0..5 = +x,+y,+z,
 *           -x,-y,-z (the device table d_J of main.cu:693-708);
 *   perms   which_clon_this_beta[sample][replica][beta] bytes
 *           (header.h:155, upd.c:3-48);
 *   energy  [sample][replica][clone] int32, = -H (main.cu:1379-1380).
 *
 * Conventions: every function returns 0 on success or a negative EA_E*
 * code; ea_last_error() gives the message of the last failure on the
 * calling thread.  Nothing calls exit().  An engine owns its device
 * buffers and one HIP stream; it is not thread-safe, several engines may
 * coexist.  All host arrays are caller-owned and copied.
 */
#ifndef EATUBE_H
#define EATUBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EA_LUT_BYTES 10240 /* s_lut_heat_bath, NUMBITSPREBUSQUEDAS=12 (header.h:108-111) */
#define EA_MAX_SAMPLES 31  /* nbits < MAXNSAMPLES (main.cu:53-54) */
#define EA_MAX_REPLICAS 16
#define EA_MAX_BETAS 255   /* permutations are bytes (header.h:155-156) */

enum {
  EA_OK = 0,
  EA_E_ARG = -1,     /* invalid argument */
  EA_E_HIP = -2,     /* HIP runtime error */
  EA_E_NOMEM = -3,   /* allocation failed */
  EA_E_STATE = -4,   /* call not valid in the current state */
  EA_E_NODEV = -5,   /* no usable gfx950 device */
};

typedef enum { EA_PBC = 0, EA_OBC = 1 } ea_bc;

typedef struct ea_engine ea_engine;

/* Engine geometry and constants; the reference fixes these at compile time
 * (PROGRAM/Makefile:3-36, header.h:14-68). */
typedef struct {
  int32_t L, Lz;           /* multiples of 8 (header.h:41-47) */
  int32_t nr;              /* replicas, power of two in 2..16 (header.h:14) */
  int32_t nt;              /* temperatures NBETAS, 2..255 */
  int32_t ns;              /* samples nbits, 1..31 */
  int32_t bc;              /* ea_bc */
  int32_t clone_threshold; /* Houdayer acts on the clone_threshold lowest-T
                              betas: lround(NT*0.5) (HoudayerXEA.cu:136) */
  uint32_t job_id;         /* isample, < 2^14 (main.cu:59-62) */
  const double *betas;     /* nt values, sorted descending (io.c:495-498) */
  const uint8_t *lut;      /* nt * EA_LUT_BYTES, LUT file body (io.c:504-534) */
} ea_config;

/* One half-sweep's counter block for one replica, as _fill_bits builds it
 * (fill_bits.h:47-68). */
typedef struct {
  uint32_t x, y, z, w;
} ea_counter;

/* s_aleatorio_HQ_64bits (random_generator.h:96-104), 2080 bytes. */
typedef struct {
  uint64_t ira[256];
  uint64_t PR, zseed, final_;
  uint8_t ip, ip1, ip2, ip3;
  uint8_t padding[4];
} ea_hq_state;

/* s_xoshiro256pp (random_generator.h:107-112), 48 bytes. */
typedef struct {
  uint64_t s[4];
  uint64_t t, final_;
} ea_xo_state;

/* ---------------------------------------------------------------- errors */
const char *ea_last_error(void);

/* -------------------------------------------------------- engine (device) */

/* Number of visible HIP devices (main.cu:64-66). */
int ea_device_count(int *count);

/* Make `device` the calling thread's current device: cudaSetDevice
 * (main.cu:68-70).  Engines bind the device current at ea_create. */
int ea_select_device(int device);

/* Allocate device state and upload constants: replaces CopyConstOnDevice /
 * CopyDataOnDevice (main.cu:634-725) and ClusterInit (HoudayerXEA.cu:123-181).
 * Binds the calling thread's current HIP device.  With EA_PT_EXACT_WALK=1 in
 * the environment at creation, every PT call of the engine (and of batches
 * made from it) decides each swap by the exact rule pair by pair, the path
 * otherwise taken only when a pair's swap threshold does not verify; the
 * decisions are the same either way, so the switch exists to test that path. */
int ea_create(const ea_config *cfg, ea_engine **out);

/* Device bytes ea_create(cfg) allocates on gfx950 (160 KiB of LDS per
 * workgroup), before the per-call inputs and the TRW history; no GPU needed.
 * No counterpart in the reference: admission by HBM footprint. */
int ea_footprint(const ea_config *cfg, uint64_t *bytes);

/* Free everything: FreeDevice (main.cu:748-768), SWclose (HoudayerXEA.cu:214). */
int ea_destroy(ea_engine *e);

/* Wait for all work queued on the engine's stream. */
int ea_sync(ea_engine *e);

/* Philox keys of the NR replicas (seed_keys, header.h:138; filled by
 * Inicia_generadores_CPU, random_generator.cu:151-156). */
int ea_set_keys(ea_engine *e, const uint32_t *keys);

/* Spins and couplings of one sample in the reference layouts: replaces the
 * spin/J uploads of CopyDataOnDevice (main.cu:683-708). */
int ea_upload_sample(ea_engine *e, int sample, const uint64_t *even,
                     const uint64_t *odd, const uint64_t *J);

/* Spin download of one sample (main.cu:251-257). */
int ea_download_spins(ea_engine *e, int sample, uint64_t *even, uint64_t *odd);

/* PT permutations which_clon_this_beta[ns][nr][nt]: send_PT_state_to_GPU
 * (main.cu:727-746) and its inverse. */
int ea_upload_perms(ea_engine *e, const uint8_t *wctb);
int ea_download_perms(ea_engine *e, uint8_t *wctb);

/* PT generators random_PT[ns][nr] (header.h:142) and statistics
 * aceptancePT / attemptsPT / aceptancePTraw [ns][nr][nt-1] (header.h:154).
 * PT runs on the device, so its state lives there. */
int ea_upload_pt_state(ea_engine *e, const ea_hq_state *st, const double *acc,
                       const double *att, const double *raw);
int ea_download_pt_state(ea_engine *e, ea_hq_state *st, double *acc,
                         double *att, double *raw);

/* Monte_Carlo (main.cu:285-349): nsweeps sweeps, each an even then an odd
 * half-sweep of all samples, replicas and clones.  blocks holds
 * [nsweeps][2][nr] counter blocks (what _fill_bits yields before each
 * half-sweep).  The random words are generated inside the sweep kernel and
 * never stored. */
int ea_sweep(ea_engine *e, int nsweeps, const ea_counter *blocks);

/* ea_sweep that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * counter blocks go through one static device array per GPU of
 * EA_NOALLOC_COUNTERS blocks (64 KiB), so nsweeps * 2 * nr must not exceed
 * it (else EA_E_ARG, nothing done).  Every engine on a GPU shares that
 * array, and a call's kernels read it until they finish: calls on one GPU
 * whose sweeps overlap (two engines' streams) race, so wait for the
 * previous call (ea_sync) first.  Results equal ea_sweep's; its launches
 * are not bracketed by ea_set_timing's events. */
#define EA_NOALLOC_COUNTERS 4096
int ea_sweep_noalloc(ea_engine *e, int nsweeps, const ea_counter *blocks);

/* Measure_Energy (main.cu:531-554): energies [ns][nr][clone] into out (may
 * be NULL: the energies stay on the device for ea_pt). */
int ea_energy(ea_engine *e, int32_t *out);

/* Site overlaps Q = sum_i s_i(r1) s_i(r2) of every replica pair r1 < r2,
 * each replica at the clone it holds at beta, into out[ns][pair][nt] with
 * the pairs in the order (0,1), (0,2), ..., (nr-2, nr-1): popcount
 * reductions over the MSC words on the device.  The reference's program
 * (PROGRAM/) has no such measure; its overlaps are computed offline from
 * the configuration files (io.c:17-280) -- this is the same sum.  The first
 * call allocates the engine's ns*pairs*nt output words on the device. */
int ea_overlap(ea_engine *e, int32_t *out);

/* Parallel_Tempering (upd.c:3-48) for every (sample, replica) chain, on the
 * device, from the energies of the last ea_energy / ea_sweep. */
int ea_pt(ea_engine *e);

/* Houdayer (main.cu:351-453) with ClusterStep (HoudayerXEA.cu:183-211).
 * pairs = {r1, r2} of replica group 0 and {r1, r2} of group 1, as drawn on
 * the host from the cluster generator (main.cu:409-421).  Cluster coins come
 * from coin_seed and the running offset (cuda_rng_offset, header.h:144),
 * which advances by V*NT*ns per group; *rng_offset is updated. */
int ea_houdayer(ea_engine *e, const int32_t pairs[4], uint64_t coin_seed,
                uint64_t *rng_offset);

/* Cluster labels of the last ea_houdayer (test aid): [2][ns][clone_threshold]
 * [V] minimum linear site index of each site's cluster.  The engine keeps
 * them only after ea_keep_labels(e, 1). */
int ea_keep_labels(ea_engine *e, int on);
int ea_download_labels(ea_engine *e, uint32_t *labels);

/* The body of main()'s `it` loop (main.cu:149-228), n_it times, without
 * returning to the host: mesfr x { houfr x { ptfr sweeps, energy, PT },
 * Houdayer }, then energy, per-beta energy sums and the TRW history.
 *   blocks  [n_it][mesfr][houfr][ptfr][2][nr]
 *   pairs   [n_it][mesfr][4]   (ignored when cluster == 0)
 *   it0     index of the first `it` inside the bin, for the history
 *   frec_trw history stride (io.c:431-434); history rows land at it/frec_trw
 * Energy sums accumulate in double per [ns][nr][beta] until
 * ea_read_measures() returns and clears them. */
typedef struct {
  int32_t n_it, mesfr, houfr, ptfr, cluster;
  int32_t it0, frec_trw;
  const ea_counter *blocks;
  const int32_t *pairs;
  uint64_t coin_seed;
  uint64_t rng_offset; /* in: offset before the first Houdayer call */
} ea_schedule;
int ea_run(ea_engine *e, const ea_schedule *sch, uint64_t *rng_offset_out);

/* Energy sums [ns][nr][nt] (by beta) and TRW history rows
 * [max_rows][ns][nr][nt] (which_beta_this_clon) accumulated by ea_run;
 * either pointer may be NULL.  Clears the sums. */
int ea_read_measures(ea_engine *e, double *esum, uint8_t *history,
                     int max_rows);

/* Kernel timing (the role of the reference's per-bin "ps/s" report,
 * timing.c:3-41, at kernel granularity): while on, every launch is
 * bracketed by HIP events on the engine's stream.  ea_read_timing() waits
 * for the stream and returns, per kernel kind, the summed milliseconds and
 * the launch count, then clears them.  Kinds: */
enum {
  EA_K_SWEEP = 0,    /* fused random words + Metropolis half-sweep */
  EA_K_ENERGY = 1,   /* standalone energy (incl. its zero fill) */
  EA_K_PT = 2,       /* parallel tempering */
  EA_K_HOUDAYER = 3, /* all passes of one cluster move */
  EA_K_MEASURE = 4,  /* overlaps (the energy sums and history are made
                        inside PT and the cluster move) */
  EA_K_KINDS = 5
};
int ea_set_timing(ea_engine *e, int on);
int ea_read_timing(ea_engine *e, double *ms, int64_t *launches);

/* ----------------------------------------------------------------- batch */

/* No counterpart in the reference, which runs one tube size per process
 * (PROGRAM/Makefile:3-36): a batch advances several engines' tubes (any
 * mix of L, Lz, boundary condition, NT and samples on one device) in the
 * same kernel launches.  One step of every tube (sweeps, PT, Houdayer and
 * the energy measures, as ea_run) is captured once in a HIP graph and the
 * graph is replayed; the arithmetic is the engines' own, so each tube ends
 * bit-identical to ea_run on that engine alone.  The engines stay usable on
 * their own between batch runs. */
typedef struct ea_batch ea_batch;
int ea_batch_create(ea_engine *const *engines, int n, ea_batch **out);
int ea_batch_destroy(ea_batch *b);

/* ea_run for every engine of the batch, schedules[i] for engine i.  The
 * schedules share mesfr, houfr and ptfr; n_it, cluster, it0, frec_trw and the
 * inputs are per tube (a tube with fewer steps idles while the others go
 * on).  Returns when all steps are done; rng_offsets_out[i] (may be NULL)
 * receives engine i's coin offset after its last Houdayer call. */
int ea_batch_run(ea_batch *b, const ea_schedule *schedules, uint64_t *rng_offsets_out);

/* Launch shape (no counterpart in the reference, whose block sizes are
 * compile-time constants, main.cu:1095-1510).  Results do not depend on it,
 * only speed does; the best shape of a tube mix is found by timing
 * (eatube.tune_shape).
 *   sweep_kind 0: each position's random words are made once and shared by
 *     the samples through LDS, sweep_positions (32 or 64) positions per
 *     1024-thread block;
 *   sweep_kind 1: one thread per position updates every sample
 *     (sweep_positions must be 64);
 *   sweep_kind 2: kind 1 with each sample's neighbour words staged in LDS
 *     tiles (sweep_positions must be 64; the tube's tile, 256 + 2*L*L/32
 *     half sites, must fit 768, else EA_E_ARG);
 *   sweep_kind 3: a whole sweep (both halves) per launch, each 512-thread
 *     block holding whole clones of several samples in LDS
 *     (sweep_positions must be 64; tubes with L*L*Lz/128 <= 512, else
 *     EA_E_ARG);
 *   sweep_kind 4: a whole sweep per launch over z-slabs of the clones,
 *     the new words written to a second spin array that the engine
 *     allocates when the kind is set (ea_footprint does not count it) and
 *     swaps with the first after every sweep (sweep_positions must be 64;
 *     tubes with a slab thickness T dividing Lz/4, Lz/4 >= T + 4,
 *     (T + 2)*L*L/32 <= 1024 and (2T + 6)*L*L/32 <= 3072 words in LDS,
 *     else EA_E_ARG);
 *   sweep_xcd 1: each XCD runs a contiguous range of sweep blocks;
 *   cluster_threads: threads of a whole-tube Houdayer workgroup (256, 512
 *     or 1024).
 * The default is {64, 1024, 0, 0}.  ea_set_shape applies to an engine's own
 * calls, ea_batch_set_shape to the batch's launches. */
typedef struct ea_shape {
  int32_t sweep_positions;
  int32_t cluster_threads;
  int32_t sweep_kind;
  int32_t sweep_xcd;
} ea_shape;
int ea_set_shape(ea_engine *e, const ea_shape *shape);
int ea_batch_set_shape(ea_batch *b, const ea_shape *shape);

/* ------------------------------------------------- host helpers (no GPU) */

/* Init_Rand_HQ_64bits (random_generator.cu:160-201) and one draw
 * (_actualiza_aleatorio_HQ_escalar, random_generator.h:130-135). */
void ea_hq_init(ea_hq_state *g, uint64_t seed);
uint64_t ea_hq_next(ea_hq_state *g);

/* generate_seeds_from_one (random_generator.cu:203-241). */
void ea_seeds_from_one(uint64_t seed, uint64_t *out, int n);

/* Inicia_generadores_CPU (random_generator.cu:94-158). */
void ea_init_generators(uint64_t seed, ea_hq_state *prc, ea_xo_state *xo,
                        uint32_t *key);

/* _fill_bits (fill_bits.h:47-68) for nhalf consecutive half-sweeps starting
 * at counter iter0 (parity alternates 0,1,0,...): out[nhalf][nr]. */
void ea_fill_counters(ea_hq_state *prc, ea_xo_state *xo, int nr, int iter0,
                      int nhalf, uint32_t job_id, ea_counter *out);

/* prepara_LUT (create_LUT.c:265-388) for one beta: EA_LUT_BYTES bytes. */
void ea_make_lut(double beta, uint8_t *out);

/* Init_u + packing_u + Init_J + packing_J (ini.c:316-730) for one sample:
 * even/odd spins (n*nr*nt words each), J words (12*n) and the six coupling
 * byte arrays Jx,Jy,Jz,Jmx,Jmy,Jmz (6*V, the couplings.bin body, io.c:284-305;
 * may be NULL).  obc_variant: 0 independent, 1 weak, 2 strong; mattis != 0
 * builds the gauge-transformed ferromagnet of ini.c:486-509.  Returns countJ
 * (>= 0) or an error code. */
int ea_init_sample(int L, int Lz, int nr, int nt, int bc, int obc_variant,
                   int mattis, uint64_t seed_J, uint64_t seed_u,
                   uint64_t *even, uint64_t *odd, uint64_t *J,
                   uint8_t *couplings);

/* Replica pairs of one Houdayer call from the cluster generator
 * (main.cu:409-421): pairs[4] = r1,r2 of group 0 then of group 1. */
int ea_houdayer_pairs(ea_hq_state *cl, int nr, int32_t pairs[4]);

/* Unpack / pack one (replica, clone) lattice between the MSC layout and
 * one byte per site (unpacking_u / packing_u, ini.c:339-433).  Packing
 * changes only that (replica, clone)'s bits. */
int ea_unpack_lattice(int L, int Lz, int nr, int nt, const uint64_t *even,
                      const uint64_t *odd, int replica, int clone,
                      uint8_t *sites);
int ea_pack_lattice(int L, int Lz, int nr, int nt, const uint8_t *sites,
                    int replica, int clone, uint64_t *even, uint64_t *odd);

/* s_data (header.h:91-106): run parameters in input.in order and master
 * seeds, 72 bytes as conf files store them (pad_ is written as 0). */
typedef struct {
  int32_t nbin, itcut, itmax, houfr, mesfr, ptfr, nbetas, flag, l;
  int32_t pad_;
  uint64_t seed_J, seed_u, seed_MC, seed_Cluster;
} ea_run_data;

/* The checkpoint record of one (sample, replica): write_conf / read_conf
 * (io.c:17-237).  ea_conf_bytes is its size (read_conf's predicted_size).
 * Encode takes the sample's even/odd arrays (n*nr*nt words each) and the
 * energies by clone; decode checks the size and the permutation, returns
 * the energies by beta, and packs the spins into even/odd for `replica`
 * (other replicas' bits are left as they are).  Output pointers of decode
 * may be NULL. */
size_t ea_conf_bytes(int L, int Lz, int nt);
int ea_conf_encode(const ea_run_data *d, int L, int Lz, int nr, int nt,
                   const double *betas, const ea_hq_state *pt,
                   const ea_hq_state *prc, const ea_hq_state *cl,
                   const ea_xo_state *xo, uint64_t coin_offset,
                   const uint8_t *wctb, const uint64_t *even,
                   const uint64_t *odd, int replica,
                   const int32_t *ener_by_clone, uint8_t *out);
int ea_conf_decode(const uint8_t *buf, size_t len, int L, int Lz, int nr,
                   int nt, ea_run_data *d, double *betas, ea_hq_state *pt,
                   ea_hq_state *prc, ea_hq_state *cl, ea_xo_state *xo,
                   uint64_t *coin_offset, uint8_t *wctb, uint64_t *even,
                   uint64_t *odd, int replica, int32_t *ener_by_beta);

#ifdef __cplusplus
}
#endif
#endif
