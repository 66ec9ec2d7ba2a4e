/*
 * ref_harness.c -- test driver that links the reference program's own host C
 * (PROGRAM/src/ini.c, io.c, upd.c, med.c, random_generator.cu, compiled from
 * where they lie under the reference checkout; see oracle/Makefile) and dumps
 * what those functions produce, so that tests/golden/make_golden.py can turn
 * it into committed fixtures.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing here is part of the engine.  The GPU
 * kernels of the reference (main.cu) are not built: this image has no CUDA.
 *
 * usage: ref_harness OUTDIR BETAS NBITS SEED_J SEED_U SEED_MC SEED_CL JOBID
 * Every dump is OUTDIR/<name>.bin (raw little-endian bytes).
 */
#define MAIN
#include "header.h"
#include "fill_bits.h"

#include <sys/stat.h>

extern char dir[];
extern char dir_rep[][NR][MAXSTR];
extern char dirsample[][MAXSTR];

static char outdir[MAXSTR];

static void dump(const char *name, const void *p, size_t sz) {
  char path[2 * MAXSTR];
  snprintf(path, sizeof path, "%s/%s.bin", outdir, name);
  FILE *f = fopen(path, "wb");
  if (!f) {
    perror(path);
    exit(2);
  }
  if (sz && fwrite(p, 1, sz, f) != sz) exit(3);
  fclose(f);
}

static void copy_file(const char *src, const char *name) {
  FILE *f = fopen(src, "rb");
  if (!f) {
    perror(src);
    exit(4);
  }
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = (char *)malloc(sz > 0 ? sz : 1);
  if (fread(buf, 1, sz, f) != (size_t)sz) exit(5);
  fclose(f);
  dump(name, buf, sz);
  free(buf);
}

int main(int argc, char **argv) {
  if (argc != 9) {
    fprintf(stderr,
            "usage: %s OUTDIR BETAS NBITS SEED_J SEED_U SEED_MC SEED_CL JOBID\n",
            argv[0]);
    return 1;
  }
  snprintf(outdir, sizeof outdir, "%s", argv[1]);
  int nbits = atoi(argv[3]);
  read_betas(argv[2]);
  memset(&data, 0, sizeof data);
  data.nbin = 1;
  data.itmax = 1;
  data.houfr = 1;
  data.mesfr = 1;
  data.ptfr = 1;
  data.nbetas = NBETAS;
  data.flag = 1;
  data.l = (Lz << 16) + L;
  data.seed_J = strtoull(argv[4], NULL, 10);
  data.seed_u = strtoull(argv[5], NULL, 10);
  data.seed_MC = strtoull(argv[6], NULL, 10);
  data.seed_Cluster = strtoull(argv[7], NULL, 10);
  uint32_t mi_jobID = (uint32_t)atoi(argv[8]);

  read_list_samples("", nbits);
  snprintf(dir, MAXSTR, "%s/run/", outdir);
  mkdir(dir, 0775);
  create_sample_path(0, nbits);
  generate_seed_vectors();

  /* the buffers that AllocHost (main.cu:559-601) would provide */
  for (int ib = 0; ib < nbits; ib++) {
    Jx[ib] = (char *)malloc(V);
    Jy[ib] = (char *)malloc(V);
    Jz[ib] = (char *)malloc(V);
    Jmx[ib] = (char *)malloc(V);
    Jmy[ib] = (char *)malloc(V);
    Jmz[ib] = (char *)malloc(V);
    h_MSC_u_even[ib] = (MYWORD *)malloc(sizeof(MYWORD) * HALF_MSC_VNRNBETAS);
    h_MSC_u_odd[ib] = (MYWORD *)malloc(sizeof(MYWORD) * HALF_MSC_VNRNBETAS);
    uu[ib] = (char ***)malloc(sizeof(char **) * NR);
    for (int r = 0; r < NR; r++) {
      uu[ib][r] = (char **)malloc(sizeof(char *) * NBETAS);
      for (int b = 0; b < NBETAS; b++) uu[ib][r][b] = (char *)malloc(V);
    }
  }

  Init(nbits);

  dump("seeds_J", seeds_J, sizeof(uint64_t) * nbits);
  dump("seeds_u", seeds_u, sizeof(uint64_t) * nbits);
  dump("seeds_MC", seeds_MC, sizeof(uint64_t) * NR);
  dump("keys", seed_keys.my_key, sizeof(uint32_t) * NR);
  dump("prc", random_PRC, sizeof(s_aleatorio_HQ_64bits) * NR);
  dump("xo", random_xoshiro256pp, sizeof(s_xoshiro256pp) * NR);
  dump("pt_init", random_PT, sizeof(s_aleatorio_HQ_64bits) * NR * nbits);
  dump("cl_init", &random_c, sizeof(s_aleatorio_HQ_64bits));
  dump("countJ", countJ0, sizeof(int) * nbits);
  {
    int blk[2] = {numThreadsPerBlock, numBlocks};
    dump("blocks", blk, sizeof blk);
  }
  for (int ib = 0; ib < nbits; ib++) {
    char nm[64];
    snprintf(nm, sizeof nm, "even_%d", ib);
    dump(nm, h_MSC_u_even[ib], sizeof(MYWORD) * HALF_MSC_VNRNBETAS);
    snprintf(nm, sizeof nm, "odd_%d", ib);
    dump(nm, h_MSC_u_odd[ib], sizeof(MYWORD) * HALF_MSC_VNRNBETAS);
    char *J6 = (char *)malloc(6 * V);
    memcpy(J6 + 0 * V, Jx[ib], V);
    memcpy(J6 + 1 * V, Jy[ib], V);
    memcpy(J6 + 2 * V, Jz[ib], V);
    memcpy(J6 + 3 * V, Jmx[ib], V);
    memcpy(J6 + 4 * V, Jmy[ib], V);
    memcpy(J6 + 5 * V, Jmz[ib], V);
    snprintf(nm, sizeof nm, "J6_%d", ib);
    dump(nm, J6, 6 * V);
    free(J6);
    MYWORD *Jw = (MYWORD *)malloc(sizeof(MYWORD) * 12 * HALF_MSC_V);
    for (int d = 0; d < DEGREE; d++)
      for (int h = 0; h < HALF_MSC_V; h++) {
        Jw[(0 * DEGREE + d) * HALF_MSC_V + h] = viciniB[h].J[ib][d];
        Jw[(1 * DEGREE + d) * HALF_MSC_V + h] = viciniN[h].J[ib][d];
      }
    snprintf(nm, sizeof nm, "Jw_%d", ib);
    dump(nm, Jw, sizeof(MYWORD) * 12 * HALF_MSC_V);
    free(Jw);
    snprintf(nm, sizeof nm, "couplings_%d", ib);
    char path[2 * MAXSTR];
    snprintf(path, sizeof path, "%scouplings.bin", dirsample[ib]);
    copy_file(path, nm);
  }
  {
    int *ng = (int *)malloc(sizeof(int) * 12 * HALF_MSC_V);
    unsigned char *rt = (unsigned char *)malloc(2 * HALF_MSC_V);
    for (int d = 0; d < DEGREE; d++)
      for (int h = 0; h < HALF_MSC_V; h++) {
        ng[(0 * DEGREE + d) * HALF_MSC_V + h] = viciniB[h].neig[d];
        ng[(1 * DEGREE + d) * HALF_MSC_V + h] = viciniN[h].neig[d];
      }
    for (int h = 0; h < HALF_MSC_V; h++) {
      rt[h] = bianchi_rotate[h];
      rt[HALF_MSC_V + h] = neri_rotate[h];
    }
    dump("neig", ng, sizeof(int) * 12 * HALF_MSC_V);
    dump("rotate", rt, 2 * HALF_MSC_V);
    dump("white", bianchi_index, sizeof(unsigned) * HALF_MSC_V);
    dump("black", neri_index, sizeof(unsigned) * HALF_MSC_V);
    free(ng);
    free(rt);
  }

  /* scalar energies of the initial lattices (med.c) */
  int *E0 = (int *)malloc(sizeof(int) * nbits * NR * NBETAS);
  for (int ib = 0; ib < nbits; ib++)
    for (int r = 0; r < NR; r++)
      for (int c = 0; c < NBETAS; c++)
        E0[(ib * NR + r) * NBETAS + c] = calculate_scalar_energy(ib, r, c);
  dump("escalar", E0, sizeof(int) * nbits * NR * NBETAS);

  /* _fill_bits over a sequence of half-sweep counters */
  {
    const int niter = 40;
    int iters[40];
    for (int k = 0; k < 32; k++) iters[k] = k;
    iters[32] = 32766;
    iters[33] = 32767;
    iters[34] = 32768;
    iters[35] = 65537;
    iters[36] = 1 << 20;
    iters[37] = (1 << 24) + 12345;
    iters[38] = 2147483000;
    iters[39] = 777;
    uint32_t *blocks = (uint32_t *)malloc(sizeof(uint32_t) * 4 * NR * niter);
    int ir;
    uint64_t entropy, temporal, parity;
    for (int k = 0; k < niter; k++) {
      int iter = iters[k];
      parity = (uint64_t)(k & 1);
      _fill_bits;
      for (ir = 0; ir < NR; ir++) {
        blocks[(k * NR + ir) * 4 + 0] = s_time_and_entropy.vec[ir].x;
        blocks[(k * NR + ir) * 4 + 1] = s_time_and_entropy.vec[ir].y;
        blocks[(k * NR + ir) * 4 + 2] = s_time_and_entropy.vec[ir].z;
        blocks[(k * NR + ir) * 4 + 3] = s_time_and_entropy.vec[ir].w;
      }
    }
    dump("fill_iters", iters, sizeof iters);
    dump("fill_blocks", blocks, sizeof(uint32_t) * 4 * NR * niter);
    free(blocks);
  }

  /* Parallel_Tempering with the scalar energies, perturbed every step */
  {
    const int nsteps = 25;
    h_Ener = (int *)malloc(sizeof(int) * nbits * NR * NBETAS);
    int *energies = (int *)malloc(sizeof(int) * nsteps * nbits * NR * NBETAS);
    uint8_t *perm = (uint8_t *)malloc(nsteps * nbits * NR * NBETAS);
    for (int st = 0; st < nsteps; st++) {
      for (int ib = 0; ib < nbits; ib++)
        for (int r = 0; r < NR; r++)
          for (int c = 0; c < NBETAS; c++) {
            int k = (ib * NR + r) * NBETAS + c;
            h_Ener[k] = E0[k] + 4 * ((st * 7 + c * 13 + r * 3 + ib) % 41 - 20);
          }
      memcpy(energies + st * nbits * NR * NBETAS, h_Ener,
             sizeof(int) * nbits * NR * NBETAS);
      for (int ib = 0; ib < nbits; ib++)
        for (int r = 0; r < NR; r++)
          Parallel_Tempering(&random_PT[ib][r], which_clon_this_beta[ib][r],
                             which_beta_this_clon[ib][r],
                             h_Ener + (r * NBETAS + ib * NRNBETAS),
                             aceptancePT[ib][r], attemptsPT[ib][r],
                             aceptancePTraw[ib][r]);
      for (int ib = 0; ib < nbits; ib++)
        for (int r = 0; r < NR; r++)
          memcpy(perm + ((st * nbits + ib) * NR + r) * NBETAS,
                 which_clon_this_beta[ib][r], NBETAS);
    }
    dump("pt_energies", energies, sizeof(int) * nsteps * nbits * NR * NBETAS);
    dump("pt_perm", perm, nsteps * nbits * NR * NBETAS);
    double *acc = (double *)malloc(sizeof(double) * 3 * nbits * NR * (NBETAS - 1));
    for (int ib = 0; ib < nbits; ib++)
      for (int r = 0; r < NR; r++)
        for (int b = 0; b < NBETAS - 1; b++) {
          int k = (ib * NR + r) * (NBETAS - 1) + b;
          acc[k] = aceptancePT[ib][r][b];
          acc[nbits * NR * (NBETAS - 1) + k] = aceptancePTraw[ib][r][b];
          acc[2 * nbits * NR * (NBETAS - 1) + k] = attemptsPT[ib][r][b];
        }
    dump("pt_acc", acc, sizeof(double) * 3 * nbits * NR * (NBETAS - 1));
    dump("pt_final", random_PT, sizeof(s_aleatorio_HQ_64bits) * NR * nbits);
    dump("wbtc", which_beta_this_clon, sizeof(which_beta_this_clon[0]) * nbits);
    free(acc);
    free(energies);
    free(perm);
    /* conf file of sample 0 / replica 0 and of the last sample / replica */
    for (int ib = 0; ib < nbits; ib++)
      for (int r = 0; r < NR; r++)
        for (int c = 0; c < NBETAS; c++)
          h_Ener[(ib * NR + r) * NBETAS + c] = E0[(ib * NR + r) * NBETAS + c];
    cuda_rng_offset = 123456789ULL;
    write_conf(7, nbits);
    char path[2 * MAXSTR];
    snprintf(path, sizeof path, "%s/conf", dir_rep[0][0]);
    copy_file(path, "conf_0_0");
    snprintf(path, sizeof path, "%s/conf", dir_rep[nbits - 1][NR - 1]);
    copy_file(path, "conf_last");
  }

  /* measurement files (io.c:307-397): evol.txt and aceptancePT.txt of two
   * bins, with the PT statistics of the steps above and energy averages made
   * up of the initial energies; history_b.dat of bins 0 and 1 and then of
   * bin 1 again, as a run resumed from the conf of bin 0 writes it after
   * the history of its lost bin 1 reached the file */
  {
    maxit_TRW = 3;
    frecTRW = 1;
    double *avg = (double *)malloc(sizeof(double) * 2 * nbits * NRNBETAS);
    for (int bin = 0; bin < 2; bin++) {
      for (int ib = 0; ib < nbits; ib++)
        for (int k = 0; k < NRNBETAS; k++) {
          const int e0 = E0[ib * NRNBETAS + k];
          average_Energy[ib][k] = (double)e0 + (double)((k * 5 + ib + bin * 3) % 11) / 7.0 -
                                  (bin ? 1e-3 * e0 : 0.0);
          avg[(bin * nbits + ib) * NRNBETAS + k] = average_Energy[ib][k];
        }
      write_measures(bin, nbits);
    }
    dump("meas_avg", avg, sizeof(double) * 2 * nbits * NRNBETAS);
    free(avg);
    unsigned char *rows =
        (unsigned char *)malloc((size_t)3 * nbits * NR * maxit_TRW * NBETAS);
    for (int call = 0; call < 3; call++) {
      const int ib_call = call == 0 ? 0 : 1;
      for (int ib = 0; ib < nbits; ib++)
        for (int r = 0; r < NR; r++)
          for (int i = 0; i < maxit_TRW; i++)
            for (int c = 0; c < NBETAS; c++) {
              const int b = (which_beta_this_clon[ib][r][c] + i + 2 * call + r) % NBETAS;
              history_betas[ib][r][i][c] = (unsigned char)b;
              rows[((((size_t)call * nbits + ib) * NR + r) * maxit_TRW + i) * NBETAS + c] =
                  (unsigned char)b;
            }
      write_history(ib_call, nbits);
    }
    dump("hist_rows", rows, (size_t)3 * nbits * NR * maxit_TRW * NBETAS);
    free(rows);
    char path[2 * MAXSTR];
    const int last[2][2] = {{0, 0}, {nbits - 1, NR - 1}};
    const char *tag[2] = {"0_0", "last"};
    for (int k = 0; k < 2; k++) {
      char nm[64];
      snprintf(path, sizeof path, "%sevol.txt", dir_rep[last[k][0]][last[k][1]]);
      snprintf(nm, sizeof nm, "evol_%s", tag[k]);
      copy_file(path, nm);
      snprintf(path, sizeof path, "%saceptancePT.txt", dir_rep[last[k][0]][last[k][1]]);
      snprintf(nm, sizeof nm, "aceptancePT_%s", tag[k]);
      copy_file(path, nm);
      snprintf(path, sizeof path, "%shistory_b.dat", dir_rep[last[k][0]][last[k][1]]);
      snprintf(nm, sizeof nm, "history_%s", tag[k]);
      copy_file(path, nm);
    }
  }

  /* cluster stream (random_c) draws and Lüscher-derived seed lists */
  {
    uint64_t draws[64];
    for (int k = 0; k < 64; k++) {
      _actualiza_aleatorio_HQ_escalar(random_c);
      draws[k] = random_c.final;
    }
    dump("cl_draws", draws, sizeof draws);
    uint64_t seeds[3] = {1ULL, 12345ULL, 0xDEADBEEFCAFEULL};
    uint64_t lists[3][MAXNSAMPLES];
    for (int k = 0; k < 3; k++) {
      uint64_t sd = seeds[k];
      generate_seeds_from_one(&sd, lists[k], MAXNSAMPLES);
    }
    dump("seed_lists", lists, sizeof lists);
  }
  return 0;
}
