// ea_tube.cpp -- the simulation program: a drop-in for the reference's
// PROGRAM/ binary (main.cu:10-283 with io.c and timing.c), driving the gfx950
// engine through the C ABI of include/eatube.h.
//
//   ea_tube [options] isample nbits beta.dat input.in LUT device [max_time [list_samples]]
//
// Arguments, input.in / betas / LUT / list formats, the output tree
// <dir>I%03d/BIT%03d/R%02d/, the conf checkpoints, evol.txt, aceptancePT.txt,
// history_b.dat, couplings.bin, seeds files, the running lock, changenbin.in,
// the signals and max_time behave as in the reference.  What the reference
// fixes at compile time (PROGRAM/Makefile:3-36) is read at run time: L, Lz
// and NBETAS from input.in (`l` = Lz<<16 | L, `nbetas`), and
//   --nr N          replicas (NR, header.h:14; default 8)
//   --obc           open boundaries along z (OBC; default PBC)
//   --weak, --strong  OBC variants (WEAK, STRONG)
//   --no-cluster    no Houdayer moves (NO_CLUSTER)
#include <csignal>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "eatube.h"

namespace {

constexpr int kMaxSamples = 32;  // MAXNSAMPLES (header.h:19)
constexpr int kMaxItTrw = 512;   // MAXIT_TRW (header.h:15)
constexpr int kMaxNbin = 4096;   // MAXNBIN (header.h:16)

// ------------------------------------------------------------------ state

struct Options {
  int nr = 8;
  int bc = EA_PBC;
  int variant = 0;  // 0 independent, 1 weak, 2 strong
  bool cluster = true;
};

Options opt;
ea_run_data data{};
std::string dir;                        // <output>I%03d/
std::vector<std::string> dirsample;     // dir + BIT%03d/
std::vector<std::vector<std::string>> dir_rep;  // dirsample + R%02d/
int nbits = 0, sample = 0, L = 0, Lz = 0, nt = 0, V = 0, n = 0;
std::vector<int> list_samples;
std::vector<double> betas;
std::vector<uint8_t> lut;
uint64_t seeds_J[kMaxSamples], seeds_u[kMaxSamples];
std::vector<uint64_t> seeds_MC;
int countJ0[kMaxSamples];
int write_seeds = 0;
int frecTRW = 1, maxit_TRW = 1;
char host_name[100];
int my_pid = 0;

// ----------------------------------------------------------------- errors

void print_data(FILE *f, const ea_run_data *d) {  // io.c:458-474
  fprintf(f, "nbin        %16d\n", d->nbin);
  fprintf(f, "itcut       %16d\n", d->itcut);
  fprintf(f, "itmax       %16d\n", d->itmax);
  fprintf(f, "houfr       %16d\n", d->houfr);
  fprintf(f, "mesfr       %16d\n", d->mesfr);
  fprintf(f, "ptfr        %16d\n", d->ptfr);
  fprintf(f, "flag        %16d\n", d->flag);
  fprintf(f, "L           %16d\n", d->l & 0xffff);
  fprintf(f, "M           %16d\n", d->l >> 16);
  fprintf(f, "nbetas      %16d\n", d->nbetas);
  fprintf(f, "Master seeds:\n");
  fprintf(f, "seed_J      %016llx\n", (unsigned long long)d->seed_J);
  fprintf(f, "seed_u      %016llx\n", (unsigned long long)d->seed_u);
  fprintf(f, "seed_MC     %016llx\n", (unsigned long long)d->seed_MC);
  fprintf(f, "seed_Cl     %016llx\n", (unsigned long long)d->seed_Cluster);
}

[[noreturn]] void print_and_exit(const char *fmt, ...) {  // io.c:1036-1044
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(1);
}

void delete_running() { remove((dir + "/running").c_str()); }

void create_error(const char *fmt, ...) {  // io.c:920-938
  FILE *f = fopen((dir + "/error").c_str(), "w");
  va_list ap;
  va_start(ap, fmt);
  if (f) {
    vfprintf(f, fmt, ap);
    print_data(f, &data);
    fclose(f);
  }
  va_end(ap);
  print_data(stderr, &data);
  delete_running();
}

void check(int rc, const char *what) {
  if (rc != EA_OK) print_and_exit("%s: %s\n", what, ea_last_error());
}

// ---------------------------------------------------------------- signals

// The reference's handler writes the error file and exits from inside the
// signal handler (io.c:906-918).  With HIP work in flight that can deadlock
// in the runtime's teardown, so here the handler only records the signal,
// and the main thread acts on it (stop_run) at the points where the engine
// is idle: after each bin's GPU work has been waited for, before its files
// are written, and between bins.  A signal that arrives while a bin runs
// thus ends the run when that bin's GPU work is done, with the same result
// as the reference: no files of that bin, the error file, no `running`,
// exit code 1.
volatile sig_atomic_t g_signal = 0;

void note_signal(int signum) { g_signal = signum; }

[[noreturn]] void stop_run(int signum) {  // io.c:906-918
  if (signum == SIGTERM) create_error("Received SIGTERM\n");
  if (signum == SIGINT) create_error("Received SIGINT\n");
  if (signum == SIGUSR1) create_error("Received SIGUSR1\n");
  print_and_exit("Aborting programma (SIGNAL=%d)\n", signum);
}

void check_signal() {
  if (g_signal) stop_run(g_signal);
}

void ignore_sigterm() {
  signal(SIGTERM, SIG_IGN);
  signal(SIGINT, SIG_IGN);
  signal(SIGUSR1, SIG_IGN);
}

void handle_sigterm() {
  signal(SIGTERM, note_signal);
  signal(SIGINT, note_signal);
  signal(SIGUSR1, note_signal);
}

// ------------------------------------------------------------ run control

void writelog(FILE *f, time_t t) {
  fprintf(f, "%s %5d %10ld %s", host_name, my_pid, (long)t, ctime(&t));
}

void create_running() {  // io.c:940-971
  gethostname(host_name, 99);
  my_pid = getpid();
  char name[2048];
  snprintf(name, sizeof name, "%srunning.%s.%d", dir.c_str(), host_name, my_pid);
  const std::string link_name = dir + "running";
  FILE *f = fopen(name, "w");
  if (!f) print_and_exit("Unable to open %s\n", name);
  writelog(f, time(nullptr));
  fclose(f);
  if (link(name, link_name.c_str()) == -1) {
    remove(name);
    print_and_exit("%s exists\n", link_name.c_str());
  }
  remove(name);
  const std::string logname = dir + "run.log";
  FILE *lg = fopen(logname.c_str(), "a");
  if (!lg) print_and_exit("Unable to open %s\n", logname.c_str());
  writelog(lg, time(nullptr));
  fclose(lg);
}

void renew_running() {  // io.c:974-1004
  const std::string name = dir + "/running";
  FILE *f = fopen(name.c_str(), "r+");
  if (!f) {
    create_error("File %s doesn't exist\n", name.c_str());
    print_and_exit("File %s doesn't exist\n", name.c_str());
  }
  char other[100] = {0};
  int pid = 0;
  if (fscanf(f, "%99s %d", other, &pid) != 2) pid = -1;
  if (strcmp(host_name, other)) {
    create_error("Host has changed from %s to %s\n", host_name, other);
    print_and_exit("Host has changed from %s to %s\n", host_name, other);
  }
  if (my_pid != pid) {
    create_error("PID has changed from %d to %d\n", my_pid, pid);
    print_and_exit("PID has changed from %d to %d\n", my_pid, pid);
  }
  rewind(f);
  writelog(f, time(nullptr));
  fclose(f);
}

void stop(int *newnbin) {  // io.c:875-895
  const std::string name = dir + "changenbin.in";
  FILE *f = fopen(name.c_str(), "r");
  if (!f) return;
  if (fscanf(f, "%d", newnbin) == EOF) {
    newnbin[0] = -1;
    printf(" File '%s' exists\n", name.c_str());
  }
  if (*newnbin > kMaxNbin) {
    create_error(" nbin must be smaller than %d\n", kMaxNbin);
    print_and_exit(" nbin must be smaller than %d\n", kMaxNbin);
  }
  fclose(f);
  printf("nbin changed to %d\n", newnbin[0]);
  fflush(stdout);
  remove(name.c_str());
}

// ------------------------------------------------------------------ timing

void measure_time() {  // timing.c:3-41
  static timespec clock_ini = {0, 0}, cpu_ini;
  static double totalspins;
  timespec clock_fin, cpu_fin;
  clock_gettime(CLOCK_MONOTONIC, &clock_fin);
  clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &cpu_fin);
  if (clock_ini.tv_sec > 0) {
    const double elapsed = (clock_fin.tv_sec - clock_ini.tv_sec) * 1e9 +
                           (clock_fin.tv_nsec - clock_ini.tv_nsec);
    const double consumed = (cpu_fin.tv_sec - cpu_ini.tv_sec) * 1e9 +
                            (cpu_fin.tv_nsec - cpu_ini.tv_nsec);
    printf("%4.0fs (%.3f ps/s, %.0f%%): ", elapsed * 1e-9, elapsed / totalspins * 1e3,
           consumed / elapsed * 100);
  } else {
    totalspins = (double)data.houfr * data.mesfr * data.ptfr * data.itmax * nt * (double)V *
                 opt.nr * nbits;
    printf("Starting simulation\n");
  }
  clock_ini = clock_fin;
  cpu_ini = cpu_fin;
}

// Stop when the next bin would not fit in max_time seconds: mean + 3 sigma
// of the bins so far (timing.c:43-74; with one bin so far sigma is 0).
int check_time(unsigned long long max_time) {
  static timespec clock_ini = {0, 0};
  static unsigned total = 0, elapsed2 = 0, count = 0;
  timespec clock_fin;
  clock_gettime(CLOCK_MONOTONIC, &clock_fin);
  if (clock_ini.tv_sec > 0 && max_time != 0) {
    const unsigned elapsed = (unsigned)(clock_fin.tv_sec - clock_ini.tv_sec);
    total += elapsed;
    elapsed2 += elapsed * elapsed;
    ++count;
    const double mean = (double)total / count;
    double sd = 0.0;
    if (count > 1) {
      const double m2 = (double)elapsed2 / count;
      sd = std::sqrt(std::max(0.0, count * (m2 - mean * mean) / (count - 1.0)));
    }
    const unsigned next = (unsigned)(mean + 3 * sd + 1);
    if ((double)max_time - total - next <= 0) return 1;
  }
  clock_ini = clock_fin;
  return 0;
}

// ---------------------------------------------------------------- inputs

void read_input(const char *name) {  // io.c:399-434 with check_data
  FILE *f = fopen(name, "r");
  if (!f) print_and_exit("Error: No file %s\n", name);
  char line[1025], path[1025] = {0};
  if (!fgets(line, sizeof line, f) || sscanf(line, "%1024s", path) != 1)
    print_and_exit("Error: No file %s\n", name);
  dir = path;
  int32_t *ints[9] = {&data.nbin, &data.itcut, &data.itmax, &data.houfr, &data.mesfr,
                      &data.ptfr, &data.nbetas, &data.flag, &data.l};
  for (auto *p : ints) {
    unsigned v = 0;
    if (fgets(line, sizeof line, f)) sscanf(line, "%u", &v);
    *p = (int32_t)v;
  }
  uint64_t *seeds[4] = {&data.seed_J, &data.seed_u, &data.seed_MC, &data.seed_Cluster};
  for (auto *p : seeds) {
    unsigned long long v = 0;
    if (fgets(line, sizeof line, f)) sscanf(line, "%llu", &v);
    *p = v;
  }
  fclose(f);
  if (data.nbin < 0 || data.itcut < 0 || data.itmax < 0 || data.houfr < 0 || data.mesfr < 0 ||
      data.ptfr < 0)
    print_and_exit("Bad data input\n");
  if (data.ptfr >= 1000)
    print_and_exit("The number(%d) of Metropolis per PT is too large\n", data.ptfr);
  L = data.l & 0xffff;
  Lz = data.l >> 16;
  nt = data.nbetas;
  if (L < 8 || Lz < 8 || L % 8 || Lz % 8)
    print_and_exit("l=%d: L=%d and Lz=%d must be multiples of 8\n", data.l, L, Lz);
  if (nt < 2 || nt > EA_MAX_BETAS) print_and_exit("nbetas=%d out of 2..%d\n", nt, EA_MAX_BETAS);
  V = L * L * Lz;
  n = V / 128;
  printf("Path: %s\n", dir.c_str());
  print_data(stdout, &data);
  fflush(stdout);
  for (frecTRW = 1; frecTRW <= data.itmax; frecTRW++)
    if (data.itmax % frecTRW == 0 && (maxit_TRW = data.itmax / frecTRW) <= kMaxItTrw) break;
}

void read_betas(const char *name) {  // io.c:476-501
  FILE *f = fopen(name, "rt");
  if (!f) print_and_exit("File  %s doesn't exist.\n", name);
  betas.assign(nt, 0.0);
  for (int b = 0; b < nt; ++b)
    if (fscanf(f, "%lf ", &betas[b]) == EOF)
      print_and_exit("File %s only has %d lines (it needs %d)\n", name, b, nt);
  double dummy;
  if (fscanf(f, "%lf", &dummy) != EOF) fprintf(stderr, "More lines in file %s\n", name);
  for (int b = 0; b + 1 < nt; ++b)
    if (betas[b] < betas[b + 1])
      print_and_exit("Error in file %s: betas don't sort from largest to smallest!\n", name);
  fclose(f);
}

void read_lut(const char *name, const char *name_betas) {  // io.c:504-534
  FILE *f = fopen(name, "rb");
  if (!f) print_and_exit("Problems opening %s\n", name);
  std::vector<double> rd(nt, 0.0);
  if (fread(rd.data(), sizeof(double) * nt, 1, f) != 1) {
    printf("betas_read[0]=%.14g\n", rd[0]);
    printf("betas_read[1]=%.14g\n", rd[1]);
    print_and_exit("Problems reading betas\n");
  }
  for (int b = 0; b < nt; ++b)
    if (std::fabs(betas[b] - rd[b]) > 1e-14)
      print_and_exit("betas[%d]=%.14g from %s, betas[%d]=%.14g from %s (difference=%.14g)\n", b,
                     betas[b], name_betas, b, rd[b], name, betas[b] - rd[b]);
  lut.assign((size_t)nt * EA_LUT_BYTES, 0);
  if (fread(lut.data(), lut.size(), 1, f) != 1) print_and_exit("Problems reading lut_heat_bath\n");
  fclose(f);
}

void read_list_samples(const char *name) {  // io.c:537-567
  list_samples.assign(nbits, 0);
  if (!name[0]) {
    for (int i = 0; i < nbits; ++i) list_samples[i] = i;
  } else {
    FILE *f = fopen(name, "rt");
    if (!f) print_and_exit("File  %s doesn't exist.\n", name);
    for (int i = 0; i < nbits; ++i)
      if (fscanf(f, "%d ", &list_samples[i]) == EOF)
        print_and_exit("File %s only has %d lines (it needs %d)\n", name, i, nbits);
    int dummy;
    if (fscanf(f, "%d", &dummy) != EOF) fprintf(stderr, "More lines in file %s\n", name);
    fclose(f);
  }
  for (int i = 0; i + 1 < nbits; ++i)
    if (list_samples[i] >= list_samples[i + 1]) {
      create_error("List of samples %s is not sorted.\n", name);
      print_and_exit("List of samples must be sorted!\n");
    }
  for (int i = 0; i < nbits; ++i)
    if (list_samples[i] < 0 || list_samples[i] >= kMaxSamples)
      print_and_exit("Sample %d of the list is out of 0..%d\n", list_samples[i], kMaxSamples - 1);
}

void create_sample_path() {  // io.c:569-589
  char buf[64];
  snprintf(buf, sizeof buf, "I%03d/", sample);
  dir += buf;
  mkdir(dir.c_str(), 0775);
  dirsample.assign(nbits, "");
  dir_rep.assign(nbits, std::vector<std::string>(opt.nr));
  for (int i = 0; i < nbits; ++i) {
    snprintf(buf, sizeof buf, "BIT%03d/", list_samples[i]);
    dirsample[i] = dir + buf;
    mkdir(dirsample[i].c_str(), 0775);
    for (int r = 0; r < opt.nr; ++r) {
      snprintf(buf, sizeof buf, "R%02d/", r);
      dir_rep[i][r] = dirsample[i] + buf;
      mkdir(dir_rep[i][r].c_str(), 0775);
    }
  }
  printf("Dirsample = %s\n", dir.c_str());
  fflush(stdout);
}

// ----------------------------------------------------------------- seeds

uint64_t urandom64() {  // lee_urandom (random_generator.cu:18-26)
  uint64_t v = 0;
  FILE *f = fopen("/dev/urandom", "r");
  if (!f || fread(&v, sizeof v, 1, f) != 1) print_and_exit("Problems reading /dev/urandom\n");
  fclose(f);
  return v;
}

uint64_t check_seed(uint64_t s) { return s ? s : urandom64(); }  // comprueba_semilla

void generate_seed_vectors() {  // ini.c:32-59
  uint64_t tmp[kMaxSamples];
  data.seed_J = check_seed(data.seed_J);
  ea_seeds_from_one(data.seed_J, tmp, kMaxSamples);
  for (int i = 0; i < nbits; ++i) seeds_J[i] = tmp[list_samples[i]];
  data.seed_u = check_seed(data.seed_u);
  ea_seeds_from_one(data.seed_u, tmp, kMaxSamples);
  for (int i = 0; i < nbits; ++i) seeds_u[i] = tmp[list_samples[i]];
  data.seed_MC = check_seed(data.seed_MC);
  seeds_MC.assign(opt.nr, 0);
  ea_seeds_from_one(data.seed_MC, seeds_MC.data(), opt.nr);
}

int get_seeds() {  // io.c:655-776
  int no_seeds = 0, no_master = 0;
  const std::string mname = dir + "/master_seeds.in";
  FILE *f = fopen(mname.c_str(), "r");
  if (!f) {
    if (data.flag == 1) {
      ++no_master;
    } else {
      sleep(1);
      if (!(f = fopen(mname.c_str(), "r"))) {
        create_error("There is not file %s, and it must exist\n", mname.c_str());
        print_and_exit("Problems reading seeds\n");
      }
    }
  }
  if (!no_master) {
    unsigned long long sJ, su, sMC, sC;
    if (fscanf(f, "%llu %llu %llu %llu", &sJ, &su, &sMC, &sC) == EOF)
      print_and_exit("Nothing in file %s\n", mname.c_str());
    fclose(f);
    auto adopt = [&](uint64_t &mine, unsigned long long read, const char *what) {
      if (mine > 0 && mine != read) {
        create_error("Master %s in input file and %s differ\n", what, mname.c_str());
        print_and_exit("Problems reading master seedJ\n");
      }
      mine = read;
    };
    adopt(data.seed_J, sJ, "seedJ");
    adopt(data.seed_u, su, "seedu");
    adopt(data.seed_MC, sMC, "seedMC");
    adopt(data.seed_Cluster, sC, "seedCluster");
  }
  const bool u_gone = data.flag == 0 && data.seed_u == 0;
  const bool mc_gone = data.flag == 0 && data.seed_MC == 0;
  generate_seed_vectors();
  if (u_gone) data.seed_u = 0;
  if (mc_gone) data.seed_MC = 0;
  memset(countJ0, 0, sizeof countJ0);
  for (int i = 0; i < nbits; ++i) {
    const std::string name = dirsample[i] + "/seeds.in";
    FILE *s = fopen(name.c_str(), "r");
    if (!s) {
      if (data.flag == 1) {
        ++no_seeds;
        continue;
      }
      sleep(1);
      if (!(s = fopen(name.c_str(), "r"))) {
        create_error("There is not file %s, and it must exist\n", name.c_str());
        print_and_exit("Problems reading seeds\n");
      }
    }
    unsigned long long sJ = 0, su = 0, sC = 0;
    std::vector<unsigned long long> sMC(opt.nr, 0);
    int nf = 0;
    nf += fscanf(s, "%llu %llu", &sJ, &su);
    for (int r = 0; r < opt.nr; ++r) nf += fscanf(s, " %llu", &sMC[r]);
    nf += fscanf(s, " %llu", &sC);
    nf += fscanf(s, " %d", &countJ0[i]);
    if (nf != opt.nr + 4) {
      create_error("File %s must have %d fields\n", name.c_str(), opt.nr + 4);
      print_and_exit("Problems reading seeds from %s\n", name.c_str());
    }
    fclose(s);
    if (seeds_J[i] != sJ) {
      create_error("Seed J generated and %s differ (bit=%d)\n", name.c_str(), list_samples[i]);
      print_and_exit("Problems reading seeds\n");
    }
    if (seeds_u[i] != su && su != 0 && !u_gone) {
      create_error("Seed u generated and %s differ (bit=%d)\n", name.c_str(), list_samples[i]);
      print_and_exit("Problems reading seeds\n");
    }
    seeds_u[i] = su;
    for (int r = 0; r < opt.nr; ++r) {
      if (seeds_MC[r] != sMC[r] && sMC[r] != 0 && !mc_gone) {
        create_error("Seed MC generated input and %s differ\n", name.c_str());
        print_and_exit("Problems reading seeds\n");
      }
      seeds_MC[r] = sMC[r];
    }
  }
  if (no_seeds == nbits || no_master == 1) return 1;
  if (no_seeds > 0) print_and_exit("Some samples have file seeds.in, others not have\n");
  return 0;
}

void check_and_prepare_simulation_backup() {  // io.c:597-653
  int there_is_conf = 0, checked = 0, itcut = 0;
  std::string last;
  for (int i = 0; i < nbits; ++i) {
    there_is_conf = 0;
    for (int r = 0; r < opt.nr; ++r) {
      last = dir_rep[i][r] + "/conf";
      FILE *f = fopen(last.c_str(), "r");
      if (!f) continue;
      ea_run_data dc{};
      if (fread(&dc, sizeof dc, 1, f) != 1) dc.itcut = -1;
      if (checked > 0 && dc.itcut != itcut) {
        create_error("Different number of iterations in samples from %s\n", dir.c_str());
        print_and_exit("Problems with conf files\n");
      }
      itcut = dc.itcut;
      ++there_is_conf;
      ++checked;
      fclose(f);
    }
    if (there_is_conf > 0 && there_is_conf != opt.nr) {
      create_error("There are less conf files than it must be in %s (%d/%d)\n",
                   dirsample[i].c_str(), there_is_conf, opt.nr);
      print_and_exit("Problems with conf files\n");
    }
  }
  if (data.flag == 2) data.flag = there_is_conf ? 0 : 1;
  if (data.flag == 1) {
    if (there_is_conf) {
      create_error("There are some confs in %s.\nIf you want to make a new run, please delete "
                   "the configuration manually\n", dir.c_str());
      print_and_exit("There are confs\n");
    }
  } else if (data.flag == 0) {
    if (!there_is_conf) {
      create_error("There is not %s, it is impossible to do a backup\n", last.c_str());
      print_and_exit("Impossible to do a backup\n");
    }
  } else {
    print_and_exit("data.flag is not recognize\n");
  }
  write_seeds = get_seeds();
}

void write_seeds_in_file() {  // io.c:778-822
  ignore_sigterm();
  const std::string m = dir + "/master_seeds.in";
  FILE *f = fopen(m.c_str(), "w");
  if (!f) {
    create_error("Impossible to open seed file to write\n");
    print_and_exit("Problems writing seeds\n");
  }
  fprintf(f, "%llu %llu %llu %llu\n", (unsigned long long)data.seed_J,
          (unsigned long long)data.seed_u, (unsigned long long)data.seed_MC,
          (unsigned long long)data.seed_Cluster);
  fclose(f);
  handle_sigterm();
  for (int i = 0; i < nbits; ++i) {
    ignore_sigterm();
    const std::string name = dirsample[i] + "/seeds.in";
    FILE *s = fopen(name.c_str(), "w");
    if (!s) {
      create_error("Impossible to open seed file to write\n");
      print_and_exit("Problems writing seeds\n");
    }
    fprintf(s, "%llu %llu", (unsigned long long)seeds_J[i], (unsigned long long)seeds_u[i]);
    for (int r = 0; r < opt.nr; ++r) fprintf(s, " %llu", (unsigned long long)seeds_MC[r]);
    fprintf(s, " %llu", (unsigned long long)data.seed_Cluster);
    fprintf(s, " %d\n", countJ0[i]);
    fclose(s);
    handle_sigterm();
  }
}

// ------------------------------------------------------------ simulation

struct Sim {
  ea_engine *e = nullptr;
  std::vector<ea_hq_state> prc, pt;  // pt[sample][replica]
  std::vector<ea_xo_state> xo;
  std::vector<uint32_t> keys;
  ea_hq_state cl{};
  uint64_t coin_offset = 0;
  std::vector<std::vector<uint64_t>> even, odd, J;
  std::vector<uint8_t> wctb;
  std::vector<double> acc, att, raw;
};

Sim S;

void init_all() {  // Init (ini.c:8-30) with Init_Random (ini.c:61-101)
  const int nr = opt.nr;
  for (int r = 0; r < nr; ++r) seeds_MC[r] = check_seed(seeds_MC[r]);
  for (int i = 0; i < nbits; ++i) {
    seeds_J[i] = check_seed(seeds_J[i]);
    seeds_u[i] = check_seed(seeds_u[i]);
  }
  data.seed_Cluster = check_seed(data.seed_Cluster);
  ea_hq_init(&S.cl, data.seed_Cluster);
  S.prc.resize(nr);
  S.xo.resize(nr);
  S.keys.resize(nr);
  for (int r = 0; r < nr; ++r) ea_init_generators(seeds_MC[r], &S.prc[r], &S.xo[r], &S.keys[r]);
  S.pt.resize((size_t)nbits * nr);
  for (int r = 0; r < nr; ++r) {
    uint64_t tmp[kMaxSamples];
    ea_seeds_from_one(seeds_MC[r], tmp, kMaxSamples);
    for (int i = 0; i < nbits; ++i) {
      const int seed_pt = (int)(uint32_t)tmp[list_samples[i]];  // `int seed_PT` (ini.c:63,90)
      ea_hq_init(&S.pt[(size_t)i * nr + r], (uint64_t)(int64_t)seed_pt);
    }
  }
  S.coin_offset = 0;
  const size_t half = (size_t)n * nr * nt;
  S.even.assign(nbits, std::vector<uint64_t>(half));
  S.odd.assign(nbits, std::vector<uint64_t>(half));
  S.J.assign(nbits, std::vector<uint64_t>((size_t)12 * n));
  std::vector<uint8_t> couplings((size_t)6 * V);
  ea_config cfg{L, Lz, nr, nt, nbits, opt.bc, 0, (uint32_t)sample, betas.data(), lut.data()};
  cfg.clone_threshold = (opt.cluster && nr >= 4) ? (int)std::floor(nt * 0.5 + 0.5) : 0;
  check(ea_create(&cfg, &S.e), "ea_create");
  check(ea_set_keys(S.e, S.keys.data()), "ea_set_keys");
  for (int i = 0; i < nbits; ++i) {
    const int countJ = ea_init_sample(L, Lz, nr, nt, opt.bc, opt.variant, 0, seeds_J[i],
                                      seeds_u[i], S.even[i].data(), S.odd[i].data(),
                                      S.J[i].data(), couplings.data());
    if (countJ < 0) print_and_exit("Problems generating J's\n");
    if (data.flag == 0 && countJ != countJ0[i]) {  // ini.c:538-549
      create_error("countJ and countJ0 differ (%d != %d --- ibit=%d)\n", countJ, countJ0[i],
                   list_samples[i]);
      print_and_exit("Problems generating J's\n");
    }
    if (countJ0[i] != 0 && countJ0[i] != countJ) {
      create_error("CountJ and countJ0 differ\n");
      print_and_exit("Problems generating J's\n");
    }
    countJ0[i] = countJ;
    const std::string cname = dirsample[i] + "/couplings.bin";  // write_couplings, io.c:284-305
    FILE *f = fopen(cname.c_str(), "wb");
    if (!f) print_and_exit("Problems opening %s\n", cname.c_str());
    fwrite(couplings.data(), 1, couplings.size(), f);
    fclose(f);
    check(ea_upload_sample(S.e, i, S.even[i].data(), S.odd[i].data(), S.J[i].data()),
          "ea_upload_sample");
  }
  S.wctb.assign((size_t)nbits * nr * nt, 0);
  for (size_t c = 0; c < (size_t)nbits * nr; ++c)
    for (int b = 0; b < nt; ++b) S.wctb[c * nt + b] = (uint8_t)b;
  S.acc.assign((size_t)nbits * nr * (nt - 1), 0.0);
  S.att = S.acc;
  S.raw = S.acc;
}

std::vector<uint8_t> slurp(const std::string &name) {
  std::vector<uint8_t> b;
  FILE *f = fopen(name.c_str(), "rb");
  if (!f) return b;
  uint8_t buf[1 << 16];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + k);
  fclose(f);
  return b;
}

void read_conf() {  // io.c:106-237
  const int nr = opt.nr;
  const size_t want = ea_conf_bytes(L, Lz, nt);
  std::vector<std::vector<int32_t>> ener(nbits, std::vector<int32_t>((size_t)nr * nt));
  for (int i = 0; i < nbits; ++i)
    for (int r = 0; r < nr; ++r) {
      const std::string name = dir_rep[i][r] + "/conf";
      const std::vector<uint8_t> b = slurp(name);
      if (b.size() != want) {
        create_error("There are only %d bytes in %s (must be %d)\n", (int)b.size(), name.c_str(),
                     (int)want);
        print_and_exit("Problems reading confs\n");
      }
      ea_run_data dc{};
      std::vector<double> bc(nt);
      std::vector<int32_t> eb(nt);
      const int rc = ea_conf_decode(b.data(), b.size(), L, Lz, nr, nt, &dc, bc.data(),
                                    &S.pt[(size_t)i * nr + r], &S.prc[r], &S.cl, &S.xo[r],
                                    &S.coin_offset, &S.wctb[((size_t)i * nr + r) * nt],
                                    S.even[i].data(), S.odd[i].data(), r, eb.data());
      if (rc) {
        create_error("There is an error in the permutatuons in %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      if (dc.nbetas != nt || dc.l != data.l || dc.itmax != data.itmax ||
          dc.houfr != data.houfr || dc.mesfr != data.mesfr || dc.ptfr != data.ptfr) {
        create_error("Different data from input and %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      if (seeds_J[i] != dc.seed_J) {
        create_error("Different seed_J generated and read in %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      if (seeds_u[i] != dc.seed_u) {
        create_error("Different seed_u generated and read in %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      if (data.seed_Cluster != dc.seed_Cluster) {
        create_error("Different data.seed_Cluster and read in %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      if (seeds_MC[r] != dc.seed_MC) {
        create_error("Different seed_MC generated and read in %s\n", name.c_str());
        print_and_exit("Problems reading confs\n");
      }
      data.itcut = dc.itcut;
      stop(&data.nbin);
      for (int b2 = 0; b2 < nt; ++b2)
        if (std::fabs(bc[b2] - betas[b2]) > 1e-12) {
          create_error("Different beta (ibeta=%d, %.12g != %.12g)\n", b2, bc[b2], betas[b2]);
          print_and_exit("Problems reading confs\n");
        }
      for (int b2 = 0; b2 < nt; ++b2)
        ener[i][(size_t)r * nt + S.wctb[((size_t)i * nr + r) * nt + b2]] = eb[b2];
    }
  // upload, then check the stored energies against the spins' (the scalar
  // energy check of read_conf, io.c:213-224: the MSC energy equals it)
  for (int i = 0; i < nbits; ++i)
    check(ea_upload_sample(S.e, i, S.even[i].data(), S.odd[i].data(), S.J[i].data()),
          "ea_upload_sample");
  check(ea_upload_perms(S.e, S.wctb.data()), "ea_upload_perms");
  std::vector<int32_t> now((size_t)nbits * nr * nt);
  check(ea_energy(S.e, now.data()), "ea_energy");
  for (int i = 0; i < nbits; ++i)
    for (int r = 0; r < nr; ++r)
      for (int b = 0; b < nt; ++b) {
        const int c = S.wctb[((size_t)i * nr + r) * nt + b];
        const int32_t got = now[((size_t)i * nr + r) * nt + c], want = ener[i][(size_t)r * nt + c];
        if (got != want) {
          create_error("Energy read and calculate differ in ibit=%d ir=%d ibeta=%d:escalar=%d, "
                       "read=%d\n", i, r, b, got, want);
          print_and_exit("Problems reading confs\n");
        }
      }
}

void write_history(int ib, const std::vector<uint8_t> &hist) {  // io.c:353-386
  const int nr = opt.nr;
  const size_t predicted = (size_t)ib * maxit_TRW * nt;
  for (int i = 0; i < nbits; ++i)
    for (int r = 0; r < nr; ++r) {
      const std::string name = dir_rep[i][r] + "history_b.dat";
      FILE *f = fopen(name.c_str(), "a");
      if (!f) {
        create_error("Problems opening history_b.dat\n");
        print_and_exit("Problems writing history\n");
      }
      const long actual = ftell(f);
      if ((size_t)actual < predicted) {
        create_error("There are less data than expected in history_b.dat\n");
        print_and_exit("Problems writing history\n");
      }
      // A longer file (rows of a bin that a resumed run repeats) keeps its
      // bytes: the reference seeks to the predicted size, but a stream
      // opened for appending writes at the end wherever it is positioned, so
      // its rows follow the stale ones; the same bytes are written here.
      for (int k = 0; k < maxit_TRW; ++k)
        fwrite(&hist[(((size_t)k * nbits + i) * nr + r) * nt], 1, (size_t)nt, f);
      fclose(f);
    }
}

void write_measures(int ib, const std::vector<double> &avg) {  // io.c:307-351
  const int nr = opt.nr;
  printf("Block = %d\n", ib);
  for (int i = 0; i < nbits; ++i) {
    printf("Sample = %d\n", list_samples[i]);
    for (int r = 0; r < nr; ++r) {
      printf("\t ir = %d: ", r);
      for (int b = 0; b < nt; ++b) printf("E[%d]=%10.8g ", b, avg[((size_t)i * nr + r) * nt + b]);
      printf("\n");
    }
    fflush(stdout);
    for (int r = 0; r < nr; ++r) {
      std::string name = dir_rep[i][r] + "evol.txt";
      FILE *f = fopen(name.c_str(), "a");
      if (!f) {
        create_error("Problems opening evol.txt file\n");
        print_and_exit("Problems writing measures\n");
      }
      fprintf(f, "%d ", ib);
      for (int b = 0; b < nt; ++b) fprintf(f, "%14.10g ", avg[((size_t)i * nr + r) * nt + b]);
      fprintf(f, "\n");
      fclose(f);
      name = dir_rep[i][r] + "aceptancePT.txt";
      if (!(f = fopen(name.c_str(), "a"))) {
        create_error("Problems opening aceptance.txt file\n");
        print_and_exit("Problems writing measures\n");
      }
      fprintf(f, "#%d\n", ib);
      for (int b = 0; b + 1 < nt; ++b) {
        const size_t k = ((size_t)i * nr + r) * (nt - 1) + b;
        fprintf(f, "%g %g %g %g\n", betas[b], S.acc[k] / S.att[k], S.raw[k] / S.att[k], S.att[k]);
      }
      fclose(f);
    }
  }
}

void write_conf(int ib) {  // io.c:17-104
  const int nr = opt.nr;
  const size_t size = ea_conf_bytes(L, Lz, nt);
  std::vector<uint8_t> buf(size);
  std::vector<int32_t> ener((size_t)nbits * nr * nt);
  check(ea_energy(S.e, ener.data()), "ea_energy");
  for (int i = 0; i < nbits; ++i)
    check(ea_download_spins(S.e, i, S.even[i].data(), S.odd[i].data()), "ea_download_spins");
  bool error = false;
  for (int i = 0; i < nbits && !error; ++i) {
    ea_run_data dc = data;
    dc.seed_J = seeds_J[i];
    dc.seed_u = seeds_u[i];
    for (int r = 0; r < nr; ++r) {
      dc.seed_MC = seeds_MC[r];
      const size_t ch = (size_t)i * nr + r;
      check(ea_conf_encode(&dc, L, Lz, nr, nt, betas.data(), &S.pt[ch], &S.prc[r], &S.cl,
                           &S.xo[r], S.coin_offset, &S.wctb[ch * nt], S.even[i].data(),
                           S.odd[i].data(), r, &ener[ch * nt], buf.data()),
            "ea_conf_encode");
      char tname[64];
      snprintf(tname, sizeof tname, "/conf.%06d", ib);
      const std::string name_t = dir_rep[i][r] + tname, name = dir_rep[i][r] + "/conf";
      FILE *f = fopen(name_t.c_str(), "wb");
      if (!f) {
        create_error("Problems opening %s\n", name.c_str());
        error = true;
        break;
      }
      if (fwrite(buf.data(), 1, size, f) != size) {
        fclose(f);
        create_error("Problems writing conf in %s\n", name.c_str());
        error = true;
        break;
      }
      fclose(f);
      remove(name.c_str());
      if (link(name_t.c_str(), name.c_str())) {
        create_error("Problems in hard-lin to %s\n", name_t.c_str());
        error = true;
        break;
      }
      struct stat st;
      stat(name.c_str(), &st);
      if (st.st_nlink != 2) {
        create_error("The number of hard-link %s is not 2\n", name.c_str());
        error = true;
        break;
      }
      if ((size_t)st.st_size != size) {
        create_error("There are only %d bytes in %s (must be %d)\n", (int)st.st_size,
                     name.c_str(), (int)size);
        error = true;
        break;
      }
    }
  }
  if (error) print_and_exit("Problems writing confs\n");
}

void usage(const char *prog) {
  printf("Usage: %s [--nr N] [--obc [--weak|--strong]] [--no-cluster] isample nbits beta.dat "
         "input.in LUT device [max_time [list_samples]]\n", prog);
  printf("\nArguments:\n");
  printf("  isample              Sample index (int)\n");
  printf("  nbits                Number of subsample (int)\n");
  printf("  beta.dat             Path to beta.dat file (string)\n");
  printf("  input.in             Path to input.in file (string)\n");
  printf("  LUT                  Path to LUT file (string)\n");
  printf("  device               Device number (int)\n");
  printf("  max_time (opt.)      Maximum time allowed (unsigned long long)\n");
  printf("                       A 0 value disable this option.\n");
  printf("  list_samples (opt.)  oPath to the list_of_samples file (string)\n");
  printf("\nOptions (compile-time defines of the reference build):\n");
  printf("  --nr N               replicas, a power of two in 2..16 (NR, default 8)\n");
  printf("  --obc                open boundaries along z (OBC; default PBC)\n");
  printf("  --weak, --strong     OBC boundary couplings (WEAK, STRONG)\n");
  printf("  --no-cluster         no Houdayer cluster moves (NO_CLUSTER)\n");
  printf("\nUse -h to display this help message.\n");
  exit(EXIT_SUCCESS);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && (!strcmp(argv[1], "-h") || !strcmp(argv[1], "--help"))) usage(argv[0]);
  std::vector<char *> pos;
  for (int k = 1; k < argc; ++k) {
    if (!strcmp(argv[k], "--nr") && k + 1 < argc) {
      opt.nr = atoi(argv[++k]);
    } else if (!strcmp(argv[k], "--obc")) {
      opt.bc = EA_OBC;
    } else if (!strcmp(argv[k], "--weak")) {
      opt.variant = 1;
    } else if (!strcmp(argv[k], "--strong")) {
      opt.variant = 2;
    } else if (!strcmp(argv[k], "--no-cluster")) {
      opt.cluster = false;
    } else {
      pos.push_back(argv[k]);
    }
  }
  if (opt.nr < 2 || opt.nr > EA_MAX_REPLICAS || (opt.nr & (opt.nr - 1)))
    print_and_exit("--nr %d: a power of two in 2..%d\n", opt.nr, EA_MAX_REPLICAS);
  if (opt.variant && opt.bc != EA_OBC) print_and_exit("--weak and --strong need --obc\n");
  unsigned long long max_time = 0;
  int device = 0;
  char name_list[1024] = {0};
  check_time(max_time);
  if (pos.size() < 6 || pos.size() > 8)
    print_and_exit("Usage: %s isample nbits beta.dat input.in LUT device [max_time "
                   "[list_samples]]\n", argv[0]);
  if (pos.size() == 8) snprintf(name_list, sizeof name_list, "%s", pos[7]);
  if (pos.size() >= 7) sscanf(pos[6], "%llu", &max_time);
  sscanf(pos[5], "%d", &device);
  sscanf(pos[1], "%d", &nbits);
  sscanf(pos[0], "%d", &sample);
  const char *name_betas = pos[2], *name_input = pos[3], *name_lut = pos[4];

  if (nbits < 1 || nbits > EA_MAX_SAMPLES)
    print_and_exit("Error: nbits out of range, 1 <= nbits <= %d\n", EA_MAX_SAMPLES);
  if (sample < 0) print_and_exit("Error: sample out of range, 0 <= sample\n");
  if ((uint32_t)sample > (1u << 14) - 1)
    print_and_exit("Our Counter-based PRNG  cannot accomodate more than %d samples\n", 1 << 14);
  int count = 0;
  ea_device_count(&count);
  if (device < 0 || device >= count) print_and_exit("device=%d count_devices=%d\n", device, count);
  check(ea_select_device(device), "ea_select_device");

  read_input(name_input);
  read_betas(name_betas);
  read_lut(name_lut, name_betas);
  read_list_samples(name_list);
  create_sample_path();
  check_and_prepare_simulation_backup();
  printf("Betas and LUT have been read correctly.\n");
  printf("Initializating J's and Spins ...\n");
  fflush(stdout);
  handle_sigterm();
  create_running();

  init_all();
  check_signal();
  if (data.flag == 0) {  // backup (io.c:824-838)
    read_conf();
    print_data(stdout, &data);
    fflush(stdout);
    if (data.itcut >= data.nbin) {
      delete_running();
      create_error("There are %d iterations in %s\n", data.itcut, dir.c_str());
      print_and_exit("Problems in backup\n");
    }
  }
  if (write_seeds) write_seeds_in_file();
  stop(&data.nbin);

  check(ea_upload_perms(S.e, S.wctb.data()), "ea_upload_perms");
  check(ea_upload_pt_state(S.e, S.pt.data(), S.acc.data(), S.att.data(), S.raw.data()),
        "ea_upload_pt_state");

  const int nr = opt.nr;
  const bool cluster = opt.cluster && nr >= 4;
  // the reference restarts the half-sweep counter without the houfr factor
  // (main.cu:140); kept so that resumed runs match its streams
  int iter = 2 * data.itcut * data.itmax * data.mesfr * data.ptfr;
  const int halves = 2 * data.itmax * data.mesfr * data.houfr * data.ptfr;
  std::vector<ea_counter> blocks((size_t)halves * nr);
  std::vector<int32_t> pairs((size_t)std::max(1, data.itmax * data.mesfr) * 4);
  std::vector<double> esum((size_t)nbits * nr * nt), avg(esum.size());
  std::vector<uint8_t> hist((size_t)maxit_TRW * nbits * nr * nt);
  measure_time();
  int ibin;
  for (ibin = data.itcut; ibin < data.nbin; ibin++) {
    check_signal();
    renew_running();
    ea_fill_counters(S.prc.data(), S.xo.data(), nr, iter, halves, (uint32_t)sample,
                     blocks.data());
    iter += halves;
    if (cluster)
      for (int k = 0; k < data.itmax * data.mesfr; ++k)
        check(ea_houdayer_pairs(&S.cl, nr, &pairs[(size_t)k * 4]), "ea_houdayer_pairs");
    ea_schedule sch{data.itmax, data.mesfr, data.houfr, data.ptfr, cluster ? 1 : 0, 0, frecTRW,
                    blocks.data(), pairs.data(), data.seed_Cluster, S.coin_offset};
    uint64_t off = S.coin_offset;
    check(ea_run(S.e, &sch, &off), "ea_run");
    S.coin_offset = off;
    check(ea_read_measures(S.e, esum.data(), hist.data(), maxit_TRW), "ea_read_measures");
    check(ea_download_pt_state(S.e, S.pt.data(), S.acc.data(), S.att.data(), S.raw.data()),
          "ea_download_pt_state");
    check(ea_download_perms(S.e, S.wctb.data()), "ea_download_perms");
    check_signal();  // the bin's GPU work is done
    measure_time();
    ignore_sigterm();
    write_history(ibin, hist);
    for (size_t k = 0; k < esum.size(); ++k) avg[k] = esum[k] / (double)data.itmax;
    write_measures(ibin, avg);
    write_conf(++data.itcut);
    stop(&data.nbin);
    handle_sigterm();
    if (check_time(max_time)) break;
  }
  check_signal();
  ignore_sigterm();
  delete_running();
  printf("Simulation I=%d, nbits=%d finished at iteration %d\n", sample, nbits, ibin);
  fflush(stdout);
  ea_destroy(S.e);
  return 0;
}
