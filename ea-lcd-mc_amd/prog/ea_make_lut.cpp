// ea_make_lut.cpp -- the LUT file of a beta set, as the reference's
// create_LUT (PROGRAM/input/create_LUT.c:44-417) writes it:
//   ea_make_lut name_betas [NUMBETAS]
// reads NUMBETAS betas (default: every value in the file; sorted from largest
// to smallest) and writes LUT_for_PRNG_nbits12_NB<NUMBETAS>.bin in the current
// directory: the betas as doubles, then one 10240-byte table per beta
// (prepara_LUT, quad precision on the host: ea_make_lut).
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "eatube.h"

[[noreturn]] static void die(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(1);
}

int main(int argc, char **argv) {
  if (argc != 2 && argc != 3) die("Usage: %s name_betas [NUMBETAS]\n", argv[0]);
  FILE *f = fopen(argv[1], "rt");
  if (!f) die(" No existe el fichero %s.\n", argv[1]);
  std::vector<double> betas;
  const int want = argc == 3 ? atoi(argv[2]) : -1;
  double b;
  while ((want < 0 || (int)betas.size() < want) && fscanf(f, "%lf ", &b) == 1) betas.push_back(b);
  if (want > 0 && (int)betas.size() < want)
    die("El fichero %s solo tiene %d lineas (necesita %d)\n", argv[1], (int)betas.size(), want);
  if (want > 0 && fscanf(f, "%lf", &b) != EOF)
    fprintf(stderr, "Hay lineas de mas en el fichero %s\n", argv[1]);
  fclose(f);
  const int nb = (int)betas.size();
  if (nb < 1) die("No betas in %s\n", argv[1]);
  for (int i = 0; i + 1 < nb; ++i)
    if (betas[i] < betas[i + 1])
      die("Error en fichero %s: betas no ordenadas de mayor a menor!\n", argv[1]);
  std::vector<uint8_t> lut((size_t)nb * EA_LUT_BYTES);
  for (int i = 0; i < nb; ++i) ea_make_lut(betas[i], lut.data() + (size_t)i * EA_LUT_BYTES);
  char name[256];
  snprintf(name, sizeof name, "LUT_for_PRNG_nbits%02d_NB%d.bin", 12, nb);
  FILE *o = fopen(name, "wb");
  if (!o) die("Problemas abriendo %s\n", name);
  if (fwrite(betas.data(), sizeof(double), nb, o) != (size_t)nb) die("problemas escribiendo betas\n");
  if (fwrite(lut.data(), 1, lut.size(), o) != lut.size())
    die("problemas escribiendo lut_heat_bath\n");
  fclose(o);
  printf("%s\n", name);
  return 0;
}
