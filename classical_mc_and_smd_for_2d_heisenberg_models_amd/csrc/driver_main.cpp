// Command-line front ends of the drop-in runs, named after the reference's
// executables (CF/0MC_compile.sh, CF/0SMD_compile.sh):
//   xMC_run  NP [--precision f32|f64|plain] [--seed S] [--extended]
//   xSMD_run    [--precision f32|f64|plain] [--extended]
// Both read the Settings_*.txt files of the current directory.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/hsmd.h"

static int parse_precision(const char *s) {
  if (!strcmp(s, "f32")) return HSMD_BATCHED_F32;
  if (!strcmp(s, "f64")) return HSMD_BATCHED_F64;
  if (!strcmp(s, "plain")) return HSMD_PLAIN_F64;
  fprintf(stderr, "unknown precision form: %s\n", s);
  exit(2);
}

int main(int argc, char **argv) {
  const char *prog = strrchr(argv[0], '/') ? strrchr(argv[0], '/') + 1 : argv[0];
  const int mc = strstr(prog, "MC") != NULL;
  int precision = HSMD_BATCHED_F32, strict = 1, np = 1;
  unsigned long long seed = 1;
  int i = 1;
  if (mc) {
    if (argc < 2) {
      fprintf(stderr, "usage: %s NP [--precision f32|f64|plain] [--seed S] [--extended]\n", prog);
      return 2;
    }
    np = atoi(argv[1]);
    i = 2;
  }
  for (; i < argc; i++) {
    if (!strcmp(argv[i], "--precision") && i + 1 < argc) precision = parse_precision(argv[++i]);
    else if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], NULL, 10);
    else if (!strcmp(argv[i], "--extended")) strict = 0;
    else { fprintf(stderr, "unknown argument: %s\n", argv[i]); return 2; }
  }
  const hsmd_status st = mc ? hsmd_run_mc(".", np, precision, seed, strict) : hsmd_run_smd(".", precision, strict);
  if (st != HSMD_OK) {
    char buf[1024];
    hsmd_last_error(buf, sizeof buf);
    fprintf(stderr, "%s failed (%d): %s\n", prog, st, buf);
    return 1;
  }
  return 0;
}
