// Command-line front end: `lmp_amd -in in.script [-log log.file]`, the
// drop-in for `lmp_serial -in in.script` (main.cpp:73, lammps.cpp option parse).
#include <cstdio>
#include <cstring>

#include "../../include/liggghts_amd.h"

int main(int argc, char **argv) {
  const char *in = nullptr, *log = nullptr;
  for (int k = 1; k < argc; k++) {
    if ((!strcmp(argv[k], "-in") || !strcmp(argv[k], "-i")) && k + 1 < argc) in = argv[++k];
    else if ((!strcmp(argv[k], "-log") || !strcmp(argv[k], "-l")) && k + 1 < argc) log = argv[++k];
    else {
      fprintf(stderr, "usage: %s -in <script> [-log <file>]\n", argv[0]);
      return 1;
    }
  }
  if (!in) {
    fprintf(stderr, "usage: %s -in <script> [-log <file>]\n", argv[0]);
    return 1;
  }
  return dem_run_script(in, log) == 0 ? 0 : 1;
}
