// Command-line drop-in for the reference executable (source/MAIN.cu:16-22):
//   mwdg2_run [-epsilon E] [-solver hwfv1|mwdg2] case.par
// MW_PRECISION=4 selects FP32 (the reference's _USE_DOUBLES=0 build).
#include <cstdio>

#include "../../include/mwdg2.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s [-epsilon E] [-solver hwfv1|mwdg2] case.par\n", argv[0]);
    return 2;
  }
  return mw_run_simulation(argc, argv) == MW_OK ? 0 : 1;
}
