// nemoh_run: the reference's preProcessor, Solver and postProcessor
// executables in one, the Solver's work on the GPU (include/nemoh_case.h).
//
//   nemoh_run [--pre] [--solve] [--post] [--mixed] [case_dir]
//
// With no stage flag all three stages run.  Without case_dir the case
// is the one ID.dat in the working directory names (its first record the
// length, its second the path, as the reference's ReadTID reads it).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "nemoh_case.h"
#include "nemoh_core.h"

static std::string case_from_id_dat() {
  std::ifstream f("ID.dat");
  int len = 0;
  std::string id;
  if (!(f >> len >> id) || len <= 0) return "";
  return id.substr(0, (size_t)len);
}

int main(int argc, char** argv) {
  int stages = 0, precision = NEMOH_FP64;
  std::string dir;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--pre")) stages |= NEMOH_STAGE_PRE;
    else if (!std::strcmp(argv[i], "--solve")) stages |= NEMOH_STAGE_SOLVE;
    else if (!std::strcmp(argv[i], "--post")) stages |= NEMOH_STAGE_POST;
    else if (!std::strcmp(argv[i], "--mixed")) precision = NEMOH_MIXED;
    else if (argv[i][0] == '-') {
      std::fprintf(stderr, "usage: %s [--pre] [--solve] [--post] [--mixed] [case_dir]\n", argv[0]);
      return 2;
    } else dir = argv[i];
  }
  if (!stages) stages = NEMOH_STAGE_PRE | NEMOH_STAGE_SOLVE | NEMOH_STAGE_POST;
  if (dir.empty()) dir = case_from_id_dat();
  if (dir.empty()) {
    std::fprintf(stderr, "no case directory given and no readable ID.dat here\n");
    return 2;
  }
  char msg[1024];
  const int r = nemoh_case_run(dir.c_str(), stages, precision, msg, sizeof msg);
  if (r) {
    std::fprintf(stderr, "nemoh_run: error %d: %s\n", r, msg);
    return 1;
  }
  std::printf(" ->%s%s%s done (%s)\n", (stages & NEMOH_STAGE_PRE) ? " preProcessor" : "",
              (stages & NEMOH_STAGE_SOLVE) ? " Solver" : "", (stages & NEMOH_STAGE_POST) ? " postProcessor" : "",
              precision == NEMOH_MIXED ? "mixed" : "FP64");
  return 0;
}
