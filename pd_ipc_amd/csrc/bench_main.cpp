// pdx_bench: the headless report as a program, in the place of the reference's
// pd_bench (app/bench/main.cpp): the same command line, the report on stdout
// (on stderr with exit status 2 when the run fails).
#include <cstdio>
#include <vector>

#include "../../include/pdx.h"

int main(int argc, char** argv) {
  std::vector<char> text(1 << 22);
  const int rc = pdx_bench_run(argc, argv, text.data(), (long long)text.size(), nullptr);
  std::fputs(text.data(), rc == PDX_OK ? stdout : stderr);
  return rc == PDX_OK ? 0 : 2;
}
