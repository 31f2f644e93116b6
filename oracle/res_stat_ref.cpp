// Driver for the reference's result fingerprint (res_stat, common-utils/res_stat.h:32-133, with
// common-utils/md5.cpp), built from the reference sources into oracle/_ref/ to pin this
// library's fingerprint. Test infrastructure only.
//   res_stat_ref <file of int32 columns, column after column> <rows> <num_cols>
// prints the reference's log, whose "Md5sum of result histogram: <hex>" line is the result.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "res_stat.h"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: res_stat_ref <columns file> <rows> <num_cols>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  const unsigned long long rows = std::strtoull(argv[2], nullptr, 10);
  const int nc = std::atoi(argv[3]);
  std::vector<std::vector<int>> data(nc, std::vector<int>(rows + 1));
  std::vector<int*> cols(nc);
  for (int c = 0; c < nc; ++c) {
    if (std::fread(data[c].data(), 4, rows, f) != rows) return 2;
    cols[c] = data[c].data();
  }
  std::fclose(f);
  res_stat<int, unsigned long long>(cols.data(), rows, nc, false);
  return 0;
}
