// gen_ycsb: drop-in for the reference generator (generation/gen_ycsb.cu:34-58).
//
//   gen_ycsb <rownum> <txcnt> <theta> <r_ratio> <out.txs>
//
// Writes <txcnt> raw YCSBTx records (16 requests, distinct keys in ascending
// order, read with probability r_ratio) drawn by the library's deterministic
// counter-based generator from the reference's seed.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gcctb.h"

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s <rownum> <txcnt> <theta> <r_ratio> <out.txs>\n", argv[0]);
    return 1;
  }
  gcctb_workload w{};
  w.rows = std::strtoull(argv[1], nullptr, 10);
  w.ntx = std::strtoull(argv[2], nullptr, 10);
  w.theta = std::atof(argv[3]);
  w.read_ratio = std::atof(argv[4]);
  w.cols = 10;
  w.len_min = w.len_max = 16;
  w.seed = 141919810;  // generation/gen_ycsb.cu:12
  std::vector<uint64_t> ops(w.ntx * 16 + 1), off(w.ntx + 1);
  uint64_t nops = 0;
  if (gcctb_generate(&w, ops.data(), off.data(), &nops) != GCCTB_OK ||
      gcctb_write_txs_file(argv[5], ops.data(), off.data(), w.ntx) != GCCTB_OK) {
    std::fprintf(stderr, "gen_ycsb: %s\n", gcctb_last_error());
    return 1;
  }
  std::printf("GEN OK\n");
  return 0;
}
