// test_ycsb: drop-in for the reference driver (benchmark_cpu/ycsb.cu:81-302).
//
//   test_ycsb <dataset> <rownum> <txcnt> <cc_scheme> <warp_density> <block_size> <batch_size> <debug>
//
// GCCTB_DEVICE=cpu runs the same configuration on the host's cores
// (gcctb_cpu_run) instead of the GPU; with debug != 0 it is then checked by
// the serial replay only.
//
// <dataset> is a reference .txs file (raw YCSBTx records, include/ycsb.cuh) or
// "gen:theta=<t>,rr=<r>,len=<l>[,lenmax=<m>][,row=1]" for a generated
// workload.  A .txs file runs like the reference: rows of 13 seeded u64 cells
// (the 104-byte Item) and every op reading or writing its whole row
// (benchmark/ycsb.cu:57,86-91).  A generated workload runs on rows of 10 u64
// cells with one cell per op, or whole rows of 13 cells with row=1.  The
// index maps the id to the row; the run uses the reference's launch
// structure.  Output: the same
// stdout block and 15-field CSV line (appended to <dataset>.txt) as
// ExecInfo::Analyse (include/runtime.cuh:169-214).  With debug != 0 the kernels
// record the event log (AddEvent, include/runtime.cuh:64-81) and the run is
// checked afterwards like the reference's ExecInfo::Check
// (src/runtime.cu:23-158): conflict graph over the events, UNCOMMITTED
// transactions, a cycle printed with its transactions and events, GOOD or
// BAD; then every read must return the latest preceding write in the
// serialization order the kernels reported (REPLAY OK / REPLAY BAD).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gcctb.h"

static const uint64_t SEED = 141919810;  // benchmark_cpu/ycsb.cu:27

static bool parse_gen(const std::string& spec, double& theta, double& rr, uint32_t& len, uint32_t& lenmax,
                      int& row_ops) {
  if (spec.rfind("gen:", 0) != 0) return false;
  theta = 0.0;
  rr = 0.9;
  len = 16;
  lenmax = 0;
  row_ops = 0;
  std::string rest = spec.substr(4);
  size_t pos = 0;
  while (pos < rest.size()) {
    size_t end = rest.find(',', pos);
    if (end == std::string::npos) end = rest.size();
    const std::string kv = rest.substr(pos, end - pos);
    const size_t eq = kv.find('=');
    if (eq != std::string::npos) {
      const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "theta") theta = std::atof(v.c_str());
      if (k == "rr") rr = std::atof(v.c_str());
      if (k == "len") len = (uint32_t)std::atoi(v.c_str());
      if (k == "lenmax") lenmax = (uint32_t)std::atoi(v.c_str());
      if (k == "row") row_ops = std::atoi(v.c_str()) != 0 ? 1 : 0;
    }
    pos = end + 1;
  }
  if (lenmax < len) lenmax = len;
  return true;
}

// ExecInfo::Analyse (include/runtime.cuh:169-214): the CSV line appended to
// <dataset>.txt and the stdout block
static bool print_result(const gcctb_stats& st, const char* dataset, const std::string& cc_type, int wd, int bs,
                         int batch_size, int txcnt) {
  std::vector<char> csv(4096), text(8192);
  if (gcctb_format_result(&st, cc_type.c_str(), wd, bs, batch_size, (uint64_t)txcnt, csv.data(), csv.size(),
                          text.data(), text.size()) != GCCTB_OK)
    return false;
  std::fstream of;
  of.open(std::string(dataset) + ".txt", std::ios::out | std::ios::app);
  of << csv.data();
  of.close();
  std::cout << text.data();
  return true;
}

static int die(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, gcctb_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 9) {
    std::fprintf(stderr,
                 "usage: %s <dataset> <rownum> <txcnt> <cc_scheme> <warp_density> <block_size> <batch_size> <debug>\n",
                 argv[0]);
    return 1;
  }
  int rownum = 0, txcnt = 0, valid_txn_bitoffset = 0, warp_cnt = 0, batch_size = 0, debug = 0;
  std::sscanf(argv[2], "%d", &rownum);
  std::sscanf(argv[3], "%d", &txcnt);
  const std::string cc_type(argv[4]);
  std::sscanf(argv[5], "%d", &valid_txn_bitoffset);
  std::sscanf(argv[6], "%d", &warp_cnt);
  std::sscanf(argv[7], "%d", &batch_size);
  std::sscanf(argv[8], "%d", &debug);
  if (warp_cnt > 32 || valid_txn_bitoffset > 5) return -1;  // benchmark_cpu/ycsb.cu:97-98
  std::cout << "TRY " << argv[4] << "," << valid_txn_bitoffset << "," << warp_cnt << "," << batch_size << "\n";

  // transactions
  std::vector<uint64_t> ops, off(txcnt + 1);
  uint64_t nops = 0;
  double theta, rr;
  uint32_t len, lenmax;
  int row_ops = 1;     // a .txs file: whole-row ops on 104-byte rows
  uint32_t cols = 13;
  if (parse_gen(argv[1], theta, rr, len, lenmax, row_ops)) {
    cols = row_ops ? 13 : 10;
    gcctb_workload w{};
    w.rows = (uint64_t)rownum;
    w.cols = cols;
    w.ntx = (uint64_t)txcnt;
    w.len_min = len;
    w.len_max = lenmax;
    w.theta = theta;
    w.read_ratio = rr;
    w.seed = SEED;
    ops.resize((size_t)txcnt * lenmax + 1);
    if (gcctb_generate(&w, ops.data(), off.data(), &nops) != GCCTB_OK) return die("generate");
  } else {
    ops.resize((size_t)txcnt * 16 + 1);
    if (gcctb_read_txs_file(argv[1], (uint64_t)txcnt, cols, SEED, ops.data(), off.data(), &nops) != GCCTB_OK)
      return die("read dataset");
  }

  gcctb_config cfg{};
  cfg.cc_scheme = argv[4];
  cfg.warp_density = valid_txn_bitoffset;
  cfg.block_size = warp_cnt;
  cfg.batch_size = (uint64_t)batch_size;
  cfg.rows = (uint64_t)rownum;
  cfg.cols = cols;
  cfg.seed = SEED;
  cfg.backoff_ticks = 20;
  cfg.row_ops = row_ops;
  // the latch schemes (slow_*) record no event log: their debug check is the
  // serial replay alone
  const bool latch_scheme = cc_type.rfind("slow_", 0) == 0;
  cfg.debug_events = debug && !latch_scheme ? 1 : 0;

  // GCCTB_DEVICE=cpu: the same run on the host's cores (gcctb_cpu_run)
  const char* dev = std::getenv("GCCTB_DEVICE");
  if (dev && std::strcmp(dev, "cpu") == 0) {
    std::vector<uint64_t> table((size_t)rownum * cols), init, reads(nops + 1), order(3 * (size_t)txcnt + 3);
    if (gcctb_cpu_seed_table(table.data(), (uint64_t)rownum, cols, SEED) != GCCTB_OK) return die("seed table");
    if (debug) init = table;
    gcctb_stats st{};
    if (gcctb_cpu_run(&cfg, 0, table.data(), ops.data(), off.data(), (uint64_t)txcnt, reads.data(), order.data(),
                      &st) != GCCTB_OK)
      return die("cpu run");
    std::fprintf(stderr, "device cpu\n");
    if (!print_result(st, argv[1], cc_type, valid_txn_bitoffset, warp_cnt, batch_size, txcnt)) return die("format result");
    if (!debug) return 0;
    uint64_t bad = 0;
    if (gcctb_check_host(&cfg, init.data(), table.data(), ops.data(), off.data(), (uint64_t)txcnt, order.data(),
                         reads.data(), &bad) != GCCTB_OK)
      return die("check");
    if (bad) std::cout << "REPLAY BAD " << bad << "\n";
    else std::cout << "REPLAY OK\n";
    return bad ? 2 : 0;
  }

  gcctb_ctx* ctx = nullptr;
  if (gcctb_create(&cfg, &ctx) != GCCTB_OK) return die("create");
  if (gcctb_seed_table(ctx) != GCCTB_OK) return die("seed table");
  std::vector<uint64_t> init;
  if (debug) {
    init.resize((size_t)rownum * cols);
    if (gcctb_fetch(ctx, init.data(), nullptr, nullptr, nullptr) != GCCTB_OK) return die("fetch");
  }
  if (gcctb_load_ops(ctx, ops.data(), off.data(), (uint64_t)txcnt) != GCCTB_OK) return die("load ops");
  gcctb_stats st{};
  const int rc = gcctb_run_plain(ctx, &st);
  if (rc != GCCTB_OK && rc != GCCTB_EKERNEL) return die("run");
  if (rc == GCCTB_EKERNEL) std::fprintf(stderr, "warning: %s\n", gcctb_last_error());

  if (!print_result(st, argv[1], cc_type, valid_txn_bitoffset, warp_cnt, batch_size, txcnt)) return die("format result");

  int status = 0;
  if (debug) {
    if (cfg.debug_events) {
      std::vector<gcctb_event> events(nops + (uint64_t)txcnt + 1);
      uint64_t nev = 0, unc = 0, cycle_len = 0;
      if (gcctb_fetch_events(ctx, events.data(), events.size(), &nev) != GCCTB_OK) return die("fetch events");
      std::vector<char> report(1 << 20);
      if (gcctb_conflict_check(events.data(), nev, (uint64_t)txcnt, &unc, nullptr, 0, &cycle_len, report.data(),
                               report.size()) != GCCTB_OK)
        return die("conflict check");
      std::cout << report.data();
    }
    uint64_t bad = 0, uncommitted = 0;
    if (gcctb_check(ctx, init.data(), &bad, &uncommitted) != GCCTB_OK) return die("check");
    if (bad) std::cout << "REPLAY BAD " << bad << "\n";
    else std::cout << "REPLAY OK\n";
    // the exit status follows the replay, which checks every read and cell
    // exactly; the conflict graph depends on stamps taken next to (not
    // atomically with) each access, which TicToc's unvalidated reads can
    // order differently from the accesses themselves
    status = (bad || uncommitted) ? 2 : 0;
  }
  gcctb_destroy(ctx);
  return status;
}
