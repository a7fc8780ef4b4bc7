// test_ycsb: drop-in for the reference driver (benchmark_cpu/ycsb.cu:81-302).
//
//   test_ycsb <dataset> <rownum> <txcnt> <cc_scheme> <warp_density> <block_size> <batch_size> <debug>
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
  cfg.debug_events = debug ? 1 : 0;
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

  // ExecInfo::Analyse (include/runtime.cuh:169-214)
  const float clock_rate = (float)st.tick_khz;
  const float inv = 1.0f / (txcnt * clock_rate * 1.0f);
  const float ts_duration = st.ts_ticks * inv;
  const float wait_duration = st.wait_ticks * inv;
  const float abort_duration = st.abort_ticks * inv;
  const float manager_duration = st.manager_ticks * inv;
  const float tot_duration = st.tot_ticks * inv;
  const float index_duration = st.index_ticks * inv;
  const float precompute_time = (float)st.pre_ms, processing_time = (float)st.proc_ms, tot_time = (float)st.tot_ms;
  std::fstream of;
  of.open(std::string(argv[1]) + ".txt", std::ios::out | std::ios::app);
  of << cc_type << "," << valid_txn_bitoffset << "," << warp_cnt << "," << batch_size << ","
     << 1000.0f / tot_time * txcnt << "," << st.aborts << "," << ts_duration << "," << wait_duration << ","
     << abort_duration << "," << manager_duration << "," << tot_duration << "," << index_duration << ","
     << precompute_time / 1000.f << "," << processing_time / 1000.f << "," << tot_time / 1000.f << "\n";
  of.close();
  std::cout << cc_type << ","
            << "valid_txn_bitoffset " << valid_txn_bitoffset << ",\n"
            << "warp_cnt " << warp_cnt << ",\n"
            << "batch_size " << batch_size << ",\n"
            << "through " << 1000.0f / tot_time * txcnt << ",\n"
            << "abort " << st.aborts << " diff " << ((1LL << 31) - 1) - ((long long)st.aborts + (long long)batch_size)
            << " avr " << st.aborts * 1.0 / batch_size << ",\n"
            << "ts_d " << ts_duration << ",\n"
            << "wait_d " << wait_duration << ",\n"
            << "abort_d " << abort_duration << ",\n"
            << "mng_d " << manager_duration << ",\n"
            << "tot_d " << tot_duration << ",\n"
            << "idx_d " << index_duration << ",\n"
            << "pre " << precompute_time / 1000.f << ",\n"
            << "proc " << processing_time / 1000.f << ",\n"
            << "tot " << tot_time / 1000.f << "\n";

  int status = 0;
  if (debug) {
    std::vector<gcctb_event> events(nops + (uint64_t)txcnt + 1);
    uint64_t nev = 0, unc = 0, cycle_len = 0;
    if (gcctb_fetch_events(ctx, events.data(), events.size(), &nev) != GCCTB_OK) return die("fetch events");
    std::vector<char> report(1 << 20);
    if (gcctb_conflict_check(events.data(), nev, (uint64_t)txcnt, &unc, nullptr, 0, &cycle_len, report.data(),
                             report.size()) != GCCTB_OK)
      return die("conflict check");
    std::cout << report.data();
    uint64_t bad = 0, uncommitted = 0;
    if (gcctb_check(ctx, init.data(), &bad, &uncommitted) != GCCTB_OK) return die("check");
    if (bad) std::cout << "REPLAY BAD " << bad << "\n";
    else std::cout << "REPLAY OK\n";
    status = (cycle_len || unc || bad || uncommitted) ? 2 : 0;
  }
  gcctb_destroy(ctx);
  return status;
}
