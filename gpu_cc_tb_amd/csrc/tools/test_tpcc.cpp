// test_tpcc: drop-in for the reference's TPC-C driver (benchmark_cpu/tpcc.cu:440-565).
//
//   test_tpcc <isno> <warehouse_cnt> <txcnt> <cc_scheme> <warp_density> <block_size> <batch_size> <debug>
//
// isno 1 runs new_order transactions, 0 payment, over warehouse_cnt
// warehouses (the seeded tables of gcctb_tpcc_seed_tables, the transactions
// of gcctb_tpcc_generate with the reference's remote ratio 0.1 unless
// GCCTB_TPCC_REMOTE gives another).  Output: the "TRY" line, then the stdout
// block and 15-field CSV line of ExecInfo::Analyse (include/runtime.cuh:169-214),
// the CSV appended to tpcc_no_wh<W>.txt or tpcc_pm_wh<W>.txt like the
// reference's name (benchmark_cpu/tpcc.cu:453-455).  With debug != 0 the run
// is re-executed serially on the host in the order the kernels reported
// (gcctb_tpcc_check) and "REPLAY OK" or "REPLAY BAD <n>" follows; new_order
// also prints the first orders' totals (tpcc.cu:340,371).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gcctb.h"

static const uint64_t SEED = 141919810;

static int die(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, gcctb_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 9) {
    std::fprintf(stderr,
                 "usage: %s <isno> <warehouse_cnt> <txcnt> <cc_scheme> <warp_density> <block_size> <batch_size> "
                 "<debug>\n",
                 argv[0]);
    return 1;
  }
  int isno = 0, warehouse_cnt = 0, txcnt = 0, valid_txn_bitoffset = 0, warp_cnt = 0, batch_size = 0, debug = 0;
  std::sscanf(argv[1], "%d", &isno);
  std::sscanf(argv[2], "%d", &warehouse_cnt);
  std::sscanf(argv[3], "%d", &txcnt);
  const std::string cc_type(argv[4]);
  std::sscanf(argv[5], "%d", &valid_txn_bitoffset);
  std::sscanf(argv[6], "%d", &warp_cnt);
  std::sscanf(argv[7], "%d", &batch_size);
  std::sscanf(argv[8], "%d", &debug);
  if (warehouse_cnt < 1 || txcnt < 0) {
    std::fprintf(stderr, "warehouse_cnt must be at least 1 and txcnt not negative\n");
    return 1;
  }
  const std::string name = std::string(isno ? "tpcc_no_wh" : "tpcc_pm_wh") + std::to_string(warehouse_cnt);
  std::cout << "TRY " << argv[4] << "," << valid_txn_bitoffset << "," << warp_cnt << "," << batch_size << "\n";

  const uint32_t W = (uint32_t)warehouse_cnt;
  const uint64_t rows = gcctb_tpcc_rows(W);
  std::vector<uint64_t> cells(rows * 5);
  std::vector<float> price(100000);
  if (gcctb_tpcc_seed_tables(W, SEED, cells.data(), price.data()) != GCCTB_OK) return die("seed tables");

  gcctb_tpcc_workload wl{};
  wl.warehouses = W;
  wl.new_order = isno ? 1 : 0;
  const char* rr = std::getenv("GCCTB_TPCC_REMOTE");
  wl.remote_ratio = rr ? std::atof(rr) : 0.1;
  wl.ntx = (uint64_t)txcnt;
  wl.seed = SEED;
  const size_t per = isno ? 16 : 3;
  std::vector<uint64_t> ops((size_t)txcnt * per + 1), off((size_t)txcnt + 1), cust((size_t)txcnt + 1);
  std::vector<uint32_t> params((size_t)txcnt * per + 1);
  uint64_t nops = 0;
  if (gcctb_tpcc_generate(&wl, ops.data(), params.data(), off.data(), cust.data(), &nops) != GCCTB_OK)
    return die("generate");

  gcctb_config cfg{};
  cfg.cc_scheme = argv[4];
  cfg.warp_density = valid_txn_bitoffset;
  cfg.block_size = warp_cnt;
  cfg.batch_size = (uint64_t)batch_size;
  cfg.rows = rows;
  cfg.cols = 5;
  cfg.seed = SEED;
  cfg.backoff_ticks = 20;
  cfg.tpcc = 1;
  gcctb_ctx* ctx = nullptr;
  if (gcctb_create(&cfg, &ctx) != GCCTB_OK) return die("create");
  if (gcctb_load_table(ctx, cells.data(), rows, 5) != GCCTB_OK) return die("load table");
  if (gcctb_load_ops(ctx, ops.data(), off.data(), (uint64_t)txcnt) != GCCTB_OK) return die("load ops");
  if (gcctb_load_params(ctx, params.data(), nops) != GCCTB_OK) return die("load params");
  gcctb_stats st{};
  const int rc = gcctb_run_plain(ctx, &st);
  if (rc != GCCTB_OK && rc != GCCTB_EKERNEL) return die("run");
  if (rc == GCCTB_EKERNEL) std::fprintf(stderr, "warning: %s\n", gcctb_last_error());

  std::vector<char> csv(4096), text(8192);
  if (gcctb_format_result(&st, cc_type.c_str(), valid_txn_bitoffset, warp_cnt, batch_size, (uint64_t)txcnt,
                          csv.data(), csv.size(), text.data(), text.size()) != GCCTB_OK)
    return die("format result");
  std::fstream of;
  of.open(name + ".txt", std::ios::out | std::ios::app);
  of << csv.data();
  of.close();
  std::cout << text.data();

  int status = 0;
  if (debug) {
    std::vector<uint64_t> fin(rows * 5), results(nops + 1), order(3 * (size_t)txcnt + 3);
    if (gcctb_fetch(ctx, fin.data(), results.data(), order.data(), nullptr) != GCCTB_OK) return die("fetch");
    uint64_t bad = 0;
    if (gcctb_tpcc_check(W, cells.data(), fin.data(), ops.data(), params.data(), off.data(), (uint64_t)txcnt,
                         order.data(), results.data(), &bad) != GCCTB_OK)
      return die("check");
    if (bad) std::cout << "REPLAY BAD " << bad << "\n";
    else std::cout << "REPLAY OK\n";
    if (isno && txcnt > 0) {
      const uint64_t show = txcnt < 4 ? (uint64_t)txcnt : 4;
      std::vector<float> total(show);
      std::vector<uint32_t> oid(show);
      if (gcctb_tpcc_new_order_totals(W, cells.data(), price.data(), ops.data(), params.data(), off.data(), show,
                                      cust.data(), results.data(), total.data(), oid.data()) != GCCTB_OK)
        return die("totals");
      for (uint64_t t = 0; t < show; ++t) std::printf("ORDER %llu o_id %u total %.2f\n", (unsigned long long)t, oid[t], total[t]);
    }
    status = (bad || st.committed != (uint64_t)txcnt) ? 2 : 0;
  }
  gcctb_destroy(ctx);
  return status;
}
