/*
 * oracle.h -- CPU oracle for the YCSB transaction-execution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load liboracle.so, and only to check or time
 * against the GPU path; the product library never links or calls it.
 *
 * Plain C11, built with gcc from this directory (see Makefile).  Every
 * function names the reference code it restates.
 */
#ifndef GCCTB_ORACLE_H
#define GCCTB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashing (splitmix64, Steele et al. OOPSLA 2014) and seeded values ---- */
uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_hash3(uint64_t seed, uint64_t a, uint64_t b);
uint64_t orc_cell_init(uint64_t seed, uint64_t row, uint32_t col, uint32_t cols);
uint64_t orc_write_value(uint64_t seed, uint64_t tx, uint32_t op);
void orc_seed_table(uint64_t seed, uint64_t rows, uint32_t cols, uint64_t* cells);

/* ---- workload generator restatement (generation/gen_ycsb.cu:76-150) ---- */
double orc_zeta(uint64_t n, double theta);
uint64_t orc_zipf_rank(uint64_t n, double theta, double zetan, double u);
/* Ops of transaction `tx` (global id); returns the op count. */
uint32_t orc_gen_tx(uint64_t rows, uint32_t cols, uint32_t len_min, uint32_t len_max, double theta,
                    double read_ratio, uint64_t seed, double zetan, uint64_t tx, uint64_t* out);

/* ---- sorted-array index: SortedArray_GPU::Find (include/index.cuh:74-91) ---- */
uint64_t orc_find(const uint32_t* sorted_keys, const uint64_t* vals, uint32_t length, uint32_t key);

/* ---- serial replay ----
 * Executes the transactions one after another in ascending (batch, key0, key1)
 * of order[3*t .. 3*t+2] (ties broken by transaction id and reported), each op
 * of a transaction in op order: a read returns the cell's current value, a
 * write stores orc_write_value(seed, t, j).  Keys map to rows through the
 * identity index (key k -> row k, keys >= rows are skipped).
 * table: rows*cols cells, updated in place; reads: one per op (writes get 0).
 * Returns the number of tied order keys (0 for a strict order), or -1 on
 * allocation failure. */
int64_t orc_replay(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                   const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads);
/* orc_replay with the work split over `threads` threads by blocks of rows
 * (a read depends only on earlier writes to its own row): same table, reads
 * and tie count. */
int64_t orc_replay_mt(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                      const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads, int threads);
/* Whole-row mode (every op reads or writes all cols cells of its row, like
 * the reference's Item copies, benchmark/ycsb.cu:57,86-91): cell c written by
 * op j of transaction t holds orc_write_cell(seed, t, j, c) = mix64(write
 * value + c); a read returns orc_row_digest of the row it saw.
 * orc_replay_rows is orc_replay_mt in that mode. */
uint64_t orc_write_cell(uint64_t seed, uint64_t tx, uint32_t op, uint32_t col);
uint64_t orc_row_digest(const uint64_t* row, uint32_t cols);
int64_t orc_replay_rows(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                        const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads, int threads);
/* Replay in plain transaction-id order (GaccO's serial order inside a batch). */
void orc_replay_txid(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                     const uint64_t* tx_off, uint64_t ntx, uint64_t* reads);

/* ---- GPUTx one-pass rank (src/gputx.cu:5-38, include/gputx.cuh:213-294) ----
 * rank: one per transaction; counts: per-rank transaction counts (length
 * returned in *nranks, caller provides ntx+1 slots).  Keys >= rows are
 * skipped. */
int orc_gputx_rank(uint64_t rows, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx,
                   uint32_t* rank, uint32_t* counts, uint64_t* nranks);

/* ---- multi-threaded CPU path (the CPU baseline): the same protocols on C11
 * atomics, one transaction per thread at a time, ops in key order.
 * scheme: 0 tpl_nw, 1 tpl_wd, 2 to, 3 mvcc, 4 silo, 5 tictoc, 6 gacco, 7 gputx.
 * Fills order (3 u64 per tx) and reads like the GPU path; returns 0 or <0. */
typedef struct orc_cpu_stats {
  double seconds;
  uint64_t aborts;
  uint64_t committed;
  int threads;
} orc_cpu_stats;
int orc_cpu_run(int scheme, uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table,
                const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, uint64_t batch_size, int threads,
                uint64_t* order, uint64_t* reads, orc_cpu_stats* stats);

/* ---- TPC-C (benchmark/tpcc.cu:36-380) ----
 * One op's update of its 5-cell row (kind 0 warehouse, 1 district, 2
 * customer for payment; 3 district, 4 stock for new_order; param the
 * amount's float bits or quantity | remote << 16); returns the op's result
 * (the field read back, or the float updated first). */
uint64_t orc_tpcc_apply(uint32_t kind, uint32_t param, uint64_t* row);
/* Serial replay of TPC-C transactions in (batch, key0, key1) order; returns
 * the number of tied keys (-1: allocation failed). */
int64_t orc_tpcc_replay(uint64_t rows, uint64_t* table, const uint64_t* ops, const uint32_t* params,
                        const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* results);

#ifdef __cplusplus
}
#endif
#endif
