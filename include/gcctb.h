/*
 * gcctb.h -- C ABI of the MI355X (gfx950) YCSB transaction-execution path.
 *
 * Plain pointers and sizes only; every call returns GCCTB_OK (0) or a negative
 * error code, with the message available from gcctb_last_error().  A context
 * owns every device allocation it makes and is used by one host thread at a
 * time.  Contexts on the same GPU are independent.
 *
 * Each entry point names the reference interface it replaces (paths are in
 * Al0ha0e/gpu_cc_tb).
 */
#ifndef GCCTB_H
#define GCCTB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCCTB_OK 0
#define GCCTB_EINVAL (-1)   /* bad argument (the reference's driver returns -1, benchmark_cpu/ycsb.cu:97-98) */
#define GCCTB_ENOMEM (-2)   /* device or host allocation failed */
#define GCCTB_EDEVICE (-3)  /* HIP runtime error, missing GPU or missing kernel image */
#define GCCTB_ESTATE (-4)   /* call out of order (e.g. run before load_ops) */
#define GCCTB_EIO (-5)      /* file could not be read */
#define GCCTB_EKERNEL (-6)  /* a kernel reported an error flag (timestamp/pool overflow) */

typedef struct gcctb_ctx gcctb_ctx;

/* Scheme and launch parameters.  cc_scheme, warp_density, block_size and
 * batch_size keep the meaning of the reference driver's argv[4..7]
 * (benchmark_cpu/ycsb.cu:90-98): a transaction owns 2^warp_density lanes, a
 * workgroup has block_size*32 threads, and transactions run in batches of
 * batch_size. */
typedef struct gcctb_config {
  const char* cc_scheme;   /* tpl_nw tpl_wd to mvcc silo tictoc gacco gputx */
  int32_t warp_density;    /* 0..5 */
  int32_t block_size;      /* 1..32 */
  uint64_t batch_size;     /* >= 1 */
  uint64_t rows;           /* table rows (<= 2^32 - 2) */
  uint32_t cols;           /* u64 columns per row (1..255) */
  uint64_t seed;           /* seeds the table and the written values */
  uint32_t backoff_ticks;  /* minimum unit of the randomized retry backoff (10 ns ticks); the unit
                              is the failed attempt's duration when that is longer */
  uint32_t persist_window; /* persistent and serving forms: at most this many transactions in
                              flight (0 = as many as the device holds) */
  int32_t persist_adaptive; /* not supported: gcctb_create returns GCCTB_EINVAL
                               for 1 (the abort-rate-driven window was taken out) */
  int32_t row_ops;          /* 1: whole-row mode, the reference's Item copies
                               (benchmark/ycsb.cu:57,86-91): every op reads or writes all cols
                               cells of its row (cols <= 16; 13 u64 = the 104-byte Item), a
                               written cell c holds mix64(write value + c), and a read returns
                               the row's digest; plain and graph forms only */
  int32_t debug_events;     /* 1: the plain and graph forms record the event log of
                               gcctb_fetch_events (the debug mode's AddEvent) */
  int32_t tpcc;             /* 1: TPC-C (benchmark/tpcc.cu): rows of the TPC-C row space
                               (cols 5, gcctb_tpcc_* below), every op an update of its row
                               with a parameter from gcctb_load_params; plain and graph forms */
} gcctb_config;

/* Result of one run.  The first block mirrors the reference's Metrics and
 * ExecInfo timings (include/runtime.cuh:28-37,114-131). */
typedef struct gcctb_stats {
  uint64_t txs;            /* transactions executed */
  uint64_t committed;      /* transactions committed (== txs) */
  uint64_t aborts;         /* aborted attempts */
  uint64_t ts_ticks, wait_ticks, abort_ticks, manager_ticks, tot_ticks, index_ticks;
  double tick_khz;         /* frequency of the tick counters */
  double pre_ms;           /* per-batch reset / preprocessing (reference "pre") */
  double proc_ms;          /* execution kernels (reference "proc") */
  double tot_ms;           /* whole timed region (reference "tot") */
  double exec_kernel_ms;   /* sum of execution-kernel durations (HIP events) */
  uint64_t batches;        /* launches of the execution kernel */
  uint64_t ranks;          /* gputx: number of rank levels */
  uint64_t same_rank_conflicts; /* gputx: adjacent conflicting same-rank accesses */
  uint64_t index_miss;     /* ops whose key was absent from the index */
  uint64_t error_flags;    /* kernel error bits (0 = none) */
  uint64_t total_ops;      /* ops executed */
  uint64_t read_ops;       /* read ops executed */
  uint64_t deferred;       /* persistent form of the lock, timestamp and OCC schemes: transactions
                              admitted after all others because they write a hot key */
  uint64_t combined;       /* persistent form of gacco: ops run by hot-row combiners (a row whose
                              queue holds at least 128 accesses is run in queue order by one
                              workgroup before the launch) */
} gcctb_stats;

/* Create a context: replaces constructing the scheme by name and its device
 * allocations (benchmark_cpu/ycsb.cu:166-195; ConcurrencyControlCPUBase ctor,
 * include/cc.cuh:23-27; TwoPL_CPU ctor, include/twopl.cuh:400-429, etc.). */
int gcctb_create(const gcctb_config* cfg, gcctb_ctx** out);
void gcctb_destroy(gcctb_ctx* ctx);
const char* gcctb_last_error(void);

/* Upload a row-major rows x cols table of u64 cells: replaces
 * Table_CPU::InitData + DB_CPU::ToGPU (include/table.cuh:67-75,263-276,
 * include/db.cuh:97-126).  rows/cols must match the config. */
int gcctb_load_table(gcctb_ctx* ctx, const uint64_t* cells, uint64_t rows, uint32_t cols);
/* Seed the table on the device: cell[r][c] = splitmix64(seed ^ (r*cols + c)). */
int gcctb_seed_table(gcctb_ctx* ctx);

/* Build the index over the id column (key i -> row i): replaces
 * Table_CPU::AddIndex(SORTED_ARRAY, 0) (include/table.cuh:95-130,
 * include/index.cuh:96-111).  The index answers Find's question (value, or
 * the length on a miss).  Over these dense keys it is direct addressing (the
 * row of key k is k, a key >= rows misses) and stores nothing; with
 * GCCTB_INDEX_HASH=1 it is the hash table that gcctb_build_index builds.
 * Called by load_table/seed_table. */
int gcctb_build_index_identity(gcctb_ctx* ctx);
/* Build the index from host (key, value) pairs (any order; keys distinct,
 * keys < 2^32-1, values < rows and distinct: a primary-key index, so that
 * distinct keys of a transaction name distinct rows; GCCTB_EINVAL otherwise).
 * A hash table of the pairs; pairs that are exactly keys 0 .. n-1 on rows of
 * the same number make the dense index of gcctb_build_index_identity.
 * A graph captured by gcctb_run_graph is dropped and captured again on its
 * next call. */
int gcctb_build_index(gcctb_ctx* ctx, const uint32_t* keys, const uint32_t* vals, uint64_t n);
/* Index probe for tests: rows_out[i] = Find(keys[i]) (include/index.cuh:74-91);
 * a miss returns the index length. */
int gcctb_probe(gcctb_ctx* ctx, const uint32_t* keys, uint64_t n, uint32_t* rows_out);

/* Upload transactions in CSR form: ops[tx_off[t] .. tx_off[t+1]) are the ops
 * of transaction t (encoding: key bits 0..31, column bits 32..39, write bit
 * 40).  Replaces the .txs load, flattening and DynamicTransactionSet_CPU
 * (benchmark_cpu/ycsb.cu:116-157, include/transaction.cuh:199-368).
 * Contract: at most 32 ops per transaction, and the keys of one transaction
 * are distinct (the reference's generator drops repeated keys,
 * generation/gen_ycsb.cu:129-136); a repeated key returns GCCTB_EINVAL. */
int gcctb_load_ops(gcctb_ctx* ctx, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx);
/* TPC-C contexts: one update parameter per loaded op (gcctb_tpcc_generate
 * writes them), loaded after gcctb_load_ops. */
int gcctb_load_params(gcctb_ctx* ctx, const uint32_t* params, uint64_t n);
/* The input checks of gcctb_load_ops on the host alone (no context, no GPU):
 * offsets, lengths, columns below cols, distinct keys per transaction, and
 * the op and transaction limits.  *nreads (may be null) gets the read count. */
int gcctb_check_ops(const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, uint32_t cols, uint64_t* nreads);

/* Run every loaded transaction with the reference's launch structure: Init0
 * splits into batches, then per batch Init (reset / preprocess) + one kernel
 * launch + synchronization (benchmark_cpu/ycsb.cu:249-291, include/cc.cuh:29-36,
 * benchmark/ycsb.cu:21-111).  The table keeps the result; run again to
 * continue on it.  A batch whose kernel reports an error (GCCTB_EKERNEL, e.g.
 * the per-wave watchdog, GCCTB_WATCHDOG_S seconds) ends the run; every CC
 * word is then cleared, so no lock or mark of an unfinished transaction
 * outlives the call (the same holds for every run and serve entry point). */
int gcctb_run_plain(gcctb_ctx* ctx, gcctb_stats* out);

/* The same kernels with the per-batch sequence captured once into a hipGraph
 * and replayed: no host round trip between batches.  Results follow the same
 * per-batch semantics as gcctb_run_plain.  At most 4096 launches (batches, or
 * GPUTx rank levels); more returns GCCTB_EINVAL. */
int gcctb_run_graph(gcctb_ctx* ctx, gcctb_stats* out);

/* Persistent batched form: the loaded transactions are one stream (batches
 * one after another); one preparation pass and one persistent launch whose
 * groups take the next transaction of the stream as soon as the current one
 * commits.  Timestamps and tickets count across batch boundaries, so the
 * serialization order spans the stream; GaccO / GPUTx results equal those of
 * the plain form (transaction-id / (rank, id) serial order). */
int gcctb_run_persistent(gcctb_ctx* ctx, gcctb_stats* out);

/* Serving form (no reference counterpart: the reference runs a fixed input
 * file, benchmark_cpu/ycsb.cu:242-291).  The loaded transactions are requests
 * that arrive in batches: batch i holds the next batch_sizes[i] transactions
 * and arrives arrival_ms[i] after the call starts (null: every batch at 0, the
 * peak-load case).  Table, index and CC state stay on the device.  Continuous
 * batching: one persistent launch runs while requests arrive; each arrived
 * batch is copied from pinned host memory, prepared on a second stream and
 * admitted into the running launch, whose groups take its transactions as
 * they free up, beside those still running.  A launch is closed and a new one
 * opened when the next batch would exceed the admission limits (transactions
 * per launch; MVCC version pool, bounded by free device memory) or before an
 * idle gap longer than a quarter of the watchdog period; the watchdog bounds
 * how long a launch waits for an admission.  tx_latency_ms (may be null)
 * receives, per transaction, its commit time minus the arrival of its batch
 * (commits are stamped with the device wall clock, mapped to host time
 * through a clock sample).  Order keys from gcctb_fetch are {launch, key0,
 * key1}.  Not available for gputx (it ranks its whole input before running). */
typedef struct gcctb_serve_config {
  uint64_t max_launch_txs;  /* admission: transactions per launch (0 = no limit) */
  uint64_t pool_cap_bytes;  /* admission, MVCC: version-pool budget, 24 bytes per write
                               (0 = the pool load_ops sized) */
} gcctb_serve_config;
typedef struct gcctb_serve_stats {
  uint64_t txs, committed, aborts;
  uint64_t launches;
  uint64_t deferred;        /* times an arrived batch waited for a running launch to drain */
  uint64_t error_flags;
  double tot_ms;            /* call start to the last completion */
  double busy_ms;           /* sum of launch durations (open to drained and restored) */
  double txps;              /* committed / tot_ms */
  double p50_ms, p99_ms, max_ms, mean_ms;  /* commit latency */
  double mean_launch_txs;
  uint64_t admitted_live;   /* batches admitted into a launch that was already running */
  uint64_t pool_nodes;      /* MVCC: versions one launch may hold (admission limit) */
} gcctb_serve_stats;
int gcctb_serve(gcctb_ctx* ctx, const gcctb_serve_config* cfg, const uint64_t* batch_sizes,
                const double* arrival_ms, uint64_t nbatches, double* tx_latency_ms, gcctb_serve_stats* out);

/* The plain form in one call that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * runs the gcctb_run_plain kernels (per batch: counter reset, one launch,
 * restore of the touched words, host synchronization) for tpl_nw, tpl_wd, to,
 * mvcc, silo and tictoc over device buffers the caller owns; keys are row ids
 * (key < rows, the reference's identity id column; other keys are index
 * misses).  Every other device word (CC words, MVCC chain heads and version
 * pool, the row of each op, per-transaction words, counters, metrics) lives
 * in one static device array of GCCTB_STATIC_SCRATCH_BYTES per GPU, held until
 * the process ends.  Limits: rows <= 1,048,576, ntx <= 131,072,
 * nops <= 1,048,576 (GCCTB_EINVAL beyond), at most 262,144 MVCC writes per
 * batch (GCCTB_EKERNEL, version-pool-full, beyond).  Inputs are not checked
 * beyond these sizes: the ops of a transaction must name distinct keys and be
 * at most 32.  Every call clears the array's state first; calls that overlap
 * on one GPU share the array and race. */
#define GCCTB_STATIC_SCRATCH_BYTES 24156672ull
typedef struct gcctb_static_run {
  const char* cc_scheme;    /* tpl_nw tpl_wd to mvcc silo tictoc */
  int32_t warp_density;     /* 0..5, as in gcctb_config */
  int32_t block_size;       /* 1..32 */
  uint64_t batch_size;      /* >= 1 */
  uint64_t* cells;          /* device: rows x cols u64 cells, updated in place */
  uint64_t rows;
  uint32_t cols;            /* 1..255 */
  uint64_t seed;            /* seeds the written values, as in gcctb_config */
  uint32_t backoff_ticks;   /* as in gcctb_config */
  const uint64_t* ops;      /* device: CSR ops, nops entries */
  const uint64_t* tx_off;   /* device: ntx + 1 offsets */
  uint64_t ntx;
  uint64_t nops;
  uint64_t* reads;          /* device, nops u64: each read op's value (write entries untouched) */
  uint64_t* order;          /* device, 2 x ntx u64: order[t] = key0, order[ntx + t] = key1; the
                               serialization order is ascending (t / batch_size, key0, key1) */
  void* stream;             /* hipStream_t to run on (null: the null stream) */
} gcctb_static_run;
int gcctb_run_plain_static(const gcctb_static_run* run, gcctb_stats* out);

/* Uneven batches for the plain and graph forms: sizes must add up to the
 * loaded transaction count (replaces the uniform split of Init0,
 * include/cc.cuh:29-35).  Loading new ops clears it. */
int gcctb_set_batches(gcctb_ctx* ctx, const uint64_t* sizes, uint64_t n);

/* Copy results back for checking.  Any pointer may be null.
 *   table:  rows*cols cells
 *   reads:  one u64 per op (value returned by a read; writes leave 0)
 *   order:  3 u64 per transaction {batch, key0, key1}: the transaction's place
 *           in the serialization order is ascending (batch, key0, key1)
 *   aborts: one u32 per transaction */
int gcctb_fetch(gcctb_ctx* ctx, uint64_t* table, uint64_t* reads, uint64_t* order, uint32_t* aborts);
/* gputx: per-transaction rank (src/gputx.cu:5-27) and per-rank counts
 * (include/gputx.cuh:283-294).  counts may be null; nranks receives the level
 * count.  After a persistent run, which needs no per-rank counts, the first
 * call that asks for them computes them on the device. */
int gcctb_fetch_ranks(gcctb_ctx* ctx, uint32_t* rank, uint32_t* counts, uint64_t* nranks);
/* Per-row CC words after a run (2PL/T-O/MVCC/Silo/TicToc: one u64 per row). */
int gcctb_fetch_cc(gcctb_ctx* ctx, uint64_t* words);

/* Serializability check of the last run (the driver's debug mode;
 * ExecInfo::Check, src/runtime.cu:129-158): re-executes the committed
 * transactions one by one in the reported order from init_table (the table as
 * it was before the run) and counts reads and final cells that disagree with
 * the GPU (*bad) and transactions that did not commit (*uncommitted). */
int gcctb_check(gcctb_ctx* ctx, const uint64_t* init_table, uint64_t* bad, uint64_t* uncommitted);

/* Event log and conflict-graph check (the debug mode's ExecInfo::Check,
 * include/runtime.cuh:39-81, src/runtime.cu:23-158).  An event is a read
 * (type 0) or write (type 1) of object oid (a row) by transaction tid, or its
 * commit (type 2); ids order the events in time (AddEvent draws them from one
 * global counter).  gcctb_conflict_check runs on the host alone: a
 * transaction is active on an object from its first access to its commit; a
 * read links the reader to every other active writer of the object, a write
 * to every other active accessor; transactions without a commit are counted
 * (and listed in the report as UNCOMMITTED), and a depth-first search reports
 * the first cycle it closes (*cycle_len transactions, the first cycle_cap of
 * them in cycle).  report (may be null) receives the text the reference's
 * check prints: SCAN OK, UNCOMMITTED lines, for a cycle the BAD TX lines and
 * its events, then GOOD or BAD. */
typedef struct gcctb_event {
  uint64_t id;
  uint64_t oid;
  int32_t tid;
  int32_t type;
} gcctb_event;
int gcctb_conflict_check(const gcctb_event* events, uint64_t n, uint64_t ntx, uint64_t* uncommitted, uint64_t* cycle,
                         uint64_t cycle_cap, uint64_t* cycle_len, char* report, uint64_t report_len);
/* The event log of the last plain or graph run of a context created with
 * debug_events = 1: every op of each transaction's committed attempt, stamped
 * when the access took effect, then the transaction's commit.  events must
 * hold nops + ntx entries; *n receives the count (GCCTB_ESTATE without a
 * logged run). */
int gcctb_fetch_events(gcctb_ctx* ctx, gcctb_event* events, uint64_t cap, uint64_t* n);

/* Workload generation (host, deterministic in (seed, tx, draw); counter-based,
 * multi-threaded).  Zipf draw follows generation/gen_ycsb.cu:76-102 with
 * n = rows and row = rank - 1; keys are distinct inside a transaction and ops
 * are sorted by key (gen_ycsb.cu:129-150).  Each op reads with probability
 * read_ratio, else writes; the column is drawn uniformly.
 * ops must hold ntx*len_max entries, tx_off ntx+1.  Returns the op count in
 * *nops. */
typedef struct gcctb_workload {
  uint64_t rows;
  uint32_t cols;
  uint64_t ntx;
  uint64_t tx_base;        /* global id of the first transaction (stream position) */
  uint32_t len_min;        /* ops per transaction: uniform in [len_min, len_max] */
  uint32_t len_max;
  double theta;            /* Zipf skew in [0, 1) */
  double read_ratio;       /* probability that an op reads */
  uint64_t seed;
  int32_t threads;         /* host threads (0 = all cores) */
} gcctb_workload;
int gcctb_generate(const gcctb_workload* w, uint64_t* ops, uint64_t* tx_off, uint64_t* nops);
/* zeta(n, theta) as the generator computes it (for tests). */
double gcctb_zeta(uint64_t n, double theta);

/* Read a reference .txs file (raw YCSBTx records of 136 bytes, include/ycsb.cuh:21-33)
 * and convert it to CSR ops; the column of op j of transaction t is
 * hash(seed, t, j) mod cols.  ops must hold 16*ntx entries. */
int gcctb_read_txs_file(const char* path, uint64_t ntx, uint32_t cols, uint64_t seed, uint64_t* ops,
                        uint64_t* tx_off, uint64_t* nops);
int gcctb_write_txs_file(const char* path, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx);

/* The driver's result block, replacing ExecInfo::Analyse
 * (include/runtime.cuh:169-214): csv gets the 15-field line the driver
 * appends to <dataset>.txt, text the stdout block, both with the reference's
 * arithmetic (float scaling of the tick sums by 1 / (txcnt * kHz), float
 * milliseconds) and iostream formatting.  Either buffer may be null;
 * GCCTB_EINVAL if one is too small (including its terminating zero). */
int gcctb_format_result(const gcctb_stats* st, const char* cc_type, int32_t valid_txn_bitoffset, int32_t warp_cnt,
                        int32_t batch_size, uint64_t txcnt, char* csv, uint64_t csv_len, char* text,
                        uint64_t text_len);

/* The CPU path: the eight schemes run by the host's cores, batch after batch
 * like gcctb_run_plain (benchmark_cpu/ycsb.cu:249-291 with the transaction
 * body of benchmark/ycsb.cu:51-106), for running the driver's configuration
 * on a CPU.  Chosen explicitly (test_ycsb: GCCTB_DEVICE=cpu); no GPU entry
 * point ever falls back to it.  cfg gives cc_scheme, batch_size, rows, cols,
 * seed and row_ops (warp_density and block_size do not apply); keys address
 * rows directly, and a key >= rows is an index miss.  table holds rows * cols
 * cells on the host (gcctb_cpu_seed_table fills it as gcctb_seed_table does
 * on the device) and is updated in place; reads gets one value per op (0 for
 * writes) and order three keys per transaction (batch, key0, key1) as
 * gcctb_fetch returns them.  threads <= 0: every core (at most
 * OMP_NUM_THREADS when that is set).  Timings are in nanosecond ticks
 * (st->tick_khz = 1e6). */
int gcctb_cpu_seed_table(uint64_t* table, uint64_t rows, uint32_t cols, uint64_t seed);
int gcctb_cpu_run(const gcctb_config* cfg, int32_t threads, uint64_t* table, const uint64_t* ops,
                  const uint64_t* tx_off, uint64_t ntx, uint64_t* reads, uint64_t* order, gcctb_stats* st);
/* The serial-replay check of gcctb_check (src/runtime.cu:129-158) on host
 * buffers, for the CPU path: *bad = reads and cells that differ from
 * re-executing the transactions one by one in `order` from init_table. */
int gcctb_check_host(const gcctb_config* cfg, const uint64_t* init_table, const uint64_t* final_table,
                     const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, const uint64_t* order,
                     const uint64_t* reads, uint64_t* bad);

/* TPC-C (benchmark/tpcc.cu, benchmark_cpu/tpcc.cu, generation/gen_tpcc*.cu):
 * payment and new_order over the tables that carry concurrency control,
 * laid out as one row space of 5 u64 cells per row (warehouses, then
 * districts, customers and stock; the layout and the field packing are
 * stated in gcctb_common.h).  A context with gcctb_config.tpcc = 1, rows =
 * gcctb_tpcc_rows(W) and cols = 5 runs them: gcctb_load_table with the rows
 * of gcctb_tpcc_seed_tables, gcctb_load_ops + gcctb_load_params with
 * gcctb_tpcc_generate's ops and parameters, then gcctb_run_plain (or
 * gcctb_run_graph).  Every op updates its row under the scheme's exclusive
 * access and leaves its result (the next_o_id or stock quantity it read, or
 * the float it updated first) in the reads of gcctb_fetch. */
typedef struct gcctb_tpcc_workload {
  uint32_t warehouses;   /* W, 1..8192 */
  int32_t new_order;     /* 1: new_order transactions, 0: payment (benchmark_cpu/tpcc.cu's isno) */
  double remote_ratio;   /* new_order: chance that an order line's stock comes from another
                            warehouse (gen_tpcc.cu's no_remote_ratio) */
  uint64_t ntx;
  uint64_t seed;
} gcctb_tpcc_workload;
uint64_t gcctb_tpcc_rows(uint32_t warehouses);
/* The initial rows (gen_tpcc_table.cu's field values, seeded draws where it
 * draws): cells holds gcctb_tpcc_rows(W) * 5 u64; item_price (may be null)
 * 100,000 floats. */
int gcctb_tpcc_seed_tables(uint32_t warehouses, uint64_t seed, uint64_t* cells, float* item_price);
/* ntx transactions as CSR ops with one parameter per op: ops and params hold
 * 3 (payment) or 16 (new_order) entries per transaction; tx_customer (may be
 * null) gets each transaction's customer row. */
int gcctb_tpcc_generate(const gcctb_tpcc_workload* w, uint64_t* ops, uint32_t* params, uint64_t* tx_off,
                        uint64_t* tx_customer, uint64_t* nops);
/* new_order's output (tpcc.cu:340,371-372): total[t] = sum of quantity x price
 * over the order's lines x (1 - discount) x (1 + w_tax + d_tax), o_id[t]
 * (may be null) = the next_o_id the order took, from a run's results. */
int gcctb_tpcc_new_order_totals(uint32_t warehouses, const uint64_t* init_cells, const float* item_price,
                                const uint64_t* ops, const uint32_t* params, const uint64_t* tx_off, uint64_t ntx,
                                const uint64_t* tx_customer, const uint64_t* results, float* total, uint32_t* o_id);
/* The debug check of a TPC-C run (the serial-replay half of ExecInfo::Check,
 * src/runtime.cu:129-158, which the reference's TPC-C driver reaches with
 * debug != 0): the transactions re-executed one by one in `order` (3 u64 per
 * transaction, as gcctb_fetch returns it) from init_cells; *bad = op results
 * and cells that differ from results and final_cells, plus transactions
 * whose order keys tie. */
int gcctb_tpcc_check(uint32_t warehouses, const uint64_t* init_cells, const uint64_t* final_cells,
                     const uint64_t* ops, const uint32_t* params, const uint64_t* tx_off, uint64_t ntx,
                     const uint64_t* order, const uint64_t* results, uint64_t* bad);

/* Device information. */
int gcctb_device_info(char* name, uint64_t name_len, double* wall_clock_khz, int32_t* cus);

/* Stable device sort of n (key, value) pairs by the low key_bits bits of
 * the key: the sort that the GaccO and GPUTx preparation runs where the
 * reference calls thrust::sort_by_key / thrust::sort
 * (include/gacco.cuh:307,331; include/gputx.cuh:252,266,281).  Keys are
 * 32-bit (key_bytes 4) or 64-bit (key_bytes 8); keys and vals are host
 * buffers, sorted in place; *ms (may be null) gets the device time of the
 * sort alone.  Pairs with equal low key bits keep their input order. */
int gcctb_sort_pairs(void* keys, uint32_t* vals, uint64_t n, int32_t key_bytes, int32_t key_bits, double* ms);

/* Memory-system ceilings of the GPU for the roofline (no reference
 * counterpart; measured with the kernels' own access forms):
 *   which 0: random 64-bit agent-scope CAS over n_words words
 *         1: random 64-bit agent-scope atomic add
 *         2: atomic add to one word (one lane per wave)
 *         3: random coherent (agent-scope) 8-byte load
 *         4: random cached 8-byte load
 *         5: streaming copy of n_words words (rate in bytes/s, read + write)
 *         6: pairs of random coherent 8-byte loads within one 128-byte line
 *         7: pairs of random coherent 8-byte loads on two lines
 *         8: load a word, load a cell in its 128-byte line, CAS the word
 *         9: the same with the cell in another line
 *        10: atomic add to 16 words that share one 128-byte line
 *        11: atomic add to 16 words on 16 lines
 *        12: atomic add to one word (one lane per wave) while the other lanes
 *            poll words of the same line with coherent loads
 *        13: the same with the polled words on other lines
 *            (rate in pairs/s for 6 and 7, reads/s for 8 and 9, adds to the
 *            one word for 12 and 13; n_words >= 16, >= 256 for 10..13)
 * threads x iters operations; *rate in operations/s, *ms the timed launch. */
int gcctb_microbench(int32_t which, uint64_t n_words, uint64_t threads, uint64_t iters, double* rate, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* GCCTB_H */
