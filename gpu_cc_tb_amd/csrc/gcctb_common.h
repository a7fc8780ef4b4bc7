// Shared host/device definitions of the YCSB transaction-execution path:
// op encoding, deterministic hashes, CC word layouts and the kernel argument
// block.  Compiled by hipcc for both the host library and the gfx950 kernels.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GC_HD __host__ __device__ __forceinline__
#else
#define GC_HD static inline
#endif

namespace gcctb {

// ---------------------------------------------------------------------------
// Scheme identifiers.  Names match the driver's cc_scheme argument
// (benchmark_cpu/ycsb.cu:166-195).
// ---------------------------------------------------------------------------
enum Scheme : int32_t {
  SCHEME_TPL_NW = 0,
  SCHEME_TPL_WD = 1,
  SCHEME_TO = 2,
  SCHEME_MVCC = 3,
  SCHEME_SILO = 4,
  SCHEME_TICTOC = 5,
  SCHEME_GACCO = 6,
  SCHEME_GPUTX = 7,
  SCHEME_COUNT = 8
};

// Longest transaction the kernels accept.  The reference's YCSBTx holds at most
// 16 requests (include/ycsb.cuh:27); the variable-length stream goes to 32.
constexpr int MAX_OPS = 32;
// Columns per row (10 x u64 cells, 80 B).
constexpr int DEFAULT_COLS = 10;

// ---------------------------------------------------------------------------
// Op encoding: one u64 per op.
//   bits  0..31  key (looked up in the sorted-array index)
//   bits 32..39  column
//   bit  40      1 = write, 0 = read
// The reference's YCSBReq is {int read; u32 key} (include/ycsb.cuh:21-25);
// the column is the per-column extension of the u64-cell table.
// ---------------------------------------------------------------------------
GC_HD uint64_t make_op(uint32_t key, uint32_t col, bool is_write) {
  return (uint64_t)key | ((uint64_t)(col & 0xffu) << 32) | ((uint64_t)(is_write ? 1 : 0) << 40);
}
GC_HD uint32_t op_key(uint64_t op) { return (uint32_t)op; }
GC_HD uint32_t op_col(uint64_t op) { return (uint32_t)(op >> 32) & 0xffu; }
GC_HD bool op_is_write(uint64_t op) { return (op >> 40) & 1u; }
// An op id (index into the op list, below 2^31) with the op's write flag in
// bit 31: the value the GaccO / GPUTx preparation sorts carry.
constexpr uint32_t OPID_WRITE = 0x80000000u;
constexpr uint32_t OPID_MASK = 0x7fffffffu;
// The key bucket of the hot-key counts (Fibonacci hash into 2^16 buckets).
constexpr uint32_t DEFER_BUCKET_BITS = 16;
GC_HD uint32_t defer_bucket(uint32_t key) { return (key * 2654435761u) >> (32 - DEFER_BUCKET_BITS); }
// need[op] of an op that a row's combiner already executed (GaccO, persistent form)
constexpr uint32_t NEED_DONE = 0xffffffffu;

// ---------------------------------------------------------------------------
// Deterministic hashing.  splitmix64 (Steele, Lea, Flood: "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
// ---------------------------------------------------------------------------
GC_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
GC_HD uint64_t splitmix64(uint64_t x) { return mix64(x + 0x9E3779B97F4A7C15ull); }
// Counter-based stream: (seed, a, b) -> u64.
GC_HD uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return splitmix64(splitmix64(seed ^ a) ^ b);
}

// Stream tags, xored into the seed so that the draws are independent.
constexpr uint64_t TAG_KEY = 0x4b45590000000000ull;    // "KEY"
constexpr uint64_t TAG_RW = 0x5257000000000000ull;     // "RW"
constexpr uint64_t TAG_COL = 0x434f4c0000000000ull;    // "COL"
constexpr uint64_t TAG_WRITE = 0x5752495445000000ull;  // "WRITE"
constexpr uint64_t TAG_LEN = 0x4c454e0000000000ull;    // "LEN"
constexpr uint64_t TAG_BACKOFF = 0x424f464600000000ull;

// Initial cell value: splitmix64(seed ^ (row * cols + col)).
GC_HD uint64_t cell_init(uint64_t seed, uint64_t row, uint32_t col, uint32_t cols) {
  return splitmix64(seed ^ (row * (uint64_t)cols + col));
}
// Value written by op `op` (position inside the transaction) of transaction
// `tx`.  It does not depend on the attempt nor on anything the transaction
// read, so every committed attempt writes the same bits.
GC_HD uint64_t write_value(uint64_t seed, uint64_t tx, uint32_t op) {
  return hash3(seed ^ TAG_WRITE, tx, op);
}

// Whole-row mode (the reference's Item copies, benchmark/ycsb.cu:57,86-91):
// every op reads or writes all cols cells of its row.  Cell c written by op
// `op` of transaction `tx` holds write_cell(seed, tx, op, c); a read returns
// the digest of the row it saw (cells folded in column order), which pins
// every cell of the row in one 64-bit result.
constexpr int ROW_MAX_COLS = 16;  // whole-row mode: at most 16 cells (128 B) per row
constexpr uint64_t DIGEST_INIT = 0x524f570000000000ull;  // "ROW"
GC_HD uint64_t write_cell(uint64_t seed, uint64_t tx, uint32_t op, uint32_t col) {
  return mix64(write_value(seed, tx, op) + col);
}
GC_HD uint64_t digest_step(uint64_t h, uint64_t cell) { return mix64(h ^ cell); }

// ---------------------------------------------------------------------------
// CC word layouts (explicit masks; the reference uses LSB-first bitfields).
// ---------------------------------------------------------------------------
// 2PL lock word.  The reference packs {shared:1, holder_cnt:31, holder:31}
// (include/twopl.cuh:10-21) and leaves stale bits in a free word.  Here bit 0
// marks EXCLUSIVE, bits 1..31 count holders, bits 32..62 hold the exclusive
// holder, and a free word is always 0: a shared lock adds one holder, the last
// shared unlock subtracts back to 0, an exclusive unlock stores 0.  An
// uncontended lock is then a single CAS that guesses the word is 0.
namespace tpl {
constexpr uint64_t EXCL = 1ull;
constexpr uint64_t CNT_ONE = 2ull;
GC_HD uint64_t excl(uint64_t w) { return w & 1u; }
GC_HD uint64_t cnt(uint64_t w) { return (w >> 1) & 0x7fffffffull; }
GC_HD uint64_t holder(uint64_t w) { return (w >> 32) & 0x7fffffffull; }
GC_HD uint64_t make_excl(uint64_t h) { return EXCL | CNT_ONE | ((h & 0x7fffffffull) << 32); }
}  // namespace tpl

// T/O and MVCC timestamp word (include/to.cuh:11-22, include/mvcc.cuh:12-23):
// bit 0 uncommitted, bits 1..31 rts, bits 32..62 wts.
namespace tsw {
GC_HD uint64_t unc(uint64_t w) { return w & 1u; }
GC_HD uint64_t rts(uint64_t w) { return (w >> 1) & 0x7fffffffull; }
GC_HD uint64_t wts(uint64_t w) { return (w >> 32) & 0x7fffffffull; }
GC_HD uint64_t make(uint64_t u, uint64_t r, uint64_t wt) {
  return (u & 1u) | ((r & 0x7fffffffull) << 1) | ((wt & 0x7fffffffull) << 32);
}
constexpr uint64_t TS_MAX = 0x7fffffffull;  // 31-bit timestamps (to.cuh:108)
// The version a word names: every bit but rts.  Placing a write mark and
// publishing a version change these bits; raising rts never does.
GC_HD uint64_t version(uint64_t w) { return w & ~(0x7fffffffull << 1); }
}  // namespace tsw

// Silo word (include/silo.cuh:17-32): bit 0 lock, bits 1..63 ts.
namespace silo {
GC_HD uint64_t locked(uint64_t w) { return w & 1u; }
GC_HD uint64_t ts(uint64_t w) { return w >> 1; }
GC_HD uint64_t make(uint64_t l, uint64_t t) { return (l & 1u) | (t << 1); }
}  // namespace silo

// TicToc word (include/tictoc.cuh:17-28): bit 0 lock, bits 1..15 delta,
// bits 16..63 wts; rts = wts + delta.
namespace tictoc {
GC_HD uint64_t locked(uint64_t w) { return w & 1u; }
GC_HD uint64_t delta(uint64_t w) { return (w >> 1) & 0x7fffull; }
GC_HD uint64_t wts(uint64_t w) { return w >> 16; }
GC_HD uint64_t rts(uint64_t w) { return wts(w) + delta(w); }
GC_HD uint64_t make(uint64_t l, uint64_t d, uint64_t wt) {
  return (l & 1u) | ((d & 0x7fffull) << 1) | (wt << 16);
}
}  // namespace tictoc

// ---------------------------------------------------------------------------
// Metrics: the reference's 7 counters (include/runtime.cuh:28-37) plus a few
// build-specific ones.  Durations are in wall-clock ticks (s_memrealtime).
// ---------------------------------------------------------------------------
struct Metrics {
  uint64_t abort;
  uint64_t ts_duration;
  uint64_t wait_duration;
  uint64_t abort_duration;
  uint64_t manager_duration;
  uint64_t tot_duration;
  uint64_t index_duration;
  uint64_t committed;
  uint64_t index_miss;
  uint64_t error_flags;
  uint64_t pad[6];
};
constexpr int METRICS_WORDS = 16;
// Device counters (u64 indices into ExecArgs::ctrs, each on its own 128-byte
// line).  T/O and MVCC only need unique timestamps, so the allocator is
// sharded: shard s (by block index) hands out s, s+TS_SHARDS, s+2*TS_SHARDS,
// ... which divides the traffic on any one address by TS_SHARDS.  Commit
// tickets must follow a single real-time order and keep one counter.
constexpr int CTR_TICKET = 0;
constexpr int CTR_POOL = 16;
constexpr int CTR_TS0 = 32;
constexpr int TS_SHARDS = 16;
constexpr int TS_SHARD_BITS = 4;
// Admission state of the persistent form: the next stream slot, and
// GATE_SHARDS copies of the gate a wave reads before taking slots (one copy
// per block index mod GATE_SHARDS, each on its own line, so the polls of
// thousands of waves spread out): [1] transactions in urgent state, [2]
// stream handed out.  The open form (serving) adds GATE_SHARDS copies of the
// admission word that the host writes while the launch runs: a u32 at the
// start of each copy's line (a host copy writes an aligned 32-bit word in one
// piece, a 64-bit word possibly in two); slots below its low 31 bits are
// admitted, bit 31 closes the launch (no more admissions).
constexpr int CTR_SLOT = CTR_TS0 + 16 * TS_SHARDS;
constexpr int CTR_GATE0 = CTR_SLOT + 16;
constexpr int GATE_SHARDS = 16;
constexpr int CTR_ADMIT0 = CTR_GATE0 + 16 * GATE_SHARDS;
constexpr int CTR_EVENT = CTR_ADMIT0 + 16 * GATE_SHARDS;  // event-log stamps (a whole run)
constexpr int CTR_WORDS = CTR_EVENT + 16;
constexpr uint32_t ADMIT_CLOSED = 1u << 31;
// Waves add their totals into one of METRIC_SLOTS 128-byte slots (by block
// index) so that the end-of-kernel atomics do not pile onto one address; the
// host sums the slots.
constexpr int METRIC_SLOTS = 256;

// error_flags bits
constexpr uint64_t ERR_TS_OVERFLOW = 1;     // 31-bit timestamp space exhausted
constexpr uint64_t ERR_POOL_OVERFLOW = 2;   // MVCC version pool full
constexpr uint64_t ERR_CHAIN_BROKEN = 4;    // MVCC chain ended before a visible version
constexpr uint64_t ERR_TXLEN = 8;           // transaction longer than the kernel's slots
constexpr uint64_t ERR_WATCHDOG = 16;       // a wave exceeded the watchdog time and gave up

// ---------------------------------------------------------------------------
// Key -> row index with Find's contract (include/index.cuh:74-91): a key's
// value, or the index length on a miss.  Keys are unique (the build rejects
// duplicates), so a hash table answers exactly what the sorted array does, in
// one 64-byte bucket read instead of a search.  nbuckets buckets of
// [8 keys][8 values]; IDX_EMPTY (0xffffffff, a reserved key) marks a free
// slot.  A key sits in the first bucket that had a free slot when it was
// inserted, along idx_bucket(key), +1, +2, ... (mod nbuckets); slots never
// empty again, so a lookup stops at the key or at a bucket with a free slot.
// ---------------------------------------------------------------------------
constexpr uint32_t IDX_MISS = 0xffffffffu;   // device-side probe result for an absent key
constexpr uint32_t IDX_EMPTY = 0xffffffffu;  // free slot
constexpr int IDX_SLOTS = 8;                 // keys per bucket
struct IndexView {
  const uint32_t* buckets;  // nbuckets * 16 u32: [8 keys][8 values]
  uint32_t nbuckets;
  uint32_t length;          // number of keys; Find's miss value
  // 1: the keys are 0 .. length-1 and key k is row k (the identity index, a
  // sorted array of dense keys): a lookup is direct addressing and reads
  // nothing; buckets is null
  uint32_t dense;
};
// Row of key in a dense index: the key itself, or a miss past the end
GC_HD uint32_t idx_dense_find(uint32_t key, uint32_t length) { return key < length ? key : IDX_MISS; }
// Home bucket: Fibonacci hashing (multiply by 2^32 / golden ratio, an odd
// constant, so distinct keys keep distinct products), mapped onto
// [0, nbuckets) by the high half of a 32 x 32 product.
GC_HD uint32_t idx_bucket(uint32_t key, uint32_t nbuckets) {
  return (uint32_t)(((uint64_t)(uint32_t)(key * 0x9E3779B1u) * nbuckets) >> 32);
}
// Buckets for n keys: about half the slots stay free.
GC_HD uint64_t idx_buckets_for(uint64_t n) { return n / 4 + 1; }

// ---------------------------------------------------------------------------
// Kernel argument block for the execution kernels.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Per-row CC state (CC words, MVCC chain heads, turn counters) sits at word
// slot(row) = (row % SLOT_WAYS) << shift | row / SLOT_WAYS, where 2^shift is
// the smallest power of two >= ceil(rows / SLOT_WAYS).  Zipf's hottest rows
// are ids 0, 1, 2, ...; packed by row id, the 16 hottest CC words (32 turn
// counters) share one 128-byte line, whose atomics then serialize and whose
// pollers slow every handoff on it.  Spread this way, rows 0..15 sit on 16
// lines 2^shift words apart.  The arrays hold SLOT_WAYS << shift words
// (below 2 * rows + 16); a shift and an or keep the address math to two
// 32-bit operations.
// ---------------------------------------------------------------------------
constexpr uint32_t SLOT_WAYS = 16;
GC_HD constexpr uint32_t slot_shift(uint64_t rows) {
  uint32_t s = 0;
  while (((uint64_t)1 << s) * SLOT_WAYS < rows) ++s;
  return s;
}
GC_HD constexpr uint64_t slot_words(uint64_t rows) { return (uint64_t)SLOT_WAYS << slot_shift(rows); }
GC_HD constexpr uint32_t row_slot(uint32_t row, uint32_t shift) { return ((row % SLOT_WAYS) << shift) | (row / SLOT_WAYS); }

struct ExecArgs {
  // table (row-major u64 cells)
  uint64_t* cells;
  uint64_t rows;
  uint32_t cols;
  uint32_t scheme;
  // index
  IndexView index;
  // transactions (CSR)
  const uint64_t* ops;
  const uint64_t* tx_off;
  const uint32_t* tx_map;  // launch slot -> transaction id (null: identity)
  // hot keys (2PL, persistent form with the deferral): sampled counts per
  // key bucket, the hot threshold, and the word that says whether any key
  // is hot (null: no hot keys)
  const uint32_t* hot_cnt;
  const uint32_t* hot_any;
  uint32_t hot_thr;
  uint64_t slot_lo;        // launch covers slots [slot_lo, slot_hi)
  uint64_t slot_hi;
  // CC state
  uint64_t* cc;            // per-row CC word, at row_slot(row, slot_shift)
  uint32_t* vprev;         // MVCC: per-row head of the version chain (node+1), at the row's slot
  uint64_t* pool;          // MVCC: version nodes, 3 u64 each
  uint64_t pool_cap;       // MVCC: node capacity
  uint32_t* counters;      // gacco/gputx: per-row completed-access counter, at the row's slot
  uint32_t slot_shift;     // slot_shift(rows)
  const uint32_t* need;    // gacco/gputx: per-op access-group start
  uint64_t* ctrs;          // ticket, pool top and timestamp shards (CTR_* word offsets)
  // outputs
  uint32_t* rows_out;      // per op: row the index returned
  const uint32_t* rows_in; // per op: row from a probe pass before the launch (null: probe in the kernel)
  uint64_t* reads;         // per op: value returned by a read
  uint32_t* obatch;        // per tx: batch id
  uint64_t* okey0;         // per tx: serialization key (primary)
  uint64_t* okey1;         // per tx: serialization key (secondary)
  uint32_t* tx_aborts;     // per tx: aborted attempts
  Metrics* metrics;
  uint64_t seed;
  uint32_t batch_id;
  uint32_t backoff_ticks;  // randomized backoff scale after an abort (0 = none)
  const uint32_t* tx_rank; // gputx: per-tx rank (okey0)
  uint64_t watchdog_ticks; // a wave gives up after this many ticks (0 = never)
  uint64_t* dbg;           // optional: 4 u64 per op, written for transactions the watchdog stopped
  uint64_t persist_groups; // persistent form: at most this many transactions in flight (0 = device-filling)
  uint32_t open;           // persistent form: slots are admitted while the launch runs (CTR_ADMIT0)
  uint32_t persist;        // 1 in persistent launches (slots from CTR_SLOT, gates, urgent state)
  uint32_t urgent_attempts; // persistent form: aborts that make a transaction urgent (0 = never)
  uint32_t backoff_exp;     // retry delay spans at most 2^backoff_exp units
  uint32_t poll_ticks;      // ordered schemes: expected ticks per handoff on a row (rest pacing)
  uint64_t* tcommit;       // optional, per tx: wall-clock tick of its commit (serving form)
  uint64_t* ev_op;         // event-log kernels: per op, the stamp of its access (0 = none)
  uint64_t* ev_commit;     // event-log kernels: per tx, the stamp of its commit (0 = none)
  const uint32_t* op_param; // TPC-C kernels: per op, the update's parameter
};

// ---------------------------------------------------------------------------
// TPC-C (benchmark/tpcc.cu, include/tpcc.cuh): payment and new_order over the
// tables that carry concurrency control, in one row space of TPCC_COLS u64
// cells per row:
//   warehouse w                 row w                    (W rows)
//   district (w, d)             W + w * 10 + d           (10 W)
//   customer (w, d, c)          11 W + (w * 10 + d) * 3000 + c   (30,000 W)
//   stock (w, i)                30,011 W + w * 100,000 + i       (100,000 W)
// Item is read-only (a price per item, no concurrency control); the insert
// tables (history, order, new_order, order_line) are derived from each
// transaction's inputs and its ops' results.  Each 32-bit field sits in one
// half of a cell (low half first):
//   warehouse  c0 {id, tax}  c1 {ytd, -}
//   district   c0 {id, w_id}  c1 {tax, ytd}  c2 {next_o_id, -}
//   customer   c0 {id, d_id}  c1 {w_id, since}  c2 {credit_lim, discount}
//              c3 {balance, ytd_payment}  c4 {payment_cnt, delivery_cnt}
//   stock      c0 {i_id, w_id}  c1 {quantity, ytd}  c2 {order_cnt, remote_cnt}
// An op reads its row, applies its update to it and writes it back while the
// scheme gives it exclusive access (write lock, uncommitted mark, locked write
// set, or its turn); the op's result is the field value the reference's
// transaction reads back (next_o_id, the stock quantity) or the float it
// updated first.
// ---------------------------------------------------------------------------
constexpr uint32_t TPCC_DISTRICTS = 10;
constexpr uint32_t TPCC_CUSTOMERS = 3000;
constexpr uint32_t TPCC_ITEMS = 100000;
constexpr uint32_t TPCC_COLS = 5;
constexpr uint32_t TPCC_MAX_LINES = 15;
GC_HD uint64_t tpcc_rows(uint64_t w) { return w * (1 + TPCC_DISTRICTS + TPCC_DISTRICTS * TPCC_CUSTOMERS + TPCC_ITEMS); }
GC_HD uint64_t tpcc_district_row(uint64_t W, uint64_t w, uint64_t d) { return W + w * TPCC_DISTRICTS + d; }
GC_HD uint64_t tpcc_customer_row(uint64_t W, uint64_t w, uint64_t d, uint64_t c) {
  return W * (1 + TPCC_DISTRICTS) + (w * TPCC_DISTRICTS + d) * TPCC_CUSTOMERS + c;
}
GC_HD uint64_t tpcc_stock_row(uint64_t W, uint64_t w, uint64_t i) {
  return W * (1 + TPCC_DISTRICTS + TPCC_DISTRICTS * TPCC_CUSTOMERS) + w * TPCC_ITEMS + i;
}
// op kinds (the op's column field, below TPCC_COLS)
enum TpccOp : uint32_t {
  TPCC_PAY_W = 0,  // warehouse ytd += amount                        (tpcc.cu:89-110)
  TPCC_PAY_D = 1,  // district ytd += amount                         (tpcc.cu:112-134)
  TPCC_PAY_C = 2,  // customer balance -= amount, ytd_payment += amount, payment_cnt += 1 (tpcc.cu:136-163)
  TPCC_NO_D = 3,   // district next_o_id += 1, result the old value  (tpcc.cu:255-283)
  TPCC_NO_S = 4,   // stock quantity / ytd / order_cnt / remote_cnt  (tpcc.cu:301-338)
};
// param: the amount's float bits (payment), or quantity | remote << 16 (stock)
GC_HD uint32_t tpcc_lo(uint64_t c) { return (uint32_t)c; }
GC_HD uint32_t tpcc_hi(uint64_t c) { return (uint32_t)(c >> 32); }
GC_HD uint64_t tpcc_pack(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }
GC_HD float tpcc_f(uint32_t bits) {
  union {
    uint32_t u;
    float f;
  } x{bits};
  return x.f;
}
GC_HD uint32_t tpcc_bits(float f) {
  union {
    float f;
    uint32_t u;
  } x{f};
  return x.u;
}
// Applies op `kind` with `param` to the row's cells r[0..TPCC_COLS) in place
// and returns the op's result.  Float updates are single float additions in
// the reference's order, so a serial replay reproduces them bit for bit.
GC_HD uint64_t tpcc_apply(uint32_t kind, uint32_t param, uint64_t* r) {
  switch (kind) {
    case TPCC_PAY_W: {
      const uint32_t old = tpcc_lo(r[1]);
      r[1] = tpcc_pack(tpcc_bits(tpcc_f(old) + tpcc_f(param)), tpcc_hi(r[1]));
      return old;
    }
    case TPCC_PAY_D: {
      const uint32_t old = tpcc_hi(r[1]);
      r[1] = tpcc_pack(tpcc_lo(r[1]), tpcc_bits(tpcc_f(old) + tpcc_f(param)));
      return old;
    }
    case TPCC_PAY_C: {
      const float amount = tpcc_f(param);
      const uint32_t bal = tpcc_lo(r[3]), ytd = tpcc_hi(r[3]);
      r[3] = tpcc_pack(tpcc_bits(tpcc_f(bal) - amount), tpcc_bits(tpcc_f(ytd) + amount));
      r[4] = tpcc_pack(tpcc_lo(r[4]) + 1u, tpcc_hi(r[4]));
      return bal;
    }
    case TPCC_NO_D: {
      const uint32_t old = tpcc_lo(r[2]);
      r[2] = tpcc_pack(old + 1u, tpcc_hi(r[2]));
      return old;
    }
    case TPCC_NO_S: {
      const uint32_t qty = param & 0xffffu, remote = param >> 16;
      const uint32_t q = tpcc_lo(r[1]);
      const uint32_t nq = q > qty + 10u ? q - qty : q - qty + 91u;
      r[1] = tpcc_pack(nq, tpcc_hi(r[1]) + qty);
      r[2] = tpcc_pack(tpcc_lo(r[2]) + 1u, tpcc_hi(r[2]) + remote);
      return q;
    }
    default:
      return 0;
  }
}

}  // namespace gcctb
