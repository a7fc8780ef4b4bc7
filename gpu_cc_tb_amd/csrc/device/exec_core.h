// Transaction-execution kernel for gfx950 (the YCSB kernel of
// benchmark/ycsb.cu:21-111, re-designed for wave64).
//
// Launch shape.  A transaction owns G = 2^warp_density consecutive lanes of a
// wave (the reference gives each transaction 2^warp_density threads and runs it
// on the first one, benchmark/ycsb.cu:32-38).  Here the G lanes cooperate: op j
// of the transaction lives on lane (j mod G) as slot (j / G), so the ops of a
// transaction are attempted in parallel and group decisions (abort, all locks
// held, validation passed) are ballots over the group's lanes.
//
// Progress on CDNA.  A wave has a single program counter, so a lane that spins
// on a word held by another lane of the same wave would never let that lane
// run.  No lane ever spins: each op step returns OK / WAIT / ABORT, and the
// wave loops in rounds over all of its groups until every transaction has
// committed.  A waiting op is simply retried in the next round.
//
// Every transaction commits exactly once; the kernel records, per transaction,
// the key of its serialization order (commit ticket, timestamp, or position)
// and, per read op, the value it returned, so that a serial replay on the CPU
// can reproduce the final table and every read.
#pragma once
#include "dev_util.h"

// Three variants are separate sets of kernels, so that the column-mode
// kernels of namespace dev carry none of their code: whole-row mode (every op
// moves its whole row, the reference's Item copies, benchmark/ycsb.cu:57,
// 86-91), the debug event log (the reference's AddEvent,
// include/runtime.cuh:64-81) and the latch schemes (the reference's slow_*
// schemes, include/slow_twopl.cuh etc.: every access to a row's CC word is
// made under the row's latch).  A translation unit that defines
// GCCTB_ROW_OPS and/or GCCTB_EVENTS, or GCCTB_LATCH, to 1 compiles this file
// into namespace dev_row, dev_evt, dev_row_evt or dev_latch.
#ifndef GCCTB_ROW_OPS
#define GCCTB_ROW_OPS 0
#endif
#ifndef GCCTB_EVENTS
#define GCCTB_EVENTS 0
#endif
#ifndef GCCTB_LATCH
#define GCCTB_LATCH 0
#endif
// TPC-C (GCCTB_TPCC, with GCCTB_ROW_OPS for whole-row versions): every op is a
// read-modify-write of its row (tpcc_apply), namespace dev_tpcc.
#ifndef GCCTB_TPCC
#define GCCTB_TPCC 0
#endif

namespace gcctb {
#if GCCTB_TPCC
namespace dev_tpcc {
using namespace dev;
#elif GCCTB_LATCH
namespace dev_latch {
using namespace dev;
#elif GCCTB_ROW_OPS && GCCTB_EVENTS
namespace dev_row_evt {
using namespace dev;
#elif GCCTB_ROW_OPS
namespace dev_row {
using namespace dev;
#elif GCCTB_EVENTS
namespace dev_evt {
using namespace dev;
#else
namespace dev {
#endif

constexpr bool kRowOps = GCCTB_ROW_OPS != 0;
constexpr bool kEvents = GCCTB_EVENTS != 0;
constexpr bool kLatch = GCCTB_LATCH != 0;
constexpr bool kTpcc = GCCTB_TPCC != 0;
static_assert(!(kLatch && (kRowOps || kEvents)), "the latch schemes run in column mode without the event log");
static_assert(!kTpcc || (kRowOps && !kEvents && !kLatch), "the TPC-C kernels move whole rows");

// per-op slot flags
constexpr uint32_t M_VALID = 1u;   // op exists and its key was found in the index
constexpr uint32_t M_WRITE = 2u;
constexpr uint32_t M_DONE = 4u;    // op finished in the current attempt
constexpr uint32_t M_HELD = 8u;    // op holds a lock / uncommitted mark
constexpr uint32_t M_HOT = 16u;    // op's key is hot (2PL locks it after the others)
constexpr uint32_t M_COL_SHIFT = 8;

// transaction phases
constexpr uint32_t PH_START = 0;   // (re)start an attempt
constexpr uint32_t PH_OPS = 1;     // per-op work (locks / reads / marks)
constexpr uint32_t PH_WRITES = 2;  // MVCC: write marks after every read finished
constexpr uint32_t PH_LOCK = 3;    // OCC: lock the write set
constexpr uint32_t PH_COMMIT = 4;

enum StepResult : uint32_t { R_OK = 0, R_WAIT = 1, R_ABORT = 2 };

// Inline CAS retries before an op reports WAIT and yields to the next round.
constexpr int CAS_TRIES = 8;

template <int G>
struct TxCtx {
  static constexpr int K = (MAX_OPS + G - 1) / G;
  uint64_t tx = 0;         // transaction id
  uint64_t op0 = 0;        // index of op 0 in the CSR op array
  uint32_t nops = 0;
  uint32_t ph = PH_START;
  uint32_t attempts = 0;   // aborted attempts
  bool live = false;
  bool done = true;
  uint64_t ts = 0;         // attempt timestamp (T/O, MVCC) or commit timestamp
  uint64_t t_start = 0, t_attempt = 0, t_resume = 0, t_round = 0;
  uint64_t m_ts = 0, m_wait = 0, m_abort = 0, m_idx = 0, m_tot = 0, m_commit = 0, m_att = 0, m_miss = 0;
  uint32_t row[K];
  uint32_t meta[K];
  uint64_t st[K];
};

// A row's CC word, MVCC chain head and turn counter (at its slot).  The
// latch schemes give each slot 16 bytes: the CC word, then the row's latch.
__device__ __forceinline__ uint64_t* cc_word(const ExecArgs& a, uint32_t row) {
  return a.cc + ((uint64_t)row_slot(row, a.slot_shift) << (kLatch ? 1 : 0));
}

// Accesses to a CC word.  In the latch schemes each read or read-modify-write
// runs under the row's latch, taken with a CAS and dropped with a store, as
// the reference's slow_* schemes guard their metadata (common::latch_lock /
// latch_unlock, include/runtime.cuh:91-112).  The access inside stays the same
// atomic one, so correctness never rests on the latch: a lane that fails
// LATCH_TRIES times to take it makes the access without it, and no lane ever
// waits on a latch beyond that (a lane spinning on a latch held by another
// lane of its own wave would wait forever).  Each attempt tests the latch
// with a load before its CAS and pauses 512 cycles after a miss.  Plain
// stores (unlock, publish) go without the latch, as in the reference.
constexpr int LATCH_TRIES = 16;
__device__ __forceinline__ uint32_t* latch_of(uint64_t* w) { return reinterpret_cast<uint32_t*>(w + 1); }
template <class F>
__device__ __forceinline__ auto latched(uint64_t* w, F f) {
  if constexpr (!kLatch) {
    return f();
  } else {
    uint32_t* l = latch_of(w);
    bool held = false;
    // test and test-and-set: a CAS only when the latch reads free, and a
    // pause after a miss, so that the waiters of a hot row do not flood its
    // line with failing CASes that the holder's release queues behind
#pragma unroll 1
    for (int t = 0; t < LATCH_TRIES && !held; ++t) {
      uint32_t free_latch = 0;
      held = ld32(l) == 0u && cas32(l, free_latch, 1u);
      if (!held) __builtin_amdgcn_s_sleep(8);
    }
    if (held) order_acquire();
    const auto r = f();
    if (held) {
      order_release();
      st32(l, 0u);
    }
    return r;
  }
}
__device__ __forceinline__ uint64_t w_ld(uint64_t* w) {
  return latched(w, [&] { return ld(w); });
}
__device__ __forceinline__ bool w_cas(uint64_t* w, uint64_t& cur, uint64_t nxt) {
  return latched(w, [&] { return cas(w, cur, nxt); });
}
__device__ __forceinline__ uint64_t w_or(uint64_t* w, uint64_t bits) {
  return latched(w, [&] { return fetch_or(w, bits); });
}
__device__ __forceinline__ uint64_t w_sub(uint64_t* w, uint64_t x) {
  return latched(w, [&] { return fetch_sub(w, x); });
}
__device__ __forceinline__ uint32_t* chain_head(const ExecArgs& a, uint32_t row) {
  return a.vprev + row_slot(row, a.slot_shift);
}
__device__ __forceinline__ uint32_t* turn_counter(const ExecArgs& a, uint32_t row) {
  return a.counters + row_slot(row, a.slot_shift);
}

__device__ __forceinline__ uint64_t* cell_ptr(const ExecArgs& a, uint32_t row, uint32_t meta) {
  return a.cells + (uint64_t)row * a.cols + (meta >> M_COL_SHIFT);
}

// The data of one op.  Column mode: the op's cell.  Whole-row mode
// (kRowOps): all cells of the row, loaded together and folded into the row
// digest, or all stored with write_cell values.
__device__ __forceinline__ uint64_t row_digest(const uint64_t* r, uint32_t cols) {
  uint64_t v[ROW_MAX_COLS];
#pragma unroll
  for (int i = 0; i < ROW_MAX_COLS; ++i) v[i] = (uint32_t)i < cols ? ld(r + i) : 0;
  uint64_t h = DIGEST_INIT;
#pragma unroll
  for (int i = 0; i < ROW_MAX_COLS; ++i)
    if ((uint32_t)i < cols) h = digest_step(h, v[i]);
  return h;
}
__device__ __forceinline__ uint64_t op_load(const ExecArgs& a, uint32_t row, uint32_t meta) {
  if constexpr (!kRowOps)
    return ld(cell_ptr(a, row, meta));
  else
    return row_digest(a.cells + (uint64_t)row * a.cols, a.cols);
}
// A write takes effect: the op's value (column mode), its whole row (row
// mode), or, in the TPC-C kernels, its update applied to the row it reads
// here, under the exclusive access the scheme holds; the update's result goes
// to reads[op], op = the op's index in the op list.
__device__ __forceinline__ void op_store(const ExecArgs& a, uint32_t row, uint32_t meta, uint64_t tx, uint32_t j,
                                         uint64_t op) {
  if constexpr (kTpcc) {
    uint64_t* r = a.cells + (uint64_t)row * TPCC_COLS;
    uint64_t v[TPCC_COLS];
#pragma unroll
    for (uint32_t i = 0; i < TPCC_COLS; ++i) v[i] = ld(r + i);
    const uint64_t res = tpcc_apply(meta >> M_COL_SHIFT, a.op_param[op], v);
#pragma unroll
    for (uint32_t i = 0; i < TPCC_COLS; ++i) st(r + i, v[i]);
    a.reads[op] = res;
  } else if constexpr (!kRowOps) {
    st(cell_ptr(a, row, meta), write_value(a.seed, tx, j));
  } else {
    uint64_t* r = a.cells + (uint64_t)row * a.cols;
#pragma unroll
    for (int i = 0; i < ROW_MAX_COLS; ++i)
      if ((uint32_t)i < a.cols) st(r + i, write_cell(a.seed, tx, j, (uint32_t)i));
  }
}

template <int G>
__device__ __forceinline__ uint32_t op_index(int k) {
  return (lane_id() & (G - 1)) + (uint32_t)k * G;
}

// Event log (kEvents): an op's event is stamped when its access takes effect,
// a transaction's commit event after all of its ops' events (the group's
// first lane stamps it after the group's ops were stamped in program order).
// Stamps come from one global counter, like AddEvent's (runtime.cuh:64-81);
// a retry overwrites the slots of the attempt before, so the log holds the
// committed attempt.
template <int G>
__device__ __forceinline__ void ev_op(const TxCtx<G>& c, const ExecArgs& a, int k) {
  if constexpr (kEvents) a.ev_op[c.op0 + op_index<G>(k)] = fetch_add(a.ctrs + CTR_EVENT, 1) + 1;
}
template <int G>
__device__ __forceinline__ uint64_t ev_draw(const ExecArgs& a) {
  if constexpr (kEvents) return fetch_add(a.ctrs + CTR_EVENT, 1) + 1;
  return 0;
}
template <int G>
__device__ __forceinline__ void ev_op_at(const TxCtx<G>& c, const ExecArgs& a, int k, uint64_t stamp) {
  if constexpr (kEvents) a.ev_op[c.op0 + op_index<G>(k)] = stamp;
}
template <int G>
__device__ __forceinline__ void ev_commit(const TxCtx<G>& c, const ExecArgs& a) {
  if constexpr (kEvents) {
    if ((lane_id() & (G - 1)) == 0) a.ev_commit[c.tx] = fetch_add(a.ctrs + CTR_EVENT, 1) + 1;
  }
}

// Abort bookkeeping, identical on every lane of the group (the wall clock is a
// scalar read, so every lane of the wave sees the same value).
//
// Retry backoff.  The lanes of a wave attempt in lockstep, so two conflicting
// transactions of one wave that restart in the same round collide again in the
// same order; an immediate retry (benchmark/ycsb.cu:99-104) can then repeat
// forever.  The restart is delayed by a random number of units in
// [0, 2^min(attempts, backoff_exp)) (backoff_exp 10 unless configured), where
// a unit is the duration of the failed attempt
// (at least backoff_ticks, at most BACKOFF_UNIT_MAX): retries spread over
// rounds, not just over ticks.  The cap matters for attempts that waited
// (wait-die, T/O reads of uncommitted rows): their duration includes the
// wait, and 1024 such units would idle a transaction for seconds.
// Urgent state (persistent form).  A transaction that has aborted
// ExecArgs::urgent_attempts times (256 by default: a safety net, far beyond
// the abort counts of a healthy run) adds itself to the urgent count of every
// gate copy until it commits; while the count is non-zero no group takes a
// new slot, so the transactions in flight drain as in a plain batch and the
// urgent one finishes.  Without it a stream that keeps refilling never lowers
// contention and a no-wait transaction can lose forever.

__device__ __forceinline__ void urgent_add(const ExecArgs& a, uint64_t delta) {
#pragma unroll 1
  for (int s = 0; s < GATE_SHARDS; ++s) fetch_add(a.ctrs + CTR_GATE0 + 16 * s + 1, delta);
}

constexpr uint64_t BACKOFF_UNIT_MAX = 5000;  // 50 us of the 100 MHz wall clock

template <int G>
__device__ __forceinline__ void abort_attempt(TxCtx<G>& c, const ExecArgs& a, uint64_t now) {
  (void)now;
  const uint64_t t = now_ticks();
  const uint64_t dur = t - c.t_attempt;
  c.attempts++;
  if (a.persist && c.attempts == a.urgent_attempts && (lane_id() & (G - 1)) == 0) urgent_add(a, 1);
  c.m_abort += dur;
  uint64_t unit = dur > a.backoff_ticks ? dur : a.backoff_ticks;
  if (unit > BACKOFF_UNIT_MAX) unit = BACKOFF_UNIT_MAX;
  const uint64_t h = hash3(a.seed ^ TAG_BACKOFF, c.tx, c.attempts);
  const uint32_t span = 1u << (c.attempts < a.backoff_exp ? c.attempts : a.backoff_exp);
  c.t_resume = t + (h % span) * unit;
  c.ph = PH_START;
}

template <int G>
__device__ __forceinline__ void finish_tx(TxCtx<G>& c, const ExecArgs& a, uint64_t now,
                                          uint64_t key0, uint64_t key1) {
  c.done = true;
  c.m_tot += now - c.t_start;
  c.m_commit += 1;
  c.m_att += c.attempts;
  if ((lane_id() & (G - 1)) == 0) {
    if (a.persist && a.urgent_attempts && c.attempts >= a.urgent_attempts)
      urgent_add(a, ~0ull);  // leave the urgent count
    a.obatch[c.tx] = a.batch_id;
    if (a.tcommit) a.tcommit[c.tx] = now_ticks();
    a.okey0[c.tx] = key0;
    a.okey1[c.tx] = key1;
    a.tx_aborts[c.tx] = c.attempts;
  }
}

// Start a new attempt: clear per-op state.
template <int G>
__device__ __forceinline__ void begin_attempt(TxCtx<G>& c, uint64_t now) {
  c.t_attempt = now;
#pragma unroll
  for (int k = 0; k < TxCtx<G>::K; ++k) c.meta[k] &= ~(M_DONE | M_HELD);
}

template <int G>
__device__ __forceinline__ bool backoff_over(const TxCtx<G>& c, uint64_t now) {
  return now >= c.t_resume;
}

// Timestamp for a starting attempt (NaiveTSAllocator::Alloc,
// include/timestamp.cuh:22-25): wave-aggregated, from the block's shard.
// Shard s of TS_SHARDS issues s, s + TS_SHARDS, ...; the allocators restart at
// 1 every batch (timestamp.cuh:76-80), so the smallest timestamp is TS_SHARDS.
template <int G>
__device__ __forceinline__ uint64_t alloc_ts(const ExecArgs& a, bool start) {
  const uint32_t shard = blockIdx.x & (TS_SHARDS - 1);
  const uint64_t local = wave_fetch_inc(a.ctrs + CTR_TS0 + 16 * shard, start && (lane_id() & (G - 1)) == 0);
  return (local << TS_SHARD_BITS) | shard;
}

// After an abort caused by a younger timestamp `seen`, move this block's
// shard past it, so that the retry's timestamp is younger than what it lost
// to.  Without it a slow shard's transactions could keep losing to the
// timestamps of a faster shard (a single counter never lags itself).
template <int G>
__device__ __forceinline__ void raise_ts_floor(const ExecArgs& a, uint64_t seen) {
  if (seen == 0 || (lane_id() & (G - 1)) != 0) return;
  const uint32_t shard = blockIdx.x & (TS_SHARDS - 1);
  __hip_atomic_fetch_max(a.ctrs + CTR_TS0 + 16 * shard, (seen >> TS_SHARD_BITS) + 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// ===========================================================================
// Two-phase locking, no-wait and wait-die (include/twopl.cuh:59-382).
// Strict 2PL: shared locks for reads, exclusive locks for writes, writes are
// applied at commit while every lock is held, locks released after.  The
// commit ticket is drawn while all locks are held, so ticket order is a valid
// serial order.
// ===========================================================================
template <bool WAIT_DIE>
struct TwoPL {
  static constexpr bool kRestInOps = false;  // t_resume paces restarts only (PH_START)
  static constexpr bool kBulkRefill = true;
  // lock_shared (twopl.cuh:233-267), continuing from an observed word `cur`
  // (the round's first CAS guessed a free word and returned it).
  static __device__ StepResult lock_shared(uint64_t* w, uint64_t self, uint64_t cur) {
#pragma unroll 1
    for (int t = 0; t < CAS_TRIES; ++t) {
      if (tpl::excl(cur)) {
        const uint64_t h = tpl::holder(cur);
        if (h == self) return R_OK;
        if (!WAIT_DIE) return R_ABORT;
        if (h < self) return R_ABORT;  // die: an older transaction holds it
        return R_WAIT;                 // wait for a younger holder
      }
      if (w_cas(w, cur, cur + tpl::CNT_ONE)) return R_OK;
    }
    return R_WAIT;
  }
  // lock_exclusive (twopl.cuh:269-303); a shared-held word always aborts
  // (twopl.cuh:285-286).
  static __device__ StepResult lock_excl(uint64_t* w, uint64_t self, uint64_t cur) {
#pragma unroll 1
    for (int t = 0; t < CAS_TRIES; ++t) {
      if (cur == 0) {
        if (w_cas(w, cur, tpl::make_excl(self))) return R_OK;
        continue;
      }
      if (!tpl::excl(cur)) return R_ABORT;
      const uint64_t h = tpl::holder(cur);
      if (h == self) return R_OK;
      if (!WAIT_DIE) return R_ABORT;
      if (h < self) return R_ABORT;
      return R_WAIT;
    }
    return R_WAIT;
  }
  // unlock_shared drops one holder (twopl.cuh:305-315), the last one leaves 0;
  // unlock_exclusive frees the word (twopl.cuh:317-322).
  static __device__ void unlock(uint64_t* w, bool excl) {
    if (excl)
      st(w, 0);
    else
      w_sub(w, tpl::CNT_ONE);
  }

  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    const uint64_t now = now_ticks();
    bool commit = false;
    if (!c.done) {
      if (c.ph == PH_START && backoff_over(c, now)) {
        begin_attempt(c, now);
        c.ph = PH_OPS;
      }
      if (c.ph == PH_OPS) {
        bool conflict = false, pending = false;
        // First lock attempt of every unheld op at once: a CAS that guesses
        // a free word, all issued before any result is examined, so an
        // uncontended transaction locks in one atomic latency.
        // Hot keys (M_HOT) are locked only once every other lock is held,
        // so a hot lock is held for the commit alone.
        bool cold = false;
#pragma unroll
        for (int k = 0; k < K; ++k) cold |= (c.meta[k] & (M_VALID | M_HELD | M_HOT)) == M_VALID;
        const bool cold_left = group_any<G>(cold);
        uint64_t cur[K];
        bool got[K], tried[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const uint32_t m = c.meta[k];
          cur[k] = 0;
          got[k] = false;
          tried[k] = (m & M_VALID) && !(m & M_HELD) && !(cold_left && (m & M_HOT));
          pending |= (m & M_VALID) && !(m & M_HELD) && !tried[k];
          if (tried[k])
            got[k] = w_cas(cc_word(a, c.row[k]), cur[k], (m & M_WRITE) ? tpl::make_excl(c.tx) : tpl::CNT_ONE);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const uint32_t m = c.meta[k];
          if (tried[k]) {
            uint64_t* w = cc_word(a, c.row[k]);
            StepResult r = R_OK;
            if (!got[k]) r = (m & M_WRITE) ? lock_excl(w, c.tx, cur[k]) : lock_shared(w, c.tx, cur[k]);
            if (r == R_OK)
              c.meta[k] = m | M_HELD;
            else if (r == R_ABORT)
              conflict = true;
            else
              pending = true;
          }
        }
        if (group_any<G>(conflict)) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (c.meta[k] & M_HELD) unlock(cc_word(a, c.row[k]), c.meta[k] & M_WRITE);
          abort_attempt(c, a, now);
        } else if (group_all<G>(!pending)) {
          c.ph = PH_COMMIT;
        } else {
          c.m_wait += now - c.t_round;
        }
      }
      commit = (c.ph == PH_COMMIT);
    }
    // Every lock is held: the cell accesses and the commit ticket both fall
    // inside that window in either order (a conflicting transaction locks
    // only after our unlocks, so its ticket is later), so they are issued
    // together and waited for once.
    uint64_t v[K];
    if (commit) {
      order_acquire();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint32_t m = c.meta[k];
        v[k] = 0;
        if (!(m & M_VALID)) continue;
        if (m & M_WRITE)
          op_store(a, c.row[k], m, c.tx, op_index<G>(k), c.op0 + op_index<G>(k));
        else
          v[k] = op_load(a, c.row[k], m);
      }
    }
    uint64_t ticket = wave_fetch_inc(a.ctrs + CTR_TICKET, commit && (lane_id() & (G - 1)) == 0);
    if (commit) {
      ticket = group_bcast64<G>(ticket);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint32_t m = c.meta[k];
        if ((m & M_VALID) && !(m & M_WRITE)) a.reads[c.op0 + op_index<G>(k)] = v[k];
      }
      if constexpr (kEvents) {
        // every lock is held: the accesses take effect now, the commit before the unlocks
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (c.meta[k] & M_VALID) ev_op<G>(c, a, k);
        ev_commit<G>(c, a);
      }
      order_release();
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (c.meta[k] & M_VALID) unlock(cc_word(a, c.row[k]), c.meta[k] & M_WRITE);
      finish_tx(c, a, now, ticket, 0);
    }
  }
};

// ===========================================================================
// Basic timestamp ordering (include/to.cuh:57-271).  Fresh timestamp per
// attempt from the wave-aggregated allocator.  Read: abort if wts > ts, wait
// while uncommitted, else read and raise rts by CAS.  Write: abort if rts > ts,
// wts > ts or uncommitted, else mark {uncommitted, ts, ts}.  Commit installs
// the data and clears the mark; abort restores the saved words.  The
// serialization order is timestamp order.
// ===========================================================================
// After a cell of version `ver` was read: leave rts >= ts on that same
// version.  `cur` is a word observed before the cell load (fresh = false) or
// after it (fresh = true).  Other readers raising rts meanwhile leave the
// version, and the value read, valid, so only a change of the version bits
// sends the caller back to re-read (false; cur then holds the current word).
__device__ __forceinline__ bool hold_version(uint64_t* w, uint64_t ts, uint64_t ver, uint64_t& cur, bool fresh) {
#pragma unroll 1
  for (int t = 0; t < CAS_TRIES; ++t) {
    if (fresh && tsw::version(cur) != ver) return false;
    if (tsw::rts(cur) >= ts) {
      if (fresh) return true;
      cur = w_ld(w);
      fresh = true;
      continue;
    }
    if (w_cas(w, cur, tsw::make(0, ts, tsw::wts(cur)))) return true;
    fresh = true;
  }
  return false;
}

// Head reads of T/O and MVCC for every pending read of the lane's slots,
// each step issued for all slots before its wait: the word; then the cell of
// every slot whose head is committed and old enough (wts <= ts); then the
// first rts raise (CAS) or re-check of each.  res[k]: R_OK (v[k] read and
// held), R_WAIT (uncommitted head, or the version moved: retry next round) or
// R_ABORT (head younger than ts: T/O aborts, MVCC walks the chain instead);
// cur[k] is the last word observed.
template <int G>
__device__ __forceinline__ void ts_head_reads(const TxCtx<G>& c, const ExecArgs& a, uint64_t ts,
                                              uint64_t (&cur)[TxCtx<G>::K], uint64_t (&v)[TxCtx<G>::K],
                                              bool (&rd)[TxCtx<G>::K], StepResult (&res)[TxCtx<G>::K],
                                              uint64_t (&stamp)[TxCtx<G>::K]) {
  constexpr int K = TxCtx<G>::K;
  bool cand[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t m = c.meta[k];
    rd[k] = (m & M_VALID) && !(m & (M_DONE | M_WRITE));
    cur[k] = rd[k] ? w_ld(cc_word(a, c.row[k])) : 0;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    res[k] = R_WAIT;
    cand[k] = false;
    if (!rd[k]) continue;
    if (tsw::unc(cur[k])) continue;
    if (tsw::wts(cur[k]) > ts) {
      res[k] = R_ABORT;
      continue;
    }
    cand[k] = true;
  }
  // event log: a head read takes effect between the word it saw committed
  // and the rts raise that holds that version (a writer that marks the row
  // later is stamped after its mark)
#pragma unroll
  for (int k = 0; k < K; ++k) stamp[k] = cand[k] ? ev_draw<G>(a) : 0;
  order_acquire();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = cand[k] ? op_load(a, c.row[k], c.meta[k]) : 0;
  order_acquire();
  uint64_t ver[K];
  bool ok[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    ver[k] = tsw::version(cur[k]);
    ok[k] = false;
    if (!cand[k]) continue;
    if (tsw::rts(cur[k]) >= ts)
      cur[k] = w_ld(cc_word(a, c.row[k]));  // nothing to raise: re-check the version
    else
      ok[k] = w_cas(cc_word(a, c.row[k]), cur[k], tsw::make(0, ts, tsw::wts(cur[k])));
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!cand[k]) continue;
    if (!ok[k]) ok[k] = hold_version(cc_word(a, c.row[k]), ts, ver[k], cur[k], true);
    res[k] = ok[k] ? R_OK : R_WAIT;
  }
}

struct TimestampOrdering {
  static constexpr bool kRestInOps = false;  // t_resume paces restarts only (PH_START)
  static constexpr bool kBulkRefill = true;
  static __device__ StepResult write_mark(uint64_t* w, uint64_t ts, uint64_t& prev, uint64_t& seen) {
    uint64_t cur = w_ld(w);
#pragma unroll 1
    for (int t = 0; t < CAS_TRIES; ++t) {
      if (tsw::rts(cur) > ts || tsw::wts(cur) > ts || tsw::unc(cur)) {
        seen = tsw::rts(cur) > tsw::wts(cur) ? tsw::rts(cur) : tsw::wts(cur);
        return R_ABORT;
      }
      prev = cur;
      if (w_cas(w, cur, tsw::make(1, ts, ts))) return R_OK;
    }
    return R_WAIT;
  }

  // Reads first, write marks second: an attempt never waits while it holds
  // an uncommitted mark, so no chain of waits forms through it.
  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    const uint64_t now = now_ticks();
    const bool start = !c.done && c.ph == PH_START && backoff_over(c, now);
    uint64_t ts = alloc_ts<G>(a, start);
    if (start) {
      c.ts = group_bcast64<G>(ts);
      c.m_ts += now_ticks() - now;
      if (c.ts > tsw::TS_MAX) {
        if ((lane_id() & (G - 1)) == 0) atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_TS_OVERFLOW);
        finish_tx(c, a, now, c.ts, 0);
        return;
      }
      begin_attempt(c, now);
      c.ph = PH_OPS;
    }
    if (c.done) return;
    uint64_t seen = 0;
    bool abort = false;
    if (c.ph == PH_OPS) {
      bool pending = false;
      uint64_t cur[K], v[K], stamp[K];
      bool rd[K];
      StepResult r[K];
      ts_head_reads<G>(c, a, c.ts, cur, v, rd, r, stamp);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!rd[k]) continue;
        if (r[k] == R_OK) {
          c.meta[k] |= M_DONE;
          a.reads[c.op0 + op_index<G>(k)] = v[k];
          ev_op_at<G>(c, a, k, stamp[k]);
        } else if (r[k] == R_ABORT) {
          // a younger write committed (to.cuh:148-152)
          const uint64_t y = tsw::wts(cur[k]);
          seen = y > seen ? y : seen;
          abort = true;
        } else {
          pending = true;
        }
      }
      if (!group_any<G>(abort)) {
        if (group_all<G>(!pending))
          c.ph = PH_WRITES;
        else
          c.m_wait += now - c.t_round;
      }
    }
    if (!abort && c.ph == PH_WRITES) {
      bool pending = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint32_t m = c.meta[k];
        if (!(m & M_VALID) || !(m & M_WRITE) || (m & M_DONE)) continue;
        const StepResult r = write_mark(cc_word(a, c.row[k]), c.ts, c.st[k], seen);
        if (r == R_OK) {
          c.meta[k] = m | M_DONE | M_HELD;
          ev_op<G>(c, a, k);
        } else if (r == R_ABORT)
          abort = true;
        else
          pending = true;
      }
      if (!group_any<G>(abort) && !group_all<G>(!pending)) return;
    }
    if (group_any<G>(abort)) {
      // rollback (to.cuh:255-270): restore the words this attempt marked
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), c.st[k]);
      raise_ts_floor<G>(a, group_max64<G>(seen));
      abort_attempt(c, a, now);
      return;
    }
    if (c.ph != PH_WRITES) return;
    // commit (to.cuh:117-134)
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) op_store(a, c.row[k], c.meta[k], c.tx, op_index<G>(k), c.op0 + op_index<G>(k));
    ev_commit<G>(c, a);
    order_release();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), tsw::make(0, c.ts, c.ts));
    finish_tx(c, a, now, c.ts, 0);
  }
};

// ===========================================================================
// Multi-version timestamp ordering (include/mvcc.cuh:61-297).
// Per row: head word {uncommitted, rts, wts} + head of a version chain.  The
// table always holds the newest committed version; a commit pushes the old
// value of the written column into a node of the version pool (device bump
// allocator).  A read at ts returns the newest version with wts <= ts and never
// aborts; it waits while the head is uncommitted (mvcc.cuh:169-203).  Writes
// abort like T/O (mvcc.cuh:240-241).  Reads run before write marks so that an
// attempt never waits while it holds an uncommitted mark.
// Node layout (mvcc_stride u64): [0] ts word of the superseded version,
// [1] previous node + 1 | column << 32, [2..] the superseded data: the old
// value of the written column, or in whole-row mode the old row (cols cells).
// ===========================================================================
constexpr uint64_t MVCC_COL_STRIDE = 3;
__device__ __forceinline__ uint64_t mvcc_stride(const ExecArgs& a) { return kRowOps ? 2 + (uint64_t)a.cols : MVCC_COL_STRIDE; }
struct Mvcc {
  static constexpr bool kRestInOps = false;  // t_resume paces restarts only (PH_START)
  static constexpr bool kBulkRefill = true;
  static __device__ StepResult read(const ExecArgs& a, uint32_t row, uint32_t col, uint64_t ts,
                                    uint64_t& val) {
    uint64_t* w = cc_word(a, row);
    const uint32_t meta = col << M_COL_SHIFT;
    const uint64_t stride = mvcc_stride(a);
    uint64_t cur = w_ld(w);
#pragma unroll 1
    for (int t = 0; t < CAS_TRIES; ++t) {
      if (tsw::unc(cur)) return R_WAIT;
      if (tsw::wts(cur) <= ts) {
        order_acquire();
        val = op_load(a, row, meta);
        order_acquire();
        if (hold_version(w, ts, tsw::version(cur), cur, false)) return R_OK;
        continue;
      }
      // Walk back to the newest version with wts <= ts.  The head value and
      // the chain head are read between two observations of the same version
      // (a writer marks the word before it installs, so no install falls in
      // between; rts raises by other readers do not matter).  Nodes below the
      // chain head never change, so the walk itself needs no validation and a
      // reader far behind a hot row's head still finishes.
      order_acquire();
      uint32_t p = ld32(chain_head(a, row));
      uint64_t v = op_load(a, row, meta);
      order_acquire();
      const uint64_t after = w_ld(w);
      if (tsw::version(after) != tsw::version(cur)) {
        cur = after;
        continue;
      }
      uint64_t cur_wts = tsw::wts(cur);
      bool broken = false;
      uint32_t last = 0;  // whole-row mode: the last node walked holds the row read
#pragma unroll 1
      while (cur_wts > ts) {
        // a link read from shared memory is an index: never follow one that
        // leaves the pool (0 ends the chain; nodes are 1..pool_cap)
        if (p == 0 || p > a.pool_cap) {
          broken = true;
          break;
        }
        const uint64_t* node = a.pool + (uint64_t)(p - 1) * stride;
        const uint64_t link = ld(node + 1);
        if (!kRowOps && (uint32_t)(link >> 32) == col) v = ld(node + 2);
        last = p;
        cur_wts = tsw::wts(ld(node + 0));
        p = (uint32_t)link;
      }
      if (broken) {
        atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_CHAIN_BROKEN);
        val = 0;
        return R_OK;
      }
      if (kRowOps && last) v = row_digest(a.pool + (uint64_t)(last - 1) * stride + 2, a.cols);
      val = v;
      return R_OK;
    }
    return R_WAIT;
  }

  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    const uint64_t now = now_ticks();
    const bool start = !c.done && c.ph == PH_START && backoff_over(c, now);
    uint64_t ts = alloc_ts<G>(a, start);
    if (start) {
      c.ts = group_bcast64<G>(ts);
      c.m_ts += now_ticks() - now;
      if (c.ts > tsw::TS_MAX) {
        if ((lane_id() & (G - 1)) == 0) atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_TS_OVERFLOW);
        finish_tx(c, a, now, c.ts, 0);
      } else {
        begin_attempt(c, now);
        c.ph = PH_OPS;
      }
    }
    bool commit = false;
    if (!c.done && c.ph == PH_OPS) {
      bool pending = false;
      uint64_t cur[K], hv[K], stamp[K];
      bool rd[K];
      StepResult r[K];
      ts_head_reads<G>(c, a, c.ts, cur, hv, rd, r, stamp);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!rd[k]) continue;
        const uint32_t m = c.meta[k];
        uint64_t v = hv[k];
        // a head younger than ts: walk the version chain (never aborts)
        StepResult rk = r[k];
        uint64_t sk = stamp[k];
        if (rk == R_ABORT) {
          // a head younger than ts: the version read is older than a
          // committed write, so its stamp only has to follow that commit
          rk = read(a, c.row[k], m >> M_COL_SHIFT, c.ts, v);
          sk = rk == R_OK ? ev_draw<G>(a) : 0;
        }
        if (rk == R_OK) {
          c.meta[k] = m | M_DONE;
          a.reads[c.op0 + op_index<G>(k)] = v;
          ev_op_at<G>(c, a, k, sk);
        } else {
          pending = true;
        }
      }
      if (group_all<G>(!pending))
        c.ph = PH_WRITES;
      else
        c.m_wait += now - c.t_round;
    }
    if (!c.done && c.ph == PH_WRITES) {
      bool abort = false, pending = false;
      uint64_t seen = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint32_t m = c.meta[k];
        if (!(m & M_VALID) || !(m & M_WRITE) || (m & M_DONE)) continue;
        uint64_t* w = cc_word(a, c.row[k]);
        uint64_t cur = w_ld(w);
        StepResult r = R_WAIT;
#pragma unroll 1
        for (int t = 0; t < CAS_TRIES; ++t) {
          if (tsw::rts(cur) > c.ts || tsw::wts(cur) > c.ts || tsw::unc(cur)) {
            const uint64_t y = tsw::rts(cur) > tsw::wts(cur) ? tsw::rts(cur) : tsw::wts(cur);
            seen = y > seen ? y : seen;
            r = R_ABORT;
            break;
          }
          c.st[k] = cur;
          if (w_cas(w, cur, tsw::make(1, c.ts, c.ts))) {
            r = R_OK;
            break;
          }
        }
        if (r == R_OK) {
          c.meta[k] = m | M_DONE | M_HELD;
          ev_op<G>(c, a, k);
        } else if (r == R_ABORT)
          abort = true;
        else
          pending = true;
      }
      if (group_any<G>(abort)) {
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), c.st[k]);
        raise_ts_floor<G>(a, group_max64<G>(seen));
        abort_attempt(c, a, now);
      } else if (group_all<G>(!pending)) {
        commit = true;
      }
    }
    // version-node allocation: one wave-wide atomic
    uint32_t nw = 0;
    if (commit) {
#pragma unroll
      for (int k = 0; k < K; ++k) nw += (c.meta[k] & M_HELD) ? 1u : 0u;
    }
    uint64_t node = wave_fetch_add(a.ctrs + CTR_POOL, nw);
    if (!commit) return;
    if (node + nw > a.pool_cap) {
      atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_POOL_OVERFLOW);
    }
    // Install, in three steps for all of the lane's writes at once (this
    // attempt holds the uncommitted mark of every row it installs on):
    // version nodes and new cells; then, after a release, the chain heads;
    // then, after another release, the committed words.  A reader that sees
    // a committed word therefore sees its chain head and every node below it.
    const uint64_t stride = mvcc_stride(a);
    uint64_t oldv[K];
    uint32_t prev[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool h = (c.meta[k] & M_HELD) != 0;
      oldv[k] = h && !kRowOps ? ld(cell_ptr(a, c.row[k], c.meta[k])) : 0;
      prev[k] = h ? ld32(chain_head(a, c.row[k])) : 0;
    }
    uint64_t nd_of[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t m = c.meta[k];
      nd_of[k] = ~0ull;
      if (!(m & M_HELD)) continue;
      if (node < a.pool_cap) {
        uint64_t* nd = a.pool + node * stride;
        st(nd + 0, c.st[k]);
        st(nd + 1, (uint64_t)prev[k] | ((uint64_t)(m >> M_COL_SHIFT) << 32));
        if constexpr (!kRowOps) {
          st(nd + 2, oldv[k]);
        } else {
          // the whole superseded row (this attempt holds the row's mark)
          const uint64_t* r = a.cells + (uint64_t)c.row[k] * a.cols;
#pragma unroll
          for (int i = 0; i < ROW_MAX_COLS; ++i)
            if ((uint32_t)i < a.cols) st(nd + 2 + i, ld(r + i));
        }
        nd_of[k] = node;
      }
      node++;
    }
    if constexpr (kRowOps) order_release();  // the old row is saved before it is overwritten
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) op_store(a, c.row[k], c.meta[k], c.tx, op_index<G>(k), c.op0 + op_index<G>(k));
    order_release();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (nd_of[k] != ~0ull) st32(chain_head(a, c.row[k]), (uint32_t)(nd_of[k] + 1));
    ev_commit<G>(c, a);
    order_release();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), tsw::make(0, c.ts, c.ts));
    finish_tx(c, a, now, c.ts, 0);
  }
};

// ===========================================================================
// OCC helpers shared by Silo and TicToc (both keep a lock in bit 0 of the
// row word).
// ===========================================================================
// Consistent snapshot of every pending read of the lane's slots: the word,
// the cell, the word again; each step is issued for all slots before its
// wait, so a round costs three memory latencies however many slots read.
// ok[k]: slot k read and the word was unchanged and unlocked; w1[k] the word.
template <int G>
__device__ __forceinline__ void occ_snapshot(const TxCtx<G>& c, const ExecArgs& a, uint64_t (&w1)[TxCtx<G>::K],
                                             uint64_t (&v)[TxCtx<G>::K], bool (&rd)[TxCtx<G>::K],
                                             bool (&ok)[TxCtx<G>::K]) {
  constexpr int K = TxCtx<G>::K;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t m = c.meta[k];
    rd[k] = (m & M_VALID) && !(m & (M_DONE | M_WRITE));
    w1[k] = rd[k] ? w_ld(cc_word(a, c.row[k])) : 0;
  }
  order_acquire();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = rd[k] ? op_load(a, c.row[k], c.meta[k]) : 0;
  order_acquire();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t w2 = rd[k] ? w_ld(cc_word(a, c.row[k])) : 0;
    ok[k] = rd[k] && w1[k] == w2 && !(w1[k] & 1u);
  }
}

// Lock every unheld write of the lane's slots with one fetch_or each, all
// issued before any result is examined.  old[k]: the word before the lock;
// got[k]: the lock was free (now ours); tried[k]: slot k was attempted.
template <int G>
__device__ __forceinline__ void occ_lock(const TxCtx<G>& c, const ExecArgs& a, uint64_t (&old)[TxCtx<G>::K],
                                         bool (&tried)[TxCtx<G>::K], bool (&got)[TxCtx<G>::K]) {
  constexpr int K = TxCtx<G>::K;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t m = c.meta[k];
    tried[k] = (m & M_VALID) && (m & M_WRITE) && !(m & M_HELD);
    old[k] = tried[k] ? w_or(cc_word(a, c.row[k]), 1ull) : 0;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) got[k] = tried[k] && !(old[k] & 1u);
}

// ===========================================================================
// Silo OCC (include/silo.cuh:70-339).
// Read: consistent snapshot (word, data, word unchanged and unlocked).
// Validation: lock the write set (no-wait), draw the commit ticket, compute
// commit_ts = max(read ts, current ts of the locked write words) + 1, abort if a
// read word changed or is locked; write phase installs data and publishes
// {unlocked, commit_ts}.  The ticket drawn between locking and read validation
// is the serialization order.
// ===========================================================================
struct Silo {
  static constexpr bool kRestInOps = false;  // t_resume paces restarts only (PH_START)
  static constexpr bool kBulkRefill = true;
  // Read: the word, then the cell (silo.cuh:223-259 loads the word a second
  // time after the data).  The commit validation re-reads every read word
  // and aborts unless it still holds the TID seen here, unlocked; a writer
  // that changed the cell in between must have locked the word and installed
  // a new TID, so a torn read never commits, and the second load (one more
  // line fetched per read) is left out.
  template <int G>
  static __device__ __forceinline__ void snapshot(const TxCtx<G>& c, const ExecArgs& a, uint64_t (&w1)[TxCtx<G>::K],
                                                  uint64_t (&v)[TxCtx<G>::K], bool (&rd)[TxCtx<G>::K],
                                                  bool (&ok)[TxCtx<G>::K]) {
    constexpr int K = TxCtx<G>::K;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t m = c.meta[k];
      rd[k] = (m & M_VALID) && !(m & (M_DONE | M_WRITE));
      w1[k] = rd[k] ? w_ld(cc_word(a, c.row[k])) : 0;
    }
    order_acquire();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      ok[k] = rd[k] && !silo::locked(w1[k]);
      v[k] = ok[k] ? op_load(a, c.row[k], c.meta[k]) : 0;
    }
  }
  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    const uint64_t now = now_ticks();
    bool validate = false;
    if (!c.done) {
      if (c.ph == PH_START && backoff_over(c, now)) {
        begin_attempt(c, now);
        c.ph = PH_OPS;
      }
      if (c.ph == PH_OPS) {
        bool pending = false;
        uint64_t w1[K], v[K];
        bool rd[K], ok[K];
        snapshot<G>(c, a, w1, v, rd, ok);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!rd[k]) continue;
          if (!ok[k]) {
            pending = true;
            continue;
          }
          c.st[k] = silo::ts(w1[k]);
          c.meta[k] |= M_DONE;
          a.reads[c.op0 + op_index<G>(k)] = v[k];
        }
        if (group_all<G>(!pending))
          c.ph = PH_LOCK;
        else
          c.m_wait += now - c.t_round;
      }
      if (c.ph == PH_LOCK) {
        bool conflict = false, pending = false;
        // lock (silo.cuh:199-215) as one fetch_or: the old word says whether
        // the lock was free (then it is ours) and carries the current TID;
        // setting an already-set bit changes nothing, so no undo is needed
        uint64_t old[K];
        bool tried[K], got[K];
        occ_lock<G>(c, a, old, tried, got);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!tried[k]) continue;
          if (got[k]) {
            c.st[k] = silo::ts(old[k]);
            c.meta[k] |= M_HELD;
            ev_op<G>(c, a, k);
          } else {
            conflict = true;
          }
        }
        if (group_any<G>(conflict)) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), silo::make(0, c.st[k]));
          abort_attempt(c, a, now);
        } else if (group_all<G>(!pending)) {
          validate = true;
        }
      }
    }
    // the ticket follows the locks: their fetch_or results were consumed
    uint64_t ticket = wave_fetch_inc(a.ctrs + CTR_TICKET, validate && (lane_id() & (G - 1)) == 0);
    if (!validate) return;
    ticket = group_bcast64<G>(ticket);
    order_acquire();
    uint64_t mx = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
      if ((c.meta[k] & M_VALID) && c.st[k] > mx) mx = c.st[k];
    const uint64_t commit_ts = group_max64<G>(mx) + 1;
    bool fail = false;
    if constexpr (kEvents) {
      // event log: a committed read is current until its validation (a
      // writer that locks the row before that makes the validation fail),
      // so it is stamped here, before the validation loads
#pragma unroll
      for (int k = 0; k < K; ++k)
        if ((c.meta[k] & M_VALID) && !(c.meta[k] & M_WRITE)) ev_op<G>(c, a, k);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t m = c.meta[k];
      if (!(m & M_VALID) || (m & M_WRITE)) continue;
      const uint64_t w = w_ld(cc_word(a, c.row[k]));
      if (silo::ts(w) != c.st[k] || silo::locked(w)) fail = true;
    }
    if (group_any<G>(fail)) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), silo::make(0, c.st[k]));
      abort_attempt(c, a, now);
      return;
    }
    // write phase (silo.cuh:323-338)
    ev_commit<G>(c, a);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) op_store(a, c.row[k], c.meta[k], c.tx, op_index<G>(k), c.op0 + op_index<G>(k));
    order_release();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), silo::make(0, commit_ts));
    finish_tx(c, a, now, ticket, 0);
  }
};

// ===========================================================================
// TicToc OCC (include/tictoc.cuh:67-369).
// Read: snapshot {wts, rts = wts + delta}.  Validation: lock the write set
// (no-wait); commit_ts = max(max over writes of rts + 1, max over reads of
// wts); each read whose rts < commit_ts is re-checked (abort if wts changed or
// it is locked by another) and its rts extended to commit_ts by CAS, shifting
// wts when delta overflows 15 bits (tictoc.cuh:302-341).  The commit ticket is
// drawn after validation while the write locks are held; (commit_ts, ticket)
// is the serialization order.
// ===========================================================================
struct Tictoc {
  static constexpr bool kRestInOps = false;  // t_resume paces restarts only (PH_START)
  static constexpr bool kBulkRefill = true;
  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    const uint64_t now = now_ticks();
    bool commit = false;
    uint64_t commit_ts = 0;
    if (!c.done) {
      if (c.ph == PH_START && backoff_over(c, now)) {
        begin_attempt(c, now);
        c.ph = PH_OPS;
      }
      if (c.ph == PH_OPS) {
        bool pending = false;
        uint64_t w1[K], v[K];
        bool rd[K], ok[K];
        occ_snapshot<G>(c, a, w1, v, rd, ok);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!rd[k]) continue;
          if (!ok[k]) {
            pending = true;
            continue;
          }
          c.st[k] = w1[k];
          c.meta[k] |= M_DONE;
          a.reads[c.op0 + op_index<G>(k)] = v[k];
          ev_op<G>(c, a, k);
        }
        if (group_all<G>(!pending))
          c.ph = PH_LOCK;
        else
          c.m_wait += now - c.t_round;
      }
      if (c.ph == PH_LOCK) {
        bool conflict = false, pending = false;
        // lock (tictoc.cuh:198-215) as one fetch_or.  A reader's rts
        // extension never CASes a locked word (tictoc.cuh:316-338), so the
        // old word stays valid until this attempt unlocks or writes.
        uint64_t old[K];
        bool tried[K], got[K];
        occ_lock<G>(c, a, old, tried, got);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!tried[k]) continue;
          if (got[k]) {
            c.st[k] = old[k];
            c.meta[k] |= M_HELD;
            ev_op<G>(c, a, k);
          } else {
            conflict = true;
          }
        }
        if (group_any<G>(conflict)) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), c.st[k]);
          abort_attempt(c, a, now);
        } else if (group_all<G>(!pending)) {
          // validation (tictoc.cuh:262-350)
          order_acquire();
          uint64_t mx = 0;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const uint32_t m = c.meta[k];
            if (!(m & M_VALID)) continue;
            const uint64_t cand = (m & M_WRITE) ? tictoc::rts(c.st[k]) + 1 : tictoc::wts(c.st[k]);
            if (cand > mx) mx = cand;
          }
          commit_ts = group_max64<G>(mx);
          bool fail = false;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const uint32_t m = c.meta[k];
            if (!(m & M_VALID) || (m & M_WRITE)) continue;
            if (tictoc::rts(c.st[k]) >= commit_ts) continue;
            uint64_t* w = cc_word(a, c.row[k]);
            // first CAS from the snapshot word: if the word is unchanged it
            // succeeds without a load; otherwise it returns the current word
            uint64_t cur = c.st[k];
#pragma unroll 1
            while (true) {
              const uint64_t wts = tictoc::wts(cur);
              const uint64_t rts = tictoc::rts(cur);
              if (wts != tictoc::wts(c.st[k]) || (rts <= commit_ts && tictoc::locked(cur))) {
                fail = true;
                break;
              }
              if (rts > commit_ts) break;
              const uint64_t delta = commit_ts - wts;
              const uint64_t shift = delta - (delta & 0x7fffull);
              const uint64_t nxt = tictoc::make(tictoc::locked(cur), delta - shift, wts + shift);
              if (w_cas(w, cur, nxt)) break;
            }
          }
          if (group_any<G>(fail)) {
#pragma unroll
            for (int k = 0; k < K; ++k)
              if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), c.st[k]);
            abort_attempt(c, a, now);
          } else {
            commit = true;
          }
        }
      }
    }
    // the ticket follows validation: its loads and CAS results were consumed
    uint64_t ticket = wave_fetch_inc(a.ctrs + CTR_TICKET, commit && (lane_id() & (G - 1)) == 0);
    if (!commit) return;
    ticket = group_bcast64<G>(ticket);
    // write phase (tictoc.cuh:352-368)
    ev_commit<G>(c, a);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) op_store(a, c.row[k], c.meta[k], c.tx, op_index<G>(k), c.op0 + op_index<G>(k));
    order_release();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (c.meta[k] & M_HELD) st(cc_word(a, c.row[k]), tictoc::make(0, 0, commit_ts));
    finish_tx(c, a, now, commit_ts, ticket);
  }
};

// ===========================================================================
// Ordered execution for GaccO (include/gacco.cuh:49-176) and GPUTx
// (include/gputx.cuh:58-146).  Preprocessing sorts every access of the batch
// by (row, order key) and gives op j its access group start need[j] (a write
// is its own group; consecutive reads share one).  counters[row] starts at the
// row's first position and counts completed accesses; an op runs once
// counters[row] >= need, then increments it.  Each row therefore sees its
// accesses in order-key order -- ascending transaction id for GaccO, (rank,
// transaction id) for GPUTx -- which makes the outcome that of the serial
// execution in that order.  No aborts.
// ===========================================================================
struct Ordered {
  template <int G>
  static __device__ void round(TxCtx<G>& c, const ExecArgs& a) {
    constexpr int K = TxCtx<G>::K;
    if (c.done) return;
    const uint64_t now = now_ticks();
    if (c.ph == PH_START) {  // c.st[k] holds each op's turn (load_tx)
      begin_attempt(c, now);
      // an op on a row that a combiner ran (GaccO, persistent form) is done
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (c.st[k] == NEED_DONE) c.meta[k] |= M_DONE;
      c.ph = PH_OPS;
    }
    if (now < c.t_resume) return;  // resting: no op can have its turn yet
    // Poll every waiting op's row counter at once (independent loads, one
    // memory latency for all slots), then run the ops whose turn has come:
    // one acquire before their cell accesses, one release before the counter
    // increments that pass the turn on.
    uint32_t seen[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t m = c.meta[k];
      seen[k] = ((m & M_VALID) && !(m & M_DONE)) ? ld32(turn_counter(a, c.row[k])) : 0u;
    }
    bool pending = false, any = false;
    bool ready[K];
    uint32_t gap = 0xffffffffu;  // fewest turns any still-waiting op is away from its own
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t m = c.meta[k];
      const bool waiting = (m & M_VALID) && !(m & M_DONE);
      ready[k] = waiting && seen[k] >= (uint32_t)c.st[k];
      pending |= waiting && !ready[k];
      any |= ready[k];
      if (waiting && !ready[k]) {
        const uint32_t d = (uint32_t)c.st[k] - seen[k];
        gap = d < gap ? d : gap;
      }
    }
    if (any) {
      order_acquire();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!ready[k]) continue;
        const uint32_t m = c.meta[k];
        const uint32_t j = op_index<G>(k);
        if (m & M_WRITE)
          op_store(a, c.row[k], m, c.tx, j, c.op0 + j);
        else
          a.reads[c.op0 + j] = op_load(a, c.row[k], m);
        ev_op<G>(c, a, k);
      }
      order_release();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!ready[k]) continue;
        fetch_add32(turn_counter(a, c.row[k]), 1u);
        c.meta[k] |= M_DONE;
      }
    }
    if (group_all<G>(!pending)) {
      const uint64_t key0 = a.tx_rank ? a.tx_rank[c.tx] : c.tx;
      ev_commit<G>(c, a);
      finish_tx(c, a, now, key0, c.tx);
    } else {
      c.m_wait += now - c.t_round;
      // An op `gap` turns short of its own cannot run before `gap` handoffs
      // on its row: rest until then (a.poll_ticks each, at most
      // POLL_REST_MAX) instead of polling every round; deep waiters on a hot
      // row then stop flooding the memory system the row's chain depends on.
      const uint64_t turns = group_min64<G>(gap);
      if (turns > 1) {
        const uint64_t rest = (turns - 1) * (uint64_t)a.poll_ticks;
        c.t_resume = now + (rest < POLL_REST_MAX ? rest : POLL_REST_MAX);
      }
    }
  }
  static constexpr bool kRestInOps = true;  // t_resume also paces the polls of PH_OPS
  // Refill at once: a done group that waited for its wave would idle behind
  // groups queued on hot rows (measured: 29% slower on the batched stream).
  static constexpr bool kBulkRefill = false;
  static constexpr uint64_t POLL_REST_MAX = 12800;     // 128 us
};

// ===========================================================================
// The kernel.
//
// Plain form (PERSIST = false): one launch per batch, one transaction per
// group, the reference's launch structure (benchmark_cpu/ycsb.cu:262-285).
// Persistent form (PERSIST = true): one launch over the whole stream; groups
// take the next transaction as soon as theirs commits (run_persistent).
// ===========================================================================
// Inputs of a launch (offsets, ops, rows, turns): plain loads, or coherent
// loads in the open form, where the host writes them while the launch runs
// and a line cached earlier by one XCD's L2 could be stale.
__device__ __forceinline__ uint64_t ld_in(const uint64_t* p, bool coherent) { return coherent ? ld(p) : *p; }
__device__ __forceinline__ uint32_t ld_in32(const uint32_t* p, bool coherent) { return coherent ? ld32(p) : *p; }

template <int G>
__device__ __forceinline__ void load_tx(TxCtx<G>& c, const ExecArgs& a, uint64_t slot, uint64_t now) {
  constexpr int K = TxCtx<G>::K;
  const bool coh = a.open != 0;
  c.live = slot < a.slot_hi;
  c.done = !c.live;
  c.ph = PH_START;
  c.attempts = 0;
  c.t_resume = 0;
  c.t_start = now;
  c.nops = 0;
  if (c.live) {
    c.tx = a.tx_map ? ld_in32(a.tx_map + slot, coh) : slot;
    c.op0 = ld_in(a.tx_off + c.tx, coh);
    c.nops = (uint32_t)(ld_in(a.tx_off + c.tx + 1, coh) - c.op0);
  }
  // Every slot's op, then every slot's row, each step issued for all slots
  // before its wait: rows from the probe pass, from the key itself (a dense
  // index), or from the hash index (all home buckets read at once; a key
  // whose home bucket is full without it walks on through index_find).
  uint64_t opv[K];
  uint32_t rowv[K];
  bool has[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    has[k] = c.live && op_index<G>(k) < c.nops;
    opv[k] = has[k] ? ld_in(a.ops + c.op0 + op_index<G>(k), coh) : 0ull;
  }
  if (a.rows_in) {
#pragma unroll
    for (int k = 0; k < K; ++k) rowv[k] = has[k] ? ld_in32(a.rows_in + c.op0 + op_index<G>(k), coh) : IDX_MISS;
  } else if (a.index.dense) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      rowv[k] = has[k] ? idx_dense_find(op_key(opv[k]), a.index.length) : IDX_MISS;
      if (has[k]) a.rows_out[c.op0 + op_index<G>(k)] = rowv[k];
    }
  } else {
    uint4 q0[K], q1[K];
    const uint32_t* bk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t key = op_key(opv[k]);
      bk[k] = a.index.buckets + (size_t)idx_bucket(key, a.index.nbuckets) * (2 * IDX_SLOTS);
      if (has[k] && key != IDX_EMPTY) {
        q0[k] = reinterpret_cast<const uint4*>(bk[k])[0];
        q1[k] = reinterpret_cast<const uint4*>(bk[k])[1];
      } else {
        q0[k] = q1[k] = make_uint4(IDX_EMPTY, IDX_EMPTY, IDX_EMPTY, IDX_EMPTY);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      rowv[k] = IDX_MISS;
      if (!has[k]) continue;
      const uint32_t key = op_key(opv[k]);
      if (key == IDX_EMPTY) continue;
      const uint32_t ks[IDX_SLOTS] = {q0[k].x, q0[k].y, q0[k].z, q0[k].w, q1[k].x, q1[k].y, q1[k].z, q1[k].w};
      int hit = -1;
      bool free_slot = false;
#pragma unroll
      for (int i = 0; i < IDX_SLOTS; ++i) {
        hit = ks[i] == key ? i : hit;
        free_slot |= ks[i] == IDX_EMPTY;
      }
      if (hit >= 0)
        rowv[k] = bk[k][IDX_SLOTS + hit];
      else if (!free_slot)
        rowv[k] = index_find(a.index, key);  // the home bucket is full: walk on
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (has[k]) a.rows_out[c.op0 + op_index<G>(k)] = rowv[k];
  }
  const bool hot = a.hot_cnt && *a.hot_any;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    c.meta[k] = 0;
    c.row[k] = 0;
    c.st[k] = 0;
    if (has[k]) {
      const uint32_t j = op_index<G>(k);
      const uint64_t op = opv[k];
      const uint32_t row = rowv[k];
      if (row < a.rows) {
        c.row[k] = row;
        c.meta[k] = M_VALID | (op_is_write(op) ? M_WRITE : 0u) | (op_col(op) << M_COL_SHIFT);
        if (hot && a.hot_cnt[defer_bucket(op_key(op))] >= a.hot_thr) c.meta[k] |= M_HOT;
        if (a.need) c.st[k] = ld_in32(a.need + c.op0 + j, coh);  // ordered schemes: the op's turn on its row
      } else {
        c.m_miss++;
      }
    }
  }
  if (c.live) {
    if (c.nops > (uint32_t)(K * G)) atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_TXLEN);
    c.m_idx += now_ticks() - now;
  }
}

// Watchdog stop: flag the error and, for diagnosis, dump the state of every op
// of the group's unfinished transaction.
template <int G>
__device__ void watchdog_stop(const TxCtx<G>& c, const ExecArgs& a) {
  constexpr int K = TxCtx<G>::K;
  if (lane_id() == 0) atomicOr((unsigned long long*)&a.metrics->error_flags, ERR_WATCHDOG);
  if (a.dbg && !c.done) {
    // per op: flags/row, the CC word through the coherent load and through an
    // atomic read, and phase/attempts
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!(c.meta[k] & M_VALID)) continue;
      uint64_t* d = a.dbg + (c.op0 + op_index<G>(k)) * 4;
      uint64_t* w = cc_word(a, c.row[k]);
      d[0] = ((uint64_t)c.meta[k] << 32) | c.row[k];
      d[1] = ld(w);
      d[2] = fetch_or(w, 0);
      d[3] = ((uint64_t)c.ph << 56) | ((uint64_t)c.attempts << 24) | 1u;
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent form.  Groups take stream slots, in stream order, from one
// wave-aggregated counter as soon as their transaction commits, so a group
// held up by a slow transaction never holds back later slots, and the lowest
// unfinished slot always belongs to a running group (which the ordered
// schemes need: they wait on lower slots only).  The grid holds the
// admission window: at most that many groups, so at most that many
// transactions in flight.
//
// Open form (a.open, the serving loop): the launch starts before its
// transactions have arrived.  The host admits them batch by batch while the
// launch runs, by raising the admission word (CTR_ADMIT0, one copy per gate
// shard) after their ops are in device memory, and closes the launch by
// setting the word's top bit once nothing more will come.  A group draws its
// next slot as in the closed form and holds it until the slot is admitted
// (then it runs it) or the launch is closed below it (then it is finished).
// Ops, offsets, rows and turns written while the launch runs are read with
// coherent loads.
//
// Exits: a group is finished when the counter has passed the end of the
// stream (or of a closed launch); a wave leaves when all its groups are
// finished, when a live transaction outlives the watchdog, when it is held
// back (urgent state) and the slot counter has not moved for the watchdog
// period, or when it waits for an admission that has not moved for the
// watchdog period.
// ---------------------------------------------------------------------------
constexpr uint64_t GATE_TICKS = 5000;    // a refilling wave re-reads its gate copy at most every 50 us
constexpr uint64_t PARK_TICKS = 20000;   // a wave with no live group: every 200 us
constexpr uint64_t ADMIT_TICKS = 1000;   // a wave whose groups wait for an admission: every 10 us
constexpr uint64_t STALL_TICKS = 100000; // held-back waves sample the slot counter every 1 ms
constexpr uint64_t NO_SLOT = ~0ull;

// One round of a scheme, inlined for groups of at least 4 lanes (at most
// TxCtx<G>::K = 8 op slots per lane).  Called out of line, the group's
// context would live in scratch memory, and every access to it would go
// through the stack: the compiler's own choice did that to some
// instantiations and halved their throughput.  Narrower groups carry up to
// 32 slots per lane, and there the compiler decides.
template <class S, int G>
__device__ __forceinline__ void scheme_round(TxCtx<G>& c, const ExecArgs& a) {
  if constexpr (G >= 4) {
    [[clang::always_inline]] S::round(c, a);
  } else {
    S::round(c, a);
  }
}

template <class S, int G>
__device__ void run_persistent(TxCtx<G>& c, const ExecArgs& a) {
  const bool leader = (lane_id() & (G - 1)) == 0;
  const uint64_t nslots = a.slot_hi - a.slot_lo;
  const bool open = a.open != 0;
  uint64_t* slot_ctr = a.ctrs + CTR_SLOT;
  const uint32_t shard = blockIdx.x & (GATE_SHARDS - 1);
  const uint64_t* gate = a.ctrs + CTR_GATE0 + 16 * shard;
  const uint32_t* admit = reinterpret_cast<const uint32_t*>(a.ctrs + CTR_ADMIT0 + 16 * shard);
  bool finished = false;
  uint64_t held = NO_SLOT;  // slot drawn and not yet running (open form: waiting for its admission)
  uint64_t now = now_ticks();
  uint64_t next_gate = now, next_stall = now;
  uint64_t urgent = 0, handed_out = 0;
  uint64_t admitted = open ? 0 : nslots;
  bool closed = !open;
  uint64_t adm_seen = 0, adm_seen_at = now;
  uint64_t seen_slot = ~0ull, seen_at = now;
  c.live = false;
  c.done = true;
  c.t_round = now;
#pragma unroll 1
  while (true) {
    // admission (the wall clock is a scalar read: `now` is wave-uniform)
    const bool need = c.done && !finished;
    bool stalled = false;
    // Bulk refill (schemes with kBulkRefill): the groups of a wave that are
    // done wait until half of the wave's groups are (or none is live), then
    // refill together, so the wave pays the refill's chain of dependent loads
    // (slot, offsets, ops and rows) once for many transactions, like a freshly
    // dispatched wave of the plain form; refilling one group at a time puts
    // that chain into most rounds (measured: Silo 19% faster on the batched
    // stream).
    const uint32_t ngw = (uint32_t)__popcll(ballot(leader));
    const uint32_t nneed = (uint32_t)__popcll(ballot(need && leader));
    const bool bulk = !S::kBulkRefill || 2 * nneed >= ngw || ballot(!c.done) == 0;
    if (bulk && ballot(need) != 0) {
      const bool read_gate = now >= next_gate;
      if (read_gate) {
        // this block's gate copy: urgent count, stream handed out; and, in
        // the open form, its copy of the admission word
        urgent = ld(gate + 1);
        handed_out = ld(gate + 2);
        if (open) {
          const uint32_t w = ld32(admit);
          admitted = w & ~ADMIT_CLOSED;
          admitted = admitted < nslots ? admitted : nslots;
          closed = (w & ADMIT_CLOSED) != 0;
          if (admitted != adm_seen) {
            adm_seen = admitted;
            adm_seen_at = now;
          }
        }
      }
      if (handed_out && need && held == NO_SLOT) finished = true;
      const bool take = need && !finished && urgent == 0 && held == NO_SLOT;
      if (ballot(take) != 0) {
        const uint64_t got = group_bcast64<G>(wave_fetch_inc(slot_ctr, take && leader));
        if (take) {
          if (got < nslots) {
            held = got;
          } else {
            finished = true;
            // the group that drew the first slot past the end tells every gate
            if (got == nslots && leader) {
#pragma unroll 1
              for (int g = 0; g < GATE_SHARDS; ++g) st(a.ctrs + CTR_GATE0 + 16 * g + 2, 1);
            }
          }
        }
      }
      if (need && held != NO_SLOT) {
        if (held < admitted) {
          if (open) order_acquire();  // the admission word before the data it admits
          load_tx(c, a, a.slot_lo + held, now_ticks());
          held = NO_SLOT;
        } else if (closed) {
          finished = true;  // the launch closed below this slot
        }
      }
      if (read_gate) {
        // groups that wait for an admission poll their copy often; refilling
        // waves now and then; waves with no live group rarely
        const bool waiting = ballot(!finished && held != NO_SLOT) != 0;
        next_gate = now + (open && waiting ? ADMIT_TICKS : ballot(!c.done) != 0 ? GATE_TICKS : PARK_TICKS);
      }
      // groups held back by the urgent state: a slot counter that has not
      // moved for the watchdog period means that no group progresses (a
      // group that installed its held slot in this round is live, not held
      // back: c.done, not the round's `need`)
      const bool parked = c.done && !finished && held == NO_SLOT && !take;
      if (ballot(parked) != 0 && now >= next_stall) {
        next_stall = now + STALL_TICKS;
        const uint64_t cur = ld(slot_ctr);
        if (cur >= nslots) {
          if (parked) finished = true;
        } else if (cur != seen_slot) {
          seen_slot = cur;
          seen_at = now;
        } else if (a.watchdog_ticks && now - seen_at > a.watchdog_ticks) {
          stalled = true;
        }
      }
      // open form: an admission that has not moved for the watchdog period
      // (the host neither admitted nor closed) ends the wait
      if (open && need && !finished && held != NO_SLOT && a.watchdog_ticks && now - adm_seen_at > a.watchdog_ticks)
        stalled = true;
    }
    if (ballot(!finished) == 0) break;
    if (ballot(stalled) != 0) {
      watchdog_stop(c, a);
      break;
    }
    if (ballot(!c.done) == 0)
      __builtin_amdgcn_s_sleep(64);  // no group of the wave is live (held back, waiting or finished)
    else if (ballot(!c.done && !((c.ph == PH_START || S::kRestInOps) && c.t_round < c.t_resume)) == 0)
      __builtin_amdgcn_s_sleep(4);  // every live group is resting
    else
      scheme_round<S, G>(c, a);
    c.t_round = now = now_ticks();
    const bool stuck = !c.done && a.watchdog_ticks && now - c.t_start > a.watchdog_ticks;
    if (ballot(stuck) != 0) {
      watchdog_stop(c, a);
      break;
    }
  }
}

// MAXT: the largest block the instantiation is launched with.  A block of
// 1,024 threads holds 4 waves per SIMD, which caps a lane at 128 registers;
// the persistent form's tuned shapes use blocks of at most 256 threads, and
// their instantiation may use up to 512 registers instead of spilling.
template <class S, int G, bool PERSIST, int MAXT = 1024>
__global__ __launch_bounds__(MAXT) void txn_exec(ExecArgs a) {
  TxCtx<G> c;
  if constexpr (PERSIST) {
    run_persistent<S, G>(c, a);
  } else {
    // plain form: one transaction per group
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    c.t_round = now_ticks();
    load_tx(c, a, a.slot_lo + gtid / G, c.t_round);
#pragma unroll 1
    while (true) {
      if (ballot(!c.done) == 0) break;
      // every live group of the wave is resting: sleep instead of spinning
      if (ballot(!c.done && !((c.ph == PH_START || S::kRestInOps) && c.t_round < c.t_resume)) == 0)
        __builtin_amdgcn_s_sleep(4);
      else
        scheme_round<S, G>(c, a);
      c.t_round = now_ticks();
      // Watchdog on the oldest live transaction of the wave; the wall clock is
      // a scalar read, so the exit is wave-uniform.
      const bool stuck = !c.done && a.watchdog_ticks && c.t_round - c.t_start > a.watchdog_ticks;
      if (ballot(stuck) != 0) {
        watchdog_stop(c, a);
        break;
      }
    }
  }
  // metrics: sums over the transactions the group ran, reduced over the wave
  const bool leader = (lane_id() & (G - 1)) == 0;
  uint64_t v[8] = {leader ? c.m_att : 0ull, leader ? c.m_ts : 0ull, leader ? c.m_wait : 0ull,
                   leader ? c.m_abort : 0ull, leader ? (c.m_tot - c.m_idx) : 0ull,
                   leader ? c.m_tot : 0ull,   leader ? c.m_idx : 0ull, leader ? c.m_commit : 0ull};
  const uint64_t misses = c.m_miss;  // per lane (each lane probes its own ops)
  // A block of block_size*32 threads can end in a half wave (lanes 0..31);
  // the last butterfly step then has no partner lanes and is skipped.
  const bool full_wave = (ballot(true) >> 32) != 0;
#pragma unroll
  for (int f = 0; f < 8; ++f) {
    uint64_t x = v[f];
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) x += shfl_xor64(x, m);
    if (full_wave) x += shfl_xor64(x, 32);
    v[f] = x;
  }
  uint64_t mm = misses;
#pragma unroll
  for (int m = 1; m < 32; m <<= 1) mm += shfl_xor64(mm, m);
  if (full_wave) mm += shfl_xor64(mm, 32);
  if (lane_id() == 0) {
    unsigned long long* M = (unsigned long long*)(a.metrics + (blockIdx.x & (METRIC_SLOTS - 1)));
    if (v[0]) atomicAdd(M + 0, v[0]);
    atomicAdd(M + 1, v[1]);
    atomicAdd(M + 2, v[2]);
    atomicAdd(M + 3, v[3]);
    atomicAdd(M + 4, v[4]);
    atomicAdd(M + 5, v[5]);
    atomicAdd(M + 6, v[6]);
    atomicAdd(M + 7, v[7]);
    if (mm) atomicAdd(M + 8, mm);
  }
}

}  // namespace dev (dev_row, dev_evt, dev_row_evt)
}  // namespace gcctb
