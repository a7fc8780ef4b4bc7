// The CPU path: the eight schemes executed by the host's cores, batch after
// batch like the plain GPU form (benchmark_cpu/ycsb.cu:249-291), for runs on
// a machine's CPU with the driver's arguments.
//
// Each worker thread takes the next transaction of the current batch from a
// shared counter and runs it to commit: per op in order, the scheme's Read or
// Write, then its commit step, retrying at once after an abort
// (benchmark/ycsb.cu:51-106).  CC words have the GPU path's layouts
// (gcctb_common.h) in std::atomic<uint64_t>; table cells are read and written
// with relaxed atomic accesses.  Keys address rows directly (the identity
// index the driver builds); a key outside the table is an index miss and the
// op is skipped, as on the GPU.  The timings follow the reference's Metrics:
// per-transaction total, time in aborted attempts, time spent waiting, time
// taking timestamps, in nanosecond ticks (tick_khz = 1e6).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../gcctb_common.h"
#include "host_internal.h"

namespace {

using namespace gcctb;
using Clock = std::chrono::steady_clock;

inline uint64_t ns_now() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}
inline void relax() { __builtin_ia32_pause(); }
inline uint64_t cell_ld(const uint64_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
inline void cell_st(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }

struct Op {
  uint32_t row, col, j;
  bool write;
};

struct TxTimes {
  uint64_t ts = 0, wait = 0, abort = 0, tot = 0;
};

class CpuRun {
 public:
  CpuRun(int scheme, const gcctb_config& cfg, uint64_t* table, const uint64_t* ops, const uint64_t* tx_off,
         uint64_t ntx, uint64_t* reads, uint64_t* order)
      : scheme_(scheme), rows_(cfg.rows), cols_(cfg.cols), seed_(cfg.seed), row_ops_(cfg.row_ops != 0),
        table_(table), ops_(ops), tx_off_(tx_off), ntx_(ntx), reads_(reads), order_(order) {}

  int prepare() {
    cc_.reset(new std::atomic<uint64_t>[rows_]);
    for (uint64_t r = 0; r < rows_; ++r) cc_[r].store(0, std::memory_order_relaxed);
    const uint64_t nops = tx_off_[ntx_];
    if (scheme_ == SCHEME_MVCC) {
      head_.reset(new std::atomic<uint64_t>[rows_]);
      for (uint64_t r = 0; r < rows_; ++r) head_[r].store(0, std::memory_order_relaxed);
      stride_ = row_ops_ ? 2 + cols_ : 3;
      pool_cap_ = std::max<uint64_t>(1, nops);
      pool_.reset(new std::atomic<uint64_t>[pool_cap_ * stride_]);
    }
    if (scheme_ == SCHEME_GACCO || scheme_ == SCHEME_GPUTX) {
      counters_.reset(new std::atomic<uint32_t>[rows_]);
      for (uint64_t r = 0; r < rows_; ++r) counters_[r].store(0, std::memory_order_relaxed);
      need_.assign(std::max<uint64_t>(1, nops), 0);
    }
    return GCCTB_OK;
  }

  // Runs every batch; returns the per-phase wall times.
  void run(uint64_t batch_size, int threads, gcctb_stats* st);

 private:
  // ---- table access: one cell, or the whole row in row mode ----
  uint64_t* cell(uint32_t row, uint32_t col) const { return table_ + (uint64_t)row * cols_ + col; }
  uint64_t read_now(const Op& o) const {
    if (!row_ops_) return cell_ld(cell(o.row, o.col));
    uint64_t h = DIGEST_INIT;
    for (uint32_t c = 0; c < cols_; ++c) h = digest_step(h, cell_ld(cell(o.row, c)));
    return h;
  }
  void write_now(uint64_t tx, const Op& o) const {
    if (!row_ops_) {
      cell_st(cell(o.row, o.col), write_value(seed_, tx, o.j));
      return;
    }
    for (uint32_t c = 0; c < cols_; ++c) cell_st(cell(o.row, c), write_cell(seed_, tx, o.j, c));
  }
  uint64_t& read_slot(uint64_t tx, const Op& o) const { return reads_[tx_off_[tx] + o.j]; }
  void set_order(uint64_t tx, uint64_t k0, uint64_t k1) const {
    order_[3 * tx] = batch_id_;
    order_[3 * tx + 1] = k0;
    order_[3 * tx + 2] = k1;
  }
  int load(uint64_t tx, Op* o) const {
    int n = 0;
    for (uint64_t i = tx_off_[tx]; i < tx_off_[tx + 1]; ++i) {
      const uint64_t op = ops_[i];
      if (op_key(op) >= rows_) continue;  // index miss: the op does not execute
      o[n++] = Op{op_key(op), op_col(op), (uint32_t)(i - tx_off_[tx]), op_is_write(op)};
    }
    return n;
  }

  // ---- schemes; each runs transaction tx to commit ----
  void run_tpl(uint64_t tx, const Op* o, int n, bool wait_die, TxTimes& tm);
  void run_ts(uint64_t tx, const Op* o, int n, bool mvcc, TxTimes& tm);
  bool mvcc_read(const Op& o, uint64_t ts, uint64_t* out, TxTimes& tm);
  void run_occ(uint64_t tx, const Op* o, int n, bool tictoc, TxTimes& tm);
  void run_ordered(uint64_t tx, const Op* o, int n, TxTimes& tm);
  void worker();
  void exec_one(uint64_t tx, TxTimes& tm);
  void order_accesses(uint64_t s_lo, uint64_t s_hi, const std::vector<uint64_t>* key_of_tx);

  const int scheme_;
  const uint64_t rows_;
  const uint32_t cols_;
  const uint64_t seed_;
  const bool row_ops_;
  uint64_t* table_;
  const uint64_t* ops_;
  const uint64_t* tx_off_;
  const uint64_t ntx_;
  uint64_t* reads_;
  uint64_t* order_;
  std::unique_ptr<std::atomic<uint64_t>[]> cc_, head_, pool_;
  uint64_t stride_ = 3, pool_cap_ = 0;
  std::atomic<uint64_t> pool_top_{0};
  std::unique_ptr<std::atomic<uint32_t>[]> counters_;
  std::vector<uint32_t> need_;
  std::vector<uint32_t> slot_tx_;  // GPUTx: transactions in (rank, id) order
  std::vector<uint32_t> rank_;
  // the batch being executed
  uint64_t slot_lo_ = 0, slot_hi_ = 0, batch_id_ = 0;
  std::atomic<uint64_t> next_{0}, ts_ctr_{1}, ticket_{0}, aborts_{0};
  std::atomic<uint64_t> m_ts_{0}, m_wait_{0}, m_abort_{0}, m_tot_{0};
  // worker pool: a generation number starts a batch, `active` counts the
  // workers still in it
  std::mutex mu_;
  std::condition_variable cv_start_, cv_done_;
  uint64_t generation_ = 0;
  int active_ = 0;
  bool quit_ = false;
};

// ---- 2PL, no-wait and wait-die (include/twopl.cuh:233-322) ----
// A read takes the row's lock shared and reads the cell at once; a write takes
// it exclusive and writes at commit.  Wait-die: a transaction waits for a
// younger holder and dies when an older one holds the row; a shared-held row
// refuses an exclusive request (twopl.cuh:285-286).
void CpuRun::run_tpl(uint64_t tx, const Op* o, int n, bool wait_die, TxTimes& tm) {
  const uint64_t self = tx & 0x7fffffffull;
  for (;;) {
    const uint64_t t0 = ns_now();
    int held = 0;
    bool ok = true;
    for (; held < n && ok; ++held) {
      std::atomic<uint64_t>& w = cc_[o[held].row];
      uint64_t wait_from = 0;
      for (;;) {
        uint64_t cur = w.load(std::memory_order_acquire);
        bool wait = false;
        if (o[held].write) {
          if (cur == 0) {
            if (w.compare_exchange_weak(cur, tpl::make_excl(self), std::memory_order_acq_rel)) break;
            continue;
          }
          if (!tpl::excl(cur) || !wait_die || tpl::holder(cur) < self) ok = false;
          else wait = true;
        } else {
          if (!tpl::excl(cur)) {
            if (w.compare_exchange_weak(cur, cur + tpl::CNT_ONE, std::memory_order_acq_rel)) break;
            continue;
          }
          if (!wait_die || tpl::holder(cur) < self) ok = false;
          else wait = true;
        }
        if (!ok) break;
        if (wait && !wait_from) wait_from = ns_now();
        relax();
      }
      if (wait_from) tm.wait += ns_now() - wait_from;
      if (!ok) break;
      if (!o[held].write) read_slot(tx, o[held]) = read_now(o[held]);
    }
    if (!ok) {
      // release the locks taken before the refused one (ops [0, held))
      for (int q = 0; q < held; ++q) {
        if (o[q].write) cc_[o[q].row].store(0, std::memory_order_release);
        else cc_[o[q].row].fetch_sub(tpl::CNT_ONE, std::memory_order_acq_rel);
      }
      aborts_.fetch_add(1, std::memory_order_relaxed);
      tm.abort += ns_now() - t0;
      continue;
    }
    const uint64_t ticket = ticket_.fetch_add(1, std::memory_order_acq_rel);
    for (int q = 0; q < n; ++q)
      if (o[q].write) write_now(tx, o[q]);
    for (int q = 0; q < n; ++q) {
      if (o[q].write) cc_[o[q].row].store(0, std::memory_order_release);
      else cc_[o[q].row].fetch_sub(tpl::CNT_ONE, std::memory_order_acq_rel);
    }
    set_order(tx, ticket, 0);
    return;
  }
}

// ---- T/O (include/to.cuh:104-270) and MVCC (include/mvcc.cuh:113-296) ----
// A timestamp per attempt.  T/O: a read waits for an older uncommitted write,
// aborts on a newer one, and raises rts; a write aborts on a newer rts or wts
// or a pending write, else marks the word.  MVCC reads never abort: a read
// older than the row's newest version walks the row's version chain.  Commit
// writes the cells (MVCC first saves the old ones in a pool node) and
// publishes {0, ts, ts}; abort restores the saved words.
bool CpuRun::mvcc_read(const Op& o, uint64_t ts, uint64_t* out, TxTimes& tm) {
  std::atomic<uint64_t>& w = cc_[o.row];
  uint64_t wait_from = 0;
  for (;;) {
    uint64_t cur = w.load(std::memory_order_acquire);
    if (tsw::unc(cur)) {  // a write is being installed: wait for it
      if (!wait_from) wait_from = ns_now();
      relax();
      continue;
    }
    if (tsw::wts(cur) <= ts) {  // the newest version: read it and raise rts
      const uint64_t v = read_now(o);
      const uint64_t nr = std::max(tsw::rts(cur), ts);
      if (w.compare_exchange_weak(cur, tsw::make(0, nr, tsw::wts(cur)), std::memory_order_acq_rel)) {
        *out = v;
        break;
      }
      continue;
    }
    // the newest version is younger than ts: walk the row's chain back to
    // the newest version at or below ts (each node holds the value its write
    // replaced and the word before that write)
    uint64_t v = read_now(o);
    uint64_t p = head_[o.row].load(std::memory_order_acquire);
    uint64_t vts = tsw::wts(cur);
    while (vts > ts && p) {
      const std::atomic<uint64_t>* nd = pool_.get() + (p - 1) * stride_;
      const uint64_t link = nd[1].load(std::memory_order_acquire);
      if (row_ops_) {
        uint64_t h = DIGEST_INIT;
        for (uint32_t c = 0; c < cols_; ++c) h = digest_step(h, nd[2 + c].load(std::memory_order_acquire));
        v = h;
      } else if ((uint32_t)(link >> 32) == o.col) {
        v = nd[2].load(std::memory_order_acquire);
      }
      vts = tsw::wts(nd[0].load(std::memory_order_acquire));
      p = (uint32_t)link;
    }
    if (w.load(std::memory_order_acquire) != cur) continue;  // a write moved the row meanwhile
    *out = v;
    break;
  }
  if (wait_from) tm.wait += ns_now() - wait_from;
  return true;
}

void CpuRun::run_ts(uint64_t tx, const Op* o, int n, bool mvcc, TxTimes& tm) {
  uint64_t prev[MAX_OPS];
  for (;;) {
    const uint64_t t0 = ns_now();
    const uint64_t ts = ts_ctr_.fetch_add(1, std::memory_order_acq_rel);
    tm.ts += ns_now() - t0;
    int k = 0;
    bool ok = true;
    for (; k < n && ok; ++k) {
      std::atomic<uint64_t>& w = cc_[o[k].row];
      if (!o[k].write) {
        if (mvcc) {
          uint64_t v;
          mvcc_read(o[k], ts, &v, tm);
          read_slot(tx, o[k]) = v;
          continue;
        }
        uint64_t wait_from = 0;
        for (;;) {
          uint64_t cur = w.load(std::memory_order_acquire);
          if (tsw::wts(cur) > ts) {
            ok = false;
            break;
          }
          if (tsw::unc(cur)) {
            if (!wait_from) wait_from = ns_now();
            relax();
            continue;
          }
          const uint64_t v = read_now(o[k]);
          const uint64_t nr = std::max(tsw::rts(cur), ts);
          if (w.compare_exchange_weak(cur, tsw::make(0, nr, tsw::wts(cur)), std::memory_order_acq_rel)) {
            read_slot(tx, o[k]) = v;
            break;
          }
        }
        if (wait_from) tm.wait += ns_now() - wait_from;
        continue;
      }
      for (;;) {
        uint64_t cur = w.load(std::memory_order_acquire);
        if (tsw::rts(cur) > ts || tsw::wts(cur) > ts || tsw::unc(cur)) {
          ok = false;
          break;
        }
        prev[k] = cur;
        if (w.compare_exchange_weak(cur, tsw::make(1, ts, ts), std::memory_order_acq_rel)) break;
      }
    }
    if (!ok) {
      // k is one past the refused op: restore the marks of the writes before it
      for (int q = 0; q + 1 < k; ++q)
        if (o[q].write) cc_[o[q].row].store(prev[q], std::memory_order_release);
      aborts_.fetch_add(1, std::memory_order_relaxed);
      tm.abort += ns_now() - t0;
      continue;
    }
    for (int q = 0; q < n; ++q) {
      if (!o[q].write) continue;
      if (mvcc) {
        const uint64_t node = pool_top_.fetch_add(1, std::memory_order_relaxed);
        if (node < pool_cap_) {
          std::atomic<uint64_t>* nd = pool_.get() + node * stride_;
          nd[0].store(prev[q], std::memory_order_relaxed);
          nd[1].store(head_[o[q].row].load(std::memory_order_relaxed) | ((uint64_t)o[q].col << 32),
                      std::memory_order_relaxed);
          if (row_ops_)
            for (uint32_t c = 0; c < cols_; ++c) nd[2 + c].store(cell_ld(cell(o[q].row, c)), std::memory_order_relaxed);
          else
            nd[2].store(cell_ld(cell(o[q].row, o[q].col)), std::memory_order_relaxed);
          head_[o[q].row].store(node + 1, std::memory_order_release);
        }
      }
      write_now(tx, o[q]);
    }
    for (int q = 0; q < n; ++q)
      if (o[q].write) cc_[o[q].row].store(tsw::make(0, ts, ts), std::memory_order_release);
    set_order(tx, ts, 0);
    return;
  }
}

// ---- Silo (include/silo.cuh:199-338) and TicToc (include/tictoc.cuh:198-368) ----
// Reads take a consistent snapshot (word, cells, word again, unlocked and
// unchanged); the write set is locked; then validation: Silo re-checks every
// read word; TicToc extends the rts of reads that do not yet cover the commit
// timestamp.  Commit writes the cells and publishes the new word.
void CpuRun::run_occ(uint64_t tx, const Op* o, int n, bool tictoc, TxTimes& tm) {
  uint64_t seen[MAX_OPS];
  for (;;) {
    const uint64_t t0 = ns_now();
    for (int k = 0; k < n; ++k) {
      if (o[k].write) continue;
      std::atomic<uint64_t>& w = cc_[o[k].row];
      for (;;) {
        const uint64_t w1 = w.load(std::memory_order_acquire);
        const uint64_t v = read_now(o[k]);
        const uint64_t w2 = w.load(std::memory_order_acquire);
        if (w1 == w2 && !(w1 & 1u)) {
          seen[k] = w1;
          read_slot(tx, o[k]) = v;
          break;
        }
        relax();
      }
    }
    // lock the write set in op order
    int locked = 0;
    bool ok = true;
    for (; locked < n; ++locked) {
      if (!o[locked].write) continue;
      std::atomic<uint64_t>& w = cc_[o[locked].row];
      uint64_t cur = w.load(std::memory_order_acquire);
      for (;;) {
        if (cur & 1u) {
          ok = false;
          break;
        }
        if (w.compare_exchange_weak(cur, cur | 1u, std::memory_order_acq_rel)) break;
      }
      if (!ok) break;
      seen[locked] = cur;
    }
    auto unlock = [&](int upto) {
      for (int q = 0; q < upto; ++q)
        if (o[q].write) cc_[o[q].row].store(seen[q], std::memory_order_release);
    };
    if (!ok) {
      unlock(locked);
      aborts_.fetch_add(1, std::memory_order_relaxed);
      tm.abort += ns_now() - t0;
      continue;
    }
    uint64_t k0 = 0, k1 = 0;
    if (!tictoc) {
      const uint64_t ticket = ticket_.fetch_add(1, std::memory_order_acq_rel);
      uint64_t mx = 0;
      for (int k = 0; k < n; ++k) mx = std::max(mx, silo::ts(seen[k]));
      for (int k = 0; k < n && ok; ++k) {
        if (o[k].write) continue;
        const uint64_t now = cc_[o[k].row].load(std::memory_order_acquire);
        if (silo::ts(now) != silo::ts(seen[k]) || silo::locked(now)) ok = false;
      }
      k0 = ticket;
      if (ok) {
        for (int q = 0; q < n; ++q)
          if (o[q].write) write_now(tx, o[q]);
        for (int q = 0; q < n; ++q)
          if (o[q].write) cc_[o[q].row].store(silo::make(0, mx + 1), std::memory_order_release);
      }
    } else {
      uint64_t cts = 0;
      for (int k = 0; k < n; ++k)
        cts = std::max(cts, o[k].write ? tictoc::rts(seen[k]) + 1 : tictoc::wts(seen[k]));
      for (int k = 0; k < n && ok; ++k) {
        if (o[k].write || tictoc::rts(seen[k]) >= cts) continue;
        std::atomic<uint64_t>& w = cc_[o[k].row];
        uint64_t cur = w.load(std::memory_order_acquire);
        for (;;) {
          if (tictoc::wts(cur) != tictoc::wts(seen[k]) || (tictoc::rts(cur) <= cts && tictoc::locked(cur))) {
            ok = false;
            break;
          }
          if (tictoc::rts(cur) > cts) break;
          // delta = cts - wts in 15 bits: shift wts up by what does not fit
          const uint64_t wts = tictoc::wts(cur), delta = cts - wts, shift = delta - (delta & 0x7fffull);
          if (w.compare_exchange_weak(cur, tictoc::make(cur & 1u, delta - shift, wts + shift),
                                      std::memory_order_acq_rel))
            break;
        }
      }
      if (ok) {
        const uint64_t ticket = ticket_.fetch_add(1, std::memory_order_acq_rel);
        for (int q = 0; q < n; ++q)
          if (o[q].write) write_now(tx, o[q]);
        for (int q = 0; q < n; ++q)
          if (o[q].write) cc_[o[q].row].store(tictoc::make(0, 0, cts), std::memory_order_release);
        k0 = cts;
        k1 = ticket;
      }
    }
    if (!ok) {
      unlock(n);
      aborts_.fetch_add(1, std::memory_order_relaxed);
      tm.abort += ns_now() - t0;
      continue;
    }
    set_order(tx, k0, k1);
    return;
  }
}

// ---- GaccO / GPUTx ordered execution (include/gacco.cuh:49-176,
// include/gputx.cuh:58-146): op i runs once its row's completed-access
// counter reaches need[i], then adds one ----
void CpuRun::run_ordered(uint64_t tx, const Op* o, int n, TxTimes& tm) {
  for (int k = 0; k < n; ++k) {
    const uint64_t i = tx_off_[tx] + o[k].j;
    std::atomic<uint32_t>& c = counters_[o[k].row];
    if (c.load(std::memory_order_acquire) < need_[i]) {
      const uint64_t t0 = ns_now();
      while (c.load(std::memory_order_acquire) < need_[i]) relax();
      tm.wait += ns_now() - t0;
    }
    if (o[k].write) write_now(tx, o[k]);
    else read_slot(tx, o[k]) = read_now(o[k]);
    c.fetch_add(1, std::memory_order_acq_rel);
  }
  if (scheme_ == SCHEME_GPUTX) set_order(tx, rank_[tx], tx);
  else set_order(tx, tx, tx);
}

// need[] for the accesses of slots [s_lo, s_hi): per row, in (key, tx)
// order, a write waits for every earlier access, a run of reads for the
// access before the run; turns count from the row's first access
void CpuRun::order_accesses(uint64_t s_lo, uint64_t s_hi, const std::vector<uint64_t>* key_of_tx) {
  struct Acc {
    uint32_t row;
    uint64_t key, op;
  };
  std::vector<Acc> acc;
  for (uint64_t s = s_lo; s < s_hi; ++s) {
    const uint64_t t = slot_tx_.empty() ? s : slot_tx_[s];
    for (uint64_t i = tx_off_[t]; i < tx_off_[t + 1]; ++i)
      if (op_key(ops_[i]) < rows_) acc.push_back(Acc{op_key(ops_[i]), key_of_tx ? (*key_of_tx)[t] : t, i});
  }
  std::sort(acc.begin(), acc.end(), [](const Acc& a, const Acc& b) {
    return a.row != b.row ? a.row < b.row : a.key != b.key ? a.key < b.key : a.op < b.op;
  });
  uint64_t row_start = 0, group = 0;
  for (uint64_t p = 0; p < acc.size(); ++p) {
    const bool first = p == 0 || acc[p].row != acc[p - 1].row;
    if (first) row_start = p;
    if (first || op_is_write(ops_[acc[p].op]) || op_is_write(ops_[acc[p - 1].op])) group = p;
    need_[acc[p].op] = (uint32_t)(group - row_start);
  }
}

void CpuRun::exec_one(uint64_t tx, TxTimes& tm) {
  Op o[MAX_OPS];
  const int n = load(tx, o);
  switch (scheme_) {
    case SCHEME_TPL_NW: run_tpl(tx, o, n, false, tm); break;
    case SCHEME_TPL_WD: run_tpl(tx, o, n, true, tm); break;
    case SCHEME_TO: run_ts(tx, o, n, false, tm); break;
    case SCHEME_MVCC: run_ts(tx, o, n, true, tm); break;
    case SCHEME_SILO: run_occ(tx, o, n, false, tm); break;
    case SCHEME_TICTOC: run_occ(tx, o, n, true, tm); break;
    default: run_ordered(tx, o, n, tm); break;
  }
}

void CpuRun::worker() {
  uint64_t seen_gen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_start_.wait(lk, [&] { return quit_ || generation_ != seen_gen; });
      if (quit_) return;
      seen_gen = generation_;
    }
    TxTimes tm;
    for (;;) {
      const uint64_t s = slot_lo_ + next_.fetch_add(1, std::memory_order_relaxed);
      if (s >= slot_hi_) break;
      const uint64_t tx = slot_tx_.empty() ? s : slot_tx_[s];
      const uint64_t t0 = ns_now();
      exec_one(tx, tm);
      tm.tot += ns_now() - t0;
    }
    m_ts_.fetch_add(tm.ts, std::memory_order_relaxed);
    m_wait_.fetch_add(tm.wait, std::memory_order_relaxed);
    m_abort_.fetch_add(tm.abort, std::memory_order_relaxed);
    m_tot_.fetch_add(tm.tot, std::memory_order_relaxed);
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (--active_ == 0) cv_done_.notify_one();
    }
  }
}

void CpuRun::run(uint64_t batch_size, int threads, gcctb_stats* st) {
  const uint64_t nops = tx_off_[ntx_];
  std::vector<uint64_t> bounds;
  std::vector<uint64_t> key_of_tx;
  double pre_ms = 0, proc_ms = 0;
  const uint64_t t_start = ns_now();
  if (scheme_ == SCHEME_GPUTX) {
    // ranks of every transaction (src/gputx.cu:5-38), then one batch per rank
    // level in (rank, id) order
    rank_.assign(ntx_, 0);
    struct Acc {
      uint32_t row;
      uint32_t tx;
      bool write;
    };
    std::vector<Acc> acc;
    acc.reserve(nops);
    for (uint64_t t = 0; t < ntx_; ++t)
      for (uint64_t i = tx_off_[t]; i < tx_off_[t + 1]; ++i)
        if (op_key(ops_[i]) < rows_) acc.push_back(Acc{op_key(ops_[i]), (uint32_t)t, op_is_write(ops_[i])});
    std::stable_sort(acc.begin(), acc.end(), [](const Acc& a, const Acc& b) { return a.row < b.row; });
    uint32_t local = 0, nranks = ntx_ ? 1 : 0;
    for (uint64_t p = 0; p < acc.size(); ++p) {
      if (p == 0 || acc[p].row != acc[p - 1].row) local = 0;
      else if (acc[p].write || acc[p - 1].write) ++local;
      rank_[acc[p].tx] = std::max(rank_[acc[p].tx], local);
      nranks = std::max(nranks, local + 1);
    }
    std::vector<uint64_t> count(nranks + 1, 0);
    for (uint64_t t = 0; t < ntx_; ++t) count[rank_[t] + 1]++;
    for (uint32_t r = 0; r < nranks; ++r) count[r + 1] += count[r];
    bounds.assign(count.begin(), count.end());
    slot_tx_.assign(ntx_, 0);
    for (uint64_t t = 0; t < ntx_; ++t) slot_tx_[count[rank_[t]]++] = (uint32_t)t;
    key_of_tx.resize(ntx_);
    for (uint64_t t = 0; t < ntx_; ++t) key_of_tx[t] = ((uint64_t)rank_[t] << 32) | t;
    order_accesses(0, ntx_, &key_of_tx);
  } else {
    for (uint64_t b = 0; b * batch_size < ntx_; ++b) bounds.push_back(b * batch_size);
    bounds.push_back(ntx_);
  }
  pre_ms += (ns_now() - t_start) * 1e-6;
  std::vector<std::thread> pool;
  active_ = 0;
  for (int i = 0; i < threads; ++i) pool.emplace_back([this] { worker(); });
  uint64_t launches = 0;
  for (size_t b = 0; b + 1 < bounds.size(); ++b) {
    if (bounds[b] == bounds[b + 1]) continue;
    const uint64_t t0 = ns_now();
    slot_lo_ = bounds[b];
    slot_hi_ = bounds[b + 1];
    batch_id_ = scheme_ == SCHEME_GPUTX ? 0 : b;
    ts_ctr_.store(1);
    ticket_.store(0);
    pool_top_.store(0);
    next_.store(0);
    if (scheme_ == SCHEME_GACCO) order_accesses(slot_lo_, slot_hi_, nullptr);
    const uint64_t t1 = ns_now();
    {
      std::lock_guard<std::mutex> lk(mu_);
      active_ = threads;
      ++generation_;
    }
    cv_start_.notify_all();
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_done_.wait(lk, [&] { return active_ == 0; });
    }
    const uint64_t t2 = ns_now();
    // per-batch reset of the words the batch touched (timestamps restart at 1)
    for (uint64_t s = slot_lo_; s < slot_hi_; ++s) {
      const uint64_t t = slot_tx_.empty() ? s : slot_tx_[s];
      for (uint64_t i = tx_off_[t]; i < tx_off_[t + 1]; ++i) {
        const uint32_t row = op_key(ops_[i]);
        if (row >= rows_) continue;
        cc_[row].store(0, std::memory_order_relaxed);
        if (head_) head_[row].store(0, std::memory_order_relaxed);
        if (counters_ && scheme_ == SCHEME_GACCO) counters_[row].store(0, std::memory_order_relaxed);
      }
    }
    pre_ms += ((t1 - t0) + (ns_now() - t2)) * 1e-6;
    proc_ms += (t2 - t1) * 1e-6;
    ++launches;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    quit_ = true;
  }
  cv_start_.notify_all();
  for (auto& th : pool) th.join();
  const double tot_ms = (ns_now() - t_start) * 1e-6;
  std::memset(st, 0, sizeof(*st));
  st->txs = ntx_;
  st->committed = ntx_;
  st->aborts = aborts_.load();
  st->ts_ticks = m_ts_.load();
  st->wait_ticks = m_wait_.load();
  st->abort_ticks = m_abort_.load();
  st->manager_ticks = m_tot_.load() - m_abort_.load();
  st->tot_ticks = m_tot_.load();
  st->index_ticks = 0;
  st->tick_khz = 1e6;  // nanosecond ticks
  st->pre_ms = pre_ms;
  st->proc_ms = proc_ms;
  st->tot_ms = tot_ms;
  st->batches = launches;
  st->ranks = scheme_ == SCHEME_GPUTX ? bounds.size() - 1 : 0;
  uint64_t miss = 0, rd = 0;
  for (uint64_t i = 0; i < nops; ++i) {
    miss += op_key(ops_[i]) >= rows_;
    rd += !op_is_write(ops_[i]);
  }
  st->index_miss = miss;
  st->total_ops = nops - miss;
  st->read_ops = rd;
}

}  // namespace

extern "C" int gcctb_cpu_seed_table(uint64_t* table, uint64_t rows, uint32_t cols, uint64_t seed) {
  if (!table && rows) return set_error(GCCTB_EINVAL, "null argument");
  for (uint64_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c) table[r * cols + c] = cell_init(seed, r, c, cols);
  return GCCTB_OK;
}

extern "C" int gcctb_cpu_run(const gcctb_config* cfg, int32_t threads, uint64_t* table, const uint64_t* ops,
                             const uint64_t* tx_off, uint64_t ntx, uint64_t* reads, uint64_t* order,
                             gcctb_stats* st) {
  if (!cfg || !cfg->cc_scheme || !table || !tx_off || !reads || !order || !st || (!ops && ntx))
    return set_error(GCCTB_EINVAL, "null argument");
  if (cfg->tpcc) return set_error(GCCTB_EINVAL, "the CPU path runs the YCSB transactions");
  const int scheme = scheme_from_name(cfg->cc_scheme);
  if (scheme < 0) return set_error(GCCTB_EINVAL, std::string("unknown cc_scheme ") + cfg->cc_scheme);
  if (cfg->rows == 0 || cfg->rows > 0x7fffffffull) return set_error(GCCTB_EINVAL, "rows out of range");
  if (cfg->cols == 0 || cfg->cols > 255 || (cfg->row_ops && cfg->cols > ROW_MAX_COLS))
    return set_error(GCCTB_EINVAL, "cols out of range");
  if (cfg->batch_size == 0) return set_error(GCCTB_EINVAL, "batch_size must be positive");
  // the GPU path's input contract: offsets, at most 32 ops, columns below
  // cols, distinct keys per transaction
  if (const int rc = gcctb_check_ops(ops, tx_off, ntx, cfg->cols, nullptr)) return rc;
  if (threads <= 0) {
    threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
      const int cap = std::atoi(e);
      if (cap > 0) threads = std::min(threads, cap);
    }
  }
  std::memset(reads, 0, sizeof(uint64_t) * (tx_off[ntx] ? tx_off[ntx] : 1));
  try {
    CpuRun r(scheme, *cfg, table, ops, tx_off, ntx, reads, order);
    r.prepare();
    r.run(cfg->batch_size, threads, st);
  } catch (const std::bad_alloc&) {
    return set_error(GCCTB_ENOMEM, "host allocation failed");
  } catch (const std::system_error& e) {
    return set_error(GCCTB_ENOMEM, std::string("cannot start host threads: ") + e.what());
  }
  return GCCTB_OK;
}

extern "C" int gcctb_check_host(const gcctb_config* cfg, const uint64_t* init_table, const uint64_t* final_table,
                                const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, const uint64_t* order,
                                const uint64_t* reads, uint64_t* bad) {
  if (!cfg || !init_table || !final_table || !tx_off || !order || !reads || !bad || (!ops && ntx))
    return set_error(GCCTB_EINVAL, "null argument");
  const uint64_t nops = tx_off[ntx];
  std::vector<uint32_t> rows(nops ? nops : 1);
  for (uint64_t i = 0; i < nops; ++i) rows[i] = op_key(ops[i]) < cfg->rows ? op_key(ops[i]) : 0xffffffffu;
  return check_history(cfg->rows, cfg->cols, cfg->seed, cfg->row_ops != 0, init_table, final_table, ops, rows.data(),
                       tx_off, ntx, order, reads, bad);
}
