// Host engine of the YCSB transaction-execution path: the context behind the C
// ABI, the per-scheme host classes (the counterpart of the reference's
// ConcurrencyControlCPUBase plugins, include/cc.cuh:18-56), and the plain
// per-batch launch loop (benchmark_cpu/ycsb.cu:242-291).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../device/launch.h"
#include "../gcctb_common.h"
#include "host_internal.h"

namespace gcctb {

static thread_local std::string g_err;
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

namespace {

struct HipFailure {
  hipError_t err;
  std::string where;
};
#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) throw HipFailure{_e, std::string(#expr) + " @" + std::to_string(__LINE__)}; \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  uint64_t n = 0;
  void alloc(uint64_t count) {
    if (p && count <= n) return;
    release();
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

int bits_for(uint64_t maxval) {  // bits needed to represent values <= maxval
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

}  // namespace

int scheme_from_name(const char* s) {
  static const char* names[SCHEME_COUNT] = {"tpl_nw", "tpl_wd", "to",     "mvcc",
                                            "silo",   "tictoc", "gacco", "gputx"};
  if (!s) return -1;
  for (int i = 0; i < SCHEME_COUNT; ++i)
    if (std::strcmp(s, names[i]) == 0) return i;
  return -1;
}

// The latch schemes slow_tpl_nw, slow_tpl_wd, slow_to, slow_mvcc, slow_silo
// and slow_tictoc (benchmark_cpu/ycsb.cu:170-190): the id of the scheme they
// guard with latches, or -1.
int latch_scheme_from_name(const char* s) {
  if (!s || std::strncmp(s, "slow_", 5) != 0) return -1;
  const int base = scheme_from_name(s + 5);
  return base >= 0 && base != SCHEME_GACCO && base != SCHEME_GPUTX ? base : -1;
}

class SchemeHost;

}  // namespace gcctb

struct gcctb_ctx {
  gcctb::Scheme scheme;
  bool latch = false;  // a slow_* scheme: CC-word accesses under per-row latches
  bool tpcc = false;   // TPC-C: every op updates its row (tpcc_apply)
  gcctb::DevBuf<uint32_t> op_param;  // TPC-C: per-op update parameters
  uint64_t params_loaded = 0;        // TPC-C: 1 + the op count the parameters were loaded for (0: none)
  gcctb_config cfg;
  int G = 1;
  int threads = 1024;
  uint64_t watchdog_ticks = 0;
  uint32_t urgent_attempts = 256;
  // persistent and serving forms: rows from a probe pass before the launch
  // (1), looked up by each group as it takes a transaction (0), or the
  // scheme's measured default (-1; GCCTB_PROBE_PASS sets 0 or 1)
  int probe_pass = -1;
  uint32_t backoff_exp = 10;
  uint32_t poll_ticks = 10;  // 0.1 us (measured at theta 0.99 with the hot rows' turn counters spread)
  // persistent form of the lock, timestamp and OCC schemes: transactions that
  // write a key touched by at least one transaction in defer_share run after
  // the others.  GCCTB_DEFER_SHARE sets it for every scheme (0 turns it
  // off); by default (-1) 16 for 2PL, T/O and MVCC, and off for Silo and
  // TicToc, which measured no gain from it
  int defer_share = -1;
  // GaccO, persistent form: a row whose access queue holds at least this many
  // accesses is run by a combiner before the launch (GCCTB_COMBINE_MIN; 0 off)
  uint32_t comb_min = 128;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  // pinned host word: the error flags of the launches since the last metrics
  // clear, copied behind each launch so the host sees them at its sync
  uint64_t* h_err = nullptr;
  // table + index
  gcctb::DevBuf<uint64_t> cells;
  gcctb::DevBuf<uint32_t> buckets;  // hash index
  gcctb::IndexView ix{};
  bool index_hash = false;  // GCCTB_INDEX_HASH=1: the identity index as a hash table
  bool table_ready = false;
  // transactions
  uint64_t ntx = 0, nops = 0, nreads = 0;
  std::vector<uint64_t> h_tx_off;
  gcctb::DevBuf<uint64_t> ops, tx_off;
  gcctb::DevBuf<uint32_t> op_tx;
  // CC state
  gcctb::DevBuf<uint64_t> cc;
  gcctb::DevBuf<uint32_t> vprev;
  gcctb::DevBuf<uint64_t> pool;
  uint64_t pool_cap = 0;
  gcctb::DevBuf<uint32_t> counters;
  gcctb::DevBuf<uint32_t> need;
  gcctb::DevBuf<uint64_t> ctrs;
  // outputs
  gcctb::DevBuf<uint32_t> rows_out;
  gcctb::DevBuf<uint64_t> reads;
  gcctb::DevBuf<uint32_t> obatch;
  gcctb::DevBuf<uint64_t> okey0, okey1;
  gcctb::DevBuf<uint32_t> tx_aborts;
  gcctb::DevBuf<gcctb::Metrics> metrics;
  // preprocessing scratch
  gcctb::DevBuf<uint32_t> sk0, sv0, sk1, sv1, flags, gstart, sort_tmp, S, SS;
  gcctb::DevBuf<uint64_t> sk64a, sk64b;
  gcctb::DevBuf<uint32_t> rank, tx_map, max_rank;
  // hot-writer deferral: sampled counts and state, slot -> transaction map;
  // in use while defer_active
  gcctb::DevBuf<uint32_t> defer_cnt, defer_map;
  bool defer_active = false;
  uint32_t defer_thr = 0;  // the pass's hot threshold (sampled accesses)
  bool hot_active = false; // the launch reads hot flags (2PL with the deferral)
  // 2PL: lock hot keys after the others (GCCTB_HOT_LAST=0 turns it off)
  bool hot_last = true;
  // GaccO combiners: queue list (first position, length) and counters
  // ([0] low half: queues listed; [1]: ops the combiners ran)
  gcctb::DevBuf<uint64_t> comb_segs, comb_ctr, comb_wv;
  uint64_t comb_cap = 0;
  gcctb::DevBuf<uint64_t> conflicts;
  // watchdog diagnostics (GCCTB_DEBUG_DUMP=<file>)
  gcctb::DevBuf<uint64_t> dbg;
  std::string dbg_path;
  uint64_t nranks = 0;
  uint64_t last_committed = 0;
  bool rows_probed = false;  // rows_out holds every op's row (persistent form)
  gcctb::DevBuf<uint64_t> tcommit;  // serving form: per-tx commit tick
  bool record_commit = false;
  // serving form: the admission stream (copies and preparation that feed a
  // running open launch), GaccO's per-row access counts of the launch, and a
  // device word for clock samples
  hipStream_t stream2 = nullptr;
  bool open_launch = false;
  gcctb::DevBuf<uint32_t> rowbase;
  gcctb::DevBuf<uint64_t> clock_word;
  // event log (debug_events): per-op access stamps, per-tx commit stamps
  gcctb::DevBuf<uint64_t> ev_op, ev_commit;
  bool events_ready = false;
  std::vector<uint32_t> h_rank_counts;
  bool rank_counts_ready = false;  // h_rank_counts matches rank[]
  std::vector<uint64_t> custom_batches;  // uneven batch sizes (empty: batch_size)
  // hipGraph of the per-batch sequence (graph form), rebuilt when inputs change
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_gen = 0, inputs_gen = 1;
  std::unique_ptr<gcctb::SchemeHost> host;
};

namespace gcctb {

// ---------------------------------------------------------------------------
// Scheme host classes.  Init0 splits the transactions into batches
// (include/cc.cuh:29-35); Init prepares one batch (reset / preprocessing,
// timed as "pre"); Launch runs the batch's execution kernel (timed as "proc");
// Post restores the per-row words the batch touched (the reference memsets the
// whole CC table per batch instead, e.g. include/to.cuh:316-321).
// ---------------------------------------------------------------------------
class SchemeHost {
 public:
  explicit SchemeHost(gcctb_ctx& c) : ctx(c) {}
  virtual ~SchemeHost() = default;
  virtual void Init0() {
    batches.clear();
    if (!ctx.custom_batches.empty()) {  // uneven batches of a stream (gcctb_set_batches)
      batches = ctx.custom_batches;
      return;
    }
    const uint64_t bs = ctx.cfg.batch_size;
    for (uint64_t st = 0; st < ctx.ntx; st += bs) batches.push_back(std::min(bs, ctx.ntx - st));
  }
  // Persistent form: prepare the whole stream as one unit (no per-batch
  // resets: timestamps and tickets keep counting across batch boundaries).
  virtual void PrepStream() { HIP_TRY(launch::batch_begin(ctx.ctrs.p, ctx.stream)); }
  // Persistent form: restore the words every transaction of the stream touched.
  virtual void PostStream() { Post(0, 0, ctx.ntx); }
  // Serving form: prepare one launch over transactions whose ops are [o0, o1).
  virtual void PrepLaunch(uint64_t o0, uint64_t o1) {
    (void)o0;
    (void)o1;
    HIP_TRY(launch::batch_begin(ctx.ctrs.p, ctx.stream));
  }
  // Serving form, open launch: reset the counters of a launch whose
  // transactions are admitted while it runs (on the launch's stream).
  virtual void PrepOpen() {
    HIP_TRY(launch::batch_begin(ctx.ctrs.p, ctx.stream));
    HIP_TRY(launch::persist_begin(ctx.ctrs.p, ctx.stream));
  }
  // Serving form, open launch: prepare the ops [o0, o1) of an admitted batch
  // on the admission stream s, before the admission word covers them (rows
  // of the probe pass where the scheme wants one).
  virtual void PrepAdmit(uint64_t o0, uint64_t o1, hipStream_t s) {
    const bool pass = ctx.probe_pass >= 0 ? ctx.probe_pass != 0 : WantsProbePass();
    if (pass && o1 > o0) HIP_TRY(launch::probe(ctx.ix, ctx.ops.p + o0, o1 - o0, ctx.rows_out.p + o0, s));
  }
  // True when the open launch reads every op's row from rows_out (PrepAdmit wrote it).
  virtual bool OpenRowsProbed() const { return ctx.probe_pass >= 0 ? ctx.probe_pass != 0 : WantsProbePass(); }
  // True when PrepStream / PrepLaunch leave the row of every op they cover in
  // rows_out, so the persistent and serving forms need no probe pass of their own.
  virtual bool ProbesRows() const { return false; }
  // Persistent and serving forms: whether a probe pass before the launch
  // beats looking rows up in the index as each transaction is taken (one
  // bucket read on the refill chain over a hash table, none over the dense
  // index).  Measured on the low-contention workload: only MVCC over a hash
  // table gains from the pass.
  virtual bool WantsProbePass() const { return false; }
  virtual void Init(int batch_id, uint64_t batch_st) {
    (void)batch_id;
    (void)batch_st;
    HIP_TRY(launch::batch_begin(ctx.ctrs.p, ctx.stream));
  }
  virtual void Launch(int batch_id, uint64_t slot_lo, uint64_t slot_hi, bool persist = false) = 0;
  virtual void Post(int batch_id, uint64_t batch_st, uint64_t n) {
    (void)batch_id;
    const uint64_t o0 = ctx.h_tx_off[batch_st], o1 = ctx.h_tx_off[batch_st + n];
    HIP_TRY(launch::reset_rows(ctx.rows_out.p + o0, o1 - o0, ctx.cfg.rows, ctx.cc.p,
                               ctx.scheme == SCHEME_MVCC ? ctx.vprev.p : nullptr, ctx.stream, ctx.latch));
  }
  // Device allocations of the scheme (GetMemSize, include/cc.cuh:39).
  virtual uint64_t GetMemSize() const { return ctx.cc.n * 8; }
  // ops that row combiners ran in the last persistent pass (GaccO)
  virtual uint64_t Combined() const { return 0; }
  std::vector<uint64_t> batches;

 protected:
  ExecArgs args(int batch_id, uint64_t slot_lo, uint64_t slot_hi) const {
    ExecArgs a{};
    a.cells = ctx.cells.p;
    a.rows = ctx.cfg.rows;
    a.cols = ctx.cfg.cols;
    a.scheme = (uint32_t)ctx.scheme;
    a.index = ctx.ix;
    a.ops = ctx.ops.p;
    a.tx_off = ctx.tx_off.p;
    a.tx_map = ctx.defer_active ? ctx.defer_map.p : nullptr;
    a.hot_cnt = nullptr;
    a.hot_any = nullptr;
    a.hot_thr = 0;
    if (ctx.hot_active) {
      a.hot_cnt = ctx.defer_cnt.p;
      a.hot_any = ctx.defer_cnt.p + launch::DEFER_BUCKETS + launch::DEFER_ANY;
      a.hot_thr = ctx.defer_thr;
    }
    a.slot_lo = slot_lo;
    a.slot_hi = slot_hi;
    a.cc = ctx.cc.p;
    a.vprev = ctx.vprev.p;
    a.pool = ctx.pool.p;
    a.pool_cap = ctx.pool_cap;
    a.counters = ctx.counters.p;
    a.op_param = ctx.op_param.p;
    a.slot_shift = slot_shift(ctx.cfg.rows);
    a.need = ctx.need.p;
    a.ctrs = ctx.ctrs.p;
    a.rows_out = ctx.rows_out.p;
    a.rows_in = ctx.rows_probed ? ctx.rows_out.p : nullptr;
    a.tcommit = ctx.record_commit ? ctx.tcommit.p : nullptr;
    a.reads = ctx.reads.p;
    a.obatch = ctx.obatch.p;
    a.okey0 = ctx.okey0.p;
    a.okey1 = ctx.okey1.p;
    a.tx_aborts = ctx.tx_aborts.p;
    a.metrics = ctx.metrics.p;
    a.seed = ctx.cfg.seed;
    a.batch_id = (uint32_t)batch_id;
    a.backoff_ticks = ctx.cfg.backoff_ticks;
    a.tx_rank = nullptr;
    a.watchdog_ticks = ctx.watchdog_ticks;
    a.urgent_attempts = ctx.urgent_attempts;
    a.backoff_exp = ctx.backoff_exp;
    a.poll_ticks = ctx.poll_ticks;
    a.dbg = ctx.dbg.p;
    a.persist_groups = ctx.cfg.persist_window;
    a.open = ctx.open_launch ? 1u : 0u;
    a.ev_op = ctx.cfg.debug_events ? ctx.ev_op.p : nullptr;
    a.ev_commit = ctx.cfg.debug_events ? ctx.ev_commit.p : nullptr;
    return a;
  }
  // Kernel variants (whole-row mode, the event log): every scheme's plain
  // kernels of a variant come from one launcher.
  bool launch_rows(const ExecArgs& a, uint64_t n, bool persist) {
    if (!ctx.cfg.row_ops && !ctx.cfg.debug_events && !ctx.latch && !ctx.tpcc) return false;
    if (persist)
      throw HipFailure{hipErrorInvalidValue, "whole-row mode, the event log and the latch schemes have no persistent kernels"};
    const int sch = (int)ctx.scheme;
    if (ctx.tpcc)
      HIP_TRY(launch::exec_tpcc(sch, ctx.G, a, n, ctx.threads, ctx.stream));
    else if (ctx.latch)
      HIP_TRY(launch::exec_latch(sch, ctx.G, a, n, ctx.threads, ctx.stream));
    else if (ctx.cfg.row_ops && ctx.cfg.debug_events)
      HIP_TRY(launch::exec_row_evt(sch, ctx.G, a, n, ctx.threads, ctx.stream));
    else if (ctx.cfg.row_ops)
      HIP_TRY(launch::exec_row(sch, ctx.G, a, n, ctx.threads, ctx.stream));
    else
      HIP_TRY(launch::exec_evt(sch, ctx.G, a, n, ctx.threads, ctx.stream));
    return true;
  }
  gcctb_ctx& ctx;
};

class TwoPLHost : public SchemeHost {
 public:
  TwoPLHost(gcctb_ctx& c, bool wd) : SchemeHost(c), wait_die(wd) {}
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_tpl(wait_die, ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
  // Locks are free again once every transaction of the batch committed; a
  // launch that stopped on an error is followed by clear_cc_state instead.
  void Post(int, uint64_t, uint64_t) override {}
  bool wait_die;
};
class TOHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_to(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
};
class MVCCHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_mvcc(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
  uint64_t GetMemSize() const override { return ctx.cc.n * 8 + ctx.vprev.n * 4 + ctx.pool.n * 8; }
  bool WantsProbePass() const override { return !ctx.ix.dense; }
};
class SiloHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_silo(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
};
class TictocHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_tictoc(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
};

// GaccO: per-batch lock-queue construction (Gacco_CPU::Init,
// include/gacco.cuh:244-339) as probe -> stable radix sort by row -> access
// groups -> need / counters, all on the device with no host round trip.
class GaccoHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  void Init(int batch_id, uint64_t batch_st) override {
    const uint64_t n_tx = batches[batch_id];
    prep_range(ctx.h_tx_off[batch_st], ctx.h_tx_off[batch_st + n_tx]);
  }
  // The persistent form queues every row's accesses of the whole stream in
  // transaction-id order: the outcome is that of the serial run in id order,
  // the same as batch after batch.
  void PrepStream() override { prep_range(0, ctx.nops, true); }
  void PostStream() override {}
  void PrepLaunch(uint64_t o0, uint64_t o1) override { prep_range(o0, o1); }
  bool ProbesRows() const override { return true; }
  // Open launch: every row's queue starts empty; each admitted batch's
  // accesses queue behind those of the batches admitted before it, in
  // transaction-id order (rowbase counts the queued accesses per row).
  void PrepOpen() override {
    SchemeHost::PrepOpen();
    HIP_TRY(launch::fill_u32(ctx.counters.p, ctx.counters.n, 0, ctx.stream));
    HIP_TRY(launch::fill_u32(ctx.rowbase.p, ctx.rowbase.n, 0, ctx.stream));
  }
  void PrepAdmit(uint64_t o0, uint64_t o1, hipStream_t s) override {
    const uint64_t n = o1 - o0;
    if (n == 0) return;
    HIP_TRY(launch::probe(ctx.ix, ctx.ops.p + o0, n, ctx.rows_out.p + o0, s, ctx.sk0.p, ctx.sv0.p, (uint32_t)o0));
    uint32_t *srow, *sop;
    HIP_TRY(launch::radix_sort_u32(ctx.sk0.p, ctx.sv0.p, ctx.sk1.p, ctx.sv1.p, n, bits_for(ctx.cfg.rows),
                                   ctx.sort_tmp.p, &srow, &sop, s));
    uint32_t* rflags = srow == ctx.sk0.p ? ctx.sk1.p : ctx.sk0.p;
    uint32_t* rstart = sop == ctx.sv0.p ? ctx.sv1.p : ctx.sv0.p;
    HIP_TRY(launch::access_group_flags(srow, sop, n, ctx.flags.p, s, rflags));
    HIP_TRY(launch::inclusive_max_u32(ctx.flags.p, ctx.gstart.p, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::inclusive_max_u32(rflags, rstart, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::access_need_open(srow, sop, ctx.gstart.p, rstart, n, ctx.cfg.rows, ctx.rowbase.p, ctx.need.p, s));
    HIP_TRY(launch::row_base_add(srow, rstart, n, ctx.cfg.rows, ctx.rowbase.p, s));
  }
  bool OpenRowsProbed() const override { return true; }
  // per-row access queues for ops [o0, o1) in (row, transaction id) order;
  // the rows also land in rows_out[o0, o1).  relative: turns count from the
  // row's first access and the counters start from one dense zero fill (the
  // whole stream at once); otherwise each touched row's counter is set to
  // its first position, which costs a random store per row but no fill of
  // the whole array (one batch at a time).
  void prep_range(uint64_t o0, uint64_t o1, bool relative = false) {
    const uint64_t n = o1 - o0;
    if (n == 0) return;
    hipStream_t s = ctx.stream;
    HIP_TRY(launch::probe(ctx.ix, ctx.ops.p + o0, n, ctx.rows_out.p + o0, s, ctx.sk0.p, ctx.sv0.p, (uint32_t)o0));
    uint32_t *srow, *sop;
    // rows are below cfg.rows; a miss (IDX_MISS) has all row bits set and sorts last
    HIP_TRY(launch::radix_sort_u32(ctx.sk0.p, ctx.sv0.p, ctx.sk1.p, ctx.sv1.p, n, bits_for(ctx.cfg.rows),
                                   ctx.sort_tmp.p, &srow, &sop, s));
    if (relative) {
      // the sort's other key and value buffers are free now: row-start flags
      // and the row start of every position
      uint32_t* rflags = srow == ctx.sk0.p ? ctx.sk1.p : ctx.sk0.p;
      uint32_t* rstart = sop == ctx.sv0.p ? ctx.sv1.p : ctx.sv0.p;
      HIP_TRY(launch::access_group_flags(srow, sop, n, ctx.flags.p, s, rflags));
      HIP_TRY(launch::inclusive_max_u32(ctx.flags.p, ctx.gstart.p, n, ctx.sort_tmp.p, s));
      HIP_TRY(launch::inclusive_max_u32(rflags, rstart, n, ctx.sort_tmp.p, s));
      // rows with long queues go to combiners (whole stream, column mode)
      const bool comb = ctx.scheme == SCHEME_GACCO && ctx.comb_min && ctx.comb_cap && ctx.cfg.cols <= 64 &&
                        o0 == 0 && !ctx.cfg.row_ops && !ctx.tpcc;
      if (comb) HIP_TRY(launch::fill_u64(ctx.comb_ctr.p, 2, 0, s));
      uint2* segs = reinterpret_cast<uint2*>(ctx.comb_segs.p);
      uint32_t* nseg = reinterpret_cast<uint32_t*>(ctx.comb_ctr.p);
      HIP_TRY(launch::access_need_rel(sop, ctx.gstart.p, rstart, n, ctx.need.p, s, srow, ctx.cfg.rows,
                                      comb ? ctx.comb_min : 0u, segs, nseg, (uint32_t)ctx.comb_cap));
      if (comb)
        HIP_TRY(launch::gacco_combine(segs, nseg, (uint32_t)ctx.comb_cap, srow, sop, ctx.ops.p, ctx.op_tx.p,
                                      ctx.tx_off.p, ctx.nops, ctx.cells.p, ctx.cfg.cols, ctx.cfg.seed, ctx.comb_wv.p,
                                      ctx.reads.p, ctx.need.p, ctx.comb_ctr.p + 1, s));
      combined_ = comb;
      HIP_TRY(launch::fill_u32(ctx.counters.p, ctx.counters.n, 0, s));
      return;
    }
    HIP_TRY(launch::access_group_flags(srow, sop, n, ctx.flags.p, s));
    HIP_TRY(launch::inclusive_max_u32(ctx.flags.p, ctx.gstart.p, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::access_need(srow, sop, ctx.gstart.p, n, 0, ctx.need.p, ctx.counters.p, ctx.cfg.rows, s));
  }
  void Launch(int b, uint64_t lo, uint64_t hi, bool persist) override {
    const ExecArgs a = args(b, lo, hi);
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_ordered(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
  void Post(int, uint64_t, uint64_t) override {}
  uint64_t GetMemSize() const override { return ctx.counters.n * 4 + ctx.need.n * 4; }
  // ops run by the combiners in the last persistent pass
  uint64_t Combined() const override {
    uint64_t v = 0;
    if (combined_) HIP_TRY(hipMemcpy(&v, ctx.comb_ctr.p + 1, 8, hipMemcpyDeviceToHost));
    return v;
  }
  bool combined_ = false;
};

// GPUTx: one-shot ranking of every transaction (GPUTx_CPU::Init0,
// include/gputx.cuh:213-299; gputx_pre2/pre3, src/gputx.cu:5-38), then one
// launch per rank level in (rank, transaction id) order.
class GPUTxHost : public SchemeHost {
 public:
  using SchemeHost::SchemeHost;
  // plain and graph forms: ranks, then one batch per rank level
  void Init0() override {
    rank_all();
    rank_counts(ctx);
    batches.clear();
    for (uint32_t c : ctx.h_rank_counts) batches.push_back(c);
  }
  // per-rank counts (gputx_pre3, src/gputx.cu:30-38), copied to the host
  static void rank_counts(gcctb_ctx& c) {
    if (c.rank_counts_ready) return;
    hipStream_t s = c.stream;
    c.h_rank_counts.assign(c.nranks, 0);
    if (c.nops && c.nranks) {
      DevBuf<uint32_t> hist;
      hist.alloc(c.nranks);
      HIP_TRY(launch::fill_u32(hist.p, c.nranks, 0, s));
      HIP_TRY(launch::rank_hist(c.rank.p, c.ntx, (uint32_t)c.nranks, hist.p, s));
      HIP_TRY(hipMemcpyAsync(c.h_rank_counts.data(), hist.p, c.nranks * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    } else if (c.nranks) {
      // no ops at all: every transaction has rank 0
      c.h_rank_counts[0] = (uint32_t)c.ntx;
    }
    c.rank_counts_ready = true;
  }
  // rank[] of every transaction, tx_map in (rank, tx) order, and need[] /
  // counters[] for execution in (row, rank, tx) order
  void rank_all() {
    ctx.rank_counts_ready = false;
    const uint64_t n = ctx.nops, ntx = ctx.ntx;
    hipStream_t s = ctx.stream;
    // per-row access lists in (row, tx) order; the sorted values are the
    // accesses' transactions (with the write flag), so the rank pass needs
    // no op -> transaction gather
    HIP_TRY(launch::probe(ctx.ix, ctx.ops.p, n, ctx.rows_out.p, s, ctx.sk0.p, ctx.sv0.p, 0, ctx.op_tx.p));
    uint32_t *srow, *stx_ops;
    // rows are below cfg.rows; a miss (IDX_MISS) has all row bits set and sorts last
    const int row_bits = bits_for(ctx.cfg.rows);
    HIP_TRY(launch::radix_sort_u32(ctx.sk0.p, ctx.sv0.p, ctx.sk1.p, ctx.sv1.p, n, row_bits, ctx.sort_tmp.p,
                                   &srow, &stx_ops, s));
    // one-pass local ranks and rank[tx] = max (src/gputx.cu:5-27)
    HIP_TRY(launch::gtx_flags(srow, stx_ops, n, ctx.cfg.rows, ctx.flags.p, ctx.gstart.p, s));
    HIP_TRY(launch::inclusive_sum_u32(ctx.flags.p, ctx.S.p, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::inclusive_max_u32(ctx.gstart.p, ctx.SS.p, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::fill_u32(ctx.rank.p, ntx, 0, s));
    HIP_TRY(launch::fill_u32(ctx.max_rank.p, 1, 0, s));
    HIP_TRY(launch::gtx_rank(stx_ops, ctx.S.p, ctx.SS.p, n, ctx.rank.p, ctx.max_rank.p, s));
    uint32_t max_rank = 0;
    HIP_TRY(hipMemcpyAsync(&max_rank, ctx.max_rank.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // no ops at all: every transaction has rank 0
    ctx.nranks = n ? (uint64_t)max_rank + 1 : (ntx ? 1 : 0);
    // transactions in (rank, tx) order: stable sort of ranks (include/gputx.cuh:281)
    const int rank_bits = std::max(1, bits_for(ctx.nranks ? ctx.nranks - 1 : 0));
    HIP_TRY(hipMemcpyAsync(ctx.sk0.p, ctx.rank.p, ntx * 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(launch::iota_u32(ctx.sv0.p, ntx, 0, s));
    uint32_t *srank, *stx;
    HIP_TRY(launch::radix_sort_u32(ctx.sk0.p, ctx.sv0.p, ctx.sk1.p, ctx.sv1.p, ntx, rank_bits, ctx.sort_tmp.p,
                                   &srank, &stx, s));
    HIP_TRY(hipMemcpyAsync(ctx.tx_map.p, stx, ntx * 4, hipMemcpyDeviceToDevice, s));
    // execution order of every row: (row, rank, tx); 32-bit composite keys
    // when row and rank fit, else 64-bit
    if (n && row_bits + rank_bits <= 32) {
      HIP_TRY(launch::gtx_exec_keys32(ctx.rows_out.p, ctx.op_tx.p, ctx.rank.p, n, rank_bits, ctx.sk0.p, s));
      HIP_TRY(launch::op_ids(ctx.ops.p, n, 0, ctx.sv0.p, s));
      uint32_t *skey, *sop2;
      HIP_TRY(launch::radix_sort_u32(ctx.sk0.p, ctx.sv0.p, ctx.sk1.p, ctx.sv1.p, n, row_bits + rank_bits,
                                     ctx.sort_tmp.p, &skey, &sop2, s));
      uint32_t* srow2 = ctx.S.p;  // free once the ranks are made
      HIP_TRY(launch::u32_key_high(skey, n, rank_bits, srow2, s));
      ordered_need(srow2, sop2);
      HIP_TRY(launch::fill_u64(ctx.conflicts.p, 1, 0, s));
      HIP_TRY(launch::gtx_conflicts32(skey, ctx.flags.p, n, ctx.conflicts.p, s));
    } else if (n) {
      HIP_TRY(launch::gtx_exec_keys(ctx.rows_out.p, ctx.op_tx.p, ctx.rank.p, n, rank_bits, ctx.sk64a.p, s));
      HIP_TRY(launch::op_ids(ctx.ops.p, n, 0, ctx.sv0.p, s));
      uint64_t* skey;
      uint32_t* sop2;
      HIP_TRY(launch::radix_sort_u64(ctx.sk64a.p, ctx.sv0.p, ctx.sk64b.p, ctx.sv1.p, n, row_bits + rank_bits,
                                     ctx.sort_tmp.p, &skey, &sop2, s));
      HIP_TRY(launch::u64_key_high(skey, n, rank_bits, ctx.sk0.p, s));
      ordered_need(ctx.sk0.p, sop2);
      HIP_TRY(launch::fill_u64(ctx.conflicts.p, 1, 0, s));
      HIP_TRY(launch::gtx_conflicts(skey, ctx.flags.p, n, ctx.conflicts.p, s));
    }
  }
  // need[] from accesses in execution order (group flags in ctx.flags, which
  // the conflict count reads afterwards).  A row occupies the same positions
  // in the (row, tx) and (row, rank, tx) orders, so ctx.SS (the row start of
  // every position, made for the ranks) makes each turn relative to its row,
  // and the counters start at zero: one dense fill instead of a random store
  // per row.
  void ordered_need(const uint32_t* srow, const uint32_t* sop) {
    const uint64_t n = ctx.nops;
    hipStream_t s = ctx.stream;
    HIP_TRY(launch::access_group_flags(srow, sop, n, ctx.flags.p, s));
    HIP_TRY(launch::inclusive_max_u32(ctx.flags.p, ctx.gstart.p, n, ctx.sort_tmp.p, s));
    HIP_TRY(launch::access_need_rel(sop, ctx.gstart.p, ctx.SS.p, n, ctx.need.p, s));
    HIP_TRY(launch::fill_u32(ctx.counters.p, ctx.counters.n, 0, s));
  }
  void Init(int, uint64_t) override {}
  // the persistent form runs every rank level in one launch: it needs no
  // per-rank counts (and no host round trip for them)
  void PrepStream() override { rank_all(); }
  bool ProbesRows() const override { return true; }
  void PostStream() override {}
  void Launch(int, uint64_t lo, uint64_t hi, bool persist) override {
    ExecArgs a = args(0, lo, hi);
    a.tx_map = ctx.tx_map.p;
    a.tx_rank = ctx.rank.p;
    if (launch_rows(a, hi - lo, persist)) return;
    HIP_TRY(launch::exec_ordered(ctx.G, a, hi - lo, ctx.threads, persist, ctx.stream));
  }
  void Post(int, uint64_t, uint64_t) override {}
  uint64_t GetMemSize() const override { return ctx.counters.n * 4 + ctx.need.n * 4 + ctx.rank.n * 8; }
};

namespace {

int hip_fail(const HipFailure& f) {
  return set_error(GCCTB_EDEVICE, std::string("HIP error: ") + hipGetErrorString(f.err) + " in " + f.where);
}

bool index_ready(const gcctb_ctx& c) { return c.ix.buckets || c.ix.dense; }

// Index over n keys: the identity (dkeys and dvals null) is a dense index,
// direct addressing with no table, unless GCCTB_INDEX_HASH=1 asks for the
// hash table over keys 0 .. n-1; other keys (device arrays) go into the hash
// table.
void build_index(gcctb_ctx& c, const uint32_t* dkeys, const uint32_t* dvals, uint64_t n) {
  const bool dense = !dkeys && !dvals && !c.index_hash;
  const uint64_t nb = dense ? 0 : idx_buckets_for(n);
  if (nb > 0xffffffffull || n > 0xffffffffull) throw HipFailure{hipErrorInvalidValue, "index too large"};
  // A captured graph holds the old bucket array in its kernels' arguments:
  // it is rebuilt on its next launch, and dropped now, before the buckets it
  // points to are freed.
  c.inputs_gen++;
  if (c.graph_exec) {
    (void)hipGraphExecDestroy(c.graph_exec);
    c.graph_exec = nullptr;
  }
  c.ix = IndexView{};
  c.buckets.release();
  if (dense) {
    c.ix.length = (uint32_t)n;
    c.ix.dense = 1;
    return;
  }
  c.buckets.alloc(nb * 2 * IDX_SLOTS);
  HIP_TRY(launch::fill_u32(c.buckets.p, nb * 2 * IDX_SLOTS, IDX_EMPTY, c.stream));
  HIP_TRY(launch::index_insert(dkeys, dvals, n, c.buckets.p, (uint32_t)nb, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
  c.ix.buckets = c.buckets.p;
  c.ix.nbuckets = (uint32_t)nb;
  c.ix.length = (uint32_t)n;
}

// Copy the error word (the kernels OR their flags into metrics slot 0) into
// pinned host memory behind the work queued so far.
void enqueue_error_copy(gcctb_ctx& c) {
  HIP_TRY(hipMemcpyAsync(c.h_err, &c.metrics.p[0].error_flags, sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream));
}

// After a launch that reported an error (a watchdog stop above all), CC words
// can still hold the locks or marks of transactions that never finished: 2PL
// releases its locks only at commit, so its per-batch Post has nothing to
// restore after a clean batch.  Every CC word, chain head and turn counter is
// cleared, which is the state outside a run.
void clear_cc_state(gcctb_ctx& c) {
  HIP_TRY(launch::fill_u64(c.cc.p, c.cc.n, 0, c.stream));
  if (c.vprev.p) HIP_TRY(launch::fill_u32(c.vprev.p, c.vprev.n, 0, c.stream));
  if (c.counters.p) HIP_TRY(launch::fill_u32(c.counters.p, c.counters.n, 0, c.stream));
  HIP_TRY(hipStreamSynchronize(c.stream));
}

void ensure_state(gcctb_ctx& c) {
  const uint64_t rows = c.cfg.rows;
  const uint64_t words = slot_words(rows);  // per-row state at each row's slot
  // the latch schemes keep the row's latch in the word after its CC word
  const uint64_t cc_words = c.latch ? 2 * words : words;
  c.cc.alloc(cc_words);
  HIP_TRY(hipMemsetAsync(c.cc.p, 0, cc_words * 8, c.stream));
  if (c.scheme == SCHEME_MVCC) {
    c.vprev.alloc(words);
    HIP_TRY(hipMemsetAsync(c.vprev.p, 0, words * 4, c.stream));
  }
  if (c.scheme == SCHEME_GACCO || c.scheme == SCHEME_GPUTX) {
    c.counters.alloc(words);
    HIP_TRY(hipMemsetAsync(c.counters.p, 0, words * 4, c.stream));
  }
  c.ctrs.alloc(CTR_WORDS);
  HIP_TRY(hipMemsetAsync(c.ctrs.p, 0, CTR_WORDS * sizeof(uint64_t), c.stream));
  c.metrics.alloc(METRIC_SLOTS);
}

}  // namespace
}  // namespace gcctb

using namespace gcctb;

extern "C" {

const char* gcctb_last_error(void) { return g_err.c_str(); }

int gcctb_create(const gcctb_config* cfg, gcctb_ctx** out) {
  if (!cfg || !out) return set_error(GCCTB_EINVAL, "null argument");
  *out = nullptr;
  const int latch_base = latch_scheme_from_name(cfg->cc_scheme);
  const int sch = latch_base >= 0 ? latch_base : scheme_from_name(cfg->cc_scheme);
  if (sch < 0) return set_error(GCCTB_EINVAL, std::string("unknown cc_scheme ") + (cfg->cc_scheme ? cfg->cc_scheme : "(null)"));
  if (latch_base >= 0 && (cfg->row_ops || cfg->debug_events))
    return set_error(GCCTB_EINVAL, "the latch schemes run in column mode without the event log");
  if (cfg->tpcc && (latch_base >= 0 || cfg->row_ops || cfg->debug_events || cfg->cols != TPCC_COLS))
    return set_error(GCCTB_EINVAL, "TPC-C runs the eight schemes on rows of 5 cells, without row_ops or the event log");
  // benchmark_cpu/ycsb.cu:97-98: warp_cnt > 32 || valid_txn_bitoffset > 5 -> -1
  if (cfg->block_size < 1 || cfg->block_size > 32 || cfg->warp_density < 0 || cfg->warp_density > 5)
    return set_error(GCCTB_EINVAL, "warp_density must be 0..5 and block_size 1..32");
  if (cfg->batch_size < 1) return set_error(GCCTB_EINVAL, "batch_size must be >= 1");
  if (cfg->rows < 1 || cfg->rows > 0xfffffffeull) return set_error(GCCTB_EINVAL, "rows out of range");
  if (cfg->cols < 1 || cfg->cols > 255) return set_error(GCCTB_EINVAL, "cols out of range");
  // The adaptive admission window is not part of the library: with Silo it
  // once made a kernel access memory out of bounds (and once return reads
  // that differ from the serial replay), its cause was never shown, and its
  // code was taken out of the kernels.
  if (cfg->persist_adaptive)
    return set_error(GCCTB_EINVAL, "persist_adaptive is not supported; use a static persist_window");
  if (cfg->row_ops && cfg->cols > (uint32_t)ROW_MAX_COLS)
    return set_error(GCCTB_EINVAL, "whole-row mode takes rows of at most 16 cells");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(GCCTB_EDEVICE, "no HIP device: the transaction kernels need an MI355X (gfx950)");
  auto* c = new (std::nothrow) gcctb_ctx();
  if (!c) return set_error(GCCTB_ENOMEM, "out of host memory");
  c->scheme = (Scheme)sch;
  c->latch = latch_base >= 0;
  c->tpcc = cfg->tpcc != 0;
  c->cfg = *cfg;
  c->cfg.cc_scheme = nullptr;
  c->G = 1 << cfg->warp_density;
  c->threads = cfg->block_size * 32;
  {
    const char* ih = std::getenv("GCCTB_INDEX_HASH");
    c->index_hash = ih && std::atoi(ih) == 1;
    // watchdog: 20 s of the 100 MHz wall clock unless GCCTB_WATCHDOG_S says otherwise
    const char* wd = std::getenv("GCCTB_WATCHDOG_S");
    const double sec = wd ? std::atof(wd) : 20.0;
    c->watchdog_ticks = sec > 0 ? (uint64_t)(sec * 1e8) : 0;
    // persistent form: aborts after which a transaction pauses admission
    // until it commits (GCCTB_URGENT_ATTEMPTS, 0 = never)
    if (const char* ua = std::getenv("GCCTB_URGENT_ATTEMPTS")) c->urgent_attempts = (uint32_t)std::atoi(ua);
    if (const char* pp = std::getenv("GCCTB_PROBE_PASS")) c->probe_pass = std::atoi(pp) != 0 ? 1 : 0;
    // retry delay: at most 2^GCCTB_BACKOFF_EXP units (0..16, default 10)
    if (const char* be = std::getenv("GCCTB_BACKOFF_EXP")) {
      const int e = std::atoi(be);
      c->backoff_exp = (uint32_t)(e < 0 ? 0 : e > 16 ? 16 : e);
    }
    // GaccO / GPUTx: assumed handoff time per turn on a row, in 10 ns ticks,
    // that paces a waiter's rest (GCCTB_POLL_TICKS, 1..1000, default 10)
    if (const char* pt = std::getenv("GCCTB_POLL_TICKS")) {
      const int t = std::atoi(pt);
      c->poll_ticks = (uint32_t)(t < 1 ? 1 : t > 1000 ? 1000 : t);
    }
    if (const char* ds = std::getenv("GCCTB_DEFER_SHARE")) {
      const int d = std::atoi(ds);
      c->defer_share = d < 0 ? 0 : d > 65536 ? 65536 : d;
    }
    if (const char* hl = std::getenv("GCCTB_HOT_LAST")) c->hot_last = std::atoi(hl) != 0;
    if (const char* cm = std::getenv("GCCTB_COMBINE_MIN")) {
      const long v = std::atol(cm);
      c->comb_min = (uint32_t)(v < 0 ? 0 : v > 0x7fffffffL ? 0x7fffffffL : v);
    }
    const char* dump = std::getenv("GCCTB_DEBUG_DUMP");
    if (dump && *dump) c->dbg_path = dump;
  }
  try {
    HIP_TRY(hipSetDevice(0));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipHostMalloc((void**)&c->h_err, sizeof(uint64_t), hipHostMallocDefault));
    *c->h_err = 0;
    switch (c->scheme) {
      case SCHEME_TPL_NW: c->host.reset(new TwoPLHost(*c, false)); break;
      case SCHEME_TPL_WD: c->host.reset(new TwoPLHost(*c, true)); break;
      case SCHEME_TO: c->host.reset(new TOHost(*c)); break;
      case SCHEME_MVCC: c->host.reset(new MVCCHost(*c)); break;
      case SCHEME_SILO: c->host.reset(new SiloHost(*c)); break;
      case SCHEME_TICTOC: c->host.reset(new TictocHost(*c)); break;
      case SCHEME_GACCO: c->host.reset(new GaccoHost(*c)); break;
      case SCHEME_GPUTX: c->host.reset(new GPUTxHost(*c)); break;
      default: break;
    }
    c->cells.alloc(cfg->rows * cfg->cols);
    ensure_state(*c);
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& f) {
    const int rc = hip_fail(f);
    gcctb_destroy(c);
    return rc;
  }
  *out = c;
  return GCCTB_OK;
}

void gcctb_destroy(gcctb_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->host.reset();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_err) (void)hipHostFree(c->h_err);
  hipStream_t s = c->stream, s2 = c->stream2;
  if (s2) (void)hipStreamSynchronize(s2);
  delete c;  // DevBuf destructors free device memory
  if (s) (void)hipStreamDestroy(s);
  if (s2) (void)hipStreamDestroy(s2);
}

int gcctb_build_index_identity(gcctb_ctx* c) {
  if (!c) return set_error(GCCTB_EINVAL, "null context");
  try {
    build_index(*c, nullptr, nullptr, c->cfg.rows);
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_build_index(gcctb_ctx* c, const uint32_t* keys, const uint32_t* vals, uint64_t n) {
  if (!c || (!keys && n) || (!vals && n)) return set_error(GCCTB_EINVAL, "null argument");
  if (n == 0 || n > 0xfffffffeull) return set_error(GCCTB_EINVAL, "index length out of range");
  std::vector<std::pair<uint32_t, uint32_t>> kv(n);
  for (uint64_t i = 0; i < n; ++i) {
    if (keys[i] == 0xffffffffu) return set_error(GCCTB_EINVAL, "key 0xffffffff is reserved");
    kv[i] = {keys[i], vals[i]};
  }
  std::sort(kv.begin(), kv.end());  // SortedArray_CPU ctor sorts (include/index.cuh:106-108)
  for (uint64_t i = 1; i < n; ++i)
    if (kv[i].first == kv[i - 1].first) return set_error(GCCTB_EINVAL, "duplicate index key");
  std::vector<uint32_t> k(n), v(n);
  for (uint64_t i = 0; i < n; ++i) {
    k[i] = kv[i].first;
    v[i] = kv[i].second;
  }
  // The kernels assume that the ops of one transaction touch distinct rows;
  // load_ops checks that their keys are distinct, so the index must map
  // distinct keys to distinct rows (a primary-key index, key -> row id).
  {
    std::vector<uint32_t> sv(v);
    std::sort(sv.begin(), sv.end());
    for (uint64_t i = 1; i < n; ++i)
      if (sv[i] == sv[i - 1]) return set_error(GCCTB_EINVAL, "two index keys map to one row");
  }
  // keys 0 .. n-1, each on the row of its own number: the dense index
  bool identity = true;
  for (uint64_t i = 0; i < n && identity; ++i) identity = k[i] == i && v[i] == i;
  try {
    if (identity) {
      build_index(*c, nullptr, nullptr, n);
      return GCCTB_OK;
    }
    DevBuf<uint32_t> dk, dv;
    dk.alloc(n);
    dv.alloc(n);
    HIP_TRY(hipMemcpyAsync(dk.p, k.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dv.p, v.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    build_index(*c, dk.p, dv.p, n);
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_probe(gcctb_ctx* c, const uint32_t* keys, uint64_t n, uint32_t* rows_out) {
  if (!c || (n && (!keys || !rows_out))) return set_error(GCCTB_EINVAL, "null argument");
  if (!index_ready(*c)) return set_error(GCCTB_ESTATE, "no index");
  if (n == 0) return GCCTB_OK;
  try {
    std::vector<uint64_t> ops(n);
    for (uint64_t i = 0; i < n; ++i) ops[i] = make_op(keys[i], 0, false);
    DevBuf<uint64_t> dops;
    DevBuf<uint32_t> drows;
    dops.alloc(n);
    drows.alloc(n);
    HIP_TRY(hipMemcpyAsync(dops.p, ops.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch::probe(c->ix, dops.p, n, drows.p, c->stream));
    HIP_TRY(hipMemcpyAsync(rows_out, drows.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // Find's contract: a miss returns the index length (include/index.cuh:89-90)
    for (uint64_t i = 0; i < n; ++i)
      if (rows_out[i] == IDX_MISS) rows_out[i] = c->ix.length;
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_load_table(gcctb_ctx* c, const uint64_t* cells, uint64_t rows, uint32_t cols) {
  if (!c || !cells) return set_error(GCCTB_EINVAL, "null argument");
  if (rows != c->cfg.rows || cols != c->cfg.cols) return set_error(GCCTB_EINVAL, "table shape differs from the config");
  try {
    HIP_TRY(hipMemcpyAsync(c->cells.p, cells, rows * cols * 8, hipMemcpyHostToDevice, c->stream));
    ensure_state(*c);
    if (!index_ready(*c)) build_index(*c, nullptr, nullptr, rows);
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  c->table_ready = true;
  return GCCTB_OK;
}

int gcctb_seed_table(gcctb_ctx* c) {
  if (!c) return set_error(GCCTB_EINVAL, "null context");
  try {
    HIP_TRY(launch::seed_table(c->cells.p, c->cfg.rows, c->cfg.cols, c->cfg.seed, c->stream));
    ensure_state(*c);
    if (!index_ready(*c)) build_index(*c, nullptr, nullptr, c->cfg.rows);
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  c->table_ready = true;
  return GCCTB_OK;
}

int gcctb_check_ops(const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, uint32_t cols, uint64_t* nreads_out) {
  if (!tx_off || (!ops && ntx)) return set_error(GCCTB_EINVAL, "null argument");
  if (ntx >= (1ull << 31)) return set_error(GCCTB_EINVAL, "at most 2^31-1 transactions (31-bit lock holder ids)");
  if (tx_off[0] != 0) return set_error(GCCTB_EINVAL, "tx_off[0] must be 0");
  uint64_t nreads = 0;
  for (uint64_t t = 0; t < ntx; ++t) {
    if (tx_off[t + 1] < tx_off[t]) return set_error(GCCTB_EINVAL, "tx_off must be non-decreasing");
    const uint64_t len = tx_off[t + 1] - tx_off[t];
    if (len > (uint64_t)MAX_OPS) return set_error(GCCTB_EINVAL, "a transaction has more than 32 ops");
    // Each op of a transaction must name its own row: the kernels hold one
    // lock, mark or turn per op, and two ops on one row would conflict with
    // each other.  The reference's generator drops repeated keys
    // (generation/gen_ycsb.cu:129-136); keys are checked here, and
    // gcctb_build_index keeps distinct keys on distinct rows.
    uint32_t keys[MAX_OPS];
    for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
      if (op_col(ops[i]) >= cols) return set_error(GCCTB_EINVAL, "op column out of range");
      nreads += op_is_write(ops[i]) ? 0 : 1;
      const uint32_t key = op_key(ops[i]);
      const uint64_t j = i - tx_off[t];
      for (uint64_t q = 0; q < j; ++q)
        if (keys[q] == key)
          return set_error(GCCTB_EINVAL, "transaction " + std::to_string(t) + " names key " + std::to_string(key) +
                                             " twice (keys must be distinct within a transaction)");
      keys[j] = key;
    }
  }
  // op ids carry a write flag in bit 31 through the GaccO / GPUTx sorts
  if (tx_off[ntx] >= (1ull << 31)) return set_error(GCCTB_EINVAL, "at most 2^31-1 ops");
  if (nreads_out) *nreads_out = nreads;
  return GCCTB_OK;
}

int gcctb_load_ops(gcctb_ctx* c, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx) {
  if (!c) return set_error(GCCTB_EINVAL, "null argument");
  uint64_t nreads = 0;
  if (const int rc = gcctb_check_ops(ops, tx_off, ntx, c->cfg.cols, &nreads)) return rc;
  const uint64_t nops = tx_off[ntx];
  uint64_t cap = 0;
  // MVCC version pool: one node per write of the whole load (the persistent
  // form keeps every version of the stream; a plain batch needs at most that)
  if (c->scheme == SCHEME_MVCC) cap = nops - nreads;
  try {
    c->ntx = ntx;
    c->nops = nops;
    c->nreads = nreads;
    c->params_loaded = 0;  // TPC-C parameters belong to the ops loaded before
    c->inputs_gen++;
    c->custom_batches.clear();
    c->h_tx_off.assign(tx_off, tx_off + ntx + 1);
    c->ops.alloc(nops);
    c->tx_off.alloc(ntx + 1);
    c->op_tx.alloc(nops);
    c->rows_out.alloc(nops);
    c->reads.alloc(nops);
    c->obatch.alloc(ntx);
    c->okey0.alloc(ntx);
    c->okey1.alloc(ntx);
    c->tx_aborts.alloc(ntx);
    if (!c->dbg_path.empty()) {
      c->dbg.alloc(nops * 4);
      HIP_TRY(hipMemsetAsync(c->dbg.p, 0, c->dbg.n * sizeof(uint64_t), c->stream));
    }
    if (nops) HIP_TRY(hipMemcpyAsync(c->ops.p, ops, nops * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tx_off.p, tx_off, (ntx + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->reads.p, 0, c->reads.n * sizeof(uint64_t), c->stream));
    HIP_TRY(hipMemsetAsync(c->tx_aborts.p, 0, c->tx_aborts.n * sizeof(uint32_t), c->stream));
    HIP_TRY(launch::op_tx(c->tx_off.p, ntx, c->op_tx.p, c->stream));
    if (c->cfg.debug_events) {
      c->ev_op.alloc(nops);
      c->ev_commit.alloc(ntx);
      c->events_ready = false;
    }
    if (c->scheme == SCHEME_MVCC) {
      // node: ts word, link, the superseded cell (or row, in whole-row mode)
      const uint64_t stride = (c->cfg.row_ops || c->tpcc) ? 2 + (uint64_t)c->cfg.cols : 3;
      c->pool_cap = cap;
      c->pool.alloc(stride * cap + stride);
      // nodes are published before they are linked, but a zeroed pool keeps
      // an unpublished node from ever holding a link that points anywhere
      HIP_TRY(hipMemsetAsync(c->pool.p, 0, c->pool.n * sizeof(uint64_t), c->stream));
    }
    if (c->scheme != SCHEME_GACCO && c->scheme != SCHEME_GPUTX && ntx) {
      c->defer_cnt.alloc(launch::DEFER_WORDS);
      c->defer_map.alloc(ntx);
    }
    if (c->scheme == SCHEME_GACCO || c->scheme == SCHEME_GPUTX) {
      // sort buffers for the whole load (the persistent form prepares it at once)
      const uint64_t n = std::max(nops, ntx);
      c->need.alloc(nops);
      c->sk0.alloc(n);
      c->sv0.alloc(n);
      c->sk1.alloc(n);
      c->sv1.alloc(n);
      c->flags.alloc(n);
      c->gstart.alloc(n);
      c->sort_tmp.alloc(std::max(launch::radix_sort_tmp_words(n), launch::scan_tmp_words(n)));
      if (c->scheme == SCHEME_GACCO && c->comb_min) {
        // at most nops / comb_min queues reach comb_min accesses
        c->comb_cap = std::min<uint64_t>(nops / c->comb_min + 1, 0x7fffffffull);
        c->comb_segs.alloc(c->comb_cap);
        c->comb_ctr.alloc(2);
        c->comb_wv.alloc(nops);  // every op's write value, for the combiners
      }
      if (c->scheme == SCHEME_GPUTX) {
        c->S.alloc(n);
        c->SS.alloc(n);
        c->sk64a.alloc(n);
        c->sk64b.alloc(n);
        c->rank.alloc(ntx);
        c->tx_map.alloc(ntx);
        c->max_rank.alloc(1);
        c->conflicts.alloc(1);
      }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_load_params(gcctb_ctx* c, const uint32_t* params, uint64_t n) {
  if (!c || (!params && n)) return set_error(GCCTB_EINVAL, "null argument");
  if (!c->tpcc) return set_error(GCCTB_EINVAL, "op parameters belong to a TPC-C context");
  if (n != c->nops) return set_error(GCCTB_EINVAL, "one parameter per loaded op");
  try {
    c->op_param.alloc(n ? n : 1);
    if (n) HIP_TRY(hipMemcpy(c->op_param.p, params, n * 4, hipMemcpyHostToDevice));
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  c->params_loaded = n + 1;
  return GCCTB_OK;
}

enum RunForm { FORM_PLAIN = 0, FORM_GRAPH = 1, FORM_PERSIST = 2 };
// Largest per-batch sequence the graph form captures (GPUTx inputs can have
// one rank level per launch by the hundred thousand).
constexpr uint64_t GRAPH_MAX_LAUNCHES = 4096;

// Totals of the per-slot metrics the kernels accumulated since the last clear.
static Metrics collect_metrics(gcctb_ctx* c) {
  Metrics m{};
  std::vector<Metrics> slots(METRIC_SLOTS);
  HIP_TRY(hipMemcpy(slots.data(), c->metrics.p, sizeof(Metrics) * METRIC_SLOTS, hipMemcpyDeviceToHost));
  for (const Metrics& x : slots) {
    m.abort += x.abort;
    m.ts_duration += x.ts_duration;
    m.wait_duration += x.wait_duration;
    m.abort_duration += x.abort_duration;
    m.manager_duration += x.manager_duration;
    m.tot_duration += x.tot_duration;
    m.index_duration += x.index_duration;
    m.committed += x.committed;
    m.index_miss += x.index_miss;
    m.error_flags |= x.error_flags;
  }
  return m;
}

static std::string error_text(uint64_t flags) {
  std::string what;
  if (flags & ERR_TS_OVERFLOW) what += " timestamp-overflow";
  if (flags & ERR_POOL_OVERFLOW) what += " version-pool-full";
  if (flags & ERR_CHAIN_BROKEN) what += " version-chain-broken";
  if (flags & ERR_TXLEN) what += " transaction-too-long";
  if (flags & ERR_WATCHDOG) what += " watchdog";
  return "kernel error flags " + std::to_string(flags) + ":" + what;
}

// One pass over every loaded transaction.
//   plain:      the reference's structure -- per batch Init, one launch, Post,
//               host synchronization (benchmark_cpu/ycsb.cu:262-285)
//   graph:      the same per-batch kernels captured once into a hipGraph and
//               replayed, with no host round trip between batches
//   persistent: one preparation of the whole stream and one persistent launch
static int run_form(gcctb_ctx* c, gcctb_stats* out, RunForm form) {
  if (!c) return set_error(GCCTB_EINVAL, "null context");
  if (!c->table_ready) return set_error(GCCTB_ESTATE, "load or seed the table first");
  if (form == FORM_PERSIST && (c->cfg.row_ops || c->cfg.debug_events || c->latch || c->tpcc))
    return set_error(GCCTB_EINVAL, "whole-row mode, the event log, the latch schemes and TPC-C run the plain and graph forms");
  // a TPC-C op reads its update's parameter: none may be missing
  if (c->tpcc && c->params_loaded != c->nops + 1)
    return set_error(GCCTB_ESTATE, "load the TPC-C op parameters (gcctb_load_params) after the ops");
  gcctb_stats st{};
  Metrics m{};
  try {
    hipStream_t s = c->stream;
    HIP_TRY(hipMemsetAsync(c->metrics.p, 0, sizeof(Metrics) * METRIC_SLOTS, s));
    if (c->cfg.debug_events) {
      // a fresh log per run: no stamps, the stamp counter at zero
      if (c->nops) HIP_TRY(launch::fill_u64(c->ev_op.p, c->nops, 0, s));
      if (c->ntx) HIP_TRY(launch::fill_u64(c->ev_commit.p, c->ntx, 0, s));
      HIP_TRY(launch::fill_u64(c->ctrs.p + CTR_EVENT, 1, 0, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    *c->h_err = 0;
    float ms = 0;
    double pre = 0, proc = 0;
    uint64_t launches = 0;
    HIP_TRY(hipEventRecord(c->ev[0], s));
    // hot-writer deferral (lock, timestamp and OCC schemes; the ordered
    // schemes wait on lower slots and keep the stream order)
    const uint32_t share = c->defer_share >= 0                                          ? (uint32_t)c->defer_share
                           : (c->scheme == SCHEME_SILO || c->scheme == SCHEME_TICTOC) ? 0u
                                                                                        : 16u;
    const bool defer = form == FORM_PERSIST && share && c->defer_map.p && c->scheme != SCHEME_GACCO &&
                       c->scheme != SCHEME_GPUTX;
    // 2PL with the deferral also locks hot keys last
    const bool hot = defer && c->hot_last && (c->scheme == SCHEME_TPL_NW || c->scheme == SCHEME_TPL_WD);
    if (form == FORM_PERSIST) {
      c->host->PrepStream();
      if (defer) {
        HIP_TRY(launch::defer_order(c->tx_off.p, c->ops.p, c->ntx, c->nops, share, c->defer_cnt.p, c->defer_map.p, s));
        c->defer_thr = launch::defer_threshold(c->ntx, c->nops, share);
      }
      // Rows: GaccO's and GPUTx's preparation already has them; for the
      // other schemes a probe pass finds every op's row first, or each group
      // looks its rows up as it takes a transaction (WantsProbePass).
      const bool pass = c->probe_pass >= 0 ? c->probe_pass != 0 : c->host->WantsProbePass();
      const bool probed = c->host->ProbesRows() || pass;
      if (!c->host->ProbesRows() && pass && c->nops)
        HIP_TRY(launch::probe(c->ix, c->ops.p, c->nops, c->rows_out.p, s));
      HIP_TRY(launch::persist_begin(c->ctrs.p, s));
      HIP_TRY(hipEventRecord(c->ev[3], s));
      c->rows_probed = probed;
      c->defer_active = defer;
      c->hot_active = hot;
      c->host->Launch(0, 0, c->ntx, true);
      c->defer_active = false;
      c->hot_active = false;
      c->rows_probed = false;
      HIP_TRY(hipEventRecord(c->ev[4], s));
      c->host->PostStream();
      enqueue_error_copy(*c);
      HIP_TRY(hipEventRecord(c->ev[5], s));
      HIP_TRY(hipEventSynchronize(c->ev[5]));
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
      pre += ms;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
      pre += ms;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
      proc += ms;
      launches = 1;
    } else if (form == FORM_GRAPH) {
      // Init0 may synchronize with the host (GPUTx ranks): it stays outside
      c->host->Init0();
      const auto& batches = c->host->batches;
      uint64_t nlaunch = 0;
      for (uint64_t n : batches) nlaunch += n ? 1 : 0;
      if (nlaunch > GRAPH_MAX_LAUNCHES)
        return set_error(GCCTB_EINVAL, "the graph form captures at most " + std::to_string(GRAPH_MAX_LAUNCHES) +
                                           " launches; this input needs " + std::to_string(nlaunch));
      if (!c->graph_exec || c->graph_gen != c->inputs_gen) {
        if (c->graph_exec) {
          (void)hipGraphExecDestroy(c->graph_exec);
          c->graph_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        uint64_t batch_st = 0;
        for (size_t i = 0; i < batches.size(); ++i) {
          const uint64_t n = batches[i];
          if (!n) continue;
          c->host->Init((int)i, batch_st);
          c->host->Launch((int)i, batch_st, batch_st + n, false);
          c->host->Post((int)i, batch_st, n);
          batch_st += n;
        }
        HIP_TRY(hipStreamEndCapture(s, &g));
        const hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(ie);
        c->graph_gen = c->inputs_gen;
      }
      for (uint64_t n : batches) launches += n ? 1 : 0;
      HIP_TRY(hipEventRecord(c->ev[3], s));
      HIP_TRY(hipGraphLaunch(c->graph_exec, s));
      HIP_TRY(hipEventRecord(c->ev[4], s));
      HIP_TRY(hipEventSynchronize(c->ev[4]));
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[3]));
      pre += ms;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
      proc += ms;  // kernels and per-batch preparation inside the graph
    } else {
    c->host->Init0();
    HIP_TRY(hipEventRecord(c->ev[1], s));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    pre += ms;
    uint64_t batch_st = 0;
    const auto& batches = c->host->batches;
    for (size_t i = 0; i < batches.size(); ++i) {
      const uint64_t n = batches[i];
      if (!n) continue;  // benchmark_cpu/ycsb.cu:266-267
      HIP_TRY(hipEventRecord(c->ev[2], s));
      c->host->Init((int)i, batch_st);
      HIP_TRY(hipEventRecord(c->ev[3], s));
      c->host->Launch((int)i, batch_st, batch_st + n, false);
      HIP_TRY(hipEventRecord(c->ev[4], s));
      c->host->Post((int)i, batch_st, n);
      enqueue_error_copy(*c);
      HIP_TRY(hipEventRecord(c->ev[5], s));
      HIP_TRY(hipEventSynchronize(c->ev[5]));
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
      pre += ms;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[4], c->ev[5]));
      pre += ms;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[3], c->ev[4]));
      proc += ms;
      batch_st += n;
      launches++;
      if (*c->h_err) break;  // the first batch that reported an error ends the run
    }
    }
    HIP_TRY(hipEventRecord(c->ev[6], s));
    HIP_TRY(hipEventSynchronize(c->ev[6]));
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[6]));
    st.tot_ms = ms;
    st.pre_ms = pre;
    st.proc_ms = proc;
    st.exec_kernel_ms = proc;
    st.batches = launches;
    m = collect_metrics(c);
    if (m.error_flags) clear_cc_state(*c);
    if (c->scheme == SCHEME_GPUTX) {
      st.ranks = c->nranks;
      if (c->nops) HIP_TRY(hipMemcpy(&st.same_rank_conflicts, c->conflicts.p, 8, hipMemcpyDeviceToHost));
    }
    if (form == FORM_PERSIST) st.combined = c->host->Combined();
    if (defer && c->ntx) {
      uint32_t dst[3] = {0, 0, 0};
      HIP_TRY(hipMemcpy(dst, c->defer_cnt.p + launch::DEFER_BUCKETS, sizeof dst, hipMemcpyDeviceToHost));
      st.deferred = dst[launch::DEFER_ANY] ? dst[launch::DEFER_NHOT] : 0;
    }
    int rate = 0;
    HIP_TRY(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, 0));
    st.tick_khz = rate;
    if ((m.error_flags & ERR_WATCHDOG) && c->dbg.p) {
      std::vector<uint64_t> d(c->dbg.n);
      HIP_TRY(hipMemcpy(d.data(), c->dbg.p, d.size() * 8, hipMemcpyDeviceToHost));
      if (FILE* f = std::fopen(c->dbg_path.c_str(), "a")) {
        std::fprintf(f, "# op tx meta row ld_word atomic_word phase attempts\n");
        for (uint64_t t = 0; t < c->ntx; ++t)
          for (uint64_t i = c->h_tx_off[t]; i < c->h_tx_off[t + 1]; ++i) {
            const uint64_t* e = &d[i * 4];
            if (!(e[3] & 1)) continue;
            std::fprintf(f, "%llu %llu %llx %llu %llx %llx %llu %llu\n", (unsigned long long)i,
                         (unsigned long long)t, (unsigned long long)(e[0] >> 32),
                         (unsigned long long)(e[0] & 0xffffffffu), (unsigned long long)e[1],
                         (unsigned long long)e[2], (unsigned long long)(e[3] >> 56),
                         (unsigned long long)((e[3] >> 24) & 0xffffffffu));
          }
        std::fclose(f);
      }
    }
  } catch (const HipFailure& f) {
    c->rows_probed = false;
    c->defer_active = false;
    c->hot_active = false;
    return hip_fail(f);
  }
  st.txs = c->ntx;
  st.committed = m.committed;
  st.aborts = m.abort;
  st.ts_ticks = m.ts_duration;
  st.wait_ticks = m.wait_duration;
  st.abort_ticks = m.abort_duration;
  st.manager_ticks = m.manager_duration;
  st.tot_ticks = m.tot_duration;
  st.index_ticks = m.index_duration;
  st.index_miss = m.index_miss;
  st.error_flags = m.error_flags;
  st.total_ops = c->nops;
  st.read_ops = c->nreads;
  c->last_committed = m.committed;
  c->events_ready = c->cfg.debug_events != 0;
  if (out) *out = st;
  if (m.error_flags) return set_error(GCCTB_EKERNEL, error_text(m.error_flags));
  if (m.committed != c->ntx) return set_error(GCCTB_EKERNEL, "not every transaction committed");
  return GCCTB_OK;
}

int gcctb_run_plain(gcctb_ctx* c, gcctb_stats* out) { return run_form(c, out, FORM_PLAIN); }
int gcctb_run_graph(gcctb_ctx* c, gcctb_stats* out) { return run_form(c, out, FORM_GRAPH); }
int gcctb_run_persistent(gcctb_ctx* c, gcctb_stats* out) { return run_form(c, out, FORM_PERSIST); }

// ---------------------------------------------------------------------------
// Serving form: continuous batching.  Requests arrive as batches at given
// times; table, index and CC state stay resident on the device.  One open
// persistent launch runs while requests arrive.  The host admits each batch
// as it arrives: on the admission stream it copies the batch's ops and
// offsets from pinned host memory into device memory, prepares them (the
// probe pass for MVCC; GaccO's per-row turns), then raises the launch's
// admission word, and the launch's groups take the admitted slots as they
// free up, next to the transactions still running.  A launch is closed (its
// admission word gets the closed bit; it drains and its touched CC words are
// restored) and a new one opened when the next batch would exceed the
// admission limits: max_launch_txs transactions per launch, and, for MVCC,
// the version pool (pool_cap_bytes at 24 bytes per version, bounded by the
// allocated pool and by free device memory), whose versions stay live until
// their launch ends.  A launch is also closed before an idle gap longer than
// a quarter of the watchdog period (its waiting waves would otherwise reach
// the watchdog).  Latency: a transaction's commit stamp (device wall clock,
// mapped to host time through a clock sample taken at the start) minus its
// batch's arrival.
// ---------------------------------------------------------------------------
namespace {

struct PinnedBuf {
  uint64_t* p = nullptr;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + (ptrdiff_t)k, v.end());
  return v[k];
}

constexpr int ADMIT_RING = 64;  // staging blocks of admission words in flight

}  // namespace

int gcctb_serve(gcctb_ctx* c, const gcctb_serve_config* sc, const uint64_t* batch_sizes, const double* arrival_ms,
                uint64_t nbatches, double* tx_latency_ms, gcctb_serve_stats* out) {
  if (!c || !sc || (nbatches && !batch_sizes)) return set_error(GCCTB_EINVAL, "null argument");
  if (!c->table_ready) return set_error(GCCTB_ESTATE, "load or seed the table first");
  if (c->scheme == SCHEME_GPUTX)
    return set_error(GCCTB_EINVAL, "gputx ranks its whole input before running; it has no serving form");
  if (c->cfg.row_ops || c->cfg.debug_events || c->latch || c->tpcc)
    return set_error(GCCTB_EINVAL, "whole-row mode, the event log, the latch schemes and TPC-C run the plain and graph forms");
  uint64_t total = 0;
  for (uint64_t i = 0; i < nbatches; ++i) {
    if (batch_sizes[i] == 0) return set_error(GCCTB_EINVAL, "empty batch");
    total += batch_sizes[i];
    if (arrival_ms && (!(arrival_ms[i] >= 0) || (i && arrival_ms[i] < arrival_ms[i - 1])))
      return set_error(GCCTB_EINVAL, "arrival times must be non-negative and non-decreasing");
  }
  if (total != c->ntx) return set_error(GCCTB_EINVAL, "batch sizes must add up to the loaded transactions");
  const uint64_t max_txs = sc->max_launch_txs ? sc->max_launch_txs : c->ntx;
  uint64_t pool_nodes = ~0ull;
  if (c->scheme == SCHEME_MVCC) {
    // admission by memory: the allocated pool, the configured budget, and
    // what free device memory would hold besides the allocated pool
    pool_nodes = c->pool_cap;
    if (sc->pool_cap_bytes) pool_nodes = std::min(pool_nodes, sc->pool_cap_bytes / 24);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      pool_nodes = std::min<uint64_t>(pool_nodes, ((uint64_t)free_b + c->pool.n * 8) / 24);
  }
  gcctb_serve_stats st{};
  st.txs = c->ntx;
  st.pool_nodes = pool_nodes;
  using clk = std::chrono::steady_clock;
  const uint64_t saved_pool_cap = c->pool_cap;
  std::vector<hipEvent_t> ring_ev(ADMIT_RING, nullptr);
  hipEvent_t ev_open = nullptr, ev_done = nullptr, ev_s2 = nullptr;
  bool launch_open = false;  // an open launch is running (closed on the error path too)
  auto cleanup = [&]() {
    c->open_launch = false;
    c->rows_probed = false;
    c->record_commit = false;
    c->pool_cap = saved_pool_cap;
    for (auto& e : ring_ev)
      if (e) (void)hipEventDestroy(e);
    if (ev_open) (void)hipEventDestroy(ev_open);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (ev_s2) (void)hipEventDestroy(ev_s2);
  };
  try {
    hipStream_t s = c->stream;
    if (!c->stream2) HIP_TRY(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    hipStream_t s2 = c->stream2;
    for (auto& e : ring_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_open, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_s2, hipEventDisableTiming));
    std::vector<bool> ring_used(ADMIT_RING, false);
    // requests wait in pinned host memory until admitted: ops and offsets
    PinnedBuf host_ops, host_off, staging;
    if (c->nops) {
      HIP_TRY(hipHostMalloc((void**)&host_ops.p, c->nops * 8, hipHostMallocDefault));
      HIP_TRY(hipMemcpy(host_ops.p, c->ops.p, c->nops * 8, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipHostMalloc((void**)&host_off.p, (c->ntx + 1) * 8, hipHostMallocDefault));
    std::memcpy(host_off.p, c->h_tx_off.data(), (c->ntx + 1) * 8);
    constexpr uint64_t ADMIT_WORDS = 16 * GATE_SHARDS;
    HIP_TRY(hipHostMalloc((void**)&staging.p, ADMIT_RING * ADMIT_WORDS * 8, hipHostMallocDefault));
    std::memset(staging.p, 0, ADMIT_RING * ADMIT_WORDS * 8);
    if (c->scheme == SCHEME_GACCO) c->rowbase.alloc(c->cfg.rows);
    c->clock_word.alloc(1);
    // per-batch first transactions and write counts (version-pool admission)
    std::vector<uint64_t> bfirst(nbatches + 1, 0), bwrites(nbatches, 0);
    for (uint64_t i = 0; i < nbatches; ++i) {
      bfirst[i + 1] = bfirst[i] + batch_sizes[i];
      for (uint64_t o = c->h_tx_off[bfirst[i]]; o < c->h_tx_off[bfirst[i + 1]]; ++o)
        bwrites[i] += op_is_write(host_ops.p[o]) ? 1 : 0;
      if (bwrites[i] > pool_nodes) {
        cleanup();
        return set_error(GCCTB_EINVAL, "a batch has more writes than the version pool holds");
      }
    }
    HIP_TRY(hipMemsetAsync(c->metrics.p, 0, sizeof(Metrics) * METRIC_SLOTS, s));
    HIP_TRY(hipStreamSynchronize(s));
    *c->h_err = 0;
    if (c->scheme == SCHEME_MVCC) c->pool_cap = std::min(c->pool_cap, pool_nodes);
    c->tcommit.alloc(std::max<uint64_t>(1, c->ntx));
    HIP_TRY(launch::fill_u64(c->tcommit.p, c->tcommit.n, 0, s));
    c->record_commit = true;
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
    const double ticks_per_ms = khz > 0 ? (double)khz : 1e5;
    const double idle_close_ms = c->watchdog_ticks ? 0.25 * (double)c->watchdog_ticks / ticks_per_ms : 1e300;
    const auto t0 = clk::now();
    auto since = [&](clk::time_point t) { return std::chrono::duration<double, std::milli>(t - t0).count(); };
    // clock sample: device tick d0 at host time h0 (midpoint of the round trip)
    double h0 = 0;
    uint64_t d0 = 0;
    {
      const auto a0 = clk::now();
      HIP_TRY(launch::clock_sample(c->clock_word.p, s));
      HIP_TRY(hipMemcpyAsync(&d0, c->clock_word.p, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      const auto a1 = clk::now();
      h0 = 0.5 * (since(a0) + since(a1));
    }
    std::vector<double> admit_ms(nbatches, 0.0);
    // the open launch
    struct Session {
      bool open = false;
      uint64_t lo = 0, hi = 0, writes = 0;
      double start_ms = 0;
    } sess;
    uint64_t ring_pos = 0, launches = 0;
    double busy = 0, launch_txs = 0;
    auto write_admit = [&](uint32_t value) {
      const uint64_t k = ring_pos++ % ADMIT_RING;
      if (ring_used[k]) HIP_TRY(hipEventSynchronize(ring_ev[k]));  // its last copy has read the block
      uint64_t* blk = staging.p + k * ADMIT_WORDS;
      for (int g = 0; g < GATE_SHARDS; ++g) blk[16 * g] = value;  // the u32 word (low half), zero above
      HIP_TRY(hipMemcpyAsync(c->ctrs.p + CTR_ADMIT0, blk, ADMIT_WORDS * 8, hipMemcpyHostToDevice, s2));
      HIP_TRY(hipEventRecord(ring_ev[k], s2));
      ring_used[k] = true;
    };
    auto open_session = [&](uint64_t lo) {
      c->host->PrepOpen();
      // the admission stream starts on this launch only after its reset
      HIP_TRY(hipEventRecord(ev_open, s));
      HIP_TRY(hipStreamWaitEvent(s2, ev_open, 0));
      c->open_launch = true;
      c->rows_probed = c->host->OpenRowsProbed();
      c->host->Launch((int)launches, lo, c->ntx, true);  // capacity: every remaining slot
      c->open_launch = false;
      c->rows_probed = false;
      sess.open = launch_open = true;
      sess.lo = sess.hi = lo;
      sess.writes = 0;
      sess.start_ms = since(clk::now());
      ++launches;
    };
    auto close_session = [&]() {
      write_admit((uint32_t)(sess.hi - sess.lo) | ADMIT_CLOSED);
      // the restore reads rows the admission stream wrote: order it after
      // that stream's work
      HIP_TRY(hipEventRecord(ev_s2, s2));
      HIP_TRY(hipStreamWaitEvent(s, ev_s2, 0));
      c->host->Post((int)launches - 1, sess.lo, sess.hi - sess.lo);
      enqueue_error_copy(*c);
      HIP_TRY(hipEventRecord(ev_done, s));
      HIP_TRY(hipEventSynchronize(ev_done));
      busy += since(clk::now()) - sess.start_ms;
      launch_txs += (double)(sess.hi - sess.lo);
      sess.open = launch_open = false;
    };
    uint64_t next = 0;
    while (next < nbatches && !*c->h_err) {
      const double due = arrival_ms ? arrival_ms[next] : 0.0;
      double now = since(clk::now());
      if (due > now) {
        // a long idle gap: close the open launch instead of leaving it waiting
        if (sess.open && due - now > idle_close_ms) {
          close_session();
          if (*c->h_err) break;
          now = since(clk::now());
        }
        if (due - now > 0.2) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((due - now - 0.1) * 1e3)));
        while (since(clk::now()) < due) {
        }
        now = since(clk::now());
      }
      // admit every batch that has arrived
      while (next < nbatches && (!arrival_ms || arrival_ms[next] <= now)) {
        const uint64_t n = batch_sizes[next];
        if (sess.open && (sess.hi - sess.lo + n > max_txs || sess.writes + bwrites[next] > pool_nodes)) {
          close_session();  // the batch waits for the running launch to drain
          st.deferred++;
          if (*c->h_err) break;
        }
        if (!sess.open)
          open_session(bfirst[next]);
        else
          st.admitted_live++;  // joins a launch that is already running
        const uint64_t lo2 = bfirst[next], hi2 = bfirst[next + 1];
        const uint64_t o0 = c->h_tx_off[lo2], o1 = c->h_tx_off[hi2];
        if (o1 > o0) HIP_TRY(hipMemcpyAsync(c->ops.p + o0, host_ops.p + o0, (o1 - o0) * 8, hipMemcpyHostToDevice, s2));
        HIP_TRY(hipMemcpyAsync(c->tx_off.p + lo2, host_off.p + lo2, (hi2 - lo2 + 1) * 8, hipMemcpyHostToDevice, s2));
        c->host->PrepAdmit(o0, o1, s2);
        sess.hi = hi2;
        sess.writes += bwrites[next];
        write_admit((uint32_t)(sess.hi - sess.lo));
        admit_ms[next] = since(clk::now());
        ++next;
        now = since(clk::now());
      }
    }
    if (sess.open) close_session();
    st.tot_ms = since(clk::now());
    // commit latency: stamp mapped to host time, never before the admission
    std::vector<double> lat(c->ntx, 0.0);
    {
      std::vector<uint64_t> tick(c->ntx);
      if (c->ntx) HIP_TRY(hipMemcpy(tick.data(), c->tcommit.p, c->ntx * 8, hipMemcpyDeviceToHost));
      for (uint64_t b = 0; b < nbatches; ++b) {
        const double arr = arrival_ms ? arrival_ms[b] : 0.0;
        for (uint64_t t = bfirst[b]; t < bfirst[b + 1]; ++t) {
          const double commit = h0 + ((double)(int64_t)(tick[t] - d0)) / ticks_per_ms;
          lat[t] = std::max(commit, admit_ms[b]) - arr;
        }
      }
    }
    const Metrics m = collect_metrics(c);
    c->record_commit = false;
    c->pool_cap = saved_pool_cap;
    if (m.error_flags) clear_cc_state(*c);
    st.committed = m.committed;
    st.aborts = m.abort;
    st.error_flags = m.error_flags;
    st.launches = launches;
    st.busy_ms = busy;
    st.mean_launch_txs = launches ? launch_txs / (double)launches : 0;
    st.txps = st.tot_ms > 0 ? (double)m.committed / (st.tot_ms / 1e3) : 0;
    double sum = 0, mx = 0;
    for (double x : lat) {
      sum += x;
      mx = std::max(mx, x);
    }
    st.mean_ms = lat.empty() ? 0 : sum / (double)lat.size();
    st.max_ms = mx;
    st.p50_ms = percentile(lat, 0.50);
    st.p99_ms = percentile(lat, 0.99);
    if (tx_latency_ms && !lat.empty()) std::memcpy(tx_latency_ms, lat.data(), lat.size() * sizeof(double));
    c->last_committed = m.committed;
  } catch (const HipFailure& f) {
    if (launch_open) {
      // close the running launch so that it drains instead of waiting for
      // the watchdog: every admission word gets the closed bit at the
      // admitted count it already holds
      uint64_t words[16 * GATE_SHARDS];
      if (hipMemcpy(words, c->ctrs.p + CTR_ADMIT0, sizeof(words), hipMemcpyDeviceToHost) == hipSuccess) {
        for (int g = 0; g < GATE_SHARDS; ++g) words[16 * g] |= (uint64_t)ADMIT_CLOSED;  // the u32's top bit
        (void)hipMemcpy(c->ctrs.p + CTR_ADMIT0, words, sizeof(words), hipMemcpyHostToDevice);
      }
      (void)hipStreamSynchronize(c->stream);
      try {
        clear_cc_state(*c);
      } catch (const HipFailure&) {
      }
    }
    cleanup();
    return hip_fail(f);
  }
  cleanup();
  if (out) *out = st;
  if (st.error_flags) return set_error(GCCTB_EKERNEL, error_text(st.error_flags));
  if (st.committed != c->ntx) return set_error(GCCTB_EKERNEL, "not every transaction committed");
  return GCCTB_OK;
}

int gcctb_set_batches(gcctb_ctx* c, const uint64_t* sizes, uint64_t n) {
  if (!c || (n && !sizes)) return set_error(GCCTB_EINVAL, "null argument");
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; ++i) total += sizes[i];
  if (n && total != c->ntx) return set_error(GCCTB_EINVAL, "batch sizes must add up to the loaded transactions");
  c->custom_batches.assign(sizes, sizes + n);
  c->inputs_gen++;
  return GCCTB_OK;
}

int gcctb_fetch(gcctb_ctx* c, uint64_t* table, uint64_t* reads, uint64_t* order, uint32_t* aborts) {
  if (!c) return set_error(GCCTB_EINVAL, "null context");
  try {
    hipStream_t s = c->stream;
    HIP_TRY(hipStreamSynchronize(s));
    if (table) HIP_TRY(hipMemcpy(table, c->cells.p, c->cfg.rows * c->cfg.cols * 8, hipMemcpyDeviceToHost));
    if (reads && c->nops) HIP_TRY(hipMemcpy(reads, c->reads.p, c->nops * 8, hipMemcpyDeviceToHost));
    if (aborts && c->ntx) HIP_TRY(hipMemcpy(aborts, c->tx_aborts.p, c->ntx * 4, hipMemcpyDeviceToHost));
    if (order && c->ntx) {
      std::vector<uint32_t> b(c->ntx);
      std::vector<uint64_t> k0(c->ntx), k1(c->ntx);
      HIP_TRY(hipMemcpy(b.data(), c->obatch.p, c->ntx * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(k0.data(), c->okey0.p, c->ntx * 8, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(k1.data(), c->okey1.p, c->ntx * 8, hipMemcpyDeviceToHost));
      for (uint64_t t = 0; t < c->ntx; ++t) {
        order[3 * t + 0] = b[t];
        order[3 * t + 1] = k0[t];
        order[3 * t + 2] = k1[t];
      }
    }
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_fetch_ranks(gcctb_ctx* c, uint32_t* rank, uint32_t* counts, uint64_t* nranks) {
  if (!c) return set_error(GCCTB_EINVAL, "null context");
  if (c->scheme != SCHEME_GPUTX) return set_error(GCCTB_ESTATE, "ranks exist only for gputx");
  try {
    if (rank && c->ntx) HIP_TRY(hipMemcpy(rank, c->rank.p, c->ntx * 4, hipMemcpyDeviceToHost));
    if (counts) {
      gcctb::GPUTxHost::rank_counts(*c);
      std::memcpy(counts, c->h_rank_counts.data(), c->h_rank_counts.size() * 4);
    }
    if (nranks) *nranks = c->nranks;
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_fetch_cc(gcctb_ctx* c, uint64_t* words) {
  if (!c || !words) return set_error(GCCTB_EINVAL, "null argument");
  try {
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the words sit at their rows' slots: gather them back into row order
    std::vector<uint64_t> slots(c->cc.n);
    HIP_TRY(hipMemcpy(slots.data(), c->cc.p, c->cc.n * 8, hipMemcpyDeviceToHost));
    const uint32_t shift = slot_shift(c->cfg.rows);
    const uint32_t wshift = c->latch ? 1 : 0;
    for (uint64_t r = 0; r < c->cfg.rows; ++r) words[r] = slots[(uint64_t)row_slot((uint32_t)r, shift) << wshift];
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_generate(const gcctb_workload* w, uint64_t* ops, uint64_t* tx_off, uint64_t* nops) {
  if (!w || !tx_off || (!ops && w->ntx)) return set_error(GCCTB_EINVAL, "null argument");
  return generate(*w, ops, tx_off, nops);
}

double gcctb_zeta(uint64_t n, double theta) { return zeta(n, theta, 0); }

int gcctb_read_txs_file(const char* path, uint64_t ntx, uint32_t cols, uint64_t seed, uint64_t* ops,
                        uint64_t* tx_off, uint64_t* nops) {
  if (!path || !ops || !tx_off) return set_error(GCCTB_EINVAL, "null argument");
  return read_txs_file(path, ntx, cols, seed, ops, tx_off, nops);
}

int gcctb_write_txs_file(const char* path, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx) {
  if (!path || !tx_off || (!ops && ntx)) return set_error(GCCTB_EINVAL, "null argument");
  return write_txs_file(path, ops, tx_off, ntx);
}

int gcctb_device_info(char* name, uint64_t name_len, double* wall_clock_khz, int32_t* cus) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(GCCTB_EDEVICE, "no HIP device");
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return set_error(GCCTB_EDEVICE, "hipGetDeviceProperties failed");
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  int rate = 0;
  (void)hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, 0);
  if (wall_clock_khz) *wall_clock_khz = rate;
  if (cus) *cus = p.multiProcessorCount;
  return GCCTB_OK;
}

int gcctb_check(gcctb_ctx* c, const uint64_t* init_table, uint64_t* bad, uint64_t* uncommitted) {
  if (!c || !init_table || !bad) return set_error(GCCTB_EINVAL, "null argument");
  if (c->tpcc) return set_error(GCCTB_EINVAL, "the check covers the YCSB ops; TPC-C runs are checked by replay");
  const uint64_t ncells = c->cfg.rows * c->cfg.cols;
  std::vector<uint64_t> fin(ncells), reads(c->nops + 1), order(3 * c->ntx + 1), ops(c->nops + 1);
  std::vector<uint32_t> rows(c->nops + 1);
  try {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(fin.data(), c->cells.p, ncells * 8, hipMemcpyDeviceToHost));
    if (c->nops) {
      HIP_TRY(hipMemcpy(reads.data(), c->reads.p, c->nops * 8, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(ops.data(), c->ops.p, c->nops * 8, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(rows.data(), c->rows_out.p, c->nops * 4, hipMemcpyDeviceToHost));
    }
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  const int rc = gcctb_fetch(c, nullptr, nullptr, order.data(), nullptr);
  if (rc != GCCTB_OK) return rc;
  check_history(c->cfg.rows, c->cfg.cols, c->cfg.seed, c->cfg.row_ops != 0, init_table, fin.data(), ops.data(),
                rows.data(), c->h_tx_off.data(), c->ntx, order.data(), reads.data(), bad);
  if (uncommitted) *uncommitted = c->ntx - std::min(c->ntx, c->last_committed);
  return GCCTB_OK;
}

int gcctb_fetch_events(gcctb_ctx* c, gcctb_event* events, uint64_t cap, uint64_t* n) {
  if (!c || !n || (cap && !events)) return set_error(GCCTB_EINVAL, "null argument");
  if (!c->cfg.debug_events || !c->events_ready) return set_error(GCCTB_ESTATE, "no logged run (create with debug_events = 1)");
  try {
    std::vector<uint64_t> evo(c->nops), evc(c->ntx), ops(c->nops);
    std::vector<uint32_t> rows(c->nops);
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->nops) {
      HIP_TRY(hipMemcpy(evo.data(), c->ev_op.p, c->nops * 8, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(ops.data(), c->ops.p, c->nops * 8, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(rows.data(), c->rows_out.p, c->nops * 4, hipMemcpyDeviceToHost));
    }
    if (c->ntx) HIP_TRY(hipMemcpy(evc.data(), c->ev_commit.p, c->ntx * 8, hipMemcpyDeviceToHost));
    uint64_t k = 0;
    for (uint64_t t = 0; t < c->ntx; ++t) {
      for (uint64_t i = c->h_tx_off[t]; i < c->h_tx_off[t + 1]; ++i) {
        if (!evo[i]) continue;
        if (k < cap) events[k] = gcctb_event{evo[i], rows[i], (int32_t)t, op_is_write(ops[i]) ? 1 : 0};
        ++k;
      }
      if (evc[t]) {
        if (k < cap) events[k] = gcctb_event{evc[t], 0, (int32_t)t, 2};
        ++k;
      }
    }
    *n = k;
    if (k > cap) return set_error(GCCTB_EINVAL, "the event buffer is too small");
  } catch (const HipFailure& f) {
    return hip_fail(f);
  }
  return GCCTB_OK;
}

int gcctb_conflict_check(const gcctb_event* events, uint64_t n, uint64_t ntx, uint64_t* uncommitted, uint64_t* cycle,
                         uint64_t cycle_cap, uint64_t* cycle_len, char* report, uint64_t report_len) {
  if ((n && !events) || (cycle_cap && !cycle)) return set_error(GCCTB_EINVAL, "null argument");
  if (ntx >= (1ull << 31)) return set_error(GCCTB_EINVAL, "too many transactions");
  return conflict_check(events, n, ntx, uncommitted, cycle, cycle_cap, cycle_len, report, report_len);
}

int gcctb_sort_pairs(void* keys, uint32_t* vals, uint64_t n, int32_t key_bytes, int32_t key_bits, double* ms) {
  if ((key_bytes != 4 && key_bytes != 8) || key_bits < 0 || key_bits > key_bytes * 8 || (n && (!keys || !vals)))
    return set_error(GCCTB_EINVAL, "bad sort arguments");
  if (n > 0xffffffffull) return set_error(GCCTB_EINVAL, "the sort takes at most 2^32 - 1 pairs");
  if (ms) *ms = 0;
  if (n == 0) return GCCTB_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(GCCTB_EDEVICE, "no HIP device");
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = GCCTB_OK;
  try {
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    DevBuf<uint32_t> v0, v1, tmp;
    v0.alloc(n);
    v1.alloc(n);
    tmp.alloc(launch::radix_sort_tmp_words(n));
    HIP_TRY(hipMemcpyAsync(v0.p, vals, n * 4, hipMemcpyHostToDevice, s));
    uint32_t* vout = nullptr;
    if (key_bytes == 4) {
      DevBuf<uint32_t> k0, k1;
      k0.alloc(n);
      k1.alloc(n);
      uint32_t* kout = nullptr;
      HIP_TRY(hipMemcpyAsync(k0.p, keys, n * 4, hipMemcpyHostToDevice, s));
      HIP_TRY(hipEventRecord(ev[0], s));
      HIP_TRY(launch::radix_sort_u32(k0.p, v0.p, k1.p, v1.p, n, key_bits, tmp.p, &kout, &vout, s));
      HIP_TRY(hipEventRecord(ev[1], s));
      HIP_TRY(hipMemcpyAsync(keys, kout, n * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(vals, vout, n * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    } else {
      DevBuf<uint64_t> k0, k1;
      k0.alloc(n);
      k1.alloc(n);
      uint64_t* kout = nullptr;
      HIP_TRY(hipMemcpyAsync(k0.p, keys, n * 8, hipMemcpyHostToDevice, s));
      HIP_TRY(hipEventRecord(ev[0], s));
      HIP_TRY(launch::radix_sort_u64(k0.p, v0.p, k1.p, v1.p, n, key_bits, tmp.p, &kout, &vout, s));
      HIP_TRY(hipEventRecord(ev[1], s));
      HIP_TRY(hipMemcpyAsync(keys, kout, n * 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(vals, vout, n * 4, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    }
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
    if (ms) *ms = t;
  } catch (const HipFailure& f) {
    rc = hip_fail(f);
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

int gcctb_microbench(int32_t which, uint64_t n_words, uint64_t threads, uint64_t iters, double* rate, double* ms) {
  if (which < 0 || which > 13 || n_words < 16 || threads == 0 || iters == 0)
    return set_error(GCCTB_EINVAL, "bad microbench arguments");
  // the hot-word modes address words up to 15 * 16 + 15
  if (which >= 10 && n_words < 256) return set_error(GCCTB_EINVAL, "hot-word microbench modes need n_words >= 256");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(GCCTB_EDEVICE, "no HIP device");
  const hipError_t e = launch::microbench(which, n_words, threads, iters, rate, ms);
  if (e != hipSuccess) return set_error(GCCTB_EDEVICE, std::string("microbench: ") + hipGetErrorString(e));
  return GCCTB_OK;
}

}  // extern "C"
