// gcctb_run_plain_static: the plain per-batch form over caller-owned device
// buffers, with every other device word it needs in one file-scope device
// array, so that a call allocates nothing (no hipMalloc / hipFree, no host
// heap).  For callers that cannot allocate: code under stream capture,
// serving loops, harnesses that forbid hipMalloc.
//
// The array (`scratch`) holds, per GPU, the CC words and MVCC chain heads of
// up to NA_MAX_ROWS rows, the MVCC version pool, the row of every op, the
// per-transaction batch and abort words, the allocator counters and the
// metrics slots.  Each GPU the library runs on keeps its own copy until the
// process ends; its device address comes from hipGetSymbolAddress, taken once
// per device with that device current.  Calls that overlap on one GPU share
// the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "../../../include/gcctb.h"
#include "../gcctb_common.h"
#include "../host/host_internal.h"
#include "dev_util.h"
#include "launch.h"

namespace gcctb {
namespace {

// Limits of one call (they size the array).
constexpr uint64_t NA_MAX_ROWS = 1ull << 20;
constexpr uint64_t NA_MAX_OPS = 1ull << 20;
constexpr uint64_t NA_MAX_TXS = 1ull << 17;
constexpr uint64_t NA_POOL_NODES = 1ull << 18;  // MVCC versions per batch

// Layout of the array, in u64 words (every region starts on a 256-byte boundary).
constexpr uint64_t al32(uint64_t w) { return (w + 31) & ~31ull; }
constexpr uint64_t OFF_CC = 0;                                          // rows x u64
constexpr uint64_t OFF_VPREV = OFF_CC + al32(NA_MAX_ROWS);              // rows x u32
constexpr uint64_t OFF_POOL = OFF_VPREV + al32(NA_MAX_ROWS / 2);        // 3 u64 per node
constexpr uint64_t OFF_ROWS = OFF_POOL + al32(3 * NA_POOL_NODES);       // ops x u32
constexpr uint64_t OFF_BATCH = OFF_ROWS + al32(NA_MAX_OPS / 2);         // txs x u32
constexpr uint64_t OFF_ABORTS = OFF_BATCH + al32(NA_MAX_TXS / 2);       // txs x u32
constexpr uint64_t OFF_CTRS = OFF_ABORTS + al32(NA_MAX_TXS / 2);        // CTR_WORDS
constexpr uint64_t OFF_METRICS = OFF_CTRS + al32(CTR_WORDS);            // METRIC_SLOTS x Metrics
constexpr uint64_t SCRATCH_WORDS = OFF_METRICS + al32((uint64_t)METRIC_SLOTS * METRICS_WORDS);
static_assert(SCRATCH_WORDS * 8 == GCCTB_STATIC_SCRATCH_BYTES, "gcctb.h states the array's size");
static_assert(slot_words(NA_MAX_ROWS) == NA_MAX_ROWS, "the rows' slots fit the CC and chain-head regions");
static_assert(sizeof(Metrics) == METRICS_WORDS * 8, "metrics slot size");

}  // namespace
}  // namespace gcctb

// The array.  In host code its name is a host-side stand-in, never a device
// address: kernels and copies get the per-device pointer below.
static __device__ __align__(256) uint64_t scratch[gcctb::SCRATCH_WORDS];

namespace gcctb {
namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mu;
uint64_t* g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of this GPU's array, looked up once per device with that
// device current (the caller's current device).
hipError_t scratch_ptr(uint64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_na_ptr[dev];
  return hipSuccess;
}

// Row of every op when keys are row ids (Find(k) == k for the reference's
// identity id column, benchmark_cpu/ycsb.cu:59): key < rows, else a miss.
__global__ void k_identity_rows(const uint64_t* ops, uint64_t n, uint64_t rows, uint32_t* rows_out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = op_key(ops[i]);
    rows_out[i] = (uint64_t)k < rows ? k : IDX_MISS;
  }
}

struct NaFail {
  hipError_t err;
  const char* where;
};
#define NA_TRY(expr)                           \
  do {                                         \
    const hipError_t _e = (expr);              \
    if (_e != hipSuccess) throw NaFail{_e, #expr}; \
  } while (0)

int scheme_index(const char* s) {
  static const char* names[] = {"tpl_nw", "tpl_wd", "to", "mvcc", "silo", "tictoc", "gacco", "gputx"};
  if (!s) return -1;
  for (int i = 0; i < 8; ++i)
    if (std::strcmp(s, names[i]) == 0) return i;
  return -1;
}

hipError_t launch_exec(int scheme, int G, const ExecArgs& a, uint64_t n, int threads, hipStream_t s) {
  switch (scheme) {
    case SCHEME_TPL_NW: return launch::exec_tpl(false, G, a, n, threads, false, s);
    case SCHEME_TPL_WD: return launch::exec_tpl(true, G, a, n, threads, false, s);
    case SCHEME_TO: return launch::exec_to(G, a, n, threads, false, s);
    case SCHEME_MVCC: return launch::exec_mvcc(G, a, n, threads, false, s);
    case SCHEME_SILO: return launch::exec_silo(G, a, n, threads, false, s);
    case SCHEME_TICTOC: return launch::exec_tictoc(G, a, n, threads, false, s);
    default: return hipErrorInvalidValue;
  }
}

// Read the metrics slots (fixed host buffer on the stack) and sum them.
Metrics sum_metrics(const uint64_t* dev_metrics, hipStream_t s) {
  uint64_t slots[METRIC_SLOTS * METRICS_WORDS];
  NA_TRY(hipMemcpyAsync(slots, dev_metrics, sizeof(slots), hipMemcpyDeviceToHost, s));
  NA_TRY(hipStreamSynchronize(s));
  Metrics m{};
  for (int i = 0; i < METRIC_SLOTS; ++i) {
    Metrics x;
    std::memcpy(&x, slots + (uint64_t)i * METRICS_WORDS, sizeof(Metrics));
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

}  // namespace
}  // namespace gcctb

using namespace gcctb;

extern "C" int gcctb_run_plain_static(const gcctb_static_run* r, gcctb_stats* out) {
  if (!r) return set_error(GCCTB_EINVAL, "null argument");
  const int sch = scheme_index(r->cc_scheme);
  if (sch < 0) return set_error(GCCTB_EINVAL, "unknown cc_scheme");
  if (sch == SCHEME_GACCO || sch == SCHEME_GPUTX)
    return set_error(GCCTB_EINVAL, "gcctb_run_plain_static runs tpl_nw, tpl_wd, to, mvcc, silo and tictoc; "
                                   "gacco and gputx need their sort buffers (gcctb_run_plain)");
  if (r->block_size < 1 || r->block_size > 32 || r->warp_density < 0 || r->warp_density > 5)
    return set_error(GCCTB_EINVAL, "warp_density must be 0..5 and block_size 1..32");
  if (r->batch_size < 1) return set_error(GCCTB_EINVAL, "batch_size must be >= 1");
  if (r->cols < 1 || r->cols > 255) return set_error(GCCTB_EINVAL, "cols out of range");
  if (r->rows < 1 || r->rows > NA_MAX_ROWS)
    return set_error(GCCTB_EINVAL, "gcctb_run_plain_static takes at most 1048576 rows (its static array's limit)");
  if (r->ntx > NA_MAX_TXS) return set_error(GCCTB_EINVAL, "gcctb_run_plain_static takes at most 131072 transactions");
  if (r->nops > NA_MAX_OPS) return set_error(GCCTB_EINVAL, "gcctb_run_plain_static takes at most 1048576 ops");
  if (r->ntx && (!r->cells || !r->ops || !r->tx_off || !r->reads || !r->order))
    return set_error(GCCTB_EINVAL, "null device buffer");
  gcctb_stats st{};
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = static_cast<hipStream_t>(r->stream);
  Metrics m{};
  uint64_t launches = 0;
  try {
    uint64_t* base = nullptr;
    NA_TRY(scratch_ptr(&base));
    uint64_t* cc = base + OFF_CC;
    uint32_t* vprev = reinterpret_cast<uint32_t*>(base + OFF_VPREV);
    uint64_t* pool = base + OFF_POOL;
    uint32_t* rows = reinterpret_cast<uint32_t*>(base + OFF_ROWS);
    uint32_t* obatch = reinterpret_cast<uint32_t*>(base + OFF_BATCH);
    uint32_t* aborts = reinterpret_cast<uint32_t*>(base + OFF_ABORTS);
    uint64_t* ctrs = base + OFF_CTRS;
    uint64_t* metrics = base + OFF_METRICS;
    // The array keeps the last call's contents: clear what must start at
    // zero (fill kernels, which stream capture records as kernel nodes).
    // per-row state sits at each row's slot (slot_words(rows) <= NA_MAX_ROWS)
    const uint64_t words = slot_words(r->rows);
    NA_TRY(launch::fill_u64(cc, words, 0, s));
    if (sch == SCHEME_MVCC) {
      NA_TRY(launch::fill_u32(vprev, words, 0, s));
      NA_TRY(launch::fill_u64(pool, 3 * NA_POOL_NODES, 0, s));
    }
    NA_TRY(launch::fill_u64(ctrs, CTR_WORDS, 0, s));
    NA_TRY(launch::fill_u64(metrics, (uint64_t)METRIC_SLOTS * METRICS_WORDS, 0, s));
    if (r->nops) {
      const unsigned blocks = (unsigned)std::min<uint64_t>((r->nops + 255) / 256, 4096);
      hipLaunchKernelGGL(k_identity_rows, dim3(blocks), dim3(256), 0, s, r->ops, r->nops, r->rows, rows);
      NA_TRY(hipGetLastError());
    }
    ExecArgs a{};
    a.cells = r->cells;
    a.rows = r->rows;
    a.cols = r->cols;
    a.scheme = (uint32_t)sch;
    a.index = IndexView{};
    a.ops = r->ops;
    a.tx_off = r->tx_off;
    a.cc = cc;
    a.vprev = vprev;
    a.slot_shift = slot_shift(r->rows);
    a.pool = pool;
    a.pool_cap = NA_POOL_NODES;
    a.ctrs = ctrs;
    a.rows_out = rows;
    a.rows_in = rows;
    a.reads = r->reads;
    a.obatch = obatch;
    a.okey0 = r->order;
    a.okey1 = r->order + r->ntx;
    a.tx_aborts = aborts;
    a.metrics = reinterpret_cast<Metrics*>(metrics);
    a.seed = r->seed;
    a.backoff_ticks = r->backoff_ticks;
    {
      // 20 s of the 100 MHz wall clock unless GCCTB_WATCHDOG_S says otherwise
      const char* wd = std::getenv("GCCTB_WATCHDOG_S");
      const double sec = wd ? std::atof(wd) : 20.0;
      a.watchdog_ticks = sec > 0 ? (uint64_t)(sec * 1e8) : 0;
    }
    a.urgent_attempts = 0;
    a.backoff_exp = 10;
    a.poll_ticks = 25;
    const int G = 1 << r->warp_density;
    const int threads = r->block_size * 32;
    // Per batch: reset the counters, one launch, restore the touched words
    // (benchmark_cpu/ycsb.cu:262-285), stop on the first batch with an error.
    // The host learns the row offsets of a batch from nothing but the op
    // count, so the restore clears the rows of every op (<= rows / 8 ops:
    // those rows; more: the whole arrays).
    uint64_t b = 0;
    for (uint64_t st0 = 0; st0 < r->ntx; st0 += r->batch_size, ++b) {
      const uint64_t n = std::min<uint64_t>(r->batch_size, r->ntx - st0);
      a.batch_id = (uint32_t)b;
      a.slot_lo = st0;
      a.slot_hi = st0 + n;
      NA_TRY(launch::batch_begin(ctrs, s));
      NA_TRY(launch_exec(sch, G, a, n, threads, s));
      if (sch != SCHEME_TPL_NW && sch != SCHEME_TPL_WD)
        NA_TRY(launch::reset_rows(rows, r->nops, r->rows, cc, sch == SCHEME_MVCC ? vprev : nullptr, s));
      uint64_t flags = 0;
      NA_TRY(hipMemcpyAsync(&flags, &reinterpret_cast<Metrics*>(metrics)->error_flags, 8, hipMemcpyDeviceToHost, s));
      NA_TRY(hipStreamSynchronize(s));
      ++launches;
      if (flags) break;
    }
    m = sum_metrics(metrics, s);
    if (m.error_flags) {
      // a stopped launch can leave locks or marks behind: clear them all
      NA_TRY(launch::fill_u64(cc, words, 0, s));
      if (sch == SCHEME_MVCC) NA_TRY(launch::fill_u32(vprev, words, 0, s));
      NA_TRY(hipStreamSynchronize(s));
    }
  } catch (const NaFail& f) {
    return set_error(GCCTB_EDEVICE, std::string("HIP error: ") + hipGetErrorString(f.err) + " in " + f.where);
  }
  st.txs = r->ntx;
  st.committed = m.committed;
  st.aborts = m.abort;
  st.ts_ticks = m.ts_duration;
  st.wait_ticks = m.wait_duration;
  st.abort_ticks = m.abort_duration;
  st.manager_ticks = m.manager_duration;
  st.tot_ticks = m.tot_duration;
  st.index_ticks = m.index_duration;
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess)
      st.tick_khz = khz;
  }
  st.tot_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  st.batches = launches;
  st.index_miss = m.index_miss;
  st.error_flags = m.error_flags;
  st.total_ops = r->nops;
  if (out) *out = st;
  if (m.error_flags) return set_error(GCCTB_EKERNEL, "kernel error flags " + std::to_string(m.error_flags));
  if (m.committed != r->ntx) return set_error(GCCTB_EKERNEL, "not every transaction committed");
  return GCCTB_OK;
}
