// Deterministic YCSB workload generation and the reference's .txs file format.
//
// The reference generator (generation/gen_ycsb.cu) draws from libc rand() in
// 256 threads, so its output is not reproducible from the seed.  This one is
// counter-based: every draw is hash3(seed ^ tag, tx, counter), so a
// transaction's ops depend only on (seed, global transaction id) and any
// number of threads gives the same bits.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../gcctb_common.h"
#include "host_internal.h"

namespace gcctb {

static inline double unit_double(uint64_t h) {
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

// zeta(n, theta) = sum_{i=1..n} (1/i)^theta (gen_ycsb.cu:76-82), summed in
// blocks of 65536 consecutive terms (each block in increasing i, block sums in
// increasing block order) so that threads can share the work and still give a
// fixed rounding.
static constexpr uint64_t ZETA_BLOCK = 65536;

static double zeta_block(uint64_t b, uint64_t n, double theta) {
  const uint64_t lo = b * ZETA_BLOCK + 1;
  const uint64_t hi = std::min<uint64_t>(n, (b + 1) * ZETA_BLOCK);
  double s = 0;
  for (uint64_t i = lo; i <= hi; ++i) s += std::pow(1.0 / (double)i, theta);
  return s;
}

double zeta(uint64_t n, double theta, int threads) {
  static std::mutex mu;
  static std::map<std::pair<uint64_t, uint64_t>, double> cache;
  uint64_t tbits;
  std::memcpy(&tbits, &theta, sizeof(tbits));
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({n, tbits});
    if (it != cache.end()) return it->second;
  }
  const uint64_t nb = (n + ZETA_BLOCK - 1) / ZETA_BLOCK;
  std::vector<double> part(nb, 0.0);
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if ((uint64_t)threads > nb) threads = (int)std::max<uint64_t>(1, nb);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      for (uint64_t b = t; b < nb; b += threads) part[b] = zeta_block(b, n, theta);
    });
  for (auto& th : pool) th.join();
  double s = 0;
  for (uint64_t b = 0; b < nb; ++b) s += part[b];
  std::lock_guard<std::mutex> g(mu);
  cache[{n, tbits}] = s;
  return s;
}

// Zipf rank in [1, n] from a uniform u in [0, 1) (gen_ycsb.cu:84-102, Gray et
// al., "Quickly generating billion-record synthetic databases", SIGMOD 1994).
struct Zipf {
  uint64_t n;
  double theta, zetan, zeta2, alpha, eta, half_pow;
  Zipf(uint64_t n_, double theta_, int threads) : n(n_), theta(theta_) {
    zetan = zeta(n, theta, threads);
    zeta2 = zeta(2, theta, 1);
    alpha = 1.0 / (1.0 - theta);
    eta = (1.0 - std::pow(2.0 / (double)n, 1.0 - theta)) / (1.0 - zeta2 / zetan);
    half_pow = std::pow(0.5, theta);
  }
  uint64_t rank(double u) const {
    const double uz = u * zetan;
    if (uz < 1.0) return 1;
    if (uz < 1.0 + half_pow) return n >= 2 ? 2 : 1;
    uint64_t r = 1 + (uint64_t)((double)n * std::pow(eta * u - eta + 1.0, alpha));
    if (r > n) r = n;
    return r;
  }
};

static uint32_t tx_len(const gcctb_workload& w, uint64_t t) {
  if (w.len_max <= w.len_min) return w.len_min;
  return w.len_min + (uint32_t)(hash3(w.seed ^ TAG_LEN, t, 0) % (uint64_t)(w.len_max - w.len_min + 1));
}

// Ops of one transaction (global id t) into out[0..L); returns L.
static uint32_t gen_tx(const gcctb_workload& w, const Zipf& z, uint64_t t, uint64_t* out) {
  const uint32_t L = tx_len(w, t);
  uint32_t keys[MAX_OPS];
  uint64_t d = 0;
  const uint64_t max_draws = 64ull * L + 64;
  for (uint32_t s = 0; s < L; ++s) {
    uint32_t key = 0;
    bool ok = false;
    while (d < max_draws) {
      const uint64_t row = z.rank(unit_double(hash3(w.seed ^ TAG_KEY, t, d++))) - 1;
      bool dup = false;
      for (uint32_t q = 0; q < s; ++q) dup |= (keys[q] == (uint32_t)row);
      if (!dup) {
        key = (uint32_t)row;
        ok = true;
        break;
      }
    }
    if (!ok) {  // draw budget exhausted: smallest row not yet taken
      for (uint32_t cand = 0;; ++cand) {
        bool dup = false;
        for (uint32_t q = 0; q < s; ++q) dup |= (keys[q] == cand);
        if (!dup) {
          key = cand;
          break;
        }
      }
    }
    keys[s] = key;
    const bool is_write = !(unit_double(hash3(w.seed ^ TAG_RW, t, s)) < w.read_ratio);
    const uint32_t col = (uint32_t)(hash3(w.seed ^ TAG_COL, t, s) % w.cols);
    out[s] = make_op(key, col, is_write);
  }
  // ops sorted by key (gen_ycsb.cu:140-150); keys are distinct
  std::sort(out, out + L, [](uint64_t a, uint64_t b) { return op_key(a) < op_key(b); });
  return L;
}

int generate(const gcctb_workload& w, uint64_t* ops, uint64_t* tx_off, uint64_t* nops) {
  if (w.rows == 0 || w.rows > 0xfffffffeull) return set_error(GCCTB_EINVAL, "rows out of range");
  if (w.cols == 0 || w.cols > 255) return set_error(GCCTB_EINVAL, "cols out of range");
  if (w.len_min < 1 || w.len_max < w.len_min || w.len_max > (uint32_t)MAX_OPS)
    return set_error(GCCTB_EINVAL, "transaction length must satisfy 1 <= len_min <= len_max <= 32");
  if ((uint64_t)w.len_max > w.rows) return set_error(GCCTB_EINVAL, "len_max exceeds rows (keys are distinct)");
  if (!(w.theta >= 0.0 && w.theta < 1.0)) return set_error(GCCTB_EINVAL, "theta must be in [0, 1)");
  if (!(w.read_ratio >= 0.0 && w.read_ratio <= 1.0)) return set_error(GCCTB_EINVAL, "read_ratio must be in [0, 1]");
  int threads = w.threads > 0 ? w.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  const Zipf z(w.rows, w.theta, threads);
  // pass 1: lengths -> offsets (cheap, single thread)
  tx_off[0] = 0;
  for (uint64_t i = 0; i < w.ntx; ++i) tx_off[i + 1] = tx_off[i] + tx_len(w, w.tx_base + i);
  if ((uint64_t)threads > w.ntx) threads = (int)std::max<uint64_t>(1, w.ntx);
  std::vector<std::thread> pool;
  const uint64_t chunk = (w.ntx + threads - 1) / std::max(1, threads);
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      const uint64_t lo = (uint64_t)t * chunk, hi = std::min<uint64_t>(w.ntx, lo + chunk);
      for (uint64_t i = lo; i < hi; ++i) gen_tx(w, z, w.tx_base + i, ops + tx_off[i]);
    });
  for (auto& th : pool) th.join();
  if (nops) *nops = tx_off[w.ntx];
  return GCCTB_OK;
}

// --- reference .txs format (include/ycsb.cuh:21-33) ---
struct RefReq {
  int32_t read;
  uint32_t key;
};
struct RefTx {
  uint64_t request_cnt;
  RefReq requests[16];
};
static_assert(sizeof(RefTx) == 136, "YCSBTx is 136 bytes");

int read_txs_file(const char* path, uint64_t ntx, uint32_t cols, uint64_t seed, uint64_t* ops,
                  uint64_t* tx_off, uint64_t* nops) {
  if (cols == 0 || cols > 255) return set_error(GCCTB_EINVAL, "cols out of range");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(GCCTB_EIO, std::string("cannot open ") + path);
  std::vector<RefTx> buf(ntx);
  const size_t got = std::fread(buf.data(), sizeof(RefTx), ntx, f);
  std::fclose(f);
  if (got != ntx) return set_error(GCCTB_EIO, "short .txs file");
  tx_off[0] = 0;
  for (uint64_t t = 0; t < ntx; ++t) {
    const uint64_t cnt = buf[t].request_cnt;
    if (cnt > 16) return set_error(GCCTB_EINVAL, "request_cnt > 16 in .txs file");
    uint64_t o = tx_off[t];
    for (uint64_t j = 0; j < cnt; ++j) {
      const RefReq& r = buf[t].requests[j];
      const uint32_t col = (uint32_t)(hash3(seed ^ TAG_COL, t, j) % cols);
      ops[o++] = make_op(r.key, col, r.read == 0);
    }
    tx_off[t + 1] = o;
  }
  if (nops) *nops = tx_off[ntx];
  return GCCTB_OK;
}

int write_txs_file(const char* path, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(GCCTB_EIO, std::string("cannot create ") + path);
  for (uint64_t t = 0; t < ntx; ++t) {
    RefTx r;
    std::memset(&r, 0, sizeof(r));
    const uint64_t cnt = tx_off[t + 1] - tx_off[t];
    if (cnt > 16) {
      std::fclose(f);
      return set_error(GCCTB_EINVAL, "the .txs format holds at most 16 ops per transaction");
    }
    r.request_cnt = cnt;
    for (uint64_t j = 0; j < cnt; ++j) {
      const uint64_t op = ops[tx_off[t] + j];
      r.requests[j].read = op_is_write(op) ? 0 : 1;
      r.requests[j].key = op_key(op);
    }
    if (std::fwrite(&r, sizeof(r), 1, f) != 1) {
      std::fclose(f);
      return set_error(GCCTB_EIO, "write failed");
    }
  }
  std::fclose(f);
  return GCCTB_OK;
}

}  // namespace gcctb
