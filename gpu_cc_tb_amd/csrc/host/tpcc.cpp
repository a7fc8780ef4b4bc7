// TPC-C inputs: the initial rows of the tables that carry concurrency control
// (generation/gen_tpcc_table.cu), the payment and new_order transactions
// (generation/gen_tpcc.cu) as ops of the TPC-C row space (gcctb_common.h),
// and new_order's output totals (benchmark/tpcc.cu:340-376).
//
// The reference draws with rand() from 256 threads at once, so its datasets
// differ from run to run; here every draw is a hash of (seed, transaction,
// draw index), so a dataset is a function of its parameters.  The
// distributions are the reference's: uniform_int(x, y) in [x, y),
// uniform_float(x, y), and NURand(A, x, y) = ((U[0, A) | (U[x, y) + 42)) %
// (y - x + 1)) + x.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../gcctb_common.h"
#include "host_internal.h"

namespace {

using namespace gcctb;

constexpr uint64_t TAG_TPCC_TX = 0x5450434358000000ull;  // "TPCCX"
constexpr uint64_t TAG_TPCC_TB = 0x5450434354000000ull;  // "TPCCT"

// the draws of one transaction (or one table row)
struct Draws {
  uint64_t seed, id, n = 0;
  uint64_t next() { return hash3(seed, id, n++); }
  // [x, y)
  uint32_t uniform_int(uint32_t x, uint32_t y) { return x + (uint32_t)(next() % (uint64_t)(y - x)); }
  float uniform_float(double x, double y) { return (float)(x + (double)(next() >> 11) * 0x1.0p-53 * (y - x)); }
  uint32_t nurand(uint32_t A, uint32_t x, uint32_t y) {
    const uint32_t r0 = uniform_int(0, A);
    const uint32_t rc = uniform_int(x, y) + 42u;
    return ((r0 | rc) % (y - x + 1)) + x;
  }
};

struct StockLine {
  uint32_t i_id, supply_w, qty;
};

}  // namespace

extern "C" uint64_t gcctb_tpcc_rows(uint32_t warehouses) { return tpcc_rows(warehouses); }

extern "C" int gcctb_tpcc_seed_tables(uint32_t warehouses, uint64_t seed, uint64_t* cells, float* item_price) {
  if (warehouses == 0 || warehouses > 8192) return set_error(GCCTB_EINVAL, "warehouses must be 1..8192");
  if (!cells) return set_error(GCCTB_EINVAL, "null argument");
  const uint64_t W = warehouses;
  std::memset(cells, 0, tpcc_rows(W) * TPCC_COLS * sizeof(uint64_t));
  auto row = [&](uint64_t r) { return cells + r * TPCC_COLS; };
  for (uint64_t w = 0; w < W; ++w) {
    Draws d{seed ^ TAG_TPCC_TB, w};
    uint64_t* c = row(w);
    c[0] = tpcc_pack((uint32_t)w, tpcc_bits(d.uniform_float(0.0, 0.2)));
    c[1] = tpcc_pack(tpcc_bits(300000.0f), 0);
    for (uint64_t di = 0; di < TPCC_DISTRICTS; ++di) {
      const uint64_t r = tpcc_district_row(W, w, di);
      Draws dd{seed ^ TAG_TPCC_TB, r};
      uint64_t* x = row(r);
      x[0] = tpcc_pack((uint32_t)di, (uint32_t)w);
      x[1] = tpcc_pack(tpcc_bits(dd.uniform_float(0.0, 0.2)), tpcc_bits(30000.0f));
      x[2] = tpcc_pack(3001u, 0);
      for (uint64_t ci = 0; ci < TPCC_CUSTOMERS; ++ci) {
        const uint64_t rc = tpcc_customer_row(W, w, di, ci);
        Draws dc{seed ^ TAG_TPCC_TB, rc};
        uint64_t* y = row(rc);
        y[0] = tpcc_pack((uint32_t)ci, (uint32_t)di);
        y[1] = tpcc_pack((uint32_t)w, 0u);
        y[2] = tpcc_pack(tpcc_bits(50000.0f), tpcc_bits(dc.uniform_float(0.0, 0.5)));
        y[3] = tpcc_pack(tpcc_bits(-10.0f), tpcc_bits(10.0f));
        y[4] = tpcc_pack(1u, 0u);
      }
    }
    for (uint64_t i = 0; i < TPCC_ITEMS; ++i) {
      const uint64_t rs = tpcc_stock_row(W, w, i);
      Draws ds{seed ^ TAG_TPCC_TB, rs};
      uint64_t* z = row(rs);
      z[0] = tpcc_pack((uint32_t)i, (uint32_t)w);
      z[1] = tpcc_pack(ds.uniform_int(10, 100), 0u);
    }
  }
  if (item_price)
    for (uint64_t i = 0; i < TPCC_ITEMS; ++i) {
      Draws di{seed ^ TAG_TPCC_TB, ~i};
      item_price[i] = di.uniform_float(1.0, 100.0);
    }
  return GCCTB_OK;
}

extern "C" int gcctb_tpcc_generate(const gcctb_tpcc_workload* wl, uint64_t* ops, uint32_t* params, uint64_t* tx_off,
                                   uint64_t* tx_customer, uint64_t* nops) {
  if (!wl || !tx_off || (wl->ntx && (!ops || !params))) return set_error(GCCTB_EINVAL, "null argument");
  if (wl->warehouses == 0 || wl->warehouses > 8192) return set_error(GCCTB_EINVAL, "warehouses must be 1..8192");
  if (!(wl->remote_ratio >= 0.0 && wl->remote_ratio <= 1.0)) return set_error(GCCTB_EINVAL, "remote_ratio in [0, 1]");
  const uint32_t W = wl->warehouses;
  uint64_t o = 0;
  tx_off[0] = 0;
  for (uint64_t t = 0; t < wl->ntx; ++t) {
    Draws d{wl->seed ^ TAG_TPCC_TX, t};
    if (!wl->new_order) {
      // payment (gen_tpcc.cu:124-144)
      const uint32_t c_w = d.uniform_int(0, W), c_d = d.uniform_int(0, TPCC_DISTRICTS);
      const uint32_t c_id = d.nurand(1023, 0, TPCC_CUSTOMERS - 1);
      uint32_t w = c_w, dist = c_d;
      if (d.uniform_int(0, 100) > 85) {
        // the customer pays through another warehouse and district
        while (W > 1 && w == c_w) w = d.uniform_int(0, W);
        dist = d.uniform_int(0, TPCC_DISTRICTS - 1);
        if (dist >= c_d) dist += 1;
      }
      const uint32_t amount = tpcc_bits(d.uniform_float(1.0, 5000.0));
      ops[o] = make_op((uint32_t)w, TPCC_PAY_W, true);
      params[o++] = amount;
      ops[o] = make_op((uint32_t)tpcc_district_row(W, w, dist), TPCC_PAY_D, true);
      params[o++] = amount;
      ops[o] = make_op((uint32_t)tpcc_customer_row(W, c_w, c_d, c_id), TPCC_PAY_C, true);
      params[o++] = amount;
      if (tx_customer) tx_customer[t] = tpcc_customer_row(W, c_w, c_d, c_id);
    } else {
      // new_order (gen_tpcc.cu:146-181)
      const uint32_t w = d.uniform_int(0, W), dist = d.uniform_int(0, TPCC_DISTRICTS);
      const uint32_t c_id = d.nurand(1023, 0, TPCC_CUSTOMERS - 1);
      const uint32_t ol_cnt = d.uniform_int(5, TPCC_MAX_LINES + 1);
      StockLine lines[TPCC_MAX_LINES];
      for (uint32_t l = 0; l < ol_cnt; ++l) {
        uint32_t i_id;
        bool seen;
        do {  // distinct items per order
          i_id = d.nurand(8191, 0, TPCC_ITEMS - 1);
          seen = false;
          for (uint32_t k = 0; k < l; ++k) seen |= lines[k].i_id == i_id;
        } while (seen);
        lines[l].i_id = i_id;
        lines[l].qty = d.uniform_int(1, 11);
        lines[l].supply_w = w;
        if (d.uniform_float(0.0, 1.0) <= (float)wl->remote_ratio && W > 1) {
          uint32_t s = d.uniform_int(0, W - 1);
          if (s >= w) s += 1;
          lines[l].supply_w = s;
        }
      }
      std::sort(lines, lines + ol_cnt, [](const StockLine& a, const StockLine& b) {
        return a.supply_w != b.supply_w ? a.supply_w < b.supply_w : a.i_id < b.i_id;
      });
      ops[o] = make_op((uint32_t)tpcc_district_row(W, w, dist), TPCC_NO_D, true);
      params[o++] = 0;
      for (uint32_t l = 0; l < ol_cnt; ++l) {
        ops[o] = make_op((uint32_t)tpcc_stock_row(W, lines[l].supply_w, lines[l].i_id), TPCC_NO_S, true);
        params[o++] = lines[l].qty | ((lines[l].supply_w != w ? 1u : 0u) << 16);
      }
      if (tx_customer) tx_customer[t] = tpcc_customer_row(W, w, dist, c_id);
    }
    tx_off[t + 1] = o;
  }
  if (nops) *nops = o;
  return GCCTB_OK;
}

extern "C" int gcctb_tpcc_new_order_totals(uint32_t warehouses, const uint64_t* init_cells, const float* item_price,
                                           const uint64_t* ops, const uint32_t* params, const uint64_t* tx_off,
                                           uint64_t ntx, const uint64_t* tx_customer, const uint64_t* results,
                                           float* total, uint32_t* o_id) {
  if (!init_cells || !item_price || !tx_off || !tx_customer || !results || !total || (ntx && (!ops || !params)))
    return set_error(GCCTB_EINVAL, "null argument");
  const uint64_t W = warehouses, stock0 = tpcc_stock_row(W, 0, 0);
  for (uint64_t t = 0; t < ntx; ++t) {
    const uint64_t o0 = tx_off[t], o1 = tx_off[t + 1];
    if (o1 <= o0 || op_col(ops[o0]) != TPCC_NO_D) return set_error(GCCTB_EINVAL, "not a new_order transaction");
    // warehouse tax, district tax and customer discount never change
    const uint64_t drow = op_key(ops[o0]);
    const uint64_t w = (drow - W) / TPCC_DISTRICTS;
    const float w_tax = tpcc_f(tpcc_hi(init_cells[w * TPCC_COLS]));
    const float d_tax = tpcc_f(tpcc_lo(init_cells[drow * TPCC_COLS + 1]));
    const float discount = tpcc_f(tpcc_hi(init_cells[tx_customer[t] * TPCC_COLS + 2]));
    float sum = 0.0f;
    for (uint64_t i = o0 + 1; i < o1; ++i) {
      const uint64_t item = (op_key(ops[i]) - stock0) % TPCC_ITEMS;
      sum += (float)(params[i] & 0xffffu) * item_price[item];
    }
    total[t] = sum * (1.0f - discount) * (1.0f + w_tax + d_tax);
    if (o_id) o_id[t] = (uint32_t)results[o0];
  }
  return GCCTB_OK;
}

extern "C" int gcctb_tpcc_check(uint32_t warehouses, const uint64_t* init_cells, const uint64_t* final_cells,
                                const uint64_t* ops, const uint32_t* params, const uint64_t* tx_off, uint64_t ntx,
                                const uint64_t* order, const uint64_t* results, uint64_t* bad) {
  if (!init_cells || !final_cells || !tx_off || !order || !bad || (ntx && (!ops || !params || !results)))
    return set_error(GCCTB_EINVAL, "null argument");
  if (warehouses == 0 || warehouses > 8192) return set_error(GCCTB_EINVAL, "warehouses must be 1..8192");
  const uint64_t rows = tpcc_rows(warehouses), ncells = rows * TPCC_COLS;
  std::vector<uint64_t> seq(ntx);
  for (uint64_t t = 0; t < ntx; ++t) seq[t] = t;
  auto key = [&](uint64_t t, int i) { return order[3 * t + i]; };
  std::sort(seq.begin(), seq.end(), [&](uint64_t x, uint64_t y) {
    for (int i = 0; i < 3; ++i)
      if (key(x, i) != key(y, i)) return key(x, i) < key(y, i);
    return x < y;
  });
  uint64_t nbad = 0;
  for (uint64_t i = 1; i < ntx; ++i)
    nbad += key(seq[i], 0) == key(seq[i - 1], 0) && key(seq[i], 1) == key(seq[i - 1], 1) &&
            key(seq[i], 2) == key(seq[i - 1], 2);
  std::vector<uint64_t> cells(init_cells, init_cells + ncells);
  for (const uint64_t t : seq)
    for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
      const uint64_t row = op_key(ops[i]);
      if (row >= rows) return set_error(GCCTB_EINVAL, "op key beyond the TPC-C rows");
      nbad += tpcc_apply(op_col(ops[i]), params[i], cells.data() + row * TPCC_COLS) != results[i];
    }
  for (uint64_t c = 0; c < ncells; ++c) nbad += cells[c] != final_cells[c];
  *bad = nbad;
  return GCCTB_OK;
}
