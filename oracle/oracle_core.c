/*
 * oracle_core.c -- CPU restatement of the YCSB path's arithmetic.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define TAG_KEY 0x4b45590000000000ull
#define TAG_RW 0x5257000000000000ull
#define TAG_COL 0x434f4c0000000000ull
#define TAG_WRITE 0x5752495445000000ull
#define TAG_LEN 0x4c454e0000000000ull

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint64_t orc_splitmix64(uint64_t x) { return mix64(x + 0x9E3779B97F4A7C15ull); }
uint64_t orc_hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return orc_splitmix64(orc_splitmix64(seed ^ a) ^ b);
}
uint64_t orc_cell_init(uint64_t seed, uint64_t row, uint32_t col, uint32_t cols) {
  return orc_splitmix64(seed ^ (row * (uint64_t)cols + col));
}
uint64_t orc_write_value(uint64_t seed, uint64_t tx, uint32_t op) {
  return orc_hash3(seed ^ TAG_WRITE, tx, op);
}
/* Whole-row mode (benchmark/ycsb.cu:57,86-91 copies the 104-byte Item): cell
 * c of a written row holds mix64(write value + c); a read returns the row's
 * cells folded from DIGEST_INIT ("ROW") by h = mix64(h ^ cell). */
#define ORC_DIGEST_INIT 0x524f570000000000ull
uint64_t orc_write_cell(uint64_t seed, uint64_t tx, uint32_t op, uint32_t col) {
  return mix64(orc_write_value(seed, tx, op) + col);
}
uint64_t orc_row_digest(const uint64_t* row, uint32_t cols) {
  uint64_t h = ORC_DIGEST_INIT;
  for (uint32_t c = 0; c < cols; ++c) h = mix64(h ^ row[c]);
  return h;
}
void orc_seed_table(uint64_t seed, uint64_t rows, uint32_t cols, uint64_t* cells) {
  for (uint64_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c) cells[r * cols + c] = orc_cell_init(seed, r, c, cols);
}

/* ---- generator (generation/gen_ycsb.cu) ---- */
static double unit_double(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

/* zeta(n, theta), gen_ycsb.cu:76-82, summed in blocks of 65536 terms */
double orc_zeta(uint64_t n, double theta) {
  double total = 0.0;
  for (uint64_t lo = 1; lo <= n; lo += 65536) {
    uint64_t hi = lo + 65535;
    if (hi > n) hi = n;
    double s = 0.0;
    for (uint64_t i = lo; i <= hi; ++i) s += pow(1.0 / (double)i, theta);
    total += s;
  }
  return total;
}

/* zipf(), gen_ycsb.cu:84-102 with n = rows */
uint64_t orc_zipf_rank(uint64_t n, double theta, double zetan, double u) {
  const double zeta2 = orc_zeta(2, theta);
  const double alpha = 1.0 / (1.0 - theta);
  const double eta = (1.0 - pow(2.0 / (double)n, 1.0 - theta)) / (1.0 - zeta2 / zetan);
  const double uz = u * zetan;
  if (uz < 1.0) return 1;
  if (uz < 1.0 + pow(0.5, theta)) return n >= 2 ? 2 : 1;
  uint64_t r = 1 + (uint64_t)((double)n * pow(eta * u - eta + 1.0, alpha));
  if (r > n) r = n;
  return r;
}

static int cmp_key(const void* a, const void* b) {
  const uint32_t x = (uint32_t)(*(const uint64_t*)a), y = (uint32_t)(*(const uint64_t*)b);
  return x < y ? -1 : (x > y ? 1 : 0);
}

uint32_t orc_gen_tx(uint64_t rows, uint32_t cols, uint32_t len_min, uint32_t len_max, double theta,
                    double read_ratio, uint64_t seed, double zetan, uint64_t tx, uint64_t* out) {
  uint32_t L = len_min;
  if (len_max > len_min) L = len_min + (uint32_t)(orc_hash3(seed ^ TAG_LEN, tx, 0) % (uint64_t)(len_max - len_min + 1));
  uint32_t keys[64];
  uint64_t d = 0;
  const uint64_t max_draws = 64ull * L + 64;
  for (uint32_t s = 0; s < L; ++s) {
    int ok = 0;
    uint32_t key = 0;
    while (d < max_draws) {
      const uint64_t row = orc_zipf_rank(rows, theta, zetan, unit_double(orc_hash3(seed ^ TAG_KEY, tx, d++))) - 1;
      int dup = 0;
      for (uint32_t q = 0; q < s; ++q) dup |= keys[q] == (uint32_t)row;
      if (!dup) {
        key = (uint32_t)row;
        ok = 1;
        break;
      }
    }
    if (!ok) {
      for (uint32_t cand = 0;; ++cand) {
        int dup = 0;
        for (uint32_t q = 0; q < s; ++q) dup |= keys[q] == cand;
        if (!dup) {
          key = cand;
          break;
        }
      }
    }
    keys[s] = key;
    const int is_write = !(unit_double(orc_hash3(seed ^ TAG_RW, tx, s)) < read_ratio);
    const uint32_t col = (uint32_t)(orc_hash3(seed ^ TAG_COL, tx, s) % cols);
    out[s] = (uint64_t)key | ((uint64_t)(col & 0xff) << 32) | ((uint64_t)is_write << 40);
  }
  qsort(out, L, sizeof(uint64_t), cmp_key);
  return L;
}

/* ---- SortedArray_GPU::Find, include/index.cuh:74-91 ---- */
uint64_t orc_find(const uint32_t* keys, const uint64_t* vals, uint32_t length, uint32_t key) {
  uint32_t st = 0, en = length;
  while (st < en) {
    const uint32_t mid = (st + en) >> 1;
    const uint32_t k = keys[mid];
    if (k == key) return vals[mid];
    if (k > key)
      en = mid;
    else
      st = mid + 1;
  }
  if (st < length && keys[st] == key) return vals[st];
  return length;
}

/* ---- serial replay ---- */
typedef struct {
  uint64_t b, k0, k1, tx;
} okey_t;
static int cmp_okey(const void* pa, const void* pb) {
  const okey_t* a = (const okey_t*)pa;
  const okey_t* b = (const okey_t*)pb;
  if (a->b != b->b) return a->b < b->b ? -1 : 1;
  if (a->k0 != b->k0) return a->k0 < b->k0 ? -1 : 1;
  if (a->k1 != b->k1) return a->k1 < b->k1 ? -1 : 1;
  if (a->tx != b->tx) return a->tx < b->tx ? -1 : 1;
  return 0;
}

static void apply_tx(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                     const uint64_t* tx_off, uint64_t t, uint64_t* reads) {
  for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
    const uint64_t op = ops[i];
    const uint64_t row = (uint32_t)op;
    const uint32_t col = (uint32_t)(op >> 32) & 0xff;
    if (row >= rows || col >= cols) {
      if (reads) reads[i] = 0;
      continue;
    }
    uint64_t* cell = table + row * cols + col;
    if ((op >> 40) & 1) {
      *cell = orc_write_value(seed, t, (uint32_t)(i - tx_off[t]));
      if (reads) reads[i] = 0;
    } else if (reads) {
      reads[i] = *cell;
    }
  }
}

int64_t orc_replay(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                   const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads) {
  okey_t* ord = (okey_t*)malloc(sizeof(okey_t) * (ntx ? ntx : 1));
  if (!ord) return -1;
  for (uint64_t t = 0; t < ntx; ++t) {
    ord[t].b = order[3 * t];
    ord[t].k0 = order[3 * t + 1];
    ord[t].k1 = order[3 * t + 2];
    ord[t].tx = t;
  }
  qsort(ord, ntx, sizeof(okey_t), cmp_okey);
  int64_t ties = 0;
  for (uint64_t i = 1; i < ntx; ++i)
    if (ord[i].b == ord[i - 1].b && ord[i].k0 == ord[i - 1].k0 && ord[i].k1 == ord[i - 1].k1) ties++;
  for (uint64_t i = 0; i < ntx; ++i) apply_tx(rows, cols, seed, table, ops, tx_off, ord[i].tx, reads);
  free(ord);
  return ties;
}

void orc_replay_txid(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                     const uint64_t* tx_off, uint64_t ntx, uint64_t* reads) {
  for (uint64_t t = 0; t < ntx; ++t) apply_tx(rows, cols, seed, table, ops, tx_off, t, reads);
}

/* Serial replay split over threads by row.  A read returns the latest
 * preceding write to its own row in the order, so the ops of different rows
 * never influence each other: thread k walks every transaction in the same
 * sorted order and applies only the ops whose row lies in its block of rows.
 * The table and the reads equal orc_replay's (tests compare the two). */
typedef struct {
  uint64_t rows, row_lo, row_hi, ntx, seed;
  uint32_t cols;
  int row_ops;
  uint64_t* table;
  const uint64_t* ops;
  const uint64_t* tx_off;
  const okey_t* ord;
  uint64_t* reads;
} replay_part_t;
static void* replay_part(void* arg) {
  const replay_part_t* p = (const replay_part_t*)arg;
  for (uint64_t k = 0; k < p->ntx; ++k) {
    const uint64_t t = p->ord[k].tx;
    for (uint64_t i = p->tx_off[t]; i < p->tx_off[t + 1]; ++i) {
      const uint64_t op = p->ops[i];
      const uint64_t row = (uint32_t)op;
      if (row < p->row_lo || row >= p->row_hi) continue;
      const uint32_t col = (uint32_t)(op >> 32) & 0xff;
      if (row >= p->rows || col >= p->cols) {
        if (p->reads) p->reads[i] = 0;
        continue;
      }
      if (p->row_ops) {
        /* whole-row mode: a write stores every cell, a read returns the digest */
        uint64_t* r = p->table + row * p->cols;
        if ((op >> 40) & 1) {
          for (uint32_t c = 0; c < p->cols; ++c) r[c] = orc_write_cell(p->seed, t, (uint32_t)(i - p->tx_off[t]), c);
          if (p->reads) p->reads[i] = 0;
        } else if (p->reads) {
          p->reads[i] = orc_row_digest(r, p->cols);
        }
        continue;
      }
      uint64_t* cell = p->table + row * p->cols + col;
      if ((op >> 40) & 1) {
        *cell = orc_write_value(p->seed, t, (uint32_t)(i - p->tx_off[t]));
        if (p->reads) p->reads[i] = 0;
      } else if (p->reads) {
        p->reads[i] = *cell;
      }
    }
  }
  return NULL;
}

static int64_t replay_parts(uint64_t rows, uint32_t cols, uint64_t seed, int row_ops, uint64_t* table,
                            const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, const uint64_t* order,
                            uint64_t* reads, int threads);

int64_t orc_replay_mt(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                      const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads, int threads) {
  return replay_parts(rows, cols, seed, 0, table, ops, tx_off, ntx, order, reads, threads);
}

int64_t orc_replay_rows(uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table, const uint64_t* ops,
                        const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* reads, int threads) {
  return replay_parts(rows, cols, seed, 1, table, ops, tx_off, ntx, order, reads, threads);
}

static int64_t replay_parts(uint64_t rows, uint32_t cols, uint64_t seed, int row_ops, uint64_t* table,
                            const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, const uint64_t* order,
                            uint64_t* reads, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  okey_t* ord = (okey_t*)malloc(sizeof(okey_t) * (ntx ? ntx : 1));
  if (!ord) return -1;
  for (uint64_t t = 0; t < ntx; ++t) {
    ord[t].b = order[3 * t];
    ord[t].k0 = order[3 * t + 1];
    ord[t].k1 = order[3 * t + 2];
    ord[t].tx = t;
  }
  qsort(ord, ntx, sizeof(okey_t), cmp_okey);
  int64_t ties = 0;
  for (uint64_t i = 1; i < ntx; ++i)
    if (ord[i].b == ord[i - 1].b && ord[i].k0 == ord[i - 1].k0 && ord[i].k1 == ord[i - 1].k1) ties++;
  /* ops whose key is not a row go to the last part, which clears their reads */
  const uint64_t span = (rows + (uint64_t)threads - 1) / (uint64_t)threads;
  pthread_t th[256];
  replay_part_t part[256];
  int started = 0;
  for (int k = 0; k < threads; ++k) {
    part[k].rows = rows;
    part[k].row_lo = (uint64_t)k * span;
    part[k].row_hi = k == threads - 1 ? ~0ull : (uint64_t)(k + 1) * span;
    part[k].ntx = ntx;
    part[k].seed = seed;
    part[k].cols = cols;
    part[k].row_ops = row_ops;
    part[k].table = table;
    part[k].ops = ops;
    part[k].tx_off = tx_off;
    part[k].ord = ord;
    part[k].reads = reads;
  }
  for (int k = 1; k < threads; ++k) {
    if (pthread_create(&th[k], NULL, replay_part, &part[k]) != 0) break;
    started = k;
  }
  /* threads that could not be started run here, after part 0 */
  replay_part(&part[0]);
  for (int k = started + 1; k < threads; ++k) replay_part(&part[k]);
  for (int k = 1; k <= started; ++k) pthread_join(th[k], NULL);
  free(ord);
  return ties;
}

/* ---- GPUTx one-pass rank ----
 * Accesses sorted by (row, tx) (include/gputx.cuh:266, GTX_AccessItem order);
 * per row, cur_rank starts at 0, every access raises rank[tx] to cur_rank, and
 * cur_rank grows by one between two neighbours of which one writes
 * (src/gputx.cu:17-27).  The reference reads the end of the last row's segment
 * past the scanned offsets; here every segment ends where the next row starts
 * (the last one at the op count).  counts = histogram of ranks
 * (src/gputx.cu:30-38); nranks = max local rank + 1 (gputx.cuh:283-294). */
int orc_gputx_rank(uint64_t rows, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx,
                   uint32_t* rank, uint32_t* counts, uint64_t* nranks) {
  /* accesses in (row, tx) order by a counting sort over rows: transactions
   * are visited in id order, so each row's list is in ascending tx */
  const uint64_t nops = tx_off[ntx];
  uint64_t* start = (uint64_t*)calloc(rows + 1, sizeof(uint64_t));
  uint32_t* acc = (uint32_t*)malloc(sizeof(uint32_t) * (nops ? nops : 1)); /* tx | write << 31 */
  if (!start || !acc) {
    free(start);
    free(acc);
    return -1;
  }
  for (uint64_t i = 0; i < nops; ++i) {
    const uint32_t row = (uint32_t)ops[i];
    if (row < rows) start[row + 1]++;
  }
  for (uint64_t r = 0; r < rows; ++r) start[r + 1] += start[r];
  const uint64_t n = start[rows];
  for (uint64_t t = 0; t < ntx; ++t)
    for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
      const uint32_t row = (uint32_t)ops[i];
      if (row >= rows) continue;
      acc[start[row]++] = (uint32_t)t | ((uint32_t)((ops[i] >> 40) & 1) << 31);
    }
  /* start[r] now holds the end of row r's list, which is row r+1's start */
  for (uint64_t t = 0; t < ntx; ++t) rank[t] = 0;
  uint32_t tot = 0;
  uint64_t st = 0;
  for (uint64_t r = 0; r < rows; ++r) {
    const uint64_t en = start[r];
    if (en == st) continue;
    uint32_t cur = 0;
    for (uint64_t i = st; i < en; ++i) {
      const uint32_t tx = acc[i] & 0x7fffffffu;
      if (rank[tx] < cur) rank[tx] = cur;
      if (i + 1 < en && ((acc[i] | acc[i + 1]) >> 31)) cur++;
    }
    if (cur + 1 > tot) tot = cur + 1;
    st = en;
  }
  free(start);
  free(acc);
  if (n == 0 && ntx) tot = 1;
  if (nranks) *nranks = tot;
  if (counts) {
    for (uint32_t r = 0; r < tot; ++r) counts[r] = 0;
    for (uint64_t t = 0; t < ntx; ++t) counts[rank[t]]++;
  }
  return 0;
}

/* ---- TPC-C (benchmark/tpcc.cu) ----
 * The updates payment and new_order make to the rows that carry concurrency
 * control, on rows of 5 u64 cells, two 32-bit fields per cell (low half
 * first): warehouse {id, tax} {ytd, -}; district {id, w_id} {tax, ytd}
 * {next_o_id, -}; customer {id, d_id} {w_id, since} {credit_lim, discount}
 * {balance, ytd_payment} {payment_cnt, delivery_cnt}; stock {i_id, w_id}
 * {quantity, ytd} {order_cnt, remote_cnt}.  kind 0..4 and the result as in
 * orc_tpcc_apply's comments. */
static float f32_of(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t u32_of(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
#define LO32(c) ((uint32_t)(c))
#define HI32(c) ((uint32_t)((c) >> 32))
#define PACK32(lo, hi) ((uint64_t)(uint32_t)(lo) | ((uint64_t)(uint32_t)(hi) << 32))

uint64_t orc_tpcc_apply(uint32_t kind, uint32_t param, uint64_t* r) {
  switch (kind) {
    case 0: { /* payment, warehouse.ytd += h_amount (tpcc.cu:102) */
      const uint32_t old = LO32(r[1]);
      const float ytd = f32_of(old) + f32_of(param);
      r[1] = PACK32(u32_of(ytd), HI32(r[1]));
      return old;
    }
    case 1: { /* payment, district.ytd += h_amount (tpcc.cu:125) */
      const uint32_t old = HI32(r[1]);
      const float ytd = f32_of(old) + f32_of(param);
      r[1] = PACK32(LO32(r[1]), u32_of(ytd));
      return old;
    }
    case 2: { /* payment, customer balance / ytd_payment / payment_cnt (tpcc.cu:152-154) */
      const uint32_t bal = LO32(r[3]);
      const float nb = f32_of(bal) - f32_of(param);
      const float ny = f32_of(HI32(r[3])) + f32_of(param);
      r[3] = PACK32(u32_of(nb), u32_of(ny));
      r[4] = PACK32(LO32(r[4]) + 1u, HI32(r[4]));
      return bal;
    }
    case 3: { /* new_order, cur_o_id = next_o_id; next_o_id++ (tpcc.cu:270-271) */
      const uint32_t old = LO32(r[2]);
      r[2] = PACK32(old + 1u, HI32(r[2]));
      return old;
    }
    case 4: { /* new_order, stock update (tpcc.cu:319-327) */
      const uint32_t qty = param & 0xffffu, remote = param >> 16;
      const uint32_t q = LO32(r[1]);
      uint32_t nq;
      if (q > qty + 10u)
        nq = q - qty;
      else
        nq = q - (qty - 91u);
      r[1] = PACK32(nq, HI32(r[1]) + qty);
      r[2] = PACK32(LO32(r[2]) + 1u, HI32(r[2]) + remote);
      return q;
    }
    default:
      return 0;
  }
}

int64_t orc_tpcc_replay(uint64_t rows, uint64_t* table, const uint64_t* ops, const uint32_t* params,
                        const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, uint64_t* results) {
  okey_t* ord = (okey_t*)malloc(sizeof(okey_t) * (ntx ? ntx : 1));
  if (!ord) return -1;
  for (uint64_t t = 0; t < ntx; ++t) {
    ord[t].b = order[3 * t];
    ord[t].k0 = order[3 * t + 1];
    ord[t].k1 = order[3 * t + 2];
    ord[t].tx = t;
  }
  qsort(ord, ntx, sizeof(okey_t), cmp_okey);
  int64_t ties = 0;
  for (uint64_t i = 1; i < ntx; ++i)
    if (ord[i].b == ord[i - 1].b && ord[i].k0 == ord[i - 1].k0 && ord[i].k1 == ord[i - 1].k1) ties++;
  for (uint64_t k = 0; k < ntx; ++k) {
    const uint64_t t = ord[k].tx;
    for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
      const uint64_t row = (uint32_t)ops[i];
      if (row >= rows) {
        results[i] = 0;
        continue;
      }
      results[i] = orc_tpcc_apply((uint32_t)(ops[i] >> 32) & 0xff, params[i], table + row * 5);
    }
  }
  free(ord);
  return ties;
}
