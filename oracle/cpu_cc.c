/*
 * cpu_cc.c -- multi-threaded CPU restatement of the eight CC protocols.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): it is the CPU baseline that
 * bench.py times beside the GPU, and a second, independent executor whose
 * results the serial replay checks like the GPU's.
 *
 * Structure follows the reference kernel per transaction (benchmark/ycsb.cu:
 * 51-106): retry loop; per op in key order Read/Write through the scheme;
 * TxEnd; immediate retry on abort.  Each host thread takes transactions of the
 * current batch from a shared counter (batches run one after another,
 * include/cc.cuh:29-35).  CC words are C11 atomics; cells are read and written
 * with relaxed atomics so that concurrent access is defined.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "oracle.h"

enum { S_TPL_NW, S_TPL_WD, S_TO, S_MVCC, S_SILO, S_TICTOC, S_GACCO, S_GPUTX };
#define MAXOPS 64

typedef struct {
  int scheme;
  uint64_t rows;
  uint32_t cols;
  uint64_t seed;
  uint64_t* table;
  const uint64_t* ops;
  const uint64_t* tx_off;
  uint64_t ntx;
  uint64_t* order;
  uint64_t* reads;
  _Atomic uint64_t* cc;      /* per-row CC word */
  _Atomic uint32_t* vprev;   /* MVCC chain head (node + 1) */
  _Atomic uint64_t* pool;    /* MVCC nodes, 3 words each */
  uint64_t pool_cap;
  _Atomic uint64_t pool_top;
  _Atomic uint32_t* counters;/* gacco/gputx per-row completed accesses */
  uint32_t* need;            /* gacco/gputx per-op group start */
  const uint32_t* slot_tx;   /* batch slot -> tx (gputx: rank order) */
  uint32_t* rank;
  /* batch being executed */
  uint64_t slot_lo, slot_hi;
  uint32_t batch_id;
  _Atomic uint64_t next;
  _Atomic uint64_t ts_ctr;
  _Atomic uint64_t ticket;
  _Atomic uint64_t aborts;
} run_t;

static inline uint64_t ld_cell(uint64_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
static inline void st_cell(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
static inline void cpu_relax(void) { __builtin_ia32_pause(); }

/* word layouts: see gcctb_common.h (same bit positions as the reference) */
#define TPL_CNT(w) (((w) >> 1) & 0x7fffffffull)
#define TPL_SH(w) ((w)&1ull)
#define TPL_HOLDER(w) (((w) >> 32) & 0x7fffffffull)
#define TPL_MAKE(sh, c, h) (((sh)&1ull) | (((c)&0x7fffffffull) << 1) | (((h)&0x7fffffffull) << 32))
#define TS_UNC(w) ((w)&1ull)
#define TS_RTS(w) (((w) >> 1) & 0x7fffffffull)
#define TS_WTS(w) (((w) >> 32) & 0x7fffffffull)
#define TS_MAKE(u, r, wt) (((u)&1ull) | (((r)&0x7fffffffull) << 1) | (((wt)&0x7fffffffull) << 32))
#define TT_DELTA(w) (((w) >> 1) & 0x7fffull)
#define TT_WTS(w) ((w) >> 16)
#define TT_MAKE(l, d, wt) (((l)&1ull) | (((d)&0x7fffull) << 1) | ((wt) << 16))

typedef struct {
  uint32_t row, col, write, j;
} op_t;

static int load_tx(const run_t* r, uint64_t t, op_t* o) {
  int n = 0;
  for (uint64_t i = r->tx_off[t]; i < r->tx_off[t + 1]; ++i) {
    const uint64_t op = r->ops[i];
    const uint32_t row = (uint32_t)op;
    if (row >= r->rows) continue;
    o[n].row = row;
    o[n].col = (uint32_t)(op >> 32) & 0xff;
    o[n].write = (uint32_t)(op >> 40) & 1;
    o[n].j = (uint32_t)(i - r->tx_off[t]);
    n++;
  }
  return n;
}
static inline uint64_t* cellp(const run_t* r, const op_t* o) { return r->table + (uint64_t)o->row * r->cols + o->col; }
static inline uint64_t* readp(const run_t* r, uint64_t t, const op_t* o) { return r->reads + r->tx_off[t] + o->j; }
static inline uint64_t wval(const run_t* r, uint64_t t, const op_t* o) { return orc_write_value(r->seed, t, o->j); }
static void set_order(run_t* r, uint64_t t, uint64_t k0, uint64_t k1) {
  r->order[3 * t] = r->batch_id;
  r->order[3 * t + 1] = k0;
  r->order[3 * t + 2] = k1;
}

/* ---- 2PL (include/twopl.cuh:233-322) ---- */
static int tpl_lock(run_t* r, uint32_t row, int excl, uint64_t self, int wd) {
  _Atomic uint64_t* w = r->cc + row;
  for (;;) {
    uint64_t cur = atomic_load(w), nxt;
    if (TPL_CNT(cur) == 0) {
      nxt = excl ? TPL_MAKE(0, 1, self) : TPL_MAKE(1, 1, TPL_HOLDER(cur));
    } else if (TPL_SH(cur) && !excl) {
      nxt = cur + 2;
    } else if (TPL_SH(cur)) {
      return 0;
    } else {
      const uint64_t h = TPL_HOLDER(cur);
      if (h == self) return 1;
      if (!wd || h < self) return 0;
      cpu_relax();
      continue;
    }
    if (atomic_compare_exchange_weak(w, &cur, nxt)) return 1;
  }
}
static void tpl_unlock(run_t* r, uint32_t row, int excl) {
  if (excl)
    atomic_store(r->cc + row, 0);
  else
    atomic_fetch_sub(r->cc + row, 2);
}
static void run_tpl(run_t* r, uint64_t t, const op_t* o, int n, int wd) {
  for (;;) {
    int k = 0, ok = 1;
    for (; k < n; ++k) {
      if (!tpl_lock(r, o[k].row, o[k].write, t, wd)) {
        ok = 0;
        break;
      }
      if (!o[k].write) *readp(r, t, &o[k]) = ld_cell(cellp(r, &o[k]));
    }
    if (!ok) {
      for (int q = 0; q < k; ++q) tpl_unlock(r, o[q].row, o[q].write);
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    const uint64_t ticket = atomic_fetch_add(&r->ticket, 1);
    for (int q = 0; q < n; ++q)
      if (o[q].write) st_cell(cellp(r, &o[q]), wval(r, t, &o[q]));
    for (int q = 0; q < n; ++q) tpl_unlock(r, o[q].row, o[q].write);
    set_order(r, t, ticket, 0);
    return;
  }
}

/* ---- T/O (include/to.cuh:104-270) ---- */
static void run_to(run_t* r, uint64_t t, const op_t* o, int n) {
  uint64_t prev[MAXOPS];
  for (;;) {
    const uint64_t ts = atomic_fetch_add(&r->ts_ctr, 1);
    int k = 0, ok = 1;
    for (; k < n && ok; ++k) {
      _Atomic uint64_t* w = r->cc + o[k].row;
      if (!o[k].write) {
        for (;;) {
          uint64_t cur = atomic_load(w);
          if (TS_WTS(cur) > ts) {
            ok = 0;
            break;
          }
          if (TS_UNC(cur)) {
            cpu_relax();
            continue;
          }
          const uint64_t v = ld_cell(cellp(r, &o[k]));
          const uint64_t rr = TS_RTS(cur) > ts ? TS_RTS(cur) : ts;
          if (atomic_compare_exchange_weak(w, &cur, TS_MAKE(0, rr, TS_WTS(cur)))) {
            *readp(r, t, &o[k]) = v;
            break;
          }
        }
      } else {
        for (;;) {
          uint64_t cur = atomic_load(w);
          if (TS_RTS(cur) > ts || TS_WTS(cur) > ts || TS_UNC(cur)) {
            ok = 0;
            break;
          }
          prev[k] = cur;
          if (atomic_compare_exchange_weak(w, &cur, TS_MAKE(1, ts, ts))) break;
        }
      }
      if (!ok) break;
    }
    if (!ok) {
      for (int q = 0; q < k; ++q)
        if (o[q].write) atomic_store(r->cc + o[q].row, prev[q]);
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    for (int q = 0; q < n; ++q)
      if (o[q].write) st_cell(cellp(r, &o[q]), wval(r, t, &o[q]));
    for (int q = 0; q < n; ++q)
      if (o[q].write) atomic_store(r->cc + o[q].row, TS_MAKE(0, ts, ts));
    set_order(r, t, ts, 0);
    return;
  }
}

/* ---- MVCC (include/mvcc.cuh:113-296): the table holds the newest version;
 * a commit pushes the old value of the written column into a pool node ---- */
static int mvcc_read(run_t* r, const op_t* o, uint64_t ts, uint64_t* out) {
  _Atomic uint64_t* w = r->cc + o->row;
  for (;;) {
    uint64_t cur = atomic_load(w);
    if (TS_UNC(cur)) {
      cpu_relax();
      continue;
    }
    if (TS_WTS(cur) <= ts) {
      const uint64_t v = ld_cell(cellp(r, o));
      const uint64_t rr = TS_RTS(cur) > ts ? TS_RTS(cur) : ts;
      if (atomic_compare_exchange_weak(w, &cur, TS_MAKE(0, rr, TS_WTS(cur)))) {
        *out = v;
        return 1;
      }
      continue;
    }
    uint64_t v = ld_cell(cellp(r, o));
    uint32_t p = atomic_load(r->vprev + o->row);
    uint64_t cw = TS_WTS(cur);
    while (cw > ts && p) {
      _Atomic uint64_t* nd = r->pool + (uint64_t)(p - 1) * 3;
      const uint64_t link = atomic_load(nd + 2);
      if ((uint32_t)(link >> 32) == o->col) v = atomic_load(nd + 1);
      cw = TS_WTS(atomic_load(nd + 0));
      p = (uint32_t)link;
    }
    if (atomic_load(w) != cur) continue;
    *out = v;
    return 1;
  }
}
static void run_mvcc(run_t* r, uint64_t t, const op_t* o, int n) {
  uint64_t prev[MAXOPS];
  for (;;) {
    const uint64_t ts = atomic_fetch_add(&r->ts_ctr, 1);
    int k = 0, ok = 1;
    for (; k < n; ++k) {
      if (!o[k].write) {
        uint64_t v;
        mvcc_read(r, &o[k], ts, &v);
        *readp(r, t, &o[k]) = v;
        continue;
      }
      _Atomic uint64_t* w = r->cc + o[k].row;
      for (;;) {
        uint64_t cur = atomic_load(w);
        if (TS_RTS(cur) > ts || TS_WTS(cur) > ts || TS_UNC(cur)) {
          ok = 0;
          break;
        }
        prev[k] = cur;
        if (atomic_compare_exchange_weak(w, &cur, TS_MAKE(1, ts, ts))) break;
      }
      if (!ok) break;
    }
    if (!ok) {
      for (int q = 0; q < k; ++q)
        if (o[q].write) atomic_store(r->cc + o[q].row, prev[q]);
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    for (int q = 0; q < n; ++q) {
      if (!o[q].write) continue;
      const uint64_t node = atomic_fetch_add(&r->pool_top, 1);
      uint64_t* cell = cellp(r, &o[q]);
      if (node < r->pool_cap) {
        _Atomic uint64_t* nd = r->pool + node * 3;
        atomic_store(nd + 0, prev[q]);
        atomic_store(nd + 1, ld_cell(cell));
        atomic_store(nd + 2, (uint64_t)atomic_load(r->vprev + o[q].row) | ((uint64_t)o[q].col << 32));
        st_cell(cell, wval(r, t, &o[q]));
        atomic_store(r->vprev + o[q].row, (uint32_t)(node + 1));
      } else {
        st_cell(cell, wval(r, t, &o[q]));
      }
    }
    for (int q = 0; q < n; ++q)
      if (o[q].write) atomic_store(r->cc + o[q].row, TS_MAKE(0, ts, ts));
    set_order(r, t, ts, 0);
    return;
  }
}

/* ---- Silo (include/silo.cuh:199-338) and TicToc (include/tictoc.cuh:198-368) ---- */
static void snapshot(run_t* r, const op_t* o, uint64_t* word, uint64_t* val) {
  _Atomic uint64_t* w = r->cc + o->row;
  for (;;) {
    const uint64_t w1 = atomic_load(w);
    const uint64_t v = ld_cell(cellp(r, o));
    const uint64_t w2 = atomic_load(w);
    if (w1 == w2 && !(w1 & 1)) {
      *word = w1;
      *val = v;
      return;
    }
    cpu_relax();
  }
}
static int lock_ws(run_t* r, const op_t* o, int n, uint64_t* seen) {
  for (int k = 0; k < n; ++k) {
    if (!o[k].write) continue;
    _Atomic uint64_t* w = r->cc + o[k].row;
    for (;;) {
      uint64_t cur = atomic_load(w);
      if (cur & 1) {
        for (int q = 0; q < k; ++q)
          if (o[q].write) atomic_store(r->cc + o[q].row, seen[q]);
        return 0;
      }
      if (atomic_compare_exchange_weak(w, &cur, cur | 1)) {
        seen[k] = cur;
        break;
      }
    }
  }
  return 1;
}
static void unlock_ws(run_t* r, const op_t* o, int n, const uint64_t* seen) {
  for (int q = 0; q < n; ++q)
    if (o[q].write) atomic_store(r->cc + o[q].row, seen[q]);
}
static void run_silo(run_t* r, uint64_t t, const op_t* o, int n) {
  uint64_t seen[MAXOPS];
  for (;;) {
    for (int k = 0; k < n; ++k)
      if (!o[k].write) {
        uint64_t v;
        snapshot(r, &o[k], &seen[k], &v);
        *readp(r, t, &o[k]) = v;
      }
    if (!lock_ws(r, o, n, seen)) {
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    const uint64_t ticket = atomic_fetch_add(&r->ticket, 1);
    uint64_t mx = 0;
    for (int k = 0; k < n; ++k)
      if ((seen[k] >> 1) > mx) mx = seen[k] >> 1;
    const uint64_t cts = mx + 1;
    int ok = 1;
    for (int k = 0; k < n && ok; ++k)
      if (!o[k].write) {
        const uint64_t w = atomic_load(r->cc + o[k].row);
        if ((w >> 1) != (seen[k] >> 1) || (w & 1)) ok = 0;
      }
    if (!ok) {
      unlock_ws(r, o, n, seen);
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    for (int q = 0; q < n; ++q)
      if (o[q].write) st_cell(cellp(r, &o[q]), wval(r, t, &o[q]));
    for (int q = 0; q < n; ++q)
      if (o[q].write) atomic_store(r->cc + o[q].row, cts << 1);
    set_order(r, t, ticket, 0);
    return;
  }
}
static void run_tictoc(run_t* r, uint64_t t, const op_t* o, int n) {
  uint64_t seen[MAXOPS];
  for (;;) {
    for (int k = 0; k < n; ++k)
      if (!o[k].write) {
        uint64_t v;
        snapshot(r, &o[k], &seen[k], &v);
        *readp(r, t, &o[k]) = v;
      }
    if (!lock_ws(r, o, n, seen)) {
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    uint64_t cts = 0;
    for (int k = 0; k < n; ++k) {
      const uint64_t c = o[k].write ? TT_WTS(seen[k]) + TT_DELTA(seen[k]) + 1 : TT_WTS(seen[k]);
      if (c > cts) cts = c;
    }
    int ok = 1;
    for (int k = 0; k < n && ok; ++k) {
      if (o[k].write || TT_WTS(seen[k]) + TT_DELTA(seen[k]) >= cts) continue;
      _Atomic uint64_t* w = r->cc + o[k].row;
      for (;;) {
        uint64_t cur = atomic_load(w);
        const uint64_t wts = TT_WTS(cur), rts = wts + TT_DELTA(cur);
        if (wts != TT_WTS(seen[k]) || (rts <= cts && (cur & 1))) {
          ok = 0;
          break;
        }
        if (rts > cts) break;
        const uint64_t delta = cts - wts, shift = delta - (delta & 0x7fff);
        if (atomic_compare_exchange_weak(w, &cur, TT_MAKE(cur & 1, delta - shift, wts + shift))) break;
      }
    }
    if (!ok) {
      unlock_ws(r, o, n, seen);
      atomic_fetch_add(&r->aborts, 1);
      continue;
    }
    const uint64_t ticket = atomic_fetch_add(&r->ticket, 1);
    for (int q = 0; q < n; ++q)
      if (o[q].write) st_cell(cellp(r, &o[q]), wval(r, t, &o[q]));
    for (int q = 0; q < n; ++q)
      if (o[q].write) atomic_store(r->cc + o[q].row, TT_MAKE(0, 0, cts));
    set_order(r, t, cts, ticket);
    return;
  }
}

/* ---- GaccO / GPUTx ordered execution: each op waits until the row's
 * completed-access counter reaches its group start ---- */
static void run_ordered(run_t* r, uint64_t t, const op_t* o, int n) {
  for (int k = 0; k < n; ++k) {
    const uint64_t i = r->tx_off[t] + o[k].j;
    _Atomic uint32_t* c = r->counters + o[k].row;
    while (atomic_load(c) < r->need[i]) cpu_relax();
    if (o[k].write)
      st_cell(cellp(r, &o[k]), wval(r, t, &o[k]));
    else
      *readp(r, t, &o[k]) = ld_cell(cellp(r, &o[k]));
    atomic_fetch_add(c, 1);
  }
  if (r->scheme == S_GPUTX)
    set_order(r, t, r->rank[t], t);
  else
    set_order(r, t, t, t);
}

static void* worker(void* arg) {
  run_t* r = (run_t*)arg;
  op_t o[MAXOPS];
  for (;;) {
    const uint64_t s = atomic_fetch_add(&r->next, 1) + r->slot_lo;
    if (s >= r->slot_hi) break;
    const uint64_t t = r->slot_tx ? r->slot_tx[s] : s;
    const int n = load_tx(r, t, o);
    switch (r->scheme) {
      case S_TPL_NW: run_tpl(r, t, o, n, 0); break;
      case S_TPL_WD: run_tpl(r, t, o, n, 1); break;
      case S_TO: run_to(r, t, o, n); break;
      case S_MVCC: run_mvcc(r, t, o, n); break;
      case S_SILO: run_silo(r, t, o, n); break;
      case S_TICTOC: run_tictoc(r, t, o, n); break;
      default: run_ordered(r, t, o, n); break;
    }
  }
  return NULL;
}

/* (row, key, op) triples sorted by (row, key, op index) give each op its
 * group start: a write waits for every earlier access of the row, a run of
 * reads for the access before the run. */
typedef struct {
  uint32_t row;
  uint64_t key;
  uint64_t op;
} acc3_t;
static int cmp_acc3(const void* pa, const void* pb) {
  const acc3_t* a = (const acc3_t*)pa;
  const acc3_t* b = (const acc3_t*)pb;
  if (a->row != b->row) return a->row < b->row ? -1 : 1;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  if (a->op != b->op) return a->op < b->op ? -1 : 1;
  return 0;
}
static int build_need(run_t* r, const uint64_t* slot_key_tx, uint64_t tx_lo_slot, uint64_t tx_hi_slot) {
  uint64_t n = 0;
  for (uint64_t s = tx_lo_slot; s < tx_hi_slot; ++s) {
    const uint64_t t = r->slot_tx ? r->slot_tx[s] : s;
    n += r->tx_off[t + 1] - r->tx_off[t];
  }
  acc3_t* a = (acc3_t*)malloc(sizeof(acc3_t) * (n ? n : 1));
  if (!a) return -1;
  uint64_t m = 0;
  for (uint64_t s = tx_lo_slot; s < tx_hi_slot; ++s) {
    const uint64_t t = r->slot_tx ? r->slot_tx[s] : s;
    for (uint64_t i = r->tx_off[t]; i < r->tx_off[t + 1]; ++i) {
      const uint32_t row = (uint32_t)r->ops[i];
      if (row >= r->rows) continue;
      a[m].row = row;
      a[m].key = slot_key_tx ? slot_key_tx[t] : t;
      a[m].op = i;
      m++;
    }
  }
  qsort(a, m, sizeof(acc3_t), cmp_acc3);
  uint32_t g = 0;
  for (uint64_t p = 0; p < m; ++p) {
    const int row_start = p == 0 || a[p].row != a[p - 1].row;
    const int w = (int)((r->ops[a[p].op] >> 40) & 1);
    if (row_start || w || ((r->ops[a[p - 1].op] >> 40) & 1)) g = (uint32_t)p;
    r->need[a[p].op] = g;
    if (row_start) atomic_store(r->counters + a[p].row, (uint32_t)p);
  }
  free(a);
  return 0;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void run_batch(run_t* r, pthread_t* th, int threads) {
  atomic_store(&r->next, 0);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, worker, r);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

int orc_cpu_run(int scheme, uint64_t rows, uint32_t cols, uint64_t seed, uint64_t* table,
                const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx, uint64_t batch_size, int threads,
                uint64_t* order, uint64_t* reads, orc_cpu_stats* stats) {
  if (scheme < 0 || scheme > S_GPUTX || batch_size == 0 || cols == 0) return -1;
  for (uint64_t t = 0; t < ntx; ++t)
    if (tx_off[t + 1] - tx_off[t] > MAXOPS) return -1;
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (threads < 1) threads = 1;
  run_t* r = (run_t*)calloc(1, sizeof(run_t));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  if (!r || !th) {
    free(r);
    free(th);
    return -2;
  }
  r->scheme = scheme;
  r->rows = rows;
  r->cols = cols;
  r->seed = seed;
  r->table = table;
  r->ops = ops;
  r->tx_off = tx_off;
  r->ntx = ntx;
  r->order = order;
  r->reads = reads;
  const uint64_t nops = tx_off[ntx];
  memset(reads, 0, sizeof(uint64_t) * (nops ? nops : 1));
  r->cc = (_Atomic uint64_t*)calloc(rows, sizeof(uint64_t));
  int rc = r->cc ? 0 : -2;
  if (!rc && scheme == S_MVCC) {
    r->vprev = (_Atomic uint32_t*)calloc(rows, sizeof(uint32_t));
    r->pool_cap = nops;
    r->pool = (_Atomic uint64_t*)calloc(3 * (nops ? nops : 1), sizeof(uint64_t));
    if (!r->vprev || !r->pool) rc = -2;
  }
  uint32_t* slot_tx = NULL;
  uint64_t* key_of_tx = NULL;
  uint64_t nbatches = 0;
  uint64_t* bounds = NULL; /* slot boundaries of the batches */
  if (!rc && (scheme == S_GACCO || scheme == S_GPUTX)) {
    r->counters = (_Atomic uint32_t*)calloc(rows, sizeof(uint32_t));
    r->need = (uint32_t*)calloc(nops ? nops : 1, sizeof(uint32_t));
    if (!r->counters || !r->need) rc = -2;
  }
  if (!rc && scheme == S_GPUTX) {
    /* rank every transaction (src/gputx.cu:5-38), order by (rank, tx) */
    r->rank = (uint32_t*)calloc(ntx ? ntx : 1, sizeof(uint32_t));
    uint32_t* counts = (uint32_t*)calloc(ntx + 2, sizeof(uint32_t));
    slot_tx = (uint32_t*)malloc(sizeof(uint32_t) * (ntx ? ntx : 1));
    key_of_tx = (uint64_t*)malloc(sizeof(uint64_t) * (ntx ? ntx : 1));
    uint64_t nr = 0;
    if (!r->rank || !counts || !slot_tx || !key_of_tx ||
        orc_gputx_rank(rows, ops, tx_off, ntx, r->rank, counts, &nr) != 0) {
      rc = -2;
    } else {
      bounds = (uint64_t*)malloc(sizeof(uint64_t) * (nr + 1));
      uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (nr + 1));
      if (!bounds || !fill) {
        rc = -2;
      } else {
        bounds[0] = 0;
        for (uint64_t k = 0; k < nr; ++k) bounds[k + 1] = bounds[k] + counts[k];
        memcpy(fill, bounds, sizeof(uint64_t) * (nr + 1));
        for (uint64_t t = 0; t < ntx; ++t) slot_tx[fill[r->rank[t]]++] = (uint32_t)t;
        for (uint64_t t = 0; t < ntx; ++t) key_of_tx[t] = ((uint64_t)r->rank[t] << 32) | t;
        nbatches = nr;
      }
      free(fill);
    }
    free(counts);
    r->slot_tx = slot_tx;
  }
  if (!rc && scheme != S_GPUTX) {
    nbatches = (ntx + batch_size - 1) / batch_size;
    bounds = (uint64_t*)malloc(sizeof(uint64_t) * (nbatches + 1));
    if (!bounds) rc = -2;
    for (uint64_t b = 0; !rc && b <= nbatches; ++b) bounds[b] = b * batch_size < ntx ? b * batch_size : ntx;
  }
  double t0 = now_s();
  if (!rc && scheme == S_GPUTX && ntx) {
    if (build_need(r, key_of_tx, 0, ntx) != 0) rc = -2;
  }
  for (uint64_t b = 0; !rc && b < nbatches; ++b) {
    if (bounds[b] == bounds[b + 1]) continue;
    r->slot_lo = bounds[b];
    r->slot_hi = bounds[b + 1];
    r->batch_id = scheme == S_GPUTX ? 0 : (uint32_t)b;
    atomic_store(&r->ts_ctr, 1);
    atomic_store(&r->ticket, 0);
    atomic_store(&r->pool_top, 0);
    if (scheme == S_GACCO && build_need(r, NULL, bounds[b], bounds[b + 1]) != 0) {
      rc = -2;
      break;
    }
    run_batch(r, th, threads);
    /* per-batch reset of the touched CC words (timestamps restart at 1) */
    if (scheme != S_GACCO && scheme != S_GPUTX) {
      for (uint64_t s = bounds[b]; s < bounds[b + 1]; ++s)
        for (uint64_t i = tx_off[s]; i < tx_off[s + 1]; ++i) {
          const uint32_t row = (uint32_t)ops[i];
          if (row >= rows) continue;
          atomic_store(r->cc + row, 0);
          if (r->vprev) atomic_store(r->vprev + row, 0);
        }
    }
  }
  const double secs = now_s() - t0;
  if (stats) {
    stats->seconds = secs;
    stats->aborts = atomic_load(&r->aborts);
    stats->committed = rc ? 0 : ntx;
    stats->threads = threads;
  }
  free((void*)r->cc);
  free((void*)r->vprev);
  free((void*)r->pool);
  free((void*)r->counters);
  free(r->need);
  free(r->rank);
  free(slot_tx);
  free(key_of_tx);
  free(bounds);
  free(th);
  free(r);
  return rc;
}
