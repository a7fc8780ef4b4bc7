// Host-side launch wrappers for the gfx950 kernels.  Every wrapper enqueues on
// the given stream and returns hipSuccess or the launch error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../gcctb_common.h"

namespace gcctb {
namespace launch {

// --- execution kernels (one translation unit per scheme family) ---
// G must be 1, 2, 4, 8, 16 or 32; threads = block threads.
hipError_t exec_tpl(bool wait_die, int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
hipError_t exec_to(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
hipError_t exec_mvcc(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
hipError_t exec_silo(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
hipError_t exec_tictoc(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
hipError_t exec_ordered(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s);
// kernel variants (plain form only), any scheme: whole-row mode, the event
// log, or both
hipError_t exec_row(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s);
hipError_t exec_evt(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s);
hipError_t exec_row_evt(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s);
// the latch schemes (slow_*), plain form, column mode: the CC arrays hold a
// 16-byte slot per row (the word, then the latch)
hipError_t exec_latch(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s);
// TPC-C kernels (plain form): every op updates its row in place (tpcc_apply)
hipError_t exec_tpcc(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s);

// --- table / index ---
hipError_t seed_table(uint64_t* cells, uint64_t rows, uint32_t cols, uint64_t seed, hipStream_t s);
// Insert n keys (keys == nullptr: key i) with values (vals == nullptr:
// value i) into a hash index whose nbuckets buckets start all free
// (every u32 IDX_EMPTY).  Keys must be unique and not IDX_EMPTY.
hipError_t index_insert(const uint32_t* keys, const uint32_t* vals, uint64_t n, uint32_t* buckets,
                        uint32_t nbuckets, hipStream_t s);
// rows_out[i] = row of ops[i]'s key; also into rows_copy when it is given;
// ids_out[i] = what op_ids(ops, n, id_base) writes, when it is given; with
// id_map, id_map[i] with the op's write flag in bit 31 instead
hipError_t probe(const IndexView& ix, const uint64_t* ops, uint64_t n, uint32_t* rows_out,
                 hipStream_t s, uint32_t* rows_copy = nullptr, uint32_t* ids_out = nullptr,
                 uint32_t id_base = 0, const uint32_t* id_map = nullptr);

// --- per-op helpers ---
hipError_t op_tx(const uint64_t* tx_off, uint64_t ntx, uint32_t* op_tx, hipStream_t s);
hipError_t iota_u32(uint32_t* v, uint64_t n, uint32_t base, hipStream_t s);
// out[i] = (base + i) | OPID_WRITE if ops[base + i] writes: the values the
// GaccO / GPUTx sorts carry, so later passes read the flag in order.
hipError_t op_ids(const uint64_t* ops, uint64_t n, uint32_t base, uint32_t* out, hipStream_t s);
hipError_t fill_u32(uint32_t* v, uint64_t n, uint32_t x, hipStream_t s);
hipError_t fill_u64(uint64_t* v, uint64_t n, uint64_t x, hipStream_t s);

// --- stable LSD radix sort (8-bit digits) of (key, value) pairs ---
// Sorts bits [0, key_bits) of the keys.  Buffers: keys/vals in, alt buffers of
// the same size; the sorted result ends in (*keys_out, *vals_out), which point
// to one of the two buffer pairs.  tmp must hold radix_sort_tmp_words(n) u32.
uint64_t radix_sort_tmp_words(uint64_t n);
hipError_t radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint64_t n,
                          int key_bits, uint32_t* tmp, uint32_t** keys_out, uint32_t** vals_out,
                          hipStream_t s);
hipError_t radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t n,
                          int key_bits, uint32_t* tmp, uint64_t** keys_out, uint32_t** vals_out,
                          hipStream_t s);

// --- scans (u32) --- tmp must hold scan_tmp_words(n) u32.
uint64_t scan_tmp_words(uint64_t n);
hipError_t inclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s);
hipError_t inclusive_max_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s);

// --- ordered-execution preprocessing ---
// From accesses sorted by (row, order), sop holding op_ids values: group-start
// flags (p or 0) and, when rowstart is given, row-start flags (p or 0).
hipError_t access_group_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint32_t* flags,
                              hipStream_t s, uint32_t* rowstart = nullptr);
// need[op] = gstart[p] + pos_base for op = sop[p] & OPID_MASK;
// counters[row_slot(srow[p])] = p + pos_base at row starts.
hipError_t access_need(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart,
                       uint64_t n, uint32_t pos_base, uint32_t* need, uint32_t* counters,
                       uint64_t nrows, hipStream_t s);
// need[op] = gstart[p] - segstart[p] for op = sop[p] & OPID_MASK: the turn
// relative to the row's first access, for counters that start at zero.
// With comb_min > 0 it also lists every row queue of at least comb_min
// accesses as (first position, length) in segs (seg_ctr counts them; at most
// seg_cap are kept).
hipError_t access_need_rel(const uint32_t* sop, const uint32_t* gstart, const uint32_t* segstart, uint64_t n,
                           uint32_t* need, hipStream_t s, const uint32_t* srow = nullptr, uint64_t nrows = 0,
                           uint32_t comb_min = 0, uint2* segs = nullptr, uint32_t* seg_ctr = nullptr,
                           uint32_t seg_cap = 0);
// GaccO, persistent form: run each listed row queue in order on the row's
// cells (cols <= 64), store its reads' results, mark its ops need = NEED_DONE
// and add the ops run to *combined.  wvals (nops u64) receives every op's
// write value first, when any queue is listed.
hipError_t gacco_combine(const uint2* segs, const uint32_t* seg_ctr, uint32_t seg_cap, const uint32_t* srow,
                         const uint32_t* sop, const uint64_t* ops, const uint32_t* op_tx, const uint64_t* tx_off,
                         uint64_t nops, uint64_t* cells, uint32_t cols, uint64_t seed, uint64_t* wvals, uint64_t* reads,
                         uint32_t* need, uint64_t* combined, hipStream_t s);
// Open form, one admitted batch sorted by (row, op) with group starts gstart
// and row starts rstart: need[op] = base[row] + gstart[p] - rstart[p]; then
// row_base_add adds the batch's accesses of each row to base[row].
hipError_t access_need_open(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart, const uint32_t* rstart,
                            uint64_t n, uint64_t nrows, const uint32_t* base, uint32_t* need, hipStream_t s);
hipError_t row_base_add(const uint32_t* srow, const uint32_t* rstart, uint64_t n, uint64_t nrows, uint32_t* base,
                        hipStream_t s);
// GPUTx: *conflicts += adjacent same-row, same-rank accesses where either
// writes, from the (row << rank_bits | rank) sorted keys and the group flags.
hipError_t gtx_conflicts(const uint64_t* skey, const uint32_t* flags, uint64_t n, uint64_t* conflicts,
                         hipStream_t s);
hipError_t gtx_conflicts32(const uint32_t* skey, const uint32_t* flags, uint64_t n, uint64_t* conflicts,
                           hipStream_t s);
// GPUTx one-pass rank (src/gputx.cu:5-27).  boundary[p] = 1 if p is not the
// first access of its row and (p-1 or p) is a write; rowstart[p] = p at row
// starts else 0.
hipError_t gtx_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint64_t nrows, uint32_t* boundary,
                     uint32_t* rowstart, hipStream_t s);
// local rank = S[p] - S[SS[p]]; rank[tx] = max over tx's accesses, where
// stx[p] is the transaction at sorted position p (write flag in bit 31);
// max_rank = max.
hipError_t gtx_rank(const uint32_t* stx, const uint32_t* S, const uint32_t* SS, uint64_t n,
                    uint32_t* rank, uint32_t* max_rank, hipStream_t s);
// composite key (row << rank_bits) | rank[op_tx[op]] for ops [0, n)
hipError_t gtx_exec_keys(const uint32_t* rows, const uint32_t* op_tx, const uint32_t* rank,
                         uint64_t n, int rank_bits, uint64_t* keys, hipStream_t s);
// the same key cut to 32 bits, for row_bits + rank_bits <= 32
hipError_t gtx_exec_keys32(const uint32_t* rows, const uint32_t* op_tx, const uint32_t* rank,
                           uint64_t n, int rank_bits, uint32_t* keys, hipStream_t s);
// out[i] = keys[i] >> shift
hipError_t u64_key_high(const uint64_t* keys, uint64_t n, int shift, uint32_t* out, hipStream_t s);
hipError_t u32_key_high(const uint32_t* keys, uint64_t n, int shift, uint32_t* out, hipStream_t s);
hipError_t rank_hist(const uint32_t* rank, uint64_t ntx, uint32_t nranks, uint32_t* hist, hipStream_t s);

// --- memory-system ceilings (microbench.hip) ---
hipError_t microbench(int which, uint64_t n_words, uint64_t threads_total, uint64_t iters, double* rate,
                      double* ms_out);

// --- hot-writer deferral (persistent form of the non-ordered schemes) ---
// map[slot] = the transaction the slot runs: the transactions that write no
// hot key from slot 0 up, those that do from slot ntx - 1 down (a key is hot
// when at least one transaction in `share` touches it, estimated from about
// DEFER_SAMPLE sampled ops).  cnt holds DEFER_WORDS u32: DEFER_BUCKETS
// sampled counts, then the state words (cnt + DEFER_BUCKETS)[DEFER_ANY] (a
// key was hot) and [DEFER_NHOT] (the number deferred); map holds ntx u32.
constexpr uint32_t DEFER_BUCKETS = 1u << DEFER_BUCKET_BITS;
constexpr uint64_t DEFER_SAMPLE = 16384;
constexpr uint32_t DEFER_ANY = 0, DEFER_NCOLD = 1, DEFER_NHOT = 2;
constexpr uint64_t DEFER_WORDS = DEFER_BUCKETS + 8;
hipError_t defer_order(const uint64_t* tx_off, const uint64_t* ops, uint64_t ntx, uint64_t nops, uint32_t share,
                       uint32_t* cnt, uint32_t* map, hipStream_t s);
// the sampled-count threshold defer_order uses for a hot key
uint32_t defer_threshold(uint64_t ntx, uint64_t nops, uint32_t share);

// --- per-batch state ---
// clear the CC words (and MVCC chain heads) of rows[0, n) at their slots
// (arrays of slot_words(nrows_table) words); latch: the CC array holds two
// words per slot (the word and the row's latch)
hipError_t reset_rows(const uint32_t* rows, uint64_t n, uint64_t nrows_table, uint64_t* cc,
                      uint32_t* vprev, hipStream_t s, bool latch = false);
hipError_t batch_begin(uint64_t* ctrs, hipStream_t s);
// admission state of a persistent launch (slot counter, gates, admission words)
hipError_t persist_begin(uint64_t* ctrs, hipStream_t s);
// *out = the device wall clock (100 MHz), written by a one-block kernel
hipError_t clock_sample(uint64_t* out, hipStream_t s);

}  // namespace launch
}  // namespace gcctb
