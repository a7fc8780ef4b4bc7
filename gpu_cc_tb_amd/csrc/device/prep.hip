// Preprocessing and maintenance kernels for gfx950: table seeding, sorted-array
// index construction and probing, a stable LSD radix sort, scans, and the
// GaccO / GPUTx access-order passes that replace the reference's Thrust calls
// (include/gacco.cuh:307-336, include/gputx.cuh:252-281).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dev_util.h"
#include "launch.h"

namespace gcctb {
namespace {

constexpr int BT = 256;  // threads per block for the streaming kernels

inline unsigned grid_for(uint64_t n, int per_block) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return (unsigned)g;
}
inline unsigned grid_stride_blocks(uint64_t n) {
  uint64_t g = (n + BT - 1) / BT;
  if (g > 65536 * 4) g = 65536 * 4;
  if (g == 0) g = 1;
  return (unsigned)g;
}

// Kernels that end in an atomic on one address run at most REDUCE_BLOCKS
// blocks and combine their waves in LDS first: one address sustains only
// ~1e8 atomics/s, so one atomic per wave of a 16M-element grid costs ~3 ms.
constexpr unsigned REDUCE_BLOCKS = 2048;
inline unsigned reduce_blocks(uint64_t n) {
  const unsigned g = grid_stride_blocks(n);
  return g < REDUCE_BLOCKS ? g : REDUCE_BLOCKS;
}
// Block-wide sum (Max = false) or max (Max = true) of one value per thread;
// the result is valid in thread 0.  Every thread of the block must call it.
template <bool Max>
__device__ __forceinline__ uint64_t block_reduce_u64(uint64_t v) {
  __shared__ uint64_t part[BT / 64];
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const uint64_t o = dev::shfl_xor64(v, m);
    v = Max ? (o > v ? o : v) : v + o;
  }
  if (dev::lane_id() == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = part[0];
#pragma unroll
    for (int w = 1; w < BT / 64; ++w) v = Max ? (part[w] > v ? part[w] : v) : v + part[w];
  }
  return v;
}

__global__ void k_seed_table(uint64_t* cells, uint64_t rows, uint32_t cols, uint64_t seed) {
  const uint64_t n = rows * cols;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = i / cols;
    cells[i] = cell_init(seed, r, (uint32_t)(i - r * cols), cols);
  }
}

// Hash index build: each key takes the first free slot along its bucket
// sequence (CAS on the key word; a slot never frees again).  The table has
// about twice the slots it needs, so the sequence is short; it is bounded by
// nbuckets in any case.
__global__ void k_index_insert(const uint32_t* keys, const uint32_t* vals, uint64_t n, uint32_t* buckets,
                               uint32_t nbuckets) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t key = keys ? keys[i] : (uint32_t)i;
    const uint32_t val = vals ? vals[i] : (uint32_t)i;
    uint32_t b = idx_bucket(key, nbuckets);
    for (uint32_t step = 0; step < nbuckets; ++step) {
      uint32_t* bk = buckets + (size_t)b * (2 * IDX_SLOTS);
      int s = 0;
      while (s < IDX_SLOTS && atomicCAS(bk + s, IDX_EMPTY, key) != IDX_EMPTY) ++s;
      if (s < IDX_SLOTS) {
        bk[IDX_SLOTS + s] = val;
        break;
      }
      b = b + 1 == nbuckets ? 0u : b + 1;
    }
  }
}

// Quad (4 consecutive lanes) primitives over DPP quad permutes; every lane
// of the quad must be active.
template <int J>
__device__ __forceinline__ uint32_t quad_get(uint32_t v) {  // lane J's value
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, J * 0x55, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t quad_sum(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);  // lanes 1,0,3,2
  v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);  // lanes 2,3,0,1
  return v;
}
__device__ __forceinline__ int quad_max(int v) {
  int o = __builtin_amdgcn_mov_dpp(v, 0xb1, 0xf, 0xf, false);
  v = o > v ? o : v;
  o = __builtin_amdgcn_mov_dpp(v, 0x4e, 0xf, 0xf, false);
  return o > v ? o : v;
}

// One bucket step of a quad-cooperative lookup: lane q holds the q-th 16
// bytes of the bucket (keys 0-3, keys 4-7, values 0-3, values 4-7).  Sets
// *val and returns true on a hit; *free_slot tells whether the bucket has a
// free slot (then the key is absent).
__device__ __forceinline__ bool quad_bucket_match(const uint4& w, uint32_t key, uint32_t q, uint32_t* val,
                                                  bool* free_slot) {
  const bool keys_here = q < 2;
  int h = -1;
  h = keys_here && w.x == key ? (int)q * 4 : h;
  h = keys_here && w.y == key ? (int)q * 4 + 1 : h;
  h = keys_here && w.z == key ? (int)q * 4 + 2 : h;
  h = keys_here && w.w == key ? (int)q * 4 + 3 : h;
  const int fr = keys_here && (w.x == IDX_EMPTY || w.y == IDX_EMPTY || w.z == IDX_EMPTY || w.w == IDX_EMPTY);
  h = quad_max(h);
  *free_slot = quad_max(fr) != 0;
  // the value of slot h sits in lane 2 + h / 4, component h % 4
  const uint32_t comp = (h & 3) == 0 ? w.x : (h & 3) == 1 ? w.y : (h & 3) == 2 ? w.z : w.w;
  const uint32_t from2 = quad_get<2>(comp), from3 = quad_get<3>(comp);
  *val = (h >> 2) ? from3 : from2;
  return h >= 0;
}

// Index probe (Find, include/index.cuh:74-91, over the hash index of
// dev::index_find).  A quad walks four probes at once: lane q loads op p0+q
// (a wave loads 64 consecutive ops), the quad shares the four keys, and lane
// q reads the q-th 16 bytes of each probe's home bucket, so one load
// instruction covers one line per probe and the four buckets are read at
// once.  A probe whose bucket is full without the key goes on bucket by
// bucket, the quad together.  A dense index (keys 0 .. length-1 on rows of
// the same number) maps each key to itself and reads no bucket.
// ids_out[i] (optional): the op's id (id_base + i), or id_map[i] when given,
// with the op's write flag in bit 31
__global__ __launch_bounds__(256) void k_probe(IndexView ix, const uint64_t* ops, uint64_t n, uint32_t* rows_out,
                                               uint32_t* rows_copy, uint32_t* ids_out, uint32_t id_base,
                                               const uint32_t* id_map) {
  if (ix.dense) {  // direct addressing: no bucket to read
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
      const uint64_t op = ops[i];
      const uint32_t r = idx_dense_find(op_key(op), ix.length);
      rows_out[i] = r;
      if (rows_copy) rows_copy[i] = r;
      if (ids_out) ids_out[i] = (id_map ? id_map[i] : id_base + (uint32_t)i) | (op_is_write(op) ? OPID_WRITE : 0u);
    }
    return;
  }
  const uint32_t q = threadIdx.x & 3u;
  const uint4* bkt = reinterpret_cast<const uint4*>(ix.buckets);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;  // a multiple of 4
  for (uint64_t p0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~3ull; p0 < n; p0 += stride) {
    const uint64_t mine = p0 + q;
    const uint64_t myop = mine < n ? ops[mine] : 0ull;
    const uint32_t mykey = mine < n ? op_key(myop) : IDX_EMPTY;
    const uint32_t key[4] = {quad_get<0>(mykey), quad_get<1>(mykey), quad_get<2>(mykey), quad_get<3>(mykey)};
    uint32_t b[4];
    uint4 w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = idx_bucket(key[j], ix.nbuckets);
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = bkt[(size_t)b[j] * 4 + q];
    uint32_t res[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint32_t r = IDX_MISS;
      if (key[j] != IDX_EMPTY) {  // the reserved key is never stored
        uint4 cur = w[j];
        uint32_t bb = b[j];
        for (uint32_t step = 1;; ++step) {
          uint32_t val;
          bool free_slot;
          if (quad_bucket_match(cur, key[j], q, &val, &free_slot)) {
            r = val;
            break;
          }
          if (free_slot || step >= ix.nbuckets) break;
          bb = bb + 1 == ix.nbuckets ? 0u : bb + 1;
          cur = bkt[(size_t)bb * 4 + q];
        }
      }
      res[j] = r;
    }
    const uint32_t r = q == 0 ? res[0] : q == 1 ? res[1] : q == 2 ? res[2] : res[3];
    if (mine < n) {
      rows_out[mine] = r;
      if (rows_copy) rows_copy[mine] = r;
      if (ids_out)
        ids_out[mine] = (id_map ? id_map[mine] : id_base + (uint32_t)mine) | (op_is_write(myop) ? OPID_WRITE : 0u);
    }
  }
}

__global__ void k_op_tx(const uint64_t* tx_off, uint64_t ntx, uint32_t* op_tx) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntx;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = tx_off[t], e = tx_off[t + 1];
    for (uint64_t i = b; i < e; ++i) op_tx[i] = (uint32_t)t;
  }
}
// Op ids for the GaccO / GPUTx sorts: base + i with the op's write flag in
// bit 31, so the passes over the sorted order read it without a random load.
__global__ void k_op_ids(const uint64_t* ops, uint64_t n, uint32_t base, uint32_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (base + (uint32_t)i) | (op_is_write(ops[base + i]) ? OPID_WRITE : 0u);
}
__global__ void k_iota(uint32_t* v, uint64_t n, uint32_t base) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = base + (uint32_t)i;
}
__global__ void k_fill32(uint32_t* v, uint64_t n, uint32_t x) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = x;
}
__global__ void k_fill64(uint64_t* v, uint64_t n, uint64_t x) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = x;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8 bits per pass: per-tile digit histograms, one scan
// of the [digit][tile] matrix, then the scatter.  A tile is RS_TILE elements;
// wave w of the block holds tile elements [w*RS_WSPAN, (w+1)*RS_WSPAN),
// lane-minor, so (wave, item, lane) order is input order.  The scatter ranks
// every element within its digit (ballots inside a wave, a per-wave LDS
// counter across its items, a prefix over waves), sorts the tile in LDS, and
// writes it out digit run by digit run, so consecutive lanes store to
// consecutive addresses.
// ---------------------------------------------------------------------------
constexpr int RS_BT = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BT * RS_ITEMS;
constexpr int RS_WAVES = RS_BT / 64;
constexpr int RS_WSPAN = RS_ITEMS * 64;

template <class KT>
__global__ __launch_bounds__(RS_BT) void k_rs_hist(const KT* keys, uint64_t n, int shift, uint32_t dmask,
                                                   uint32_t* hist, uint32_t nb) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  uint32_t d[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint64_t e = base + (uint64_t)i * RS_BT + threadIdx.x;
    d[i] = e < n ? (uint32_t)(keys[e] >> shift) & dmask : 256u;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i)
    if (d[i] < 256u) atomicAdd(&h[d[i]], 1u);
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// offs[d * nb + b]: where tile b's first element of digit d goes; the digit
// is (key >> shift) & dmask (dmask < 0xff only on a last, narrower pass)
template <class KT>
__global__ __launch_bounds__(RS_BT) void k_rs_scatter(const KT* kin, const uint32_t* vin, KT* kout,
                                                      uint32_t* vout, uint64_t n, int shift, uint32_t dmask,
                                                      const uint32_t* offs, uint32_t nb) {
  __shared__ KT skey[RS_TILE];
  __shared__ uint32_t sval[RS_TILE];
  __shared__ uint32_t wcnt[RS_WAVES][256];  // digit counts per wave, then each wave's start in the digit
  __shared__ uint32_t dstart[256];          // tile-local start of each digit
  __shared__ uint32_t gbase[256];           // global start of each digit's run from this tile
  __shared__ uint32_t wsum[RS_WAVES];
  const uint32_t t = threadIdx.x, lane = dev::lane_id(), wave = t / 64;
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  const uint32_t cnt = n - base < (uint64_t)RS_TILE ? (uint32_t)(n - base) : (uint32_t)RS_TILE;
#pragma unroll
  for (int w = 0; w < RS_WAVES; ++w) wcnt[w][t] = 0;
  KT k[RS_ITEMS];
  uint32_t v[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const uint32_t e = wave * RS_WSPAN + i * 64 + lane;
    k[i] = e < cnt ? kin[base + e] : (KT)0;
    v[i] = e < cnt ? vin[base + e] : 0u;
  }
  __syncthreads();
  // rank within (wave, digit): earlier items of the wave, then lower lanes
  uint32_t r[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const bool valid = wave * RS_WSPAN + i * 64 + lane < cnt;
    const uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
    uint64_t peers = dev::ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = dev::ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below = (uint32_t)__popcll(peers & dev::lanemask_lt());
    uint32_t old = 0;
    if (valid && below == 0) old = atomicAdd(&wcnt[wave][d], (uint32_t)__popcll(peers));
    const int leader = peers ? __ffsll((unsigned long long)peers) - 1 : (int)lane;
    r[i] = (uint32_t)__shfl((int)old, leader, 64) + below;
  }
  __syncthreads();
  // digit t: each wave's start within the digit, the digit's tile total, and
  // the tile-local start of the digit (exclusive scan over digits)
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < RS_WAVES; ++w) {
    const uint32_t c = wcnt[w][t];
    wcnt[w][t] = tot;
    tot += c;
  }
  uint32_t x = tot;
#pragma unroll
  for (int dd = 1; dd < 64; dd <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)x, dd, 64);
    if (lane >= (uint32_t)dd) x += o;
  }
  if (lane == 63) wsum[wave] = x;
  gbase[t] = offs[(uint64_t)t * nb + blockIdx.x];
  __syncthreads();
  uint32_t carry = 0;
  for (uint32_t w = 0; w < wave; ++w) carry += wsum[w];
  dstart[t] = carry + x - tot;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    if (wave * RS_WSPAN + i * 64 + lane < cnt) {
      const uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
      const uint32_t lp = dstart[d] + wcnt[wave][d] + r[i];
      skey[lp] = k[i];
      sval[lp] = v[i];
    }
  }
  __syncthreads();
  for (uint32_t j = t; j < cnt; j += RS_BT) {
    const KT key = skey[j];
    const uint32_t d = (uint32_t)(key >> shift) & dmask;
    const uint32_t pos = gbase[d] + (j - dstart[d]);
    if (pos < n) {  // always true when the counts agree; keeps a logic error off foreign memory
      kout[pos] = key;
      vout[pos] = sval[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Scans over u32 with + or max.  Three kernels: block totals, a single-block
// scan of the totals, and the block-local scan with its carry-in.
// ---------------------------------------------------------------------------
constexpr int SC_BT = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BT * SC_ITEMS;

struct OpAdd {
  static __device__ __forceinline__ uint32_t id() { return 0u; }
  static __device__ __forceinline__ uint32_t f(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMax {
  static __device__ __forceinline__ uint32_t id() { return 0u; }
  static __device__ __forceinline__ uint32_t f(uint32_t a, uint32_t b) { return a > b ? a : b; }
};

// inclusive scan across the block of one value per thread; returns inclusive,
// writes the block total to *total.
template <class Op>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[SC_BT / 64];
  const uint32_t lane = dev::lane_id(), wave = threadIdx.x / 64;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)x, d, 64);
    if (lane >= (uint32_t)d) x = Op::f(o, x);
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t carry = Op::id();
  for (uint32_t w = 0; w < wave; ++w) carry = Op::f(carry, wsum[w]);
  uint32_t t = Op::id();
  for (int w = 0; w < SC_BT / 64; ++w) t = Op::f(t, wsum[w]);
  *total = t;
  __syncthreads();
  return Op::f(carry, x);
}

// Tile totals.  The operators are commutative, so the tile is read striped
// across the block (element i * SC_BT + t by thread t: every load coalesced).
template <class Op>
__global__ __launch_bounds__(SC_BT) void k_scan_reduce(const uint32_t* in, uint64_t n, uint32_t* sums) {
  const uint64_t base = (uint64_t)blockIdx.x * SC_TILE + threadIdx.x;
  uint32_t acc = Op::id();
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    const uint64_t e = base + (uint64_t)i * SC_BT;
    acc = Op::f(acc, e < n ? in[e] : Op::id());
  }
  uint32_t total;
  block_scan<Op>(acc, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
// exclusive scan of sums[0..nb) in place (single block)
template <class Op>
__global__ __launch_bounds__(SC_BT) void k_scan_top(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t tincl[SC_BT];
  const uint32_t per = (nb + SC_BT - 1) / SC_BT;
  const uint32_t b = threadIdx.x * per;
  uint32_t acc = Op::id();
  for (uint32_t i = 0; i < per; ++i)
    if (b + i < nb) acc = Op::f(acc, sums[b + i]);
  uint32_t total;
  const uint32_t incl = block_scan<Op>(acc, &total);
  tincl[threadIdx.x] = incl;
  __syncthreads();
  uint32_t c = threadIdx.x == 0 ? Op::id() : tincl[threadIdx.x - 1];
  for (uint32_t i = 0; i < per; ++i) {
    if (b + i < nb) {
      const uint32_t v = sums[b + i];
      sums[b + i] = c;
      c = Op::f(c, v);
    }
  }
}
// Tile scans.  The tile moves between global memory and LDS striped across
// the block (coalesced); thread t scans its SC_ITEMS consecutive elements out
// of LDS.  Element j sits at LDS word j + j / 16, so the 16 consecutive words
// of thread t start at 17 t and a wave's reads fall in distinct banks.  The
// whole tile is loaded before any of it is stored, so in may equal out.
template <class Op, bool INCLUSIVE>
__global__ __launch_bounds__(SC_BT) void k_scan_down(const uint32_t* in, uint32_t* out, uint64_t n,
                                                     const uint32_t* sums) {
  static_assert(SC_ITEMS == 16, "the LDS padding assumes 16 items per thread");
  __shared__ uint32_t tile[SC_TILE + SC_TILE / 16];
  __shared__ uint32_t tincl[SC_BT];
  const uint32_t t = threadIdx.x;
  const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    const uint32_t j = (uint32_t)i * SC_BT + t;
    tile[j + (j >> 4)] = base + j < n ? in[base + j] : Op::id();
  }
  __syncthreads();
  uint32_t vals[SC_ITEMS];
  uint32_t acc = Op::id();
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    vals[i] = tile[t * 17 + i];
    acc = Op::f(acc, vals[i]);
  }
  uint32_t total;
  const uint32_t incl = block_scan<Op>(acc, &total);
  // exclusive prefix for this thread = block carry + (incl of previous threads)
  tincl[t] = incl;
  __syncthreads();
  uint32_t c = Op::f(sums[blockIdx.x], t == 0 ? Op::id() : tincl[t - 1]);
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    if (INCLUSIVE) {
      c = Op::f(c, vals[i]);
      tile[t * 17 + i] = c;
    } else {
      tile[t * 17 + i] = c;
      c = Op::f(c, vals[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    const uint32_t j = (uint32_t)i * SC_BT + t;
    if (base + j < n) out[base + j] = tile[j + (j >> 4)];
  }
}

template <class Op, bool INCLUSIVE>
hipError_t scan_impl(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t nb = (n + SC_TILE - 1) / SC_TILE;
  hipLaunchKernelGGL((k_scan_reduce<Op>), dim3((unsigned)nb), dim3(SC_BT), 0, s, in, n, tmp);
  hipLaunchKernelGGL((k_scan_top<Op>), dim3(1), dim3(SC_BT), 0, s, tmp, (uint32_t)nb);
  hipLaunchKernelGGL((k_scan_down<Op, INCLUSIVE>), dim3((unsigned)nb), dim3(SC_BT), 0, s, in, out, n,
                     tmp);
  return hipGetLastError();
}

template <class KT>
hipError_t radix_impl(KT* k0, uint32_t* v0, KT* k1, uint32_t* v1, uint64_t n, int key_bits,
                      uint32_t* tmp, KT** kout, uint32_t** vout, hipStream_t s) {
  KT* ka = k0;
  uint32_t* va = v0;
  KT* kb = k1;
  uint32_t* vb = v1;
  if (n > 1 && key_bits > 0) {
    const uint64_t nb = (n + RS_TILE - 1) / RS_TILE;
    uint32_t* hist = tmp;
    uint32_t* stmp = tmp + 256 * nb;
    for (int shift = 0; shift < key_bits; shift += 8) {
      const uint32_t dmask = key_bits - shift >= 8 ? 0xffu : (1u << (key_bits - shift)) - 1u;
      hipLaunchKernelGGL((k_rs_hist<KT>), dim3((unsigned)nb), dim3(RS_BT), 0, s, ka, n, shift, dmask, hist,
                         (uint32_t)nb);
      hipError_t e = scan_impl<OpAdd, false>(hist, hist, 256 * nb, stmp, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((k_rs_scatter<KT>), dim3((unsigned)nb), dim3(RS_BT), 0, s, ka, va, kb, vb, n,
                         shift, dmask, hist, (uint32_t)nb);
      KT* tk = ka;
      ka = kb;
      kb = tk;
      uint32_t* tv = va;
      va = vb;
      vb = tv;
    }
  }
  *kout = ka;
  *vout = va;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ordered-execution passes.
// ---------------------------------------------------------------------------
// rowstart (optional): p at the first access of a row, else 0
__global__ void k_access_group_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint32_t* flags,
                                     uint32_t* rowstart) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const bool row0 = (p == 0) || srow[p] != srow[p - 1];
    const bool start = row0 || ((sop[p] | sop[p - 1]) & OPID_WRITE) != 0u;
    flags[p] = start ? (uint32_t)p : 0u;
    if (rowstart) rowstart[p] = row0 ? (uint32_t)p : 0u;
  }
}
__global__ void k_access_need(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart,
                              uint64_t n, uint32_t pos_base, uint32_t* need, uint32_t* counters,
                              uint64_t nrows) {
  const uint32_t shift = slot_shift(nrows);
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    need[sop[p] & OPID_MASK] = gstart[p] + pos_base;
    // an index miss probes to the index length; such ops never execute
    if (((p == 0) || srow[p] != srow[p - 1]) && srow[p] < nrows)
      counters[row_slot(srow[p], shift)] = (uint32_t)p + pos_base;
  }
}

__global__ void k_access_need_rel(const uint32_t* sop, const uint32_t* gstart, const uint32_t* segstart,
                                  uint64_t n, uint32_t* need, const uint32_t* srow, uint64_t nrows,
                                  uint32_t comb_min, uint2* segs, uint32_t* seg_ctr, uint32_t seg_cap) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    need[sop[p] & OPID_MASK] = gstart[p] - segstart[p];
    // the last access of a row queue at least comb_min long lists the queue
    // for a combiner
    if (comb_min && (p + 1 == n || segstart[p + 1] != segstart[p])) {
      const uint32_t len = (uint32_t)p + 1 - segstart[p];
      if (len >= comb_min && srow[p] < nrows) {
        const uint32_t i = atomicAdd(seg_ctr, 1u);
        if (i < seg_cap) segs[i] = make_uint2(segstart[p], len);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// GaccO, persistent form: a row whose access queue (its accesses in
// transaction-id order) holds at least comb_min accesses is run by one
// workgroup, the row's combiner, before the launch.  The combiner keeps the
// row's cells in LDS and applies the queue in order, so every read returns
// the value of the last write to its column before it in transaction-id
// order, or the cell's value before the stream, and the row ends with the
// last write of each column: what the queue's turns produce one handoff at a
// time.  Its ops are marked done (need = NEED_DONE) and the launch skips
// them; no other op touches the row.  Column mode, cols <= 64.
//
// Steps of CB_BT queue entries, one per thread, all 16 waves at once:
//  1. every thread gathers its entry (column, write flag, written value);
//  2. each wave resolves its own 64 entries: per column written among them
//     (one ballot each), a read finds the last write below it, and the last
//     write of each column is kept as the wave's summary;
//  3. lane c of wave 0 carries column c through the 16 summaries in order,
//     which gives every wave the value each column holds at its start;
//  4. the reads with no write below them in their wave take that value, and
//     every thread stores its read's result and the done mark.
// Written values come precomputed per op (k_gacco_wvals), so the gather is
// two dependent loads deep.
constexpr int CB_BT = 1024;
constexpr int CB_WAVES = CB_BT / 64;
// wv[op] = the value op would write, for every op (the combiners read those
// of their writes); nothing to do when no queue is listed
__global__ void k_gacco_wvals(const uint32_t* op_tx, const uint64_t* tx_off, uint64_t nops, uint64_t seed,
                              const uint32_t* seg_ctr, uint64_t* wv) {
  if (*seg_ctr == 0) return;  // the same for every thread of the grid
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nops;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = op_tx[i];
    wv[i] = write_value(seed, t, (uint32_t)(i - tx_off[t]));
  }
}
__global__ __launch_bounds__(CB_BT) void k_gacco_combine(const uint2* segs, const uint32_t* seg_ctr, uint32_t seg_cap,
                                                         const uint32_t* srow, const uint32_t* sop, const uint64_t* ops,
                                                         const uint64_t* wvals, uint64_t* cells, uint32_t cols,
                                                         uint64_t* reads, uint32_t* need, uint64_t* combined) {
  __shared__ uint64_t val[CB_BT];            // written values, by entry
  __shared__ uint64_t sum[CB_WAVES][64];     // per wave: the last value it wrote to each column
  __shared__ uint64_t cin[CB_WAVES][64];     // per wave: each column's value at its first entry
  __shared__ uint64_t wmask[CB_WAVES];       // per wave: the columns it wrote
  __shared__ uint64_t cell[64];              // the row between steps
  const uint32_t nseg = min(*seg_ctr, seg_cap);
  const uint32_t tid = threadIdx.x, w = tid / 64, lane = tid % 64;
  uint64_t done_ops = 0;
  for (uint32_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const uint2 seg = segs[sg];
    const uint32_t p0 = seg.x, len = seg.y;
    uint64_t* rc = cells + (uint64_t)srow[p0] * cols;
    if (tid < cols) cell[tid] = rc[tid];
    for (uint32_t b = 0; b < len; b += CB_BT) {
      const uint32_t i = b + tid;
      const bool valid = i < len;
      uint32_t opid = 0, col = 0;
      bool isw = false;
      uint64_t wv = 0;
      if (valid) {
        const uint32_t so = sop[p0 + i];
        opid = so & OPID_MASK;
        isw = (so & OPID_WRITE) != 0;
        col = op_col(ops[opid]);
        if (isw) wv = wvals[opid];
      }
      int src = -1;  // the lane (of this wave) whose write this read sees
      uint64_t written = 0;
      uint64_t wm = dev::ballot(valid && isw);
      while (wm) {  // the columns written among the wave's 64 entries
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)col, __ffsll((unsigned long long)wm) - 1);
        const uint64_t wc = dev::ballot(valid && isw && col == c);
        if (valid && col == c) {
          if (isw) {
            if ((wc >> lane) >> 1 == 0) sum[w][c] = wv;  // the wave's last write of c
          } else {
            const uint64_t below = wc & dev::lanemask_lt();
            if (below) src = 63 - __clzll((long long)below);
          }
        }
        written |= 1ull << c;
        wm &= ~wc;
      }
      val[tid] = wv;
      if (lane == 0) wmask[w] = written;
      __syncthreads();
      if (tid < cols) {
        uint64_t cur = cell[tid];
        for (int x = 0; x < CB_WAVES; ++x) {
          cin[x][tid] = cur;
          if ((wmask[x] >> tid) & 1ull) cur = sum[x][tid];
        }
        cell[tid] = cur;
      }
      __syncthreads();
      if (valid) {
        if (!isw) reads[opid] = src >= 0 ? val[w * 64 + (uint32_t)src] : cin[w][col];
        need[opid] = NEED_DONE;
      }
      __syncthreads();  // val / sum / cin / wmask are rewritten by the next step
    }
    if (tid < cols) rc[tid] = cell[tid];
    done_ops += len;
    __syncthreads();  // cell is reloaded for the next row
  }
  if (tid == 0 && done_ops) atomicAdd((unsigned long long*)combined, (unsigned long long)done_ops);
}

// Open form (serving), one admitted batch at a time, accesses sorted by
// (row, op): the turn counts the accesses to the row that earlier batches of
// the launch queued (base[row]) plus the group start within the batch.
__global__ void k_access_need_open(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart,
                                   const uint32_t* rstart, uint64_t n, uint64_t nrows, const uint32_t* base,
                                   uint32_t* need) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = srow[p];
    if (r >= nrows) continue;  // an index miss never executes
    need[sop[p] & OPID_MASK] = base[r] + (gstart[p] - rstart[p]);
  }
}
// base[row] += the batch's accesses to the row, written once per row (at the
// row's last position); run after k_access_need_open.
__global__ void k_row_base_add(const uint32_t* srow, const uint32_t* rstart, uint64_t n, uint64_t nrows,
                               uint32_t* base) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = srow[p];
    if (r >= nrows) continue;
    if (p + 1 == n || srow[p + 1] != r) base[r] += (uint32_t)(p - rstart[p] + 1);
  }
}

// GPUTx same-rank conflicts: adjacent accesses of one row with equal rank
// where either writes.  In the (row << rank_bits | rank) sorted keys, same
// row and rank means equal keys; flags[p] (access_group_flags) is non-zero
// at a non-row-start position exactly when either access writes.
template <class KT>
__global__ void k_gtx_conflicts(const KT* skey, const uint32_t* flags, uint64_t n,
                                unsigned long long* conflicts) {
  uint64_t local = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; p < n;
       p += (uint64_t)gridDim.x * blockDim.x)
    local += (skey[p] == skey[p - 1] && flags[p] != 0u) ? 1u : 0u;
  // one atomic per block (launched with reduce_blocks)
  local = block_reduce_u64<false>(local);
  if (threadIdx.x == 0 && local) atomicAdd(conflicts, (unsigned long long)local);
}

__global__ void k_gtx_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint64_t nrows,
                            uint32_t* boundary, uint32_t* rowstart) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const bool start = (p == 0) || srow[p] != srow[p - 1];
    rowstart[p] = start ? (uint32_t)p : 0u;
    // index misses (IDX_MISS sorts last) never execute and raise no rank
    boundary[p] = (!start && srow[p] < nrows &&
                   ((sop[p] | sop[p - 1]) & OPID_WRITE) != 0u) ? 1u : 0u;
  }
}
// stx[p]: the transaction of sorted position p (write flag in bit 31)
__global__ void k_gtx_rank(const uint32_t* stx, const uint32_t* S, const uint32_t* SS, uint64_t n,
                           uint32_t* rank, uint32_t* max_rank) {
  uint32_t mx = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t local = S[p] - S[SS[p]];
    if (local) atomicMax(&rank[stx[p] & OPID_MASK], local);  // ranks start at 0
    mx = local > mx ? local : mx;
  }
  // one atomic per block (launched with reduce_blocks)
  mx = (uint32_t)block_reduce_u64<true>(mx);
  if (threadIdx.x == 0 && mx) atomicMax(max_rank, mx);
}
template <class KT>
__global__ void k_gtx_exec_keys(const uint32_t* rows, const uint32_t* op_tx, const uint32_t* rank,
                                uint64_t n, int rank_bits, KT* keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    keys[i] = (KT)(((uint64_t)rows[i] << rank_bits) | rank[op_tx[i]]);
}
template <class KT>
__global__ void k_key_high(const KT* keys, uint64_t n, int shift, uint32_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)(keys[i] >> shift);
}
// Histogram of ranks (gputx_pre3, src/gputx.cu:30-38).  There are usually few
// ranks, so each block counts into LDS and adds its non-zero bins once.
constexpr uint32_t RANK_LDS_BINS = 8192;
__global__ __launch_bounds__(256) void k_rank_hist(const uint32_t* rank, uint64_t ntx, uint32_t nranks,
                                                   uint32_t* hist) {
  __shared__ uint32_t h[RANK_LDS_BINS];
  const bool in_lds = nranks <= RANK_LDS_BINS;
  if (in_lds)
    for (uint32_t b = threadIdx.x; b < nranks; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rank[i];
    if (in_lds)
      atomicAdd(&h[r], 1u);
    else
      atomicAdd(&hist[r], 1u);
  }
  __syncthreads();
  if (in_lds)
    for (uint32_t b = threadIdx.x; b < nranks; b += blockDim.x)
      if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ void k_reset_rows(const uint32_t* rows, uint64_t n, uint64_t nrows_table, uint32_t shift,
                             uint32_t wshift, uint64_t* cc, uint32_t* vprev) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rows[i];
    if (r >= nrows_table) continue;
    const uint32_t w = row_slot(r, shift);
    if (cc) dev::st(cc + ((uint64_t)w << wshift), 0);
    if (vprev) dev::st32(vprev + w, 0);
  }
}
__global__ void k_batch_begin(uint64_t* ctrs) {
  if (threadIdx.x == 0) {
    dev::st(ctrs + CTR_TICKET, 0);  // commit tickets
    dev::st(ctrs + CTR_POOL, 0);    // MVCC pool bump pointer
    // timestamp allocators restart at 1 (timestamp.cuh:76-80)
    for (int s = 0; s < TS_SHARDS; ++s) dev::st(ctrs + CTR_TS0 + 16 * s, 1);
  }
}

// Admission state of one persistent launch: slot counter, gate copies and
// (open form) admission words.
__global__ void k_persist_begin(uint64_t* ctrs) {
  const int i = threadIdx.x;
  if (i == 0) dev::st(ctrs + CTR_SLOT, 0);
  if (i < GATE_SHARDS) {
    dev::st(ctrs + CTR_GATE0 + 16 * i + 1, 0);
    dev::st(ctrs + CTR_GATE0 + 16 * i + 2, 0);
    dev::st(ctrs + CTR_ADMIT0 + 16 * i, 0);  // the u32 admission word and its neighbour
  }
}

// Device wall clock now (for mapping commit stamps to host time).
__global__ void k_clock_sample(uint64_t* out) {
  if (threadIdx.x == 0) dev::st(out, wall_clock64());
}

// ---------------------------------------------------------------------------
// Hot-writer deferral (persistent form of the lock, timestamp and OCC
// schemes).  A transaction that writes a hot key conflicts with every
// concurrent transaction touching that key, and under high skew those writers
// hold up (2PL) or abort (OCC, T/O) the far larger number of transactions that
// only read the hot keys.  The launch therefore takes every other transaction
// first and the hot writers last.  A key is hot when its bucket (Fibonacci
// hash of the key) holds at least `thr` of a sample of the ops (every
// stride-th op); thr is scaled so that a key touched by at least one
// transaction in `share` qualifies.  The deferral only reorders admission:
// each scheme still reports its own serialization order.

// sampled access counts per bucket; the lanes of a wave that share a bucket
// add together (one atomic per distinct bucket of the wave)
__global__ __launch_bounds__(BT) void k_defer_count(const uint64_t* ops, uint64_t nsamp, uint64_t stride,
                                                    uint32_t* cnt) {
  for (uint64_t base = (uint64_t)blockIdx.x * BT; base < nsamp; base += (uint64_t)gridDim.x * BT) {
    const uint64_t i = base + threadIdx.x;
    const bool v = i < nsamp;
    const uint32_t b = v ? defer_bucket(op_key(ops[i * stride])) : 0u;
    uint64_t act = dev::ballot(v);
    while (act) {
      const int l = __ffsll((unsigned long long)act) - 1;
      const uint32_t lb = dev::shfl32(b, l);
      const uint64_t same = dev::ballot(b == lb) & act;
      if ((int)dev::lane_id() == l) atomicAdd(cnt + lb, (uint32_t)__popcll(same));
      act &= ~same;
    }
  }
}

// any bucket hot?  (st was cleared with the counts; every hit writes 1)
__global__ __launch_bounds__(BT) void k_defer_any(const uint32_t* cnt, uint32_t thr, uint32_t* st) {
  const uint32_t i = blockIdx.x * BT + threadIdx.x;
  if (i < launch::DEFER_BUCKETS && cnt[i] >= thr) st[launch::DEFER_ANY] = 1u;
}

// map[slot] = transaction.  Nothing hot: the identity.  Otherwise a
// transaction that writes a hot key is deferred: the others fill the slots
// from 0 up, the deferred ones from ntx - 1 down, each class in the order its
// blocks claim their ranges (one pair of atomics per block step);
// st[DEFER_NHOT] ends as the number deferred.
__global__ __launch_bounds__(BT) void k_defer_order(const uint64_t* tx_off, const uint64_t* ops, uint64_t ntx,
                                                    const uint32_t* cnt, uint32_t thr, uint32_t* st, uint32_t* map) {
  __shared__ uint32_t wc[BT / 64], wh[BT / 64];
  __shared__ uint32_t bc, bh;
  const bool any = st[launch::DEFER_ANY] != 0;  // the same for every thread of the grid
  const uint32_t w = threadIdx.x / 64;
  for (uint64_t base = (uint64_t)blockIdx.x * BT; base < ntx; base += (uint64_t)gridDim.x * BT) {
    const uint64_t t = base + threadIdx.x;
    const bool v = t < ntx;
    if (!any) {
      if (v) map[t] = (uint32_t)t;
      continue;
    }
    bool h = false;
    if (v) {
      const uint64_t o1 = tx_off[t + 1];
      for (uint64_t o = tx_off[t]; o < o1; ++o) {
        const uint64_t op = ops[o];
        h |= op_is_write(op) && cnt[defer_bucket(op_key(op))] >= thr;
      }
    }
    const uint64_t mh = dev::ballot(h), mc = dev::ballot(v && !h);
    if (dev::lane_id() == 0) {
      wc[w] = (uint32_t)__popcll(mc);
      wh[w] = (uint32_t)__popcll(mh);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t sc = 0, sh = 0;
      for (int i = 0; i < BT / 64; ++i) {
        const uint32_t x = wc[i], y = wh[i];
        wc[i] = sc;
        wh[i] = sh;
        sc += x;
        sh += y;
      }
      bc = sc ? atomicAdd(st + launch::DEFER_NCOLD, sc) : 0u;
      bh = sh ? atomicAdd(st + launch::DEFER_NHOT, sh) : 0u;
    }
    __syncthreads();
    const uint64_t below = dev::lanemask_lt();
    if (h)
      map[ntx - 1 - (bh + wh[w] + (uint32_t)__popcll(mh & below))] = (uint32_t)t;
    else if (v)
      map[bc + wc[w] + (uint32_t)__popcll(mc & below)] = (uint32_t)t;
    __syncthreads();  // wc / wh / bc / bh are rewritten by the next step
  }
}

}  // namespace

namespace launch {

#define GC_LAUNCH(kern, n, ...)                                                    \
  do {                                                                            \
    hipLaunchKernelGGL(kern, dim3(grid_stride_blocks(n)), dim3(BT), 0, s, __VA_ARGS__); \
    return hipGetLastError();                                                     \
  } while (0)
#define GC_LAUNCH_REDUCE(kern, n, ...)                                            \
  do {                                                                            \
    hipLaunchKernelGGL(kern, dim3(reduce_blocks(n)), dim3(BT), 0, s, __VA_ARGS__); \
    return hipGetLastError();                                                     \
  } while (0)

hipError_t seed_table(uint64_t* cells, uint64_t rows, uint32_t cols, uint64_t seed, hipStream_t s) {
  GC_LAUNCH(k_seed_table, rows * cols, cells, rows, cols, seed);
}
hipError_t index_insert(const uint32_t* keys, const uint32_t* vals, uint64_t n, uint32_t* buckets,
                        uint32_t nbuckets, hipStream_t s) {
  GC_LAUNCH(k_index_insert, n, keys, vals, n, buckets, nbuckets);
}
hipError_t probe(const IndexView& ix, const uint64_t* ops, uint64_t n, uint32_t* rows_out,
                 hipStream_t s, uint32_t* rows_copy, uint32_t* ids_out, uint32_t id_base, const uint32_t* id_map) {
  GC_LAUNCH(k_probe, n, ix, ops, n, rows_out, rows_copy, ids_out, id_base, id_map);
}
hipError_t op_tx(const uint64_t* tx_off, uint64_t ntx, uint32_t* op_tx, hipStream_t s) {
  GC_LAUNCH(k_op_tx, ntx, tx_off, ntx, op_tx);
}
hipError_t iota_u32(uint32_t* v, uint64_t n, uint32_t base, hipStream_t s) {
  GC_LAUNCH(k_iota, n, v, n, base);
}
hipError_t op_ids(const uint64_t* ops, uint64_t n, uint32_t base, uint32_t* out, hipStream_t s) {
  GC_LAUNCH(k_op_ids, n, ops, n, base, out);
}
hipError_t fill_u32(uint32_t* v, uint64_t n, uint32_t x, hipStream_t s) { GC_LAUNCH(k_fill32, n, v, n, x); }
uint32_t defer_threshold(uint64_t ntx, uint64_t nops, uint32_t share) {
  const uint64_t stride = nops > DEFER_SAMPLE ? nops / DEFER_SAMPLE : 1;
  const uint64_t want = (ntx + (uint64_t)share * stride - 1) / ((uint64_t)share * stride);
  return (uint32_t)(want < 2 ? 2 : want > 0xffffffffull ? 0xffffffffull : want);
}
hipError_t defer_order(const uint64_t* tx_off, const uint64_t* ops, uint64_t ntx, uint64_t nops, uint32_t share,
                       uint32_t* cnt, uint32_t* map, hipStream_t s) {
  if (ntx == 0) return hipSuccess;
  // about DEFER_SAMPLE sampled ops
  const uint64_t stride = nops > DEFER_SAMPLE ? nops / DEFER_SAMPLE : 1;
  const uint64_t nsamp = nops / stride;
  const uint32_t thr = defer_threshold(ntx, nops, share);
  uint32_t* st = cnt + DEFER_BUCKETS;
  hipLaunchKernelGGL(k_fill32, dim3(grid_stride_blocks(DEFER_WORDS)), dim3(BT), 0, s, cnt, (uint64_t)DEFER_WORDS, 0u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (nsamp) {
    hipLaunchKernelGGL(k_defer_count, dim3(grid_stride_blocks(nsamp)), dim3(BT), 0, s, ops, nsamp, stride, cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_defer_any, dim3(DEFER_BUCKETS / BT), dim3(BT), 0, s, cnt, thr, st);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_defer_order, dim3(reduce_blocks(ntx)), dim3(BT), 0, s, tx_off, ops, ntx, cnt, thr, st, map);
  return hipGetLastError();
}
hipError_t fill_u64(uint64_t* v, uint64_t n, uint64_t x, hipStream_t s) { GC_LAUNCH(k_fill64, n, v, n, x); }

uint64_t scan_tmp_words(uint64_t n) { return (n + SC_TILE - 1) / SC_TILE + 64; }
uint64_t radix_sort_tmp_words(uint64_t n) {
  const uint64_t nb = (n + RS_TILE - 1) / RS_TILE;
  return 256 * nb + scan_tmp_words(256 * nb) + 64;
}
hipError_t radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint64_t n,
                          int key_bits, uint32_t* tmp, uint32_t** keys_out, uint32_t** vals_out,
                          hipStream_t s) {
  return radix_impl<uint32_t>(k0, v0, k1, v1, n, key_bits, tmp, keys_out, vals_out, s);
}
hipError_t radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint64_t n,
                          int key_bits, uint32_t* tmp, uint64_t** keys_out, uint32_t** vals_out,
                          hipStream_t s) {
  return radix_impl<uint64_t>(k0, v0, k1, v1, n, key_bits, tmp, keys_out, vals_out, s);
}
hipError_t inclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s) {
  return scan_impl<OpAdd, true>(in, out, n, tmp, s);
}
hipError_t inclusive_max_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s) {
  return scan_impl<OpMax, true>(in, out, n, tmp, s);
}
hipError_t access_group_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint32_t* flags,
                              hipStream_t s, uint32_t* rowstart) {
  GC_LAUNCH(k_access_group_flags, n, srow, sop, n, flags, rowstart);
}
hipError_t access_need(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart,
                       uint64_t n, uint32_t pos_base, uint32_t* need, uint32_t* counters,
                       uint64_t nrows, hipStream_t s) {
  GC_LAUNCH(k_access_need, n, srow, sop, gstart, n, pos_base, need, counters, nrows);
}
hipError_t access_need_rel(const uint32_t* sop, const uint32_t* gstart, const uint32_t* segstart, uint64_t n,
                           uint32_t* need, hipStream_t s, const uint32_t* srow, uint64_t nrows, uint32_t comb_min,
                           uint2* segs, uint32_t* seg_ctr, uint32_t seg_cap) {
  GC_LAUNCH(k_access_need_rel, n, sop, gstart, segstart, n, need, srow, nrows, comb_min, segs, seg_ctr, seg_cap);
}
hipError_t gacco_combine(const uint2* segs, const uint32_t* seg_ctr, uint32_t seg_cap, const uint32_t* srow,
                         const uint32_t* sop, const uint64_t* ops, const uint32_t* op_tx, const uint64_t* tx_off,
                         uint64_t nops, uint64_t* cells, uint32_t cols, uint64_t seed, uint64_t* wvals, uint64_t* reads,
                         uint32_t* need, uint64_t* combined, hipStream_t s) {
  if (seg_cap == 0 || cols == 0 || cols > 64 || nops == 0) return hipSuccess;
  // a capped grid: with no queue listed every block leaves at once
  hipLaunchKernelGGL(k_gacco_wvals, dim3(reduce_blocks(nops)), dim3(BT), 0, s, op_tx, tx_off, nops, seed, seg_ctr,
                     wvals);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t grid = seg_cap < 512 ? seg_cap : 512;
  hipLaunchKernelGGL(k_gacco_combine, dim3(grid), dim3(CB_BT), 0, s, segs, seg_ctr, seg_cap, srow, sop, ops, wvals,
                     cells, cols, reads, need, combined);
  return hipGetLastError();
}
hipError_t gtx_conflicts(const uint64_t* skey, const uint32_t* flags, uint64_t n, uint64_t* conflicts,
                         hipStream_t s) {
  GC_LAUNCH_REDUCE(k_gtx_conflicts<uint64_t>, n, skey, flags, n, (unsigned long long*)conflicts);
}
hipError_t gtx_conflicts32(const uint32_t* skey, const uint32_t* flags, uint64_t n, uint64_t* conflicts,
                           hipStream_t s) {
  GC_LAUNCH_REDUCE(k_gtx_conflicts<uint32_t>, n, skey, flags, n, (unsigned long long*)conflicts);
}
hipError_t gtx_flags(const uint32_t* srow, const uint32_t* sop, uint64_t n, uint64_t nrows, uint32_t* boundary,
                     uint32_t* rowstart, hipStream_t s) {
  GC_LAUNCH(k_gtx_flags, n, srow, sop, n, nrows, boundary, rowstart);
}
hipError_t gtx_rank(const uint32_t* stx, const uint32_t* S, const uint32_t* SS, uint64_t n,
                    uint32_t* rank, uint32_t* max_rank, hipStream_t s) {
  GC_LAUNCH_REDUCE(k_gtx_rank, n, stx, S, SS, n, rank, max_rank);
}
hipError_t gtx_exec_keys(const uint32_t* rows, const uint32_t* op_tx, const uint32_t* rank,
                         uint64_t n, int rank_bits, uint64_t* keys, hipStream_t s) {
  GC_LAUNCH(k_gtx_exec_keys<uint64_t>, n, rows, op_tx, rank, n, rank_bits, keys);
}
hipError_t gtx_exec_keys32(const uint32_t* rows, const uint32_t* op_tx, const uint32_t* rank,
                           uint64_t n, int rank_bits, uint32_t* keys, hipStream_t s) {
  GC_LAUNCH(k_gtx_exec_keys<uint32_t>, n, rows, op_tx, rank, n, rank_bits, keys);
}
hipError_t u64_key_high(const uint64_t* keys, uint64_t n, int shift, uint32_t* out, hipStream_t s) {
  GC_LAUNCH(k_key_high<uint64_t>, n, keys, n, shift, out);
}
hipError_t u32_key_high(const uint32_t* keys, uint64_t n, int shift, uint32_t* out, hipStream_t s) {
  GC_LAUNCH(k_key_high<uint32_t>, n, keys, n, shift, out);
}
hipError_t rank_hist(const uint32_t* rank, uint64_t ntx, uint32_t nranks, uint32_t* hist, hipStream_t s) {
  uint64_t blocks = (ntx + 4095) / 4096;
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_rank_hist, dim3((unsigned)blocks), dim3(256), 0, s, rank, ntx, nranks, hist);
  return hipGetLastError();
}
hipError_t reset_rows(const uint32_t* rows, uint64_t n, uint64_t nrows_table, uint64_t* cc,
                      uint32_t* vprev, hipStream_t s, bool latch) {
  // Outside a run every CC word and chain head is zero, so clearing the whole
  // arrays leaves the same state.  That streams rows x 12 bytes instead of one
  // scattered line per op: the cheaper form once n >= rows / 8.  (A latch
  // array is zero outside a critical section, so its fill changes nothing.)
  const uint32_t wshift = latch ? 1 : 0;
  if (n * 8 >= nrows_table) {
    if (cc) {
      const hipError_t e = fill_u64(cc, slot_words(nrows_table) << wshift, 0, s);
      if (e != hipSuccess) return e;
    }
    if (vprev) return fill_u32(vprev, slot_words(nrows_table), 0, s);
    return hipSuccess;
  }
  GC_LAUNCH(k_reset_rows, n, rows, n, nrows_table, slot_shift(nrows_table), wshift, cc, vprev);
}
hipError_t access_need_open(const uint32_t* srow, const uint32_t* sop, const uint32_t* gstart, const uint32_t* rstart,
                            uint64_t n, uint64_t nrows, const uint32_t* base, uint32_t* need, hipStream_t s) {
  if (n == 0) return hipSuccess;
  GC_LAUNCH(k_access_need_open, n, srow, sop, gstart, rstart, n, nrows, base, need);
}
hipError_t row_base_add(const uint32_t* srow, const uint32_t* rstart, uint64_t n, uint64_t nrows, uint32_t* base,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  GC_LAUNCH(k_row_base_add, n, srow, rstart, n, nrows, base);
}
hipError_t persist_begin(uint64_t* ctrs, hipStream_t s) {
  hipLaunchKernelGGL(k_persist_begin, dim3(1), dim3(64), 0, s, ctrs);
  return hipGetLastError();
}
hipError_t clock_sample(uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_clock_sample, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

hipError_t batch_begin(uint64_t* ctrs, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, s, ctrs);
  return hipGetLastError();
}

}  // namespace launch
}  // namespace gcctb
