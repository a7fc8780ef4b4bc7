// Memory-system ceilings of the GPU, measured with the same access forms the
// transaction kernels use: agent-scope 64-bit CAS / add to random words,
// same-address atomics, coherent (sc1) and cached random loads, and a
// streaming copy for the HBM bandwidth.  The transaction path is bound by
// these, not by arithmetic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_util.h"
#include "launch.h"

namespace gcctb {
namespace {

__device__ __forceinline__ uint64_t rnd(uint64_t s, uint64_t i) { return splitmix64(s ^ (i * 0x9E3779B97F4A7C15ull)); }

__global__ __launch_bounds__(256) void mb_cas_random(uint64_t* words, uint64_t n, uint64_t iters, uint64_t seed,
                                                      unsigned long long* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) {
    uint64_t* w = words + rnd(seed, t * iters + i) % n;
    uint64_t cur = dev::ld(w);
    dev::cas(w, cur, cur + 1);
    acc += cur;
  }
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
__global__ __launch_bounds__(256) void mb_add_random(uint64_t* words, uint64_t n, uint64_t iters, uint64_t seed,
                                                      unsigned long long* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) acc += dev::fetch_add(words + rnd(seed, t * iters + i) % n, 1);
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
__global__ __launch_bounds__(256) void mb_add_same(uint64_t* words, uint64_t iters, unsigned long long* sink) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i)
    if (dev::lane_id() == 0) acc += dev::fetch_add(words, 1);
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
__global__ __launch_bounds__(256) void mb_load_random(const uint64_t* words, uint64_t n, uint64_t iters, uint64_t seed,
                                                       int coherent, unsigned long long* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) {
    const uint64_t* p = words + rnd(seed, t * iters + i) % n;
    acc += coherent ? dev::ld(p) : *p;
  }
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
// Pairs of coherent 8-byte loads issued back to back: in one 128-byte line
// (offsets 0 and 80, like a row's cell and a word stored beside it) or in two
// random lines.  Their ratio says whether the second access of a line is free.
__global__ __launch_bounds__(256) void mb_load_pairs(const uint64_t* words, uint64_t n, uint64_t iters, uint64_t seed,
                                                      int same_line, unsigned long long* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t lines = n / 16;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) {
    const uint64_t r = rnd(seed, t * iters + i);
    const uint64_t* p = words + (r % lines) * 16;
    const uint64_t* q = same_line ? p + 10 : words + (rnd(seed ^ 0x51, t * iters + i) % lines) * 16 + 10;
    acc += dev::ld(p) + dev::ld(q);
  }
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
// A timestamp read: load a word, load a cell, CAS the word, with the cell in
// the word's 128-byte line or in another line.
__global__ __launch_bounds__(256) void mb_ts_read(uint64_t* words, uint64_t n, uint64_t iters, uint64_t seed,
                                                   int same_line, unsigned long long* sink) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t lines = n / 16;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) {
    uint64_t* w = words + (rnd(seed, t * iters + i) % lines) * 16 + 15;
    const uint64_t* c = same_line ? w - 15 : words + (rnd(seed ^ 0x51, t * iters + i) % lines) * 16;
    uint64_t cur = dev::ld(w);
    acc += dev::ld(c);
    dev::cas(w, cur, cur + 1);
  }
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
// Hot words: every lane adds to one of 16 words, which share one 128-byte
// line (spread = 0) or sit on 16 lines (spread = 1).  The ratio says whether
// the atomics of different words in one line serialize like those of one
// word (hot Zipf rows are neighbours in a dense per-row array).
__global__ __launch_bounds__(256) void mb_add_hot16(uint64_t* words, uint64_t iters, int spread,
                                                     unsigned long long* sink) {
  const uint32_t w = (blockIdx.x * 7 + threadIdx.x) & 15u;
  uint64_t* p = words + (spread ? (uint64_t)w * 16 : (uint64_t)w);
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) acc += dev::fetch_add(p, 1);
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
// One hot word under atomics (lane 0 of every wave adds to word 0) while the
// other lanes poll with coherent loads: words 1..15 of the same line
// (spread = 0) or words on 15 other lines (spread = 1).  Rate: adds to word 0.
__global__ __launch_bounds__(256) void mb_add_polled(uint64_t* words, uint64_t iters, int spread,
                                                      unsigned long long* sink) {
  const uint32_t l = dev::lane_id();
  const uint32_t w = 1u + (blockIdx.x * 5 + threadIdx.x) % 15u;
  const uint64_t* q = words + (spread ? (uint64_t)w * 16 : (uint64_t)w);
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; ++i) {
    if (l == 0)
      acc += dev::fetch_add(words, 1);
    else
      acc += dev::ld(q);
  }
  if (acc == 0x5eed) atomicAdd(sink, 1ull);
}
__global__ __launch_bounds__(256) void mb_copy(const uint4* src, uint4* dst, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

namespace launch {

// which: 0 random CAS, 1 random add, 2 same-address add (one lane per wave),
// 3 random coherent load, 4 random cached load, 5 copy (n words -> n words),
// 6 coherent load pairs in one 128-byte line, 7 coherent load pairs in two
// lines, 8 / 9 load word + load cell + CAS word with the cell in the word's
// line / in another line, 10 / 11 adds to 16 words in one line / on 16 lines,
// 12 / 13 adds to one word while other lanes poll words of its line / of
// other lines.  Returns the rate in operations/s (bytes/s for the copy, read
// + write; pairs for 6 and 7, reads for 8 and 9, adds to the hot word for 12
// and 13).
hipError_t microbench(int which, uint64_t n_words, uint64_t threads_total, uint64_t iters, double* rate,
                      double* ms_out) {
  uint64_t* buf = nullptr;
  uint64_t* buf2 = nullptr;
  unsigned long long* sink = nullptr;
  hipError_t e = hipMalloc(&buf, n_words * 8);
  if (e != hipSuccess) return e;
  if (which == 5 && (e = hipMalloc(&buf2, n_words * 8)) != hipSuccess) {
    (void)hipFree(buf);
    return e;
  }
  (void)hipMalloc(&sink, 8);
  (void)hipMemset(buf, 0, n_words * 8);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const unsigned blocks = (unsigned)((threads_total + 255) / 256);
  auto run = [&]() {
    switch (which) {
      case 0: hipLaunchKernelGGL(mb_cas_random, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, sink); break;
      case 1: hipLaunchKernelGGL(mb_add_random, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, sink); break;
      case 2: hipLaunchKernelGGL(mb_add_same, dim3(blocks), dim3(256), 0, 0, buf, iters, sink); break;
      case 3: hipLaunchKernelGGL(mb_load_random, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 1, sink); break;
      case 4: hipLaunchKernelGGL(mb_load_random, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 0, sink); break;
      case 6: hipLaunchKernelGGL(mb_load_pairs, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 1, sink); break;
      case 7: hipLaunchKernelGGL(mb_load_pairs, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 0, sink); break;
      case 8: hipLaunchKernelGGL(mb_ts_read, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 1, sink); break;
      case 9: hipLaunchKernelGGL(mb_ts_read, dim3(blocks), dim3(256), 0, 0, buf, n_words, iters, 7ull, 0, sink); break;
      case 10: hipLaunchKernelGGL(mb_add_hot16, dim3(blocks), dim3(256), 0, 0, buf, iters, 0, sink); break;
      case 11: hipLaunchKernelGGL(mb_add_hot16, dim3(blocks), dim3(256), 0, 0, buf, iters, 1, sink); break;
      case 12: hipLaunchKernelGGL(mb_add_polled, dim3(blocks), dim3(256), 0, 0, buf, iters, 0, sink); break;
      case 13: hipLaunchKernelGGL(mb_add_polled, dim3(blocks), dim3(256), 0, 0, buf, iters, 1, sink); break;
      default: hipLaunchKernelGGL(mb_copy, dim3(2048), dim3(256), 0, 0, (const uint4*)buf, (uint4*)buf2, n_words / 2); break;
    }
  };
  run();  // warm up
  (void)hipEventRecord(a, 0);
  run();
  (void)hipEventRecord(b, 0);
  e = hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  if (e == hipSuccess) e = hipGetLastError();
  double ops;
  if (which == 5)
    ops = 2.0 * 8.0 * (double)n_words;
  else if (which == 2 || which == 12 || which == 13)
    ops = (double)((threads_total + 63) / 64) * (double)iters;
  else
    ops = (double)blocks * 256.0 * (double)iters;
  if (rate) *rate = ops / (ms / 1e3);
  if (ms_out) *ms_out = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(buf);
  if (buf2) (void)hipFree(buf2);
  (void)hipFree(sink);
  return e;
}

}  // namespace launch
}  // namespace gcctb
