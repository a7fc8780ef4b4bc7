// Launch helper shared by the per-scheme translation units: one kernel
// instantiation per lanes-per-transaction G in {1, 2, 4, 8, 16, 32} and per
// form (plain per-batch / persistent).
#pragma once
#include "exec_core.h"
#include "launch.h"

namespace gcctb {
#if GCCTB_TPCC
namespace dev_tpcc {
#elif GCCTB_LATCH
namespace dev_latch {
#elif GCCTB_ROW_OPS && GCCTB_EVENTS
namespace dev_row_evt {
#elif GCCTB_ROW_OPS
namespace dev_row {
#elif GCCTB_EVENTS
namespace dev_evt {
#else
namespace dev {
#endif

template <class S, int G, bool P, int MAXT>
hipError_t launch_one(const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s) {
  if (threads > MAXT) return hipErrorInvalidValue;
  const uint64_t total = nslots * (uint64_t)G;
  uint64_t blocks = (total + threads - 1) / threads;
  if (P) {
    // a persistent grid holds no more blocks than the device keeps resident,
    // so that every group is running and the slot order guarantees progress
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&txn_exec<S, G, P, MAXT>),
                                                       threads, 0);
    if (e != hipSuccess) return e;
    uint64_t resident = (uint64_t)(per_cu > 0 ? per_cu : 1) * (uint64_t)(cus > 0 ? cus : 1);
    // The open form runs beside the admission copies and preparation kernels
    // that feed it, on another stream: it leaves half of the block slots free
    // for them (a full device would hold those kernels back until it ends).
    if (a.open && resident > 1) resident /= 2;
    if (blocks > resident) blocks = resident;
    // admission window: fewer transactions in flight under heavy contention
    if (a.persist_groups) {
      const uint64_t wblocks = (a.persist_groups * G + threads - 1) / threads;
      if (wblocks < blocks) blocks = wblocks < 1 ? 1 : wblocks;
    }
  }
  if (blocks == 0 || blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((txn_exec<S, G, P, MAXT>), dim3((unsigned)blocks), dim3((unsigned)threads), 0, s, a);
  return hipGetLastError();
}

template <class S>
hipError_t launch_exec_g(int G, const ExecArgs& args, uint64_t nslots, int threads, bool persist, hipStream_t s) {
  if (nslots == 0) return hipSuccess;
  if (threads <= 0 || threads > 1024) return hipErrorInvalidValue;
  ExecArgs a = args;
  a.persist = persist ? 1u : 0u;
  // persistent launches of at most 256 threads per block take the
  // instantiation bounded by that block size (more registers, no spills);
  // whole-row mode, the event log and the latch schemes have the plain
  // (per-batch) kernels only
#define GC_CASE(g)                                                                      \
  case g:                                                                               \
    if constexpr (kRowOps || kEvents || kLatch) {                                       \
      return persist ? hipErrorInvalidValue : launch_one<S, g, false, 1024>(a, nslots, threads, s); \
    } else {                                                                            \
      return persist ? (threads <= 256 ? launch_one<S, g, true, 256>(a, nslots, threads, s)    \
                                       : launch_one<S, g, true, 1024>(a, nslots, threads, s))  \
                     : launch_one<S, g, false, 1024>(a, nslots, threads, s);            \
    }
  switch (G) {
    GC_CASE(1)
    GC_CASE(2)
    GC_CASE(4)
    GC_CASE(8)
    GC_CASE(16)
    GC_CASE(32)
    default: return hipErrorInvalidValue;
  }
#undef GC_CASE
}

}  // namespace dev (dev_row, dev_evt, dev_row_evt, dev_latch, dev_tpcc)
}  // namespace gcctb
