// Launch helper shared by the per-scheme translation units: one kernel
// instantiation per lanes-per-transaction G in {1, 2, 4, 8, 16, 32}.
#pragma once
#include "exec_core.h"
#include "launch.h"

namespace gcctb {
namespace dev {

template <class S>
hipError_t launch_exec_g(int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s) {
  if (nslots == 0) return hipSuccess;
  const uint64_t total = nslots * (uint64_t)G;
  const uint64_t blocks = (total + threads - 1) / threads;
  if (blocks == 0 || blocks > 0x7fffffffull || threads <= 0 || threads > 1024)
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block((unsigned)threads);
  switch (G) {
    case 1: hipLaunchKernelGGL((txn_exec<S, 1>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((txn_exec<S, 2>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((txn_exec<S, 4>), grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL((txn_exec<S, 8>), grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL((txn_exec<S, 16>), grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL((txn_exec<S, 32>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dev
}  // namespace gcctb
