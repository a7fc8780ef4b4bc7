// Execution kernels for Ordered (include/gacco.cuh and include/gputx.cuh).
#include "exec_launch.h"
namespace gcctb {
namespace launch {
hipError_t exec_ordered(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s) {
  return dev::launch_exec_g<dev::Ordered>(G, a, nslots, threads, persist, s);
}
}  // namespace launch
}  // namespace gcctb
