// Execution kernels for Silo (include/silo.cuh).
#include "exec_launch.h"
namespace gcctb {
namespace launch {
hipError_t exec_silo(int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s) {
  return dev::launch_exec_g<dev::Silo>(G, a, nslots, threads, persist, s);
}
}  // namespace launch
}  // namespace gcctb
