// 2PL no-wait / wait-die execution kernels (include/twopl.cuh).
#include "exec_launch.h"
namespace gcctb {
namespace launch {
hipError_t exec_tpl(bool wait_die, int G, const ExecArgs& a, uint64_t nslots, int threads, bool persist, hipStream_t s) {
  return wait_die ? dev::launch_exec_g<dev::TwoPL<true>>(G, a, nslots, threads, persist, s)
                  : dev::launch_exec_g<dev::TwoPL<false>>(G, a, nslots, threads, persist, s);
}
}  // namespace launch
}  // namespace gcctb
