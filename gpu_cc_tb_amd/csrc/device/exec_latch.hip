// The latch schemes (the reference's slow_tpl_nw, slow_tpl_wd, slow_to,
// slow_mvcc, slow_silo, slow_tictoc; include/slow_*.cuh): the same protocols
// with every access to a row's CC word made under the row's latch.  Plain
// form, column mode.
#define GCCTB_LATCH 1
#include "exec_launch.h"
namespace gcctb {
namespace launch {
hipError_t exec_latch(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s) {
  using namespace dev_latch;
  switch (scheme) {
    case SCHEME_TPL_NW: return launch_exec_g<TwoPL<false>>(G, a, nslots, threads, false, s);
    case SCHEME_TPL_WD: return launch_exec_g<TwoPL<true>>(G, a, nslots, threads, false, s);
    case SCHEME_TO: return launch_exec_g<TimestampOrdering>(G, a, nslots, threads, false, s);
    case SCHEME_MVCC: return launch_exec_g<Mvcc>(G, a, nslots, threads, false, s);
    case SCHEME_SILO: return launch_exec_g<Silo>(G, a, nslots, threads, false, s);
    case SCHEME_TICTOC: return launch_exec_g<Tictoc>(G, a, nslots, threads, false, s);
    default: return hipErrorInvalidValue;  // GaccO and GPUTx have no latch form
  }
}
}  // namespace launch
}  // namespace gcctb
