// Launcher of a kernel variant (whole-row mode and / or the event log), for
// every scheme, plain form: the translation unit that includes this file
// defines GCCTB_ROW_OPS / GCCTB_EVENTS and the function name.
#include "exec_launch.h"
namespace gcctb {
namespace launch {
hipError_t GCCTB_VARIANT_FN(int scheme, int G, const ExecArgs& a, uint64_t nslots, int threads, hipStream_t s) {
  using namespace GCCTB_VARIANT_NS;
  switch (scheme) {
    case SCHEME_TPL_NW: return launch_exec_g<TwoPL<false>>(G, a, nslots, threads, false, s);
    case SCHEME_TPL_WD: return launch_exec_g<TwoPL<true>>(G, a, nslots, threads, false, s);
    case SCHEME_TO: return launch_exec_g<TimestampOrdering>(G, a, nslots, threads, false, s);
    case SCHEME_MVCC: return launch_exec_g<Mvcc>(G, a, nslots, threads, false, s);
    case SCHEME_SILO: return launch_exec_g<Silo>(G, a, nslots, threads, false, s);
    case SCHEME_TICTOC: return launch_exec_g<Tictoc>(G, a, nslots, threads, false, s);
    case SCHEME_GACCO:
    case SCHEME_GPUTX: return launch_exec_g<Ordered>(G, a, nslots, threads, false, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace launch
}  // namespace gcctb
