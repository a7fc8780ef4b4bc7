// Event-log execution kernels (plain form, the debug mode's AddEvent,
// include/runtime.cuh:64-81): every access and commit is stamped.
#define GCCTB_EVENTS 1
#define GCCTB_VARIANT_NS dev_evt
#define GCCTB_VARIANT_FN exec_evt
#include "exec_variant.h"
