// Whole-row mode with the event log (plain form).
#define GCCTB_ROW_OPS 1
#define GCCTB_EVENTS 1
#define GCCTB_VARIANT_NS dev_row_evt
#define GCCTB_VARIANT_FN exec_row_evt
#include "exec_variant.h"
