// Whole-row mode execution kernels (plain form): every op reads or writes all
// cells of its row, like the reference's 104-byte Item copies
// (benchmark/ycsb.cu:57,86-91).
#define GCCTB_ROW_OPS 1
#define GCCTB_VARIANT_NS dev_row
#define GCCTB_VARIANT_FN exec_row
#include "exec_variant.h"
