// TPC-C execution kernels (plain form): every op is a read-modify-write of
// one row of the TPC-C row space (tpcc_apply), all eight schemes
// (benchmark/tpcc.cu:36-380).
#define GCCTB_ROW_OPS 1
#define GCCTB_TPCC 1
#define GCCTB_VARIANT_NS dev_tpcc
#define GCCTB_VARIANT_FN exec_tpcc
#include "exec_variant.h"
