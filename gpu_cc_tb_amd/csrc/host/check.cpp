// Serializability check behind the driver's debug flag.
//
// The reference records an event per read/write/commit, builds a conflict
// graph and searches it for a cycle (src/runtime.cu:23-158).  Here the kernels
// report, for every transaction, its place in the serialization order and, for
// every read, the value it returned.  A history is conflict-serializable in
// that order exactly when re-executing the transactions one by one in that
// order, from the table as it was before the run, returns the same value for
// every read and leaves the same table; this routine does that re-execution on
// the host and counts the reads and cells that disagree.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../gcctb_common.h"
#include "host_internal.h"

namespace gcctb {

int check_history(uint64_t rows, uint32_t cols, uint64_t seed, const uint64_t* init_table,
                  const uint64_t* final_table, const uint64_t* ops, const uint32_t* op_rows,
                  const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, const uint64_t* reads,
                  uint64_t* bad) {
  std::vector<uint64_t> idx(ntx);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
    for (int f = 0; f < 3; ++f)
      if (order[3 * a + f] != order[3 * b + f]) return order[3 * a + f] < order[3 * b + f];
    return a < b;
  });
  std::vector<uint64_t> table(init_table, init_table + rows * cols);
  uint64_t mismatches = 0;
  for (uint64_t t : idx) {
    for (uint64_t i = tx_off[t]; i < tx_off[t + 1]; ++i) {
      const uint32_t row = op_rows[i];
      const uint32_t col = op_col(ops[i]);
      if (row >= rows || col >= cols) continue;
      uint64_t& cell = table[(uint64_t)row * cols + col];
      if (op_is_write(ops[i]))
        cell = write_value(seed, t, (uint32_t)(i - tx_off[t]));
      else if (reads[i] != cell)
        mismatches++;
    }
  }
  for (uint64_t c = 0; c < rows * cols; ++c)
    if (table[c] != final_table[c]) mismatches++;
  *bad = mismatches;
  return GCCTB_OK;
}

}  // namespace gcctb
