// Internal declarations of the host library.
#pragma once
#include <string>
#include "../../../include/gcctb.h"

namespace gcctb {

// Records the message for gcctb_last_error() and returns `code`.
int set_error(int code, const std::string& msg);
// Scheme id of a cc_scheme name, or -1.
int scheme_from_name(const char* s);

double zeta(uint64_t n, double theta, int threads);
int generate(const gcctb_workload& w, uint64_t* ops, uint64_t* tx_off, uint64_t* nops);
int read_txs_file(const char* path, uint64_t ntx, uint32_t cols, uint64_t seed, uint64_t* ops,
                  uint64_t* tx_off, uint64_t* nops);
int write_txs_file(const char* path, const uint64_t* ops, const uint64_t* tx_off, uint64_t ntx);

}  // namespace gcctb

namespace gcctb {
int check_history(uint64_t rows, uint32_t cols, uint64_t seed, bool row_ops, const uint64_t* init_table,
                  const uint64_t* final_table, const uint64_t* ops, const uint32_t* op_rows,
                  const uint64_t* tx_off, uint64_t ntx, const uint64_t* order, const uint64_t* reads,
                  uint64_t* bad);
}  // namespace gcctb

namespace gcctb {
int conflict_check(const gcctb_event* ev, uint64_t n, uint64_t ntx, uint64_t* uncommitted, uint64_t* cycle,
                   uint64_t cycle_cap, uint64_t* cycle_len, char* report, uint64_t report_len);
}  // namespace gcctb
