// The driver's result block: the CSV line appended to <dataset>.txt and the
// stdout block of ExecInfo::Analyse (include/runtime.cuh:169-214), with the
// same arithmetic types (the per-op tick sums scaled by a float
// 1 / (txcnt * clock_rate) whose denominator is an integer product, the
// timings as float milliseconds) and the same iostream formatting (default
// float notation, six significant digits), so the text matches byte for byte.
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include "host_internal.h"

namespace {

int copy_out(const std::string& s, char* dst, uint64_t cap, const char* what) {
  if (!dst) return GCCTB_OK;
  if (s.size() + 1 > cap) return gcctb::set_error(GCCTB_EINVAL, std::string(what) + " buffer too small");
  std::memcpy(dst, s.c_str(), s.size() + 1);
  return GCCTB_OK;
}

}  // namespace

extern "C" int gcctb_format_result(const gcctb_stats* st, const char* cc_type, int32_t valid_txn_bitoffset,
                                   int32_t warp_cnt, int32_t batch_size, uint64_t txcnt, char* csv,
                                   uint64_t csv_len, char* text, uint64_t text_len) {
  if (!st || !cc_type) return gcctb::set_error(GCCTB_EINVAL, "null argument");
  if (txcnt == 0 || !(st->tick_khz >= 1.0)) return gcctb::set_error(GCCTB_EINVAL, "txcnt and tick_khz must be positive");
  // clock_rate is an integer count of kHz in the reference (clock_t)
  const uint64_t clock_rate = (uint64_t)st->tick_khz;
  const float inv = 1.0f / ((float)(txcnt * clock_rate) * 1.0f);
  const float ts_d = st->ts_ticks * inv, wait_d = st->wait_ticks * inv, abort_d = st->abort_ticks * inv;
  const float mng_d = st->manager_ticks * inv, tot_d = st->tot_ticks * inv, idx_d = st->index_ticks * inv;
  const float pre = (float)st->pre_ms, proc = (float)st->proc_ms, tot = (float)st->tot_ms;
  const float through = 1000.0f / tot * (float)txcnt;
  std::ostringstream line;
  line << cc_type << "," << valid_txn_bitoffset << "," << warp_cnt << "," << batch_size << "," << through << ","
       << st->aborts << "," << ts_d << "," << wait_d << "," << abort_d << "," << mng_d << "," << tot_d << ","
       << idx_d << "," << pre / 1000.f << "," << proc / 1000.f << "," << tot / 1000.f << "\n";
  std::ostringstream out;
  out << cc_type << ","
      << "valid_txn_bitoffset " << valid_txn_bitoffset << ",\n"
      << "warp_cnt " << warp_cnt << ",\n"
      << "batch_size " << batch_size << ",\n"
      << "through " << through << ",\n"
      << "abort " << st->aborts << " diff " << ((1LL << 31) - 1) - ((long long)st->aborts + (long long)batch_size)
      << " avr " << st->aborts * 1.0 / batch_size << ",\n"
      << "ts_d " << ts_d << ",\n"
      << "wait_d " << wait_d << ",\n"
      << "abort_d " << abort_d << ",\n"
      << "mng_d " << mng_d << ",\n"
      << "tot_d " << tot_d << ",\n"
      << "idx_d " << idx_d << ",\n"
      << "pre " << pre / 1000.f << ",\n"
      << "proc " << proc / 1000.f << ",\n"
      << "tot " << tot / 1000.f << "\n";
  if (copy_out(line.str(), csv, csv_len, "csv") != GCCTB_OK) return GCCTB_EINVAL;
  return copy_out(out.str(), text, text_len, "text");
}
