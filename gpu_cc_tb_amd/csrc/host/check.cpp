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
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <vector>

#include "../gcctb_common.h"
#include "host_internal.h"

namespace gcctb {

int check_history(uint64_t rows, uint32_t cols, uint64_t seed, bool row_ops, const uint64_t* init_table,
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
      const uint32_t j = (uint32_t)(i - tx_off[t]);
      if (row >= rows || col >= cols) continue;
      uint64_t* r = &table[(uint64_t)row * cols];
      if (row_ops) {
        // whole-row mode: a write stores every cell, a read saw the row's digest
        if (op_is_write(ops[i])) {
          for (uint32_t q = 0; q < cols; ++q) r[q] = write_cell(seed, t, j, q);
        } else {
          uint64_t h = DIGEST_INIT;
          for (uint32_t q = 0; q < cols; ++q) h = digest_step(h, r[q]);
          if (reads[i] != h) mismatches++;
        }
        continue;
      }
      uint64_t& cell = r[col];
      if (op_is_write(ops[i]))
        cell = write_value(seed, t, j);
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

// ---------------------------------------------------------------------------
// Conflict-graph check of an event history (ExecInfo::Check,
// src/runtime.cu:23-158): events in id order; a transaction is active on an
// object from its first access until its commit event.  A read adds an edge
// from the reader to every other transaction active on the object that wrote
// it; a write adds an edge to every other transaction active on the object
// (src/runtime.cu:23-71).  Transactions without a commit event are listed
// (check0, :73-78); a depth-first search over the graph reports the first
// cycle it closes, with the transactions on it (check / backtrack, :80-127).
// The search keeps an explicit stack instead of recursing, so long chains of
// edges cannot exhaust the host stack.
// ---------------------------------------------------------------------------
namespace gcctb {
namespace {

struct Report {
  char* buf;
  uint64_t cap;
  uint64_t len = 0;
  void add(const std::string& s) {
    if (!buf || cap == 0) return;
    for (char ch : s) {
      if (len + 1 >= cap) break;
      buf[len++] = ch;
    }
    buf[len] = 0;
  }
};

}  // namespace

int conflict_check(const gcctb_event* ev, uint64_t n, uint64_t ntx, uint64_t* uncommitted, uint64_t* cycle,
                   uint64_t cycle_cap, uint64_t* cycle_len, char* report, uint64_t report_len) {
  Report rep{report, report_len};
  if (report && report_len) report[0] = 0;
  std::vector<uint64_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return ev[a].id < ev[b].id; });
  std::vector<std::vector<uint64_t>> graph(ntx), opset(ntx);
  std::vector<char> committed(ntx, 0);
  std::map<uint64_t, std::set<uint64_t>> active;  // object -> transactions active on it
  for (uint64_t k = 0; k < n; ++k) {
    const gcctb_event& e = ev[idx[k]];
    if (e.tid < 0 || (uint64_t)e.tid >= ntx) return set_error(GCCTB_EINVAL, "event with a transaction id out of range");
    const uint64_t t = (uint64_t)e.tid;
    if (e.type == 2) {
      committed[t] = 1;
      for (uint64_t j : opset[t]) active[ev[j].oid].erase(t);
      continue;
    }
    opset[t].push_back(idx[k]);
    std::set<uint64_t>& act = active[e.oid];
    for (uint64_t u : act) {
      if (u == t) continue;
      bool edge = e.type != 0;  // a write conflicts with every active accessor
      if (!edge)
        for (uint64_t j : opset[u])
          if (ev[j].oid == e.oid && ev[j].type == 1) {
            edge = true;  // a read conflicts with an active writer
            break;
          }
      if (edge) graph[t].push_back(u);
    }
    act.insert(t);
  }
  for (auto& g : graph) {
    std::sort(g.begin(), g.end());
    g.erase(std::unique(g.begin(), g.end()), g.end());
  }
  rep.add("SCAN OK\n");
  uint64_t unc = 0;
  for (uint64_t t = 0; t < ntx; ++t)
    if (!committed[t]) {
      rep.add("UNCOMMITTED " + std::to_string(t) + "\n");
      unc++;
    }
  // depth-first search: done = finished without a cycle (visit1), on_path =
  // on the current path (visit2)
  std::vector<char> done(ntx, 0), on_path(ntx, 0);
  std::vector<std::pair<uint64_t, size_t>> path;  // (transaction, next edge)
  std::vector<uint64_t> cyc;
  for (uint64_t s = 0; s < ntx && cyc.empty(); ++s) {
    if (done[s]) continue;
    path.push_back({s, 0});
    on_path[s] = 1;
    while (!path.empty() && cyc.empty()) {
      auto& top = path.back();
      const uint64_t u = top.first;
      if (top.second < graph[u].size()) {
        const uint64_t v = graph[u][top.second++];
        if (done[v]) continue;
        if (on_path[v]) {
          // backtrack: the path from v to u closes the cycle
          size_t i = path.size();
          while (i > 0 && path[i - 1].first != v) --i;
          for (size_t q = path.size(); q-- > (i ? i - 1 : 0);) cyc.push_back(path[q].first);
          break;
        }
        path.push_back({v, 0});
        on_path[v] = 1;
      } else {
        on_path[u] = 0;
        done[u] = 1;
        path.pop_back();
      }
    }
    for (auto& p : path) on_path[p.first] = 0;
    path.clear();
  }
  if (!cyc.empty()) {
    rep.add("!!!!!!!!! " + std::to_string(cyc.back()) + "\n");
    std::vector<uint64_t> bad;
    for (uint64_t t : cyc) {
      rep.add("BAD TX " + std::to_string(t) + "\n");
      std::string nx;
      for (uint64_t v : graph[t]) nx += std::to_string(v) + " ";
      rep.add(nx + "\n");
      for (uint64_t j : opset[t]) bad.push_back(j);
    }
    std::sort(bad.begin(), bad.end(), [&](uint64_t a, uint64_t b) { return ev[a].id < ev[b].id; });
    rep.add("--------------------------BAD ST-----------------------\n");
    for (uint64_t j : bad)
      rep.add(std::to_string(ev[j].id) + " tid " + std::to_string(ev[j].tid) + " oid" + std::to_string(ev[j].oid) +
              " " + (ev[j].type == 0 ? "r" : "w") + "\n");
    rep.add("--------------------------BAD EN-----------------------\n");
  }
  rep.add(cyc.empty() ? "GOOD\n" : "BAD\n");
  if (uncommitted) *uncommitted = unc;
  if (cycle_len) *cycle_len = cyc.size();
  if (cycle)
    for (uint64_t i = 0; i < cyc.size() && i < cycle_cap; ++i) cycle[i] = cyc[i];
  return GCCTB_OK;
}

}  // namespace gcctb
