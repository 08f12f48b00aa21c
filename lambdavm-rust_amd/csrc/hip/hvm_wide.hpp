// hvm_wide.hpp — the 64-bit-port evaluator (HVM_HIP_WIDE), host interface.
//
// The reference's ports are 32 bits (3-bit tag, 29-bit value), which caps a
// net at 2^29 nodes and 2^29 vars (src/hvm.cu:138-147; paper/HVM2.typst:
// 1212-1222 names 64-bit ports and 128-bit nodes as the way out, with no
// 128-bit atomics on GPUs).  This evaluator widens ports to 64 bits (tag in
// bits 0-2, value above) and nodes to two 64-bit halves, and indexes slots
// with 32-bit integers: up to 2^32 − 2 node and var slots (64 GiB of nodes,
// 32 GiB of vars).  Numbers keep the reference's 29-bit Numb in the value,
// so every result, ITRS and kind count equals run-c's.
//
// Node protocol without 128-bit atomics: a slot is free iff its first half
// is 0 (a live node's first port is never 0).  Only the lane that holds the
// redex pointing at a node takes it, so a take is two coherent loads; the
// slot is freed by zeroing the first half once both loads have returned.
// Allocation claims a free slot by seeing its first half 0 (in the wave's own
// slice) or by a 64-bit compare-and-swap on it (in the shared pool).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "hvm_dev.hpp"

struct HvmWideCtl {
  uint32_t n_in, n_out, in_head, err;
  uint32_t parity, max_steps, turns, done;
  uint32_t idle_waves, idle_quit, idle_base;
  uint32_t surplus;   // some wave of the current turn held more redexes than two per lane
  uint32_t pad[4];
};

struct HvmWideArgs {
  unsigned long long* node;      // 2 halves per slot, slot `cap` spare
  unsigned long long* vars;      // 1 per slot, slot 0 is ROOT
  uint32_t node_cap, vars_cap;
  uint32_t node_slice, vars_slice;
  uint32_t node_pool_lo, vars_pool_lo;
  const HvmDevDefX* defs;
  const unsigned long long* pairs;
  uint32_t ndefs;
  unsigned long long* bag[2];    // 2 words per redex
  uint32_t bag_cap;
  HvmWideCtl* ctl;
  uint32_t nwaves;
  HvmWaveState* wave_state;
  HvmDevStats* stats;
};

// Host side of one wide net (owned by a context created with HVM_HIP_WIDE).
struct HvmWide {
  HvmWideArgs a{};
  hipStream_t stream = nullptr;
  uint32_t turn_steps = 0, idle_pct = 25;
  std::string err;

  bool create(uint64_t node_cap, uint64_t vars_cap, uint32_t bag_cap, uint32_t waves, int cus,
              hipStream_t s);
  void destroy();
  // Boots @main, runs turns to the normal form (or until `deadline`, a
  // steady-clock time in seconds, 0 = none).  Fills the counters of `st`.
  // Returns 0 or an HVM_HIP_E_* code (err says why).
  int normalize(const HvmDevDefX* defs, const unsigned long long* pairs, uint32_t ndefs, double deadline,
                HvmDevStats& st, uint32_t& turns, uint32_t& launches);
  // Prints the net reachable from ROOT the way the reference does
  // (src/hvm.c:1618-1699), naming REFs by `names`.
  bool print_root(const std::vector<std::string>& names, std::string& out);

 private:
  double deadline = 0;
  uint32_t turns = 0, launches = 0;
  bool copy_failed = false;   // a readback copy failed (sticky until the next normalize)
  bool ok(hipError_t e, const char* what);
  bool boot(unsigned long long a_port, unsigned long long b_port);
  int run();
  unsigned long long vars_load(unsigned long long val);
  void node_load(unsigned long long loc, unsigned long long& fst, unsigned long long& snd);
  unsigned long long peek(unsigned long long port);
  int expand(unsigned long long port, unsigned long long& got);
};
