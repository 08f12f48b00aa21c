// hvm_model.cpp — CPU model of the HIP turn kernel's concurrency protocol.
// TEST INFRASTRUCTURE ONLY: never linked into or loaded by the product.
//
// The HIP evaluator (lambdavm-rust_amd/csrc/hip/hvm_eval.hip, turn_body with
// the HBM net) runs one redex per lane per step in waves of lanes that move
// in lockstep, and many waves at once.  Its correctness rests on a protocol:
//   * zero means free: a wave hands out the zero slots of its own slice
//     without an atomic, and claims slots of the shared pool by
//     compare-and-swap;
//   * a node is taken (exchanged with zero) by the one lane holding the redex
//     that points at it, before or after the step's allocation;
//   * vars are linked with the reference's lock-free linker
//     (src/hvm.c:784-846), up to four links of one interaction in flight;
//   * redexes move through per-lane deques, in-wave sharing and the HBM bags.
// This model executes the same steps on the CPU, with the memory operations
// of all waves interleaved in a seeded random order (each wave is a fiber that
// yields to a random scheduler before every operation on the net), and
// records every event that the protocol says cannot happen: a held var whose
// slot reads zero (free), a node take that finds no node, a link that does
// not finish within the kernel's round cap, a slot handed out twice.
// Results (ITRS, per-kind counts, the normal form) are compared with the
// oracle by tests/test_model.py.
//
// Variants select the protocol choices under study (hvm_model_cfg.variant).
#include <ucontext.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../lambdavm-rust_amd/csrc/hvm_core.h"

typedef uint64_t u64;
typedef uint32_t u32;

extern "C" {

typedef struct hvm_model_cfg {
  uint32_t waves;        // waves per turn
  uint32_t lanes;        // lanes per wave (<= 64; also the allocator window)
  uint32_t deq;          // deque entries per lane (power of two)
  uint32_t tbl;          // per-wave allocation table entries (nodes and vars each)
  uint32_t node_cap;     // node slots (slot 0 reserved)
  uint32_t vars_cap;     // var slots
  uint32_t root_vars;    // var slots [0, root_vars) are roots
  uint32_t bag_cap;
  uint32_t turn_steps;   // steps per turn (0: 4096)
  uint32_t idle_pct;     // idle-quit share (0: 25; > 100: never)
  uint32_t seed;
  uint32_t switch_pct;   // % of operations at which the scheduler switches waves
  uint32_t variant;      // HVM_MODEL_* bits
  uint32_t link_cap;     // rounds before a link batch reports a loop (0: 2^16)
  uint64_t max_ops;      // abandon after this many net operations (0: none)
} hvm_model_cfg;

typedef struct hvm_model_out {
  uint64_t itrs;
  uint64_t kind[10];
  uint32_t err;          // HVM_ERR_* flags as the kernel raises them
  uint32_t turns;
  uint64_t zero_publish;   // LK_PUBLISH exchange returned 0 (held var found free)
  uint64_t zero_enter;     // LK_ENTER exchange returned 0
  uint64_t take_zero;      // a take found 0 (retry)
  uint64_t putback;        // nodes put back after a retry
  uint64_t clobber;        // a put-back overwrote a node created after the take
  uint64_t double_alloc;   // a node slot handed out while it holds a live node
  uint64_t max_rounds;     // longest link batch, in rounds
  uint64_t full_scans;     // allocations that moved on from the wave's slice to the shared pool
  uint64_t pool_claims;    // slots claimed in the shared pool
  uint64_t defers;         // lane-steps that waited for allocation-table space
  uint64_t ops;            // net operations performed
  uint64_t first_event_op; // op index of the first anomaly (0: none)
  uint32_t first_event;    // its kind: 1 zero publish, 2 zero enter, 3 take zero, 4 link cap, 5 clobber
  uint32_t first_err_turn; // turn of the first error flag
  uint64_t first_err_op;   // op index at which the first error flag was raised
  uint64_t pool_gives;     // redexes handed to the mid-turn pool (HVM_MODEL_POOL)
  uint64_t pool_takes;     // redexes taken from it by waves that ran dry
  char* result;            // printed normal form of each root, ';'-separated
} hvm_model_out;

// Variant bits
#define HVM_MODEL_XCHG_TWO 1u    // two-take rules (ANNI, COMM) take by exchange (else by load, zeroed at commit)
#define HVM_MODEL_LATE_TAKES 2u  // every take after the step's allocation (no early takes)
#define HVM_MODEL_OVERLAP_SCAN 4u // the allocator's scan covers len + one window (revisits its first window)
#define HVM_MODEL_TAKE_BEFORE_FIT 8u // every lane with a take rule takes (by exchange) before knowing
                                  // whether its step fits the tables; lanes that then wait put
                                  // their nodes back
#define HVM_MODEL_XCHG_ONE 16u   // single-take rules and a CALL's copy take by exchange (else by
                                  // load, zeroed at commit)
#define HVM_MODEL_NO_RECYCLE 32u // slots of nodes taken by load are zeroed at commit (else they go
                                  // on the wave's recycling stack, as HVM_NODE_RECYCLE does)
#define HVM_MODEL_RECYCLE_KEEP 64u // a planted mistake: slots handed out from the recycling stack
                                  // stay on it (for the test that the double-allocation check works)
#define HVM_MODEL_VRECYCLE_KEEP 128u // the same mistake on the var stack
#define HVM_MODEL_ENTER_LOAD 512u // the link loop's ENTER loads B (a build with HVM_ENTER_LOAD;
                                  // else it exchanges NONE into B, as by default)
#define HVM_MODEL_POOL 1024u     // the mid-turn pool (a build with HVM_MID_POOL): busy waves hand
                                 // surplus to the output bag at their idle-quit looks, a wave that
                                 // runs dry takes it back by exchanges with 0; input-bag reads clear
                                 // their slots, so an unwritten slot reads 0
#define HVM_MODEL_VRECYCLE 256u  // var slots freed by links go on the wave's var stack (a build
                                  // with HVM_VAR_RECYCLE > 0; off by default, as in the kernel)

int hvm_model_run(const uint32_t* book, size_t n_words, const uint32_t* entries, uint32_t n_nets,
                  const hvm_model_cfg* cfg, hvm_model_out* out);
void hvm_model_free(void* p);
}

namespace {

enum { ERR_OOM_NODE = 1, ERR_OOM_VARS = 2, ERR_BAG = 4, ERR_LOOP = 8, ERR_BAD_PORT = 32 };
enum { K_CALL = 0, K_COPY, K_VOID, K_ERAS, K_ANNI, K_COMM, K_OPER1, K_OPER2, K_SWIT0, K_SWITN, K_N };

struct Def {
  u32 safe = 0, vlen = 0, root = 0;
  std::vector<u64> rbag, node;
};

// ---------------------------------------------------------------------------
// Fibers: one per wave of a turn, switched by a seeded random scheduler.
// ---------------------------------------------------------------------------

struct Rng {
  u64 s;
  explicit Rng(u64 seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
  u64 next() {
    s += 0x9E3779B97F4A7C15ull;
    u64 z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  u32 below(u32 n) { return n ? (u32)(next() % n) : 0; }
};

struct Model;
struct Fiber {
  ucontext_t ctx;
  std::vector<char> stack;
  bool done = false;
  Model* m = nullptr;
  u32 w = 0;
};

static void fiber_main(unsigned lo, unsigned hi);

struct LaneSt {
  u32 bot = 0, top = 0;
  std::vector<u64> deq;
};

struct WaveState { u32 node_cur = 0, vars_cur = 0, node_ovf = 0, vars_ovf = 0; };

struct Model {
  hvm_model_cfg C;
  hvm_model_out* O;
  std::vector<Def> defs;
  std::vector<u64> node;
  std::vector<u32> vars;
  // slot generation: incremented whenever a slot is handed out, to detect
  // a put-back that overwrites a node created after the take
  std::vector<u32> node_gen;
  // a node slot holds a live node: set by a store of a node, cleared when the
  // node is taken (a zero store, an exchange, or a push on a recycling stack)
  std::vector<uint8_t> live;
  // a var slot is handed out and not yet freed
  std::vector<uint8_t> vlive;
  // per wave: var slots freed by links and not yet handed out again (the
  // kernel's 64-entry stack, scaled to the model's lanes)
  std::vector<std::vector<u32>> vrec;
  size_t vrec_cap = 0;
  bool vrecycle = false;
  std::vector<u64> bag[2];
  // control block
  u32 n_in = 0, n_out = 0, in_head = 0, err = 0, parity = 0, max_steps = 0, idle_waves = 0, idle_quit = 0;
  u32 idle_cfg = 0, surplus = 0, out_take = 0;
  u32 node_slice = 0, vars_slice = 0, node_pool_lo = 0, vars_pool_lo = 0, node_pool_hi = 0, vars_pool_hi = 0;
  std::vector<WaveState> ws;
  Rng rng;
  ucontext_t main_ctx;
  std::vector<Fiber> fib;
  Fiber* cur = nullptr;
  u64 kind[K_N] = {0};
  bool abandoned = false;
  // optional log of var-slot operations: (op, wave, what, slot, old, new)
  struct VEv { u64 op; u32 w, what, slot, old, nw; };
  std::vector<VEv> vlog;
  bool tracing = false;
  u32 bad_slot = 0xFFFFFFFFu;
  void vtrace(u32 what, u32 slot, u32 old, u32 nw) {
    if (tracing) vlog.push_back({O->ops, cur ? cur->w : 0xFFFFFFFFu, what, slot, old, nw});
  }

  explicit Model(const hvm_model_cfg& c, hvm_model_out* o) : C(c), O(o), rng(c.seed) {}

  // ---- scheduling ----
  void yield() {
    O->ops++;
    if (C.max_ops && O->ops > C.max_ops) abandoned = true;
    if (cur && rng.below(100) < C.switch_pct) swapcontext(&cur->ctx, &main_ctx);
  }
  void event(u32 k) { if (!O->first_event_op) { O->first_event_op = O->ops; O->first_event = k; } }
  void raise(u32 e) { if (!err && e) O->first_err_op = O->ops; err |= e; }

  // ---- net accessors (every one is a point where another wave may run) ----
  u32 nslot(u32 loc) const { return std::min(loc, C.node_cap); }
  u32 vslot(u32 val) const { return std::min(val == 0x1FFFFFFFu ? 0u : val, C.vars_cap); }
  u64 node_xchg(u32 loc, u64 v) { yield(); u64 o = node[nslot(loc)]; node[nslot(loc)] = v; live[nslot(loc)] = v != 0; return o; }
  u64 node_load(u32 loc) { yield(); return node[nslot(loc)]; }
  void node_store(u32 loc, u64 v) { yield(); node[nslot(loc)] = v; live[nslot(loc)] = v != 0; }
  void handed_out(u32 slot) {
    node_gen[slot]++;
    if (live[slot]) { O->double_alloc++; event(6); }
  }
  u32 vars_xchg(u32 val, u32 x) { yield(); u32 o = vars[vslot(val)]; vars[vslot(val)] = x; vtrace(1, vslot(val), o, x); return o; }
  u32 vars_load(u32 val) { yield(); return vars[vslot(val)]; }
  void vars_store(u32 val, u32 x) {
    yield();
    vtrace(2, vslot(val), vars[vslot(val)], x);
    vars[vslot(val)] = x;
    if (x == 0) vlive[vslot(val)] = 0;
  }
  void var_handed_out(u32 slot) {
    if (vlive[slot]) { O->double_alloc++; event(6); }
    vlive[slot] = 1;
  }
  // a var that a link owns is freed: onto the wave's stack, or zeroed
  void var_release(u32 val) {
    const u32 i = vslot(val);
    std::vector<u32>& r = vrec[cur->w];
    if (vrecycle && i >= C.root_vars && i < C.vars_cap && r.size() < vrec_cap) {
      vlive[i] = 0;
      r.push_back(i);
      return;
    }
    vars_store(val, 0);
  }
  template <typename T> bool cas(std::vector<T>& buf, u32 i, T expect, T v) {
    yield();
    if (buf[i] != expect) return false;
    if ((void*)&buf == (void*)&vars) vtrace(3, i, (u32)expect, (u32)v);
    buf[i] = v;
    return true;
  }

  std::vector<u32> order(u32 n) {
    std::vector<u32> o(n);
    for (u32 i = 0; i < n; ++i) o[i] = i;
    for (u32 i = n; i > 1; --i) std::swap(o[i - 1], o[rng.below(i)]);
    return o;
  }

  // ---- the turn body of wave w (hvm_eval.hip: turn_body<false, false>) ----
  void push_redex(LaneSt& ln, u64 r, u64& spills);
  void link_lanes(std::vector<std::vector<std::pair<u32, u32>>>& batches, std::vector<LaneSt>& lanes,
                  std::vector<u32>& voids, u64& spills);
  bool wave_alloc(u32 w, u32 n_lo, u32 v_lo, u32 need_n, u32 need_v, std::vector<u32>& tn, std::vector<u32>& tv,
                  bool pre, const std::vector<u64>& pre_n, const std::vector<u32>& pre_v, u32& oom,
                  std::vector<u32>& rec);
  void turn(u32 w);
  bool run_turn();
};

static void fiber_main(unsigned lo, unsigned hi) {
  Fiber* f = (Fiber*)(((uintptr_t)hi << 32) | (uintptr_t)lo);
  f->m->turn(f->w);
  f->done = true;
  swapcontext(&f->ctx, &f->m->main_ctx);
}

void Model::push_redex(LaneSt& ln, u64 r, u64& spills) {
  if (ln.top - ln.bot == C.deq) {
    u64 old = ln.deq[ln.bot & (C.deq - 1)];
    ln.bot++;
    u32 pos = n_out++;
    if (pos < C.bag_cap) bag[parity ^ 1u][pos] = old;
    else raise(ERR_BAG);
    spills++;
  }
  ln.deq[ln.top & (C.deq - 1)] = r;
  ln.top++;
}

// The linker of one step: every lane's pairs, in batches of up to 4 run as
// the kernel's link_batch (src/hvm.c:796-842 per pair), the lanes' rounds
// interleaved.  A batch ends with its new redexes pushed in pair order.
void Model::link_lanes(std::vector<std::vector<std::pair<u32, u32>>>& batches, std::vector<LaneSt>& lanes,
                       std::vector<u32>& voids, u64& spills) {
  enum { DONE = 0, ENTER = 1, PUBLISH = 2, ORIENT = 3 };
  const u32 L = (u32)lanes.size();
  const u32 cap = C.link_cap ? C.link_cap : (1u << 16);
  struct B { u32 pa[4], pb[4], st[4], n = 0, got = 0, rounds = 0; size_t next = 0; bool active = false; };
  std::vector<B> b(L);
  auto load_batch = [&](u32 l) {
    B& x = b[l];
    x.active = false;
    if (x.next >= batches[l].size()) return;
    x.n = std::min<size_t>(4, batches[l].size() - x.next);
    for (u32 i = 0; i < 4; ++i) x.st[i] = DONE;
    for (u32 i = 0; i < x.n; ++i) {
      x.pa[i] = batches[l][x.next + i].first;
      x.pb[i] = batches[l][x.next + i].second;
      x.st[i] = ORIENT;
    }
    x.next += x.n;
    x.got = 0;
    x.rounds = 0;
    x.active = true;
  };
  for (u32 l = 0; l < L; ++l) load_batch(l);
  for (;;) {
    bool any = false;
    for (u32 l = 0; l < L; ++l) any |= b[l].active;
    if (!any) break;
    // orient
    for (u32 l = 0; l < L; ++l) {
      B& x = b[l];
      if (!x.active) continue;
      for (u32 i = 0; i < x.n; ++i) {
        if (x.st[i] != ORIENT) continue;
        if (hvm_get_tag(x.pa[i]) != HVM_VAR && hvm_get_tag(x.pb[i]) == HVM_VAR) std::swap(x.pa[i], x.pb[i]);
        if (hvm_get_tag(x.pa[i]) != HVM_VAR) { x.st[i] = DONE; x.got |= 1u << i; }
        else x.st[i] = hvm_get_tag(x.pb[i]) == HVM_VAR ? (u32)ENTER : (u32)PUBLISH;
      }
    }
    // one exchange per unfinished pair, all lanes' in a random order
    std::vector<std::pair<u32, u32>> ops;
    for (u32 l = 0; l < L; ++l)
      if (b[l].active)
        for (u32 i = 0; i < b[l].n; ++i)
          if (b[l].st[i] == ENTER || b[l].st[i] == PUBLISH) ops.push_back({l, i});
    for (u32 k = (u32)ops.size(); k > 1; --k) std::swap(ops[k - 1], ops[rng.below(k)]);
    std::vector<u32> res(L * 4, 0);
    for (auto& op : ops) {
      B& x = b[op.first];
      u32 i = op.second;
      if (x.st[i] == ENTER)
        res[op.first * 4 + i] = (C.variant & HVM_MODEL_ENTER_LOAD) ? vars_load(hvm_get_val(x.pb[i]))
                                                                   : vars_xchg(hvm_get_val(x.pb[i]), HVM_NONE);
      else res[op.first * 4 + i] = vars_xchg(hvm_get_val(x.pa[i]), x.pb[i]);
    }
    // results, frees
    for (u32 l = 0; l < L; ++l) {
      B& x = b[l];
      if (!x.active) continue;
      bool busy = false;
      for (u32 i = 0; i < x.n; ++i) {
        u32 r = res[l * 4 + i];
        if (x.st[i] == ENTER) {
          if (r == 0) { O->zero_enter++; if (!O->first_event_op) bad_slot = vslot(hvm_get_val(x.pb[i])); event(2); }
          if (r == HVM_NONE || r == 0) {
            x.st[i] = PUBLISH;
          } else {
            var_release(hvm_get_val(x.pb[i]));
            x.pb[i] = r;
            x.st[i] = hvm_get_tag(r) == HVM_VAR ? (u32)ENTER : (u32)PUBLISH;
          }
        } else if (x.st[i] == PUBLISH) {
          if (r == 0) { O->zero_publish++; if (!O->first_event_op) bad_slot = vslot(hvm_get_val(x.pa[i])); event(1); }
          if (r == HVM_NONE) {
            x.st[i] = DONE;
          } else {
            var_release(hvm_get_val(x.pa[i]));
            x.pa[i] = r;
            x.st[i] = ORIENT;
          }
        }
        busy |= x.st[i] != DONE;
      }
      x.rounds++;
      if (busy && x.rounds >= cap) { raise(ERR_LOOP); event(4); busy = false; }
      if (!busy) {
        O->max_rounds = std::max<u64>(O->max_rounds, x.rounds);
        for (u32 i = 0; i < x.n; ++i) {
          if (!(x.got & (1u << i))) continue;
          if (hvm_get_tag(x.pa[i]) - 1u < 3u && hvm_get_tag(x.pb[i]) - 1u < 3u) voids[l]++;
          else push_redex(lanes[l], hvm_new_pair(x.pa[i], x.pb[i]), spills);
        }
        load_batch(l);
      }
    }
  }
}

// Wave-cooperative allocation (hvm_eval.hip: wave_alloc2): windows of
// `lanes` slots, own slice first (zero = free), then the shared pool (claim
// by compare-and-swap).
bool Model::wave_alloc(u32 w, u32 n_lo, u32 v_lo, u32 need_n, u32 need_v, std::vector<u32>& tn, std::vector<u32>& tv,
                       bool pre, const std::vector<u64>& pre_n, const std::vector<u32>& pre_v, u32& oom,
                       std::vector<u32>& rec) {
  struct Scan { u32 lo, len, cur, got, need, scanned; bool pool; u32 cursor, slice_next; };
  const u32 L = C.lanes;
  const bool fix_scan = !(C.variant & HVM_MODEL_OVERLAP_SCAN);
  const u32 slice_windows = (C.variant & HVM_MODEL_OVERLAP_SCAN) ? 0xFFFFFFu : 8u;
  WaveState& cs = ws[w];
  // the wave's recycled node slots first (the top of its stack)
  const u32 got_n = std::min<u32>(need_n, (u32)rec.size());
  for (u32 i = 0; i < got_n; ++i) {
    tn[i] = rec[rec.size() - got_n + i];
    handed_out(tn[i]);
  }
  if (!(C.variant & HVM_MODEL_RECYCLE_KEEP)) rec.resize(rec.size() - got_n);
  std::vector<u32>& vr = vrec[w];
  const u32 got_v = std::min<u32>(need_v, (u32)vr.size());
  for (u32 i = 0; i < got_v; ++i) {
    tv[i] = vr[vr.size() - got_v + i];
    var_handed_out(tv[i]);
  }
  if (!(C.variant & HVM_MODEL_VRECYCLE_KEEP)) vr.resize(vr.size() - got_v);
  Scan sn = {n_lo, node_slice, cs.node_cur, got_n, need_n, 0, false, cs.node_cur, cs.node_cur};
  Scan sv = {v_lo, vars_slice, cs.vars_cur, got_v, need_v, 0, false, cs.vars_cur, cs.vars_cur};
  auto idx = [&](Scan& s, u32 lane, u32& rel) { rel = (s.cur - s.lo + lane) % s.len; return s.lo + rel; };
  auto window = [&](auto& buf, Scan& s, std::vector<u32>& tbl, u32 reserved, auto claim, auto* vals) {
    std::vector<u32> slot(L), rel(L);
    std::vector<bool> cand(L);
    for (u32 l = 0; l < L; ++l) slot[l] = idx(s, l, rel[l]);
    // a window never reaches past the range's length from where the scan
    // began: slots handed out earlier in this call are still zero
    for (u32 l : order(L)) cand[l] = vals[l] == 0 && slot[l] >= reserved && (fix_scan ? s.scanned + l < s.len : true);
    if (s.pool)
      for (u32 l : order(L))
        if (cand[l]) cand[l] = cas(buf, slot[l], (typename std::decay_t<decltype(buf)>::value_type)0, (typename std::decay_t<decltype(buf)>::value_type)claim);
    u32 want = s.need - s.got, r = 0, nfree = 0;
    for (u32 l = 0; l < L; ++l) nfree += cand[l];
    for (u32 l = 0; l < L; ++l) {
      if (!cand[l]) continue;
      if (r < want) {
        tbl[s.got + r] = slot[l];
        if (s.pool) O->pool_claims++;
        if (&buf == (void*)&node) handed_out(slot[l]);
        else { vtrace(4, slot[l], (u32)vals[l], s.pool); var_handed_out(slot[l]); }
        if (r == want - 1) s.cursor = s.lo + ((rel[l] + 1) % s.len);
      } else if (s.pool) {
        yield();
        if ((void*)&buf == (void*)&vars) vtrace(5, slot[l], (u32)buf[slot[l]], 0);
        buf[slot[l]] = 0;   // claimed beyond the request: back to the pool
      }
      r++;
    }
    if (nfree >= want) {
      s.got = s.need;
    } else {
      s.got += nfree;
      s.cur = s.lo + ((s.cur - s.lo + L) % s.len);
      s.scanned += L;
    }
  };
  auto advance = [&](Scan& s, u32 pool_lo, u32 pool_len, u32 pool_cur) {
    // the slice: at most slice_windows windows per allocation (the kernel's
    // HVM_SLICE_SCAN), continued from there by the next one
    const u32 lim = s.pool ? s.len : std::min(s.len, slice_windows * L);
    if (s.got >= s.need || s.scanned < (fix_scan ? lim : s.len + L)) return true;
    if (s.pool || pool_len < L) return false;
    O->full_scans++;
    s.slice_next = s.cur;
    s.pool = true;
    s.lo = pool_lo;
    s.len = pool_len;
    s.cur = pool_cur;
    s.scanned = 0;
    return true;
  };
  while (sn.got < sn.need || sv.got < sv.need) {
    const bool pn = sn.got < sn.need, pv = sv.got < sv.need;
    std::vector<u64> vn(L, 1);
    std::vector<u32> vv(L, 1);
    std::vector<u32> in(L), iv(L), rr(L);
    for (u32 l = 0; l < L; ++l) { in[l] = idx(sn, l, rr[l]); iv[l] = idx(sv, l, rr[l]); }
    if (pre) {
      if (pn) vn = pre_n;
      if (pv) vv = pre_v;
    } else {
      for (u32 l : order(L)) {
        if (pn) vn[l] = node_load(in[l]);
        if (pv) vv[l] = vars_load(iv[l]);
      }
    }
    pre = false;
    if (pn) {
      window(node, sn, tn, 1u, ~0ull, vn.data());
      if (!advance(sn, node_pool_lo, node_pool_hi - node_pool_lo, cs.node_ovf)) { oom = ERR_OOM_NODE; return false; }
    }
    if (pv) {
      window(vars, sv, tv, C.root_vars, (u32)HVM_NONE, vv.data());
      if (!advance(sv, vars_pool_lo, vars_pool_hi - vars_pool_lo, cs.vars_ovf)) { oom = ERR_OOM_VARS; return false; }
    }
  }
  if (need_n) { if (sn.pool) { cs.node_ovf = sn.cursor; cs.node_cur = sn.slice_next; } else cs.node_cur = sn.cursor; }
  if (need_v) { if (sv.pool) { cs.vars_ovf = sv.cursor; cs.vars_cur = sv.slice_next; } else cs.vars_cur = sv.cursor; }
  return true;
}

static u32 adjust_port(u32 p, const u32* tn, const u32* tv) {
  if (hvm_is_nod(p)) return hvm_new_port(hvm_get_tag(p), tn[hvm_get_val(p)]);
  if (hvm_is_var(p)) return hvm_new_port(HVM_VAR, tv[hvm_get_val(p)]);
  return p;
}
static u64 adjust_pair(u64 p, const u32* tn, const u32* tv) {
  return hvm_new_pair(adjust_port(hvm_get_fst(p), tn, tv), adjust_port(hvm_get_snd(p), tn, tv));
}

void Model::turn(u32 w) {
  const u32 L = C.lanes, W = C.waves;
  const u32 TB = C.tbl;
  const bool early = !(C.variant & HVM_MODEL_LATE_TAKES);
  std::vector<LaneSt> lanes(L);
  for (auto& ln : lanes) ln.deq.assign(C.deq, 0);
  std::vector<u32> voids(L, 0);
  u64 spills = 0;
  const u32 n_lo = 1 + w * node_slice, v_lo = C.root_vars + w * vars_slice;
  WaveState& s = ws[w];
  if (s.node_cur < n_lo || s.node_cur >= n_lo + node_slice) s.node_cur = n_lo;
  if (s.vars_cur < v_lo || s.vars_cur >= v_lo + vars_slice) s.vars_cur = v_lo;
  if (s.node_ovf < node_pool_lo || s.node_ovf >= node_pool_hi)
    s.node_ovf = node_pool_lo + (u32)(((u64)(node_pool_hi - node_pool_lo) * w) / W);
  if (s.vars_ovf < vars_pool_lo || s.vars_ovf >= vars_pool_hi)
    s.vars_ovf = vars_pool_lo + (u32)(((u64)(vars_pool_hi - vars_pool_lo) * w) / W);
  const u32 nin = n_in;
  std::vector<u64>& bag_in = bag[parity];
  std::vector<u64>& bag_out = bag[parity ^ 1u];
  const bool pool = (C.variant & HVM_MODEL_POOL) != 0;
  const bool clear_in = pool;
  // an input-bag read (an exchange with 0 when the pool keeps the bags clear)
  auto take_in = [&](u32 i) { u64 x = bag_in[i]; if (clear_in) bag_in[i] = 0; return x; };
  // node slots taken by load and not yet handed out again (the kernel's
  // 4 * 64-entry stack, scaled to the model's lanes)
  const bool recycle = early && !(C.variant & HVM_MODEL_NO_RECYCLE);
  const size_t rec_cap = 4 * (size_t)L;
  std::vector<u32> rec;
  for (u32 l = 0; l < L; ++l) {
    u32 i = w + l * W;
    if (i < nin && i < L * W) {
      u64 x = take_in(i);
      if (!pool || x) push_redex(lanes[l], x, spills);
    }
  }
  bool more = nin > L * W;
  std::vector<u32> tbl_n(TB + 8), tbl_v(TB + 8);
  for (u32 step = 0;; ++step) {
    if (abandoned) break;
    if (step >= max_steps) break;
    if (more) {
      std::vector<u32> empty;
      for (u32 l = 0; l < L; ++l) if (lanes[l].top == lanes[l].bot) empty.push_back(l);
      if (!empty.empty()) {
        yield();
        u32 base = in_head;
        in_head += (u32)empty.size();
        if (base >= nin) more = false;
        for (u32 r = 0; r < empty.size(); ++r)
          if (base + r < nin) {
            u64 x = take_in(base + r);
            if (!pool || x) push_redex(lanes[empty[r]], x, spills);
          }
      }
    }
    {  // in-wave sharing: needy lanes take the oldest redex of rich lanes
      std::vector<u32> needy, rich;
      for (u32 l = 0; l < L; ++l) {
        u32 n = lanes[l].top - lanes[l].bot;
        if (n == 0) needy.push_back(l);
        else if (n >= 2) rich.push_back(l);
      }
      size_t k = std::min(needy.size(), rich.size());
      for (size_t i = 0; i < k; ++i) {
        LaneSt& src = lanes[rich[i]];
        u64 r = src.deq[src.bot & (C.deq - 1)];
        src.bot++;
        push_redex(lanes[needy[i]], r, spills);
      }
    }
    bool anyhas = false;
    for (auto& ln : lanes) anyhas |= ln.top != ln.bot;
    if (!anyhas) {
      if (more) continue;
      if (pool) {
        // before counting as idle: a wave's worth of handed-over redexes
        yield();
        const u32 out_n = std::min(n_out, C.bag_cap);
        if (out_n > out_take) {
          // claim only what waits (the kernel's compare-and-swap)
          const u32 t = out_take;
          const u32 k = std::min(L, out_n - t);
          out_take += k;
          yield();
          bool got = false;
          for (u32 r = 0; r < k; ++r) {
            const u32 i = t + r;
            const u64 x = bag_out[i];
            bag_out[i] = 0;                    // the exchange with 0
            if (x) { push_redex(lanes[r], x, spills); got = true; O->pool_takes++; }
          }
          if (got) continue;
        }
      }
      yield();
      idle_waves++;
      break;
    }
    if ((step & 15) == 15) {
      yield();
      // a steady turn (no idle quit) gets the configured idle quit back once
      // some wave holds more redexes than two per lane (the kernel's rule)
      u32 iq = idle_quit;
      if (!iq && idle_cfg) {
        u32 mine = 0;
        for (auto& ln : lanes) mine += ln.top - ln.bot;
        if (mine > 2 * L) surplus = 1;
        if (surplus) iq = idle_cfg;
      }
      if (iq && idle_waves >= iq) break;
      if (pool && n_out <= out_take + 4 * L && n_out < C.bag_cap / 2) {
        // hand the oldest redex of every lane holding three or more to the
        // pool; the stores land well after the reservation (other waves may
        // claim and read the slots meanwhile)
        std::vector<u64> give;
        for (auto& ln : lanes)
          if (ln.top - ln.bot >= 3) { give.push_back(ln.deq[ln.bot & (C.deq - 1)]); ln.bot++; }
        if (!give.empty()) {
          const u32 base = n_out;
          n_out += (u32)give.size();
          for (int k = 0; k < 4; ++k) yield();
          for (size_t i = 0; i < give.size(); ++i) {
            if (base + i < C.bag_cap) { bag_out[base + i] = give[i]; O->pool_gives++; }
            else raise(ERR_BAG);
          }
        }
      }
      if ((step & 63) == 63 && err) break;
    }
    // ---- P0: peek + classify ----
    struct St {
      bool has = false, run = false, copy = false, take_a = false, take_b = false, oper1 = false;
      u32 a = 0, b = 0, rule = 0, nn = 0, nv = 0, off_n = 0, off_v = 0, cv = 0, ga = 0, gb = 0;
      u64 NA = 0, NB = 0;
      const Def* d = nullptr;
    };
    std::vector<St> st(L);
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      LaneSt& ln = lanes[l];
      x.has = ln.top != ln.bot;
      if (!x.has) continue;
      u64 rx = ln.deq[(ln.top - 1) & (C.deq - 1)];
      u32 a = hvm_get_fst(rx), b = hvm_get_snd(rx);
      u32 rule = hvm_get_rule(a, b);
      if (hvm_get_tag(a) == HVM_REF && hvm_get_tag(b) == HVM_VAR &&
          (b == HVM_ROOT || (hvm_get_val(b) != 0 && hvm_get_val(b) < C.root_vars))) rule = HVM_CALL;
      else if (hvm_should_swap(a, b)) std::swap(a, b);
      x.a = a; x.b = b; x.rule = rule;
      x.take_a = rule == HVM_ANNI || rule == HVM_COMM;
      x.take_b = rule == HVM_ERAS || rule == HVM_ANNI || rule == HVM_COMM || rule == HVM_OPER || rule == HVM_SWIT;
      if (rule == HVM_COMM) { x.nn = 4; x.nv = 4; }
      else if (rule == HVM_OPER) x.nn = 1;
      else if (rule == HVM_SWIT) x.nn = hvm_get_u24(hvm_get_val(a)) == 0 ? 1 : 2;
      if (rule == HVM_CALL) {
        u32 fid = hvm_get_val(a) & 0xFFFFFFFu;
        x.d = &defs[fid < defs.size() ? fid : 0];
      }
    }
    // reservation of the lanes whose rule is not CALL
    u32 on = 0, ov = 0;
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      u32 nn = x.rule != HVM_CALL ? x.nn : 0, nv = x.rule != HVM_CALL ? x.nv : 0;
      x.off_n = on; x.off_v = ov;
      on += x.has ? nn : 0; ov += x.has ? nv : 0;
      x.run = x.has && x.rule != HVM_CALL && x.off_n + x.nn <= TB && x.off_v + x.nv <= TB;
    }
    const bool before_fit = (C.variant & HVM_MODEL_TAKE_BEFORE_FIT) != 0;
    auto two = [&](const St& x) { return x.take_a && x.take_b; };
    // early takes read their nodes with loads (zeroed at commit) unless a
    // variant takes them by exchange
    auto peek = [&](const St& x) {
      return early && !before_fit && (C.variant & (two(x) ? HVM_MODEL_XCHG_TWO : HVM_MODEL_XCHG_ONE)) == 0;
    };
    // the round trip of the early takes and the allocator's first windows
    std::vector<u64> pre_n(L, 1);
    std::vector<u32> pre_v(L, 1);
    {
      std::vector<u32> ops;   // 0..L-1: take a, L..2L-1: take b, 2L..3L-1: node window, 3L..: var window
      for (u32 l = 0; l < L; ++l) {
        const bool tk = early && (st[l].run || (before_fit && st[l].has && st[l].rule != HVM_CALL));
        if (tk && st[l].take_a) ops.push_back(l);
        if (tk && st[l].take_b) ops.push_back(L + l);
        ops.push_back(2 * L + l);
        ops.push_back(3 * L + l);
      }
      for (u32 k = (u32)ops.size(); k > 1; --k) std::swap(ops[k - 1], ops[rng.below(k)]);
      for (u32 op : ops) {
        u32 l = op % L, kind_ = op / L;
        St& x = st[l];
        if (kind_ == 0) { x.ga = node_gen[nslot(hvm_get_val(x.a))]; x.NA = peek(x) ? node_load(hvm_get_val(x.a)) : node_xchg(hvm_get_val(x.a), 0); }
        if (kind_ == 1) { x.gb = node_gen[nslot(hvm_get_val(x.b))]; x.NB = peek(x) ? node_load(hvm_get_val(x.b)) : node_xchg(hvm_get_val(x.b), 0); }
        if (kind_ == 2) { u32 rel = (s.node_cur - n_lo + l) % node_slice; pre_n[l] = node_load(n_lo + rel); }
        if (kind_ == 3) { u32 rel = (s.vars_cur - v_lo + l) % vars_slice; pre_v[l] = vars_load(v_lo + rel); }
      }
    }
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      if (x.has && x.rule == HVM_CALL) {
        if ((x.d->safe & 1u) && hvm_get_tag(x.b) == HVM_DUP) x.copy = true;
        else { x.nn = (u32)x.d->node.size(); x.nv = x.d->vlen; }
      }
    }
    // ---- P1: CALL lanes reserve behind the others ----
    {
      u32 base_n = 0, base_v = 0;
      for (u32 l = 0; l < L; ++l)
        if (st[l].run) { base_n = std::max(base_n, st[l].off_n + st[l].nn); base_v = std::max(base_v, st[l].off_v + st[l].nv); }
      u32 cn = 0, cv = 0;
      for (u32 l = 0; l < L; ++l) {
        St& x = st[l];
        if (x.rule != HVM_CALL) continue;
        u32 nn = x.has ? x.nn : 0, nv = x.has ? x.nv : 0;
        if (x.has) {
          x.off_n = base_n + cn; x.off_v = base_v + cv;
          x.run = x.off_n + x.nn <= TB && x.off_v + x.nv <= TB;
        }
        cn += nn; cv += nv;
      }
    }
    u32 need_n = 0, need_v = 0;
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      if (x.run) { need_n = std::max(need_n, x.off_n + x.nn); need_v = std::max(need_v, x.off_v + x.nv); }
      if (x.run) lanes[l].top--;
      if (x.has && !x.run) O->defers++;
    }
    // lanes that took before knowing that they fit, and wait: nodes go back
    if (before_fit)
      for (u32 l : order(L)) {
        St& x = st[l];
        if (x.run || !x.has || x.rule == HVM_CALL) continue;
        if (x.take_a && x.NA) { O->putback++; if (node_gen[nslot(hvm_get_val(x.a))] != x.ga) { O->clobber++; event(5); } node_store(hvm_get_val(x.a), x.NA); }
        if (x.take_b && x.NB) { O->putback++; if (node_gen[nslot(hvm_get_val(x.b))] != x.gb) { O->clobber++; event(5); } node_store(hvm_get_val(x.b), x.NB); }
        x.NA = x.NB = 0;
      }
    // ---- late takes: a safe def's copy (all takes without early takes) ----
    {
      std::vector<u32> ops;
      for (u32 l = 0; l < L; ++l) {
        St& x = st[l];
        if (!x.run) continue;
        if (x.rule == HVM_CALL && x.copy) { x.take_b = true; ops.push_back(L + l); }
        if (!early && x.take_a) ops.push_back(l);
        if (!early && x.take_b && x.rule != HVM_CALL) ops.push_back(L + l);
      }
      if (early) {
        for (u32 k = (u32)ops.size(); k > 1; --k) std::swap(ops[k - 1], ops[rng.below(k)]);
        for (u32 op : ops) {
          St& x = st[op % L];
          x.gb = node_gen[nslot(hvm_get_val(x.b))];
          x.NB = peek(x) ? node_load(hvm_get_val(x.b)) : node_xchg(hvm_get_val(x.b), 0);
        }
        ops.clear();
      }
      // ---- P2: allocation ----
      if (need_n || need_v) {
        u32 oom = 0;
        if (!wave_alloc(w, n_lo, v_lo, need_n, need_v, tbl_n, tbl_v, early, pre_n, pre_v, oom, rec)) {
          raise(oom);
          for (u32 l = 0; l < L; ++l) if (st[l].run) lanes[l].top++;
          break;
        }
      }
      if (!early) {
        for (u32 k = (u32)ops.size(); k > 1; --k) std::swap(ops[k - 1], ops[rng.below(k)]);
        for (u32 op : ops) {
          St& x = st[op % L];
          if (op < L) { x.ga = node_gen[nslot(hvm_get_val(x.a))]; x.NA = node_xchg(hvm_get_val(x.a), 0); }
          else { x.gb = node_gen[nslot(hvm_get_val(x.b))]; x.NB = node_xchg(hvm_get_val(x.b), 0); }
        }
      }
    }
    // retries: a node this redex points to is not there
    for (u32 l : order(L)) {
      St& x = st[l];
      if (!x.run) continue;
      const u32* tn = tbl_n.data() + x.off_n;
      const u32* tv = tbl_v.data() + x.off_v;
      if ((x.take_a && x.NA == 0) || (x.take_b && x.NB == 0)) {
        O->take_zero++;
        event(3);
        const bool peeked = peek(x);
        if (!peeked && x.take_a && x.NA) {
          O->putback++;
          if (node_gen[nslot(hvm_get_val(x.a))] != x.ga) { O->clobber++; event(5); }
          node_store(hvm_get_val(x.a), x.NA);
        }
        if (!peeked && x.take_b && x.NB) {
          O->putback++;
          if (node_gen[nslot(hvm_get_val(x.b))] != x.gb) { O->clobber++; event(5); }
          node_store(hvm_get_val(x.b), x.NB);
        }
        for (u32 i = 0; i < x.nn; ++i) node_store(tn[i], 0);
        for (u32 i = 0; i < x.nv; ++i) vars_store(tv[i], 0);
        lanes[l].top++;
        x.run = false;
      }
    }
    // peeked nodes are taken now: their slots go on the wave's recycling
    // stack in lane order (zeroed once it is full), or are zeroed
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      if (!(x.run && peek(x))) continue;
      for (u32 side = 0; side < 2; ++side) {
        if (!(side ? x.take_b : x.take_a)) continue;
        const u32 loc = hvm_get_val(side ? x.b : x.a);
        if (recycle && loc < C.node_cap && rec.size() < rec_cap) { live[loc] = 0; rec.push_back(loc); }
        else node_store(loc, 0);
      }
    }
    // ---- P4: creation ----
    for (u32 l : order(L)) {
      St& x = st[l];
      if (!x.run) continue;
      const u32* tn = tbl_n.data() + x.off_n;
      const u32* tv = tbl_v.data() + x.off_v;
      u32 B1 = hvm_get_fst(x.NB), B2 = hvm_get_snd(x.NB);
      if (x.rule == HVM_OPER) {
        x.oper1 = hvm_get_tag(B1) == HVM_NUM;
        if (x.oper1) { x.cv = hvm_operate(hvm_get_val(x.a), hvm_get_val(B1)); node_store(tn[0], 0); }
      }
      if (x.rule == HVM_CALL && !x.copy) {
        for (size_t i = 0; i < x.d->node.size(); ++i) node_store(tn[i], adjust_pair(x.d->node[i], tn, tv));
        for (u32 i = 0; i < x.d->vlen; ++i) vars_store(tv[i], HVM_NONE);
      } else if (x.rule == HVM_COMM) {
        for (u32 i = 0; i < 4; ++i) vars_store(tv[i], HVM_NONE);
        node_store(tn[0], hvm_new_pair(hvm_new_port(HVM_VAR, tv[0]), hvm_new_port(HVM_VAR, tv[1])));
        node_store(tn[1], hvm_new_pair(hvm_new_port(HVM_VAR, tv[2]), hvm_new_port(HVM_VAR, tv[3])));
        node_store(tn[2], hvm_new_pair(hvm_new_port(HVM_VAR, tv[0]), hvm_new_port(HVM_VAR, tv[2])));
        node_store(tn[3], hvm_new_pair(hvm_new_port(HVM_VAR, tv[1]), hvm_new_port(HVM_VAR, tv[3])));
      } else if (x.rule == HVM_OPER && !x.oper1) {
        node_store(tn[0], hvm_new_pair(x.a, B2));
      } else if (x.rule == HVM_SWIT) {
        u32 av = hvm_get_u24(hvm_get_val(x.a));
        if (av == 0) node_store(tn[0], hvm_new_pair(B2, hvm_new_port(HVM_ERA, 0)));
        else {
          node_store(tn[0], hvm_new_pair(hvm_new_port(HVM_ERA, 0), hvm_new_port(HVM_CON, tn[1])));
          node_store(tn[1], hvm_new_pair(hvm_new_port(HVM_NUM, hvm_new_u24(av - 1)), B2));
        }
      }
    }
    // ---- P6: links ----
    std::vector<std::vector<std::pair<u32, u32>>> batches(L);
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      if (!x.run) continue;
      const u32* tn = tbl_n.data() + x.off_n;
      const u32* tv = tbl_v.data() + x.off_v;
      auto& bt = batches[l];
      u32 B1 = hvm_get_fst(x.NB), B2 = hvm_get_snd(x.NB);
      switch (x.rule) {
        case HVM_LINK: bt.push_back({x.a, x.b}); break;
        case HVM_CALL:
          if (x.copy) { bt.push_back({x.a, B1}); bt.push_back({x.a, B2}); }
          else {
            for (u64 p : x.d->rbag) { u64 q = adjust_pair(p, tn, tv); bt.push_back({hvm_get_fst(q), hvm_get_snd(q)}); }
            bt.push_back({adjust_port(x.d->root, tn, tv), x.b});
          }
          break;
        case HVM_ERAS: bt.push_back({x.a, B1}); bt.push_back({x.a, B2}); break;
        case HVM_ANNI: bt.push_back({hvm_get_fst(x.NA), B1}); bt.push_back({hvm_get_snd(x.NA), B2}); break;
        case HVM_COMM:
          bt.push_back({hvm_new_port(hvm_get_tag(x.b), tn[0]), hvm_get_fst(x.NA)});
          bt.push_back({hvm_new_port(hvm_get_tag(x.b), tn[1]), hvm_get_snd(x.NA)});
          bt.push_back({hvm_new_port(hvm_get_tag(x.a), tn[2]), B1});
          bt.push_back({hvm_new_port(hvm_get_tag(x.a), tn[3]), B2});
          break;
        case HVM_OPER:
          if (x.oper1) bt.push_back({hvm_new_port(HVM_NUM, x.cv), B2});
          else bt.push_back({B1, hvm_new_port(HVM_OPR, tn[0])});
          break;
        case HVM_SWIT: bt.push_back({hvm_new_port(HVM_CON, tn[0]), B1}); break;
        default: break;
      }
    }
    link_lanes(batches, lanes, voids, spills);
    // ---- P7: counting ----
    for (u32 l = 0; l < L; ++l) {
      St& x = st[l];
      if (!x.run) continue;
      switch (x.rule) {
        case HVM_CALL: kind[x.copy ? K_COPY : K_CALL]++; break;
        case HVM_VOID: kind[K_VOID]++; break;
        case HVM_ERAS: kind[K_ERAS]++; break;
        case HVM_ANNI: kind[K_ANNI]++; break;
        case HVM_COMM: kind[K_COMM]++; break;
        case HVM_OPER: kind[x.oper1 ? K_OPER1 : K_OPER2]++; break;
        case HVM_SWIT: kind[hvm_get_u24(hvm_get_val(x.a)) == 0 ? K_SWIT0 : K_SWITN]++; break;
        default: break;
      }
    }
  }
  // ---- end of turn: recycled slots not handed out are freed ----
  for (u32 slot : rec) node_store(slot, 0);
  rec.clear();
  for (u32 slot : vrec[w]) vars_store(slot, 0);
  vrec[w].clear();
  // ---- end of turn: deques to the output bag, unpulled input carried ----
  for (auto& ln : lanes)
    for (u32 i = ln.bot; i != ln.top; ++i) {
      u32 pos = n_out++;
      if (pos < C.bag_cap) bag[parity ^ 1u][pos] = ln.deq[i & (C.deq - 1)];
      else raise(ERR_BAG);
    }
  while (more) {
    yield();
    u32 base = in_head;
    in_head += L;
    if (base >= nin) break;
    for (u32 i = base; i < std::min(nin, base + L); ++i) {
      u32 pos = n_out++;
      if (pos < C.bag_cap) bag[parity ^ 1u][pos] = take_in(i);
      else raise(ERR_BAG);
    }
  }
  for (u32 v : voids) kind[K_VOID] += v;
  (void)spills;
}

// Runs one turn: every wave a fiber, interleaved by the scheduler.
bool Model::run_turn() {
  const u32 W = C.waves;
  fib.assign(W, Fiber());
  for (u32 w = 0; w < W; ++w) {
    Fiber& f = fib[w];
    f.m = this;
    f.w = w;
    f.stack.assign(1 << 18, 0);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack.data();
    f.ctx.uc_stack.ss_size = f.stack.size();
    f.ctx.uc_link = nullptr;
    uintptr_t p = (uintptr_t)&f;
    makecontext(&f.ctx, (void (*)())fiber_main, 2, (unsigned)(p & 0xFFFFFFFFu), (unsigned)(p >> 32));
  }
  u32 alive = W;
  while (alive) {
    u32 k = rng.below(alive);
    u32 pick = 0;
    for (u32 w = 0; w < W; ++w) {
      if (fib[w].done) continue;
      if (k-- == 0) { pick = w; break; }
    }
    cur = &fib[pick];
    swapcontext(&main_ctx, &cur->ctx);
    cur = nullptr;
    if (fib[pick].done) alive--;
  }
  return true;
}

// Prints the tree reachable from a port (the printer of the HIP runtime).
static void print_port(const Model& m, u32 p, std::string& out, int depth) {
  if (depth > 4096) { out += "<deep>"; return; }
  for (int it = 0; it < 1 << 20 && hvm_get_tag(p) == HVM_VAR; ++it) {
    u32 v = m.vars[m.vslot(hvm_get_val(p))];
    if (v == HVM_NONE || v == 0) break;
    p = v;
  }
  switch (hvm_get_tag(p)) {
    case HVM_VAR: { char b[16]; snprintf(b, sizeof b, "x%x", hvm_get_val(p)); out += b; break; }
    case HVM_REF: out += "@" + std::to_string(hvm_get_val(p) & 0xFFFFFFFu); break;
    case HVM_ERA: out += "*"; break;
    case HVM_NUM: out += std::to_string(hvm_get_u24(hvm_get_val(p))); break;
    default: {
      u64 nd = m.node[m.nslot(hvm_get_val(p))];
      u32 t = hvm_get_tag(p);
      out += t == HVM_CON ? "(" : t == HVM_DUP ? "{" : t == HVM_OPR ? "$(" : "?(";
      print_port(m, hvm_get_fst(nd), out, depth + 1);
      out += " ";
      print_port(m, hvm_get_snd(nd), out, depth + 1);
      out += t == HVM_DUP ? "}" : ")";
    }
  }
}

static bool load_book(const u32* buf, size_t n, std::vector<Def>& defs) {
  size_t i = 0;
  if (n < 1) return false;
  u32 nd = buf[i++];
  defs.assign(nd, Def());
  for (u32 k = 0; k < nd; ++k) {
    if (i + 70 > n) return false;
    u32 fid = buf[i++];
    if (fid >= nd) return false;
    Def& d = defs[fid];
    i += 64;
    d.safe = buf[i++] != 0;
    u32 rlen = buf[i++], nlen = buf[i++];
    d.vlen = buf[i++];
    d.root = buf[i++];
    if (i + 2ull * (rlen + nlen) > n) return false;
    d.rbag.resize(rlen);
    d.node.resize(nlen);
    memcpy(d.rbag.data(), buf + i, 8ull * rlen); i += 2ull * rlen;
    memcpy(d.node.data(), buf + i, 8ull * nlen); i += 2ull * nlen;
  }
  return true;
}

}  // namespace

extern "C" int hvm_model_run(const uint32_t* book, size_t n_words, const uint32_t* entries, uint32_t n_nets,
                             const hvm_model_cfg* cfg, hvm_model_out* out) {
  if (!book || !cfg || !out || n_nets == 0) return 1;
  memset(out, 0, sizeof *out);
  Model m(*cfg, out);
  m.tracing = getenv("HVM_MODEL_TRACE") != nullptr;
  hvm_model_cfg& C = m.C;
  if (!C.waves || !C.lanes || C.lanes > 64 || !C.deq || (C.deq & (C.deq - 1))) return 1;
  if (!C.tbl) C.tbl = 256;
  if (!C.bag_cap) C.bag_cap = 1u << 22;
  if (!C.root_vars) C.root_vars = 1 + n_nets;
  if (C.root_vars < 1 + n_nets) return 1;
  if (!load_book(book, n_words, m.defs) || m.defs.empty()) return 2;
  for (u32 j = 0; j < n_nets; ++j) if (entries[j] >= m.defs.size()) return 2;
  for (const Def& d : m.defs) if (d.node.size() > C.tbl || d.vlen > C.tbl) return 2;
  m.node.assign((size_t)C.node_cap + 1, 0);
  m.vars.assign((size_t)C.vars_cap + 1, 0);
  m.node_gen.assign((size_t)C.node_cap + 1, 0);
  m.live.assign((size_t)C.node_cap + 1, 0);
  m.vlive.assign((size_t)C.vars_cap + 1, 0);
  m.vrec.assign(C.waves, std::vector<u32>());
  m.vrec_cap = (size_t)C.lanes;
  m.vrecycle = (C.variant & HVM_MODEL_VRECYCLE) && !(C.variant & (HVM_MODEL_LATE_TAKES | HVM_MODEL_NO_RECYCLE));
  m.bag[0].assign(C.bag_cap, 0);
  m.bag[1].assign(C.bag_cap, 0);
  m.ws.assign(C.waves, WaveState());
  // layout as the runtime's (hvm_runtime.hip: init)
  m.node_slice = (u32)(((u64)(C.node_cap - 1) * 3 / 4) / C.waves);
  m.vars_slice = (u32)(((u64)(C.vars_cap - C.root_vars) * 3 / 4) / C.waves);
  m.node_pool_lo = 1 + C.waves * m.node_slice;
  m.vars_pool_lo = C.root_vars + C.waves * m.vars_slice;
  m.node_pool_hi = C.node_cap - std::min<u32>(1u << 16, (C.node_cap - m.node_pool_lo) / 2);
  m.vars_pool_hi = C.vars_cap - std::min<u32>(1u << 16, (C.vars_cap - m.vars_pool_lo) / 2);
  if (m.node_slice < 4 * C.lanes || m.vars_slice < 4 * C.lanes) return 1;
  // boot: one redex per net, root vars NONE
  m.vars[0] = n_nets == 1 ? HVM_NONE : 0;
  for (u32 j = 0; j < n_nets; ++j) {
    u32 root = n_nets == 1 ? HVM_ROOT : hvm_new_port(HVM_VAR, 1 + j);
    if (n_nets > 1) m.vars[1 + j] = HVM_NONE;
    m.bag[0][j] = hvm_new_pair(hvm_new_port(HVM_REF, entries[j]), root);
  }
  m.n_in = n_nets;
  m.in_head = std::min<u32>(n_nets, C.waves * C.lanes);
  const u32 pct = C.idle_pct ? (C.idle_pct > 100 ? 0u : C.idle_pct) : 25u;
  const u32 idle_cfg = pct ? std::max<u32>(1u, (C.waves * pct) / 100) : 0u;
  m.max_steps = C.turn_steps ? C.turn_steps : 4096;
  m.idle_quit = idle_cfg;
  m.idle_cfg = idle_cfg;
  for (u32 t = 0; t < (1u << 20); ++t) {
    m.n_out = 0;
    m.out_take = 0;
    m.idle_waves = 0;
    m.surplus = 0;
    m.run_turn();
    out->turns++;
    if (m.abandoned || m.err) break;
    // between turns (hvm_eval.hip: hvm_between_kernel, plain scheduler)
    u32 n = m.n_out;
    if (n > C.bag_cap) { m.err |= ERR_BAG; break; }
    const bool flat = 2 * n <= 3 * m.n_in;
    bool steady = n <= C.waves / 4 && flat;
    m.n_in = n;
    m.in_head = std::min(n, C.waves * C.lanes);
    m.parity ^= 1u;
    if (n == 0) break;
    m.idle_quit = steady ? 0u : idle_cfg;
    // the flat-frontier rule: the waves the spread leaves empty count as
    // idle at once, the share applies to the n waves that got a redex
    if (!steady && flat && n < C.waves && idle_cfg)
      m.idle_quit = (C.waves - n) + std::max<u32>(1u, (u32)(((u64)idle_cfg * n) / C.waves));
  }
  if (getenv("HVM_MODEL_TRACE") && m.bad_slot != 0xFFFFFFFFu) {
    static const char* what[] = {"", "xchg", "store", "cas", "alloc", "release"};
    fprintf(stderr, "slot %u (slice/pool bounds: vars slice %u pool [%u,%u))\n", m.bad_slot, m.vars_slice, m.vars_pool_lo, m.vars_pool_hi);
    for (auto& e : m.vlog)
      if (e.slot == m.bad_slot) fprintf(stderr, "  op %llu wave %d %s old %08x new %08x\n", (unsigned long long)e.op, (int)e.w, what[e.what], e.old, e.nw);
  }
  out->err = m.err;
  for (int k = 0; k < K_N; ++k) { out->kind[k] = m.kind[k]; out->itrs += m.kind[k]; }
  std::string res;
  for (u32 j = 0; j < n_nets; ++j) {
    if (j) res += ";";
    print_port(m, n_nets == 1 ? HVM_ROOT : hvm_new_port(HVM_VAR, 1 + j), res, 0);
  }
  out->result = (char*)malloc(res.size() + 1);
  memcpy(out->result, res.c_str(), res.size() + 1);
  return m.abandoned ? 3 : 0;
}

extern "C" void hvm_model_free(void* p) { free(p); }
