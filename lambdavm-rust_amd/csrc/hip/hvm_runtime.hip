// hvm_runtime.hip — host side of the MI355X evaluator: net and book in HBM,
// the turn loop, the run-cu IO driver's readback, result printing, and the
// C ABI declared in include/hvm_hip.h.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cerrno>
#include <mutex>
#include <cstdio>
#include <ctime>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/hvm_hip.h"
#include "../front/hvm_front.hpp"
#include "hvm_dev.hpp"
#include "hvm_wide.hpp"

extern "C" hipError_t hvm_launch_turn(const HvmTurnArgs* a, hipStream_t s);
extern "C" hipError_t hvm_launch_gather(const HvmGatherArgs* g, hipStream_t s);
extern "C" hipError_t hvm_launch_between(const HvmTurnArgs* a, uint32_t clock_khz, hipStream_t s);
extern "C" uint32_t hvm_plain_action_host(uint32_t n, uint32_t nwaves);
extern "C" uint32_t hvm_turn_policy_host(uint32_t n, uint32_t nwaves);
extern "C" uint32_t hvm_turn_mid_pool(void);
extern "C" size_t hvm_turn_lds_bytes(int local);
extern "C" int hvm_turn_waves_per_cu(int local);

namespace {

typedef unsigned long long u64;
typedef uint32_t u32;

static std::string g_create_error;

static char* dup_cstr(const std::string& s) {
  char* p = (char*)malloc(s.size() + 1);
  memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct hvm_hip_ctx {
  hvm_hip_cfg cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 0;
  // net
  u64* node = nullptr;
  u32* vars = nullptr;
  u32 node_cap = 0, vars_cap = 0;
  u32 nwaves = 0, node_slice = 0, vars_slice = 0;
  u32 node_pool_lo = 0, vars_pool_lo = 0, node_pool_hi = 0, vars_pool_hi = 0;
  // bags + control
  u64* bag[2] = {nullptr, nullptr};
  u32 bag_cap = 0;
  HvmCtl* ctl = nullptr;       // turn control block
  HvmWaveState* wave_state = nullptr;
  u64* lspill = nullptr;
  HvmDevStats* stats = nullptr;
  // scheduler
  HvmSched* sched = nullptr;
  unsigned short* d_params = nullptr;  // MLP weights (fp16)
  float* d_trace = nullptr;
  HvmSched host_sched{};
  uint32_t clock_khz = 100000;
  // book
  hvm::Book book;
  HvmDevDefX* d_defs = nullptr;
  u64* d_pairs = nullptr;
  u32 ndefs = 0;
  bool loaded = false;
  // the 64-bit-port evaluator's net (HVM_HIP_WIDE), instead of the above
  HvmWide* wide = nullptr;
  // per-eval bookkeeping
  double kernel_ms = 0, launch_ms = 0;
  u32 turns = 0, normalizes = 0, launches = 0;
  std::string last_error;
  // No-allocation mode (hvm_hip_run_static): every device buffer is carved
  // from a caller-provided region (the library's static scratch array); the
  // default stream, no events, no graphs.
  bool static_mode = false;
  unsigned char* sbase = nullptr;
  size_t scap = 0, soff = 0;
  unsigned char* s_book = nullptr;     // book records + pairs
  size_t s_book_bytes = 0;
  unsigned char* s_gather = nullptr;   // readback lists
  u32 s_gather_cap = 0;
  void* carve(size_t bytes) {
    size_t at = (soff + 255) & ~(size_t)255;
    if (at + bytes > scap) return nullptr;
    soff = at + bytes;
    return sbase + at;
  }

  bool fail(const std::string& m) { last_error = m; return false; }
  bool ck(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
    return true;
  }

  // Limits of the no-allocation mode and the layout of its region: the net
  // (HVM_HIP_STATIC_SLOTS nodes and vars), two bags of HVM_HIP_STATIC_BAG
  // redexes, the book, the readback lists and the small control blocks.
  static constexpr size_t kStaticBook = 4u << 20;
  static constexpr u32 kStaticGather = 1u << 18;

  bool init_static(const hvm_hip_cfg* c, void* base, size_t bytes) {
    static_mode = true;
    sbase = (unsigned char*)base;
    scap = bytes;
    soff = 0;
    if (c) cfg = *c;
    if (cfg.flags & (HVM_HIP_WIDE | HVM_HIP_LOCAL_NET | HVM_HIP_GRAPHS))
      return fail("hvm_hip_run_static: HVM_HIP_WIDE, HVM_HIP_LOCAL_NET and HVM_HIP_GRAPHS need allocations");
    if (cfg.node_cap > HVM_HIP_STATIC_SLOTS || cfg.vars_cap > HVM_HIP_STATIC_SLOTS)
      return fail("hvm_hip_run_static: node_cap/vars_cap above HVM_HIP_STATIC_SLOTS (" +
                  std::to_string(HVM_HIP_STATIC_SLOTS) + ")");
    if (cfg.bag_cap > HVM_HIP_STATIC_BAG)
      return fail("hvm_hip_run_static: bag_cap above HVM_HIP_STATIC_BAG (" + std::to_string(HVM_HIP_STATIC_BAG) + ")");
    if (!cfg.node_cap) cfg.node_cap = HVM_HIP_STATIC_SLOTS;
    if (!cfg.vars_cap) cfg.vars_cap = HVM_HIP_STATIC_SLOTS;
    if (!cfg.bag_cap) cfg.bag_cap = HVM_HIP_STATIC_BAG;
    return init(&cfg);
  }

  bool init(const hvm_hip_cfg* c) {
    if (c) cfg = *c;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail("no HIP device");
    if (!ck(hipSetDevice(0), "hipSetDevice")) return false;
    hipDeviceProp_t prop;
    if (!ck(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties")) return false;
    cus = prop.multiProcessorCount;
    uint64_t ncap = cfg.node_cap ? cfg.node_cap : (1ull << 28);
    uint64_t vcap = cfg.vars_cap ? cfg.vars_cap : (1ull << 28);
    if (cfg.flags & HVM_HIP_WIDE) {
      // 64-bit ports: capacities up to 2^32 - 2 slots, a net of its own
      if (cfg.flags & HVM_HIP_LOCAL_NET) return fail("HVM_HIP_WIDE and HVM_HIP_LOCAL_NET exclude each other");
      if (!ck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
      wide = new HvmWide();
      wide->turn_steps = cfg.turn_steps;
      wide->idle_pct = cfg.idle_pct ? cfg.idle_pct : 25u;
      if (!wide->create(ncap, vcap, cfg.bag_cap ? cfg.bag_cap : (1u << 24), cfg.waves, cus, stream))
        return fail(wide->err);
      node_cap = wide->a.node_cap;
      vars_cap = wide->a.vars_cap;
      nwaves = wide->a.nwaves;
      return true;
    }
    if (ncap > (1ull << 29) || vcap > (1ull << 29)) return fail("node_cap/vars_cap exceed 2^29");
    const bool local = (cfg.flags & HVM_HIP_LOCAL_NET) != 0;
    if (local && (ncap > HVM_LOCAL_BASE || vcap > HVM_LOCAL_BASE))
      return fail("node_cap/vars_cap exceed 2^28, the limit of the local-net evaluator");
    node_cap = (u32)ncap;
    vars_cap = (u32)vcap;
    // Every redex the evaluator pushes holds the principal port of at least
    // one live node (VOID redexes are counted where they are created), so a
    // bag of node_cap entries plus one per batched root never overflows.
    bag_cap = cfg.bag_cap ? cfg.bag_cap : node_cap + 1 + HVM_HIP_MAX_NETS;
    if (cfg.waves) {
      nwaves = cfg.waves;
    } else {
      // Fill the chip: as many one-wave workgroups as can be resident per CU
      // (registers and LDS), so that every wave of a turn runs at once.
      int occ = hvm_turn_waves_per_cu(local ? 1 : 0);
      size_t lds = hvm_turn_lds_bytes(local ? 1 : 0);
      size_t per_cu = occ > 0 ? (size_t)occ : (160u * 1024u) / (lds ? lds : 1);
      if (per_cu < 1) per_cu = 1;
      if (per_cu > 16) per_cu = 16;
      nwaves = (u32)(cus * per_cu);
    }
    // node slot 0 and var slots [0, 1 + HVM_HIP_MAX_NETS) (ROOT and the roots
    // of batched nets) are outside every wave's slice
    const u32 root_vars = 1 + HVM_HIP_MAX_NETS;
    if (vars_cap <= root_vars) return fail("vars_cap too small for the root var slots");
    // three quarters of each buffer is split into per-wave slices; the rest
    // is a shared overflow pool for waves whose slice is full
    node_slice = (u32)(((u64)(node_cap - 1) * 3 / 4) / nwaves);
    vars_slice = (u32)(((u64)(vars_cap - root_vars) * 3 / 4) / nwaves);
    node_pool_lo = 1 + nwaves * node_slice;
    vars_pool_lo = root_vars + nwaves * vars_slice;
    // the top of each buffer is the host's, for the nodes and vars the IO
    // driver creates (up to 2^16 slots, half the overflow pool at most)
    node_pool_hi = node_cap - std::min<u32>(1u << 16, (node_cap - node_pool_lo) / 2);
    vars_pool_hi = vars_cap - std::min<u32>(1u << 16, (vars_cap - vars_pool_lo) / 2);
    if (node_slice < 4 * HVM_WAVE || vars_slice < 4 * HVM_WAVE) return fail("net too small for the wave count");
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) == hipSuccess && khz > 0) clock_khz = (uint32_t)khz;
    memset(&host_sched, 0, sizeof host_sched);
    host_sched.mode = HVM_SCHED_PLAIN;
    if (static_mode) {
      // every buffer from the region; the default stream; no events
      node = (u64*)carve(((size_t)node_cap + 1) * 8);
      vars = (u32*)carve(((size_t)vars_cap + 1) * 4);
      bag[0] = (u64*)carve((size_t)bag_cap * 8);
      bag[1] = (u64*)carve((size_t)bag_cap * 8);
      ctl = (HvmCtl*)carve(sizeof(HvmCtl));
      wave_state = (HvmWaveState*)carve(nwaves * sizeof(HvmWaveState));
      stats = (HvmDevStats*)carve(sizeof(HvmDevStats) * HVM_STAT_STRIPES);
      sched = (HvmSched*)carve(sizeof(HvmSched));
      d_params = (unsigned short*)carve(HVM_MLP_PARAMS * sizeof(unsigned short));
      s_book = (unsigned char*)carve(kStaticBook);
      s_book_bytes = kStaticBook;
      s_gather_cap = kStaticGather;
      s_gather = (unsigned char*)carve((size_t)kStaticGather * (4 + 8 + 4 + 4 + 4) + 16);
      if (!node || !vars || !bag[0] || !bag[1] || !ctl || !wave_state || !stats || !sched || !d_params || !s_book ||
          !s_gather)
        return fail("hvm_hip_run_static: the static region is too small for this configuration");
      return true;
    }
    if (!ck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
    if (!ck(hipEventCreate(&ev0), "hipEventCreate") || !ck(hipEventCreate(&ev1), "hipEventCreate")) return false;
    // one spare slot past each capacity: the evaluator's accesses through an
    // out-of-range port land there (and raise HVM_ERR_BAD_PORT)
    if (!ck(hipMalloc(&node, ((size_t)node_cap + 1) * 8), "hipMalloc(node)")) return false;
    if (!ck(hipMalloc(&vars, ((size_t)vars_cap + 1) * 4), "hipMalloc(vars)")) return false;
    if (!ck(hipMalloc(&bag[0], (size_t)bag_cap * 8), "hipMalloc(bag)")) return false;
    if (!ck(hipMalloc(&bag[1], (size_t)bag_cap * 8), "hipMalloc(bag)")) return false;
    if (!ck(hipMalloc(&ctl, sizeof(HvmCtl)), "hipMalloc(ctl)")) return false;
    if (!ck(hipMalloc(&wave_state, nwaves * sizeof(HvmWaveState)), "hipMalloc(wave_state)")) return false;
    if (local && !ck(hipMalloc(&lspill, (size_t)nwaves * HVM_LSPILL * 8), "hipMalloc(lspill)")) return false;
    if (!ck(hipMalloc(&stats, sizeof(HvmDevStats) * HVM_STAT_STRIPES), "hipMalloc(stats)")) return false;
    if (!ck(hipMalloc(&sched, sizeof(HvmSched)), "hipMalloc(sched)")) return false;
    if (!ck(hipMalloc(&d_params, HVM_MLP_PARAMS * sizeof(unsigned short)), "hipMalloc(params)")) return false;
    return true;
  }

  void release() {
    if (static_mode) return;   // the region is not ours; no stream, events or graphs
    if (wide) { wide->destroy(); delete wide; wide = nullptr; }
    // teardown: a failing free or destroy has nothing left to undo
    for (void* p : {(void*)node, (void*)vars, (void*)bag[0], (void*)bag[1], (void*)ctl, (void*)wave_state,
                    (void*)lspill, (void*)stats, (void*)sched, (void*)d_params, (void*)d_trace, (void*)d_defs,
                    (void*)d_pairs})
      if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    drop_graphs();
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bool load(const u32* buf, size_t n_words) {
    std::string err;
    hvm::Book b;
    if (!buf) return fail("failed to load book: NULL buffer");
    if (!hvm::load_buffer(buf, n_words, b, err)) return fail(err);
    if (b.defs.empty()) return fail("failed to load book: no defs");
    // Compiled CALLs: templates pre-reduced at load, with the static
    // interactions of each def counted at its every expansion.
    std::vector<hvm::SpecCounts> spec;
    if (cfg.flags & HVM_HIP_COMPILED_CALLS) {
      for (const hvm::Def& d : b.defs)
        if (d.node.size() > HVM_GPU_DEF_MAX || d.vars > HVM_GPU_DEF_MAX)
          return fail("def '" + d.name + "' exceeds the evaluator's per-def limit of " +
                      std::to_string(HVM_GPU_DEF_MAX) + " nodes/vars");
      hvm::Book sb = hvm::specialize_book(b, HVM_HIP_SPEC_NODES, HVM_GPU_DEF_MAX, spec);
      // the printer names REFs by the original book; names are unchanged
      b = std::move(sb);
    }
    std::vector<HvmDevDef> defs(b.defs.size());
    std::vector<u64> pairs;
    for (size_t f = 0; f < b.defs.size(); ++f) {
      const hvm::Def& d = b.defs[f];
      if (d.node.size() > HVM_GPU_DEF_MAX || d.vars > HVM_GPU_DEF_MAX)
        return fail("def '" + d.name + "' exceeds the evaluator's per-def limit of " +
                    std::to_string(HVM_GPU_DEF_MAX) + " nodes/vars");
      // Every template port must stay inside its def, so that instantiation
      // only touches slots the allocator handed out.
      auto port_ok = [&](u32 p) {
        hvm_tag t = hvm_get_tag(p);
        u32 v = hvm_get_val(p);
        if (t >= HVM_CON) return v < d.node.size();
        if (t == HVM_VAR) return v < d.vars;
        if (t == HVM_REF) return (v & 0xFFFFFFFu) < b.defs.size();
        return true;
      };
      bool ok = port_ok(d.root);
      for (u64 p : d.rbag) ok = ok && port_ok(hvm_get_fst(p)) && port_ok(hvm_get_snd(p));
      for (u64 p : d.node) ok = ok && port_ok(hvm_get_fst(p)) && port_ok(hvm_get_snd(p));
      if (!ok) return fail("def '" + d.name + "' has a port outside its own nodes/vars");
      HvmDevDef& x = defs[f];
      x.rbag_off = (u32)pairs.size();
      pairs.insert(pairs.end(), d.rbag.begin(), d.rbag.end());
      x.node_off = (u32)pairs.size();
      pairs.insert(pairs.end(), d.node.begin(), d.node.end());
      x.rlen = (u32)d.rbag.size();
      x.nlen = (u32)d.node.size();
      x.vlen = d.vars;
      x.root = d.root;
      x.safe = d.safe ? 1 : 0;
      x.call_bytes = 8 * x.nlen + 4 * x.vlen + 8 * (x.rlen + 1);
      if (!spec.empty() && (spec[f].call || spec[f].anni)) {
        // bit 0 safe, bits 8-15 static CALLs, bits 16-31 static ANNIs
        x.safe |= (spec[f].call << 8) | (spec[f].anni << 16);
        x.call_bytes = spec[f].bytes;
      }
    }
    if (pairs.empty()) pairs.push_back(0);
    // records with the first template pairs inline (HvmDevDefX)
    std::vector<HvmDevDefX> defx(defs.size());
    for (size_t f = 0; f < defs.size(); ++f) {
      HvmDevDefX& x = defx[f];
      memset(&x, 0, sizeof x);
      x.h = defs[f];
      for (u32 i = 0; i < HVM_PRE_NODE && i < x.h.nlen; ++i) x.node[i] = pairs[x.h.node_off + i];
      for (u32 i = 0; i < HVM_PRE_RBAG && i < x.h.rlen; ++i) x.rbag[i] = pairs[x.h.rbag_off + i];
    }
    if (static_mode) {
      const size_t dbytes = (defx.size() * sizeof(HvmDevDefX) + 255) & ~(size_t)255;
      if (dbytes + pairs.size() * 8 > s_book_bytes)
        return fail("hvm_hip_run_static: the book exceeds the static region's " + std::to_string(s_book_bytes) +
                    " bytes for books");
      d_defs = (HvmDevDefX*)s_book;
      d_pairs = (u64*)(s_book + dbytes);
    } else {
      if (d_defs) { (void)hipFree(d_defs); d_defs = nullptr; }
      if (d_pairs) { (void)hipFree(d_pairs); d_pairs = nullptr; }
      if (!ck(hipMalloc(&d_defs, defx.size() * sizeof(HvmDevDefX)), "hipMalloc(defs)")) return false;
      if (!ck(hipMalloc(&d_pairs, pairs.size() * 8), "hipMalloc(pairs)")) return false;
    }
    if (!ck(hipMemcpy(d_defs, defx.data(), defx.size() * sizeof(HvmDevDefX), hipMemcpyHostToDevice), "copy defs")) return false;
    if (!ck(hipMemcpy(d_pairs, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice), "copy pairs")) return false;
    ndefs = (u32)defs.size();
    book = std::move(b);
    loaded = true;
    return true;
  }

  // --- small host accessors (the reference's gnet_* helpers, src/hvm.cu:1842-1881)
  // A failed copy is sticky: later reads return 0 and the readback that
  // used them fails (expand and the IO driver check copy_failed).
  bool copy_failed = false;
  void copy_ck(hipError_t e) {
    if (e != hipSuccess && !copy_failed) { copy_failed = true; ck(e, "readback copy"); }
  }
  u64 node_load(u32 loc) {
    u64 v = 0;
    if (loc < node_cap && !copy_failed) copy_ck(hipMemcpy(&v, node + loc, 8, hipMemcpyDeviceToHost));
    return v;
  }
  u32 vidx(u32 val) const { return val == HVM_ROOT_VAL ? 0u : val; }
  u32 vars_load(u32 val) {
    u32 v = 0;
    u32 i = vidx(val);
    if (i < vars_cap && !copy_failed) copy_ck(hipMemcpy(&v, vars + i, 4, hipMemcpyDeviceToHost));
    return v;
  }
  void vars_store(u32 val, u32 x) {
    u32 i = vidx(val);
    if (i < vars_cap && !copy_failed) copy_ck(hipMemcpy(vars + i, &x, 4, hipMemcpyHostToDevice));
  }
  u32 peek(u32 port) {
    for (int it = 0; it < (1 << 24) && hvm_get_tag(port) == HVM_VAR && !copy_failed; ++it) {
      u32 v = vars_load(hvm_get_val(port));
      if (v == HVM_NONE || v == 0) break;
      port = v;
    }
    return port;
  }

  bool reset_heap() {
    if (!ck(hipMemsetAsync(node, 0, ((size_t)node_cap + 1) * 8, stream), "clear node")) return false;
    if (!ck(hipMemsetAsync(vars, 0, ((size_t)vars_cap + 1) * 4, stream), "clear vars")) return false;
    if (!ck(hipMemsetAsync(wave_state, 0, nwaves * sizeof(HvmWaveState), stream), "clear cursors")) return false;
    if (!ck(hipMemsetAsync(ctl, 0, sizeof(HvmCtl), stream), "clear ctl")) return false;
    if (hvm_turn_mid_pool()) {
      // the mid-turn pool reads a zero slot as one not yet written
      if (!ck(hipMemsetAsync(bag[0], 0, (size_t)bag_cap * 8, stream), "clear bag")) return false;
      if (!ck(hipMemsetAsync(bag[1], 0, (size_t)bag_cap * 8, stream), "clear bag")) return false;
    }
    return ck(hipStreamSynchronize(stream), "reset sync");
  }

  // The evaluator's counters, summed over their copies (HVM_STAT_STRIPES).
  bool read_stats(HvmDevStats& s) {
    static_assert(sizeof(HvmDevStats) % 8 == 0, "counters are u64");
    std::vector<HvmDevStats> all(HVM_STAT_STRIPES);
    memset(&s, 0, sizeof s);
    if (!ck(hipMemcpy(all.data(), stats, sizeof(HvmDevStats) * HVM_STAT_STRIPES, hipMemcpyDeviceToHost),
            "stats copy"))
      return false;
    unsigned long long* sum = (unsigned long long*)&s;
    for (const HvmDevStats& x : all) {
      const unsigned long long* v = (const unsigned long long*)&x;
      for (size_t k = 0; k < sizeof(HvmDevStats) / 8; ++k) sum[k] += v[k];
    }
    return true;
  }

  bool reset_net() {
    copy_failed = false;
    if (!ck(hipMemsetAsync(stats, 0, sizeof(HvmDevStats) * HVM_STAT_STRIPES, stream), "clear stats")) return false;
    return reset_heap();
  }

  // Scheduler state for a new normalisation: no previous decision, counters
  // clear; the trace keeps accumulating until the next eval starts.
  bool sched_start(u32 n_first, bool new_eval) {
    u32 len = 0;
    if (!new_eval && !ck(hipMemcpy(&len, &sched->trace_len, 4, hipMemcpyDeviceToHost), "trace len")) return false;
    HvmSched s = host_sched;
    s.trace_len = len;
    s.last_clock = 0;
    s.last_action = hvm_plain_action_host(n_first, nwaves);
    s.params = d_params;
    s.trace = d_trace;
    memset(s.qavg, 0, sizeof s.qavg);
    s.qn = 0;
    // the counters' sums so far (a later normalisation of the same
    // evaluation, or a later group of a batch, continues the totals)
    memset(&s.prev, 0, sizeof s.prev);
    if (!new_eval) {
      HvmDevStats st;
      if (!read_stats(st)) return false;
      hvm_turn_view(st, s.prev);
    }
    return ck(hipMemcpy(sched, &s, sizeof s, hipMemcpyHostToDevice), "sched upload");
  }

  bool set_scheduler(u32 mode, u32 seed, u32 explore_pct, const unsigned short* params, u32 trace_cap) {
    if (mode > HVM_SCHED_FIXED) return fail("unknown scheduler mode");
    if (mode == HVM_SCHED_LEARNED && !params) return fail("learned scheduler needs parameters");
    if (params && !ck(hipMemcpy(d_params, params, HVM_MLP_PARAMS * sizeof(unsigned short), hipMemcpyHostToDevice),
                      "params upload")) return false;
    if (static_mode && trace_cap) return fail("hvm_hip_run_static: no scheduler trace without allocations");
    if (trace_cap != host_sched.trace_cap) {
      if (d_trace) { (void)hipFree(d_trace); d_trace = nullptr; }
      if (trace_cap && !ck(hipMalloc(&d_trace, (size_t)trace_cap * HVM_TRACE_REC * sizeof(float)), "hipMalloc(trace)"))
        return false;
    }
    host_sched.mode = mode;
    host_sched.seed = seed;
    host_sched.explore_pct = explore_pct;
    host_sched.trace_cap = trace_cap;
    return true;
  }

  int read_trace(float* recs, u32 cap, u32* n) {
    u32 len = 0;
    if (!ck(hipMemcpy(&len, &sched->trace_len, 4, hipMemcpyDeviceToHost), "trace len")) return HVM_HIP_E_DEVICE;
    if (len > host_sched.trace_cap) len = host_sched.trace_cap;
    u32 k = len < cap ? len : cap;
    if (k && recs && d_trace &&
        !ck(hipMemcpy(recs, d_trace, (size_t)k * HVM_TRACE_REC * sizeof(float), hipMemcpyDeviceToHost), "trace copy"))
      return HVM_HIP_E_DEVICE;
    if (n) *n = len;
    return HVM_HIP_OK;
  }

  bool eval_started = false;  // first boot of an eval resets the trace

  // Boots `redex` (src/hvm.c:1204-1207): ROOT := NONE, bag := [redex].
  bool boot(u64 redex) {
    if (!sched_start(1, !eval_started)) return false;
    eval_started = true;
    u32 none = HVM_NONE;
    if (!ck(hipMemcpyAsync(vars, &none, 4, hipMemcpyHostToDevice, stream), "boot root")) return false;
    if (!ck(hipMemcpyAsync(bag[0], &redex, 8, hipMemcpyHostToDevice, stream), "boot redex")) return false;
    HvmCtl c;
    memset(&c, 0, sizeof c);
    c.n_in = 1;
    c.in_head = 1;
    c.parity = 0;
    c.fixed_steps = std::min(cfg.turn_steps, 1u << 20);
    c.idle_quit = c.idle_back = turn_args().idle_quit;
    c.max_steps = cfg.turn_steps ? std::min(cfg.turn_steps, 1u << 20) : hvm_turn_policy_host(1, nwaves);
    c.turns = 0;
    c.done = 0;
    host_ctl = c;
    if (!ck(hipMemcpyAsync(ctl, &host_ctl, sizeof(HvmCtl), hipMemcpyHostToDevice, stream), "boot ctl")) return false;
    return ck(hipStreamSynchronize(stream), "boot sync");
  }

  HvmCtl host_ctl{};
  std::vector<hipEvent_t> tev;  // a start/stop event pair per turn of a chunk
  // HVM_HIP_GRAPHS: one instantiated graph per chunk length (2, 4, ..., 64)
  // holding the chunk's (event, turn, event, between) launches; rebuilt when
  // the turn arguments change
  hipGraphExec_t chunk_graph[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  HvmTurnArgs graph_args{};

  void drop_graphs() {
    for (hipGraphExec_t& g : chunk_graph) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }
  // Issues one chunk of `chunk` turns: captured once into a graph and
  // replayed, or launched directly.
  bool launch_chunk(const HvmTurnArgs& a, u32 chunk) {
    const bool graphs = (cfg.flags & HVM_HIP_GRAPHS) != 0;
    const bool events = !graphs && !static_mode;
    auto direct = [&](hipStream_t st) {
      for (u32 i = 0; i < chunk; ++i) {
        if (events && !ck(hipEventRecord(tev[2 * i], st), "record turn start")) return false;
        if (!ck(hvm_launch_turn(&a, st), "launch turn")) return false;
        if (events && !ck(hipEventRecord(tev[2 * i + 1], st), "record turn end")) return false;
        if (!ck(hvm_launch_between(&a, clock_khz, st), "launch between")) return false;
      }
      return true;
    };
    if (!graphs) return direct(stream);
    if (memcmp(&graph_args, &a, sizeof a) != 0) {
      drop_graphs();
      graph_args = a;
    }
    u32 k = 0;
    while ((2u << k) < chunk) ++k;
    if (!chunk_graph[k]) {
      hipGraph_t g = nullptr;
      if (!ck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture")) return false;
      bool ok = direct(stream);
      hipError_t e = hipStreamEndCapture(stream, &g);
      if (!ok) { if (g) (void)hipGraphDestroy(g); return false; }
      if (!ck(e, "end capture")) return false;
      e = hipGraphInstantiate(&chunk_graph[k], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);   // the instantiated graph no longer needs it
      if (!ck(e, "instantiate graph")) return false;
    }
    // the graph holds kernels only; the chunk is timed as a whole
    if (!ck(hipEventRecord(tev[0], stream), "record chunk start")) return false;
    if (!ck(hipGraphLaunch(chunk_graph[k], stream), "launch graph")) return false;
    return ck(hipEventRecord(tev[1], stream), "record chunk end");
  }

  HvmTurnArgs turn_args() const {
    HvmTurnArgs a;
    memset(&a, 0, sizeof a);   // padding too: launch_chunk compares argument blocks
    a.node = node; a.vars = vars;
    a.node_cap = node_cap; a.vars_cap = vars_cap;
    a.node_slice = node_slice; a.vars_slice = vars_slice;
    a.node_pool_lo = node_pool_lo; a.vars_pool_lo = vars_pool_lo;
    a.node_pool_hi = node_pool_hi; a.vars_pool_hi = vars_pool_hi;
    a.defs = d_defs; a.pairs = d_pairs; a.ndefs = ndefs;
    a.bag[0] = bag[0]; a.bag[1] = bag[1];
    a.bag_cap = bag_cap;
    a.ctl = ctl;
    a.nwaves = nwaves;
    a.wave_state = wave_state;
    a.stats = stats;
    a.root_vars = 1 + HVM_HIP_MAX_NETS;
    a.sched = sched;
    a.flags = (cfg.flags & HVM_HIP_FENCE_RELEASE) ? 0u : HVM_TURN_WAITCNT_FENCE;
    if (cfg.flags & HVM_HIP_PROFILE) a.flags |= HVM_TURN_PROFILE;
    if (cfg.flags & HVM_HIP_LOCAL_NET) a.flags |= HVM_TURN_LOCAL;
    else if (cfg.flags & HVM_HIP_COMPILED_CALLS) a.flags |= HVM_TURN_FAST;
    a.lspill = lspill;
    // a turn ends early once idle_pct% of the waves ran out of work (default 25%)
    u32 pct = cfg.idle_pct ? (cfg.idle_pct > 100 ? 0u : cfg.idle_pct) : 25u;
    a.idle_quit = pct ? std::max<u32>(1u, (u32)(((u64)nwaves * pct) / 100)) : 0u;
    return a;
  }

  double deadline = 0;  // steady-clock seconds; 0 = no limit

  // The turn loop (replaces gnet_normalize, src/hvm.cu:1810-1839).  Turns are
  // launched in chunks of (turn, between) kernel pairs with no host round trip
  // inside a chunk; the host reads the control block once per chunk.  Turns
  // launched after the normal form is reached return immediately.
  bool normalize() {
    normalizes++;
    HvmTurnArgs a = turn_args();
    u32 chunk = 2;
    const u32 max_chunk = 64;
    if (!static_mode && tev.size() < 2 * max_chunk) {
      for (size_t i = tev.size(); i < 2 * max_chunk; ++i) {
        hipEvent_t e;
        if (!ck(hipEventCreate(&e), "hipEventCreate")) return false;
        tev.push_back(e);
      }
    }
    u32 turns_before = host_ctl.turns;
    for (;;) {
      if (deadline > 0 && now_s() > deadline) return fail("time limit exceeded");
      const double t_chunk = now_s();
      if (!launch_chunk(a, chunk)) return false;
      if (!ck(hipMemcpyAsync(&host_ctl, ctl, sizeof(HvmCtl), hipMemcpyDeviceToHost, stream), "ctl readback")) return false;
      if (!ck(hipStreamSynchronize(stream), "turn sync")) return false;
      u32 ran = host_ctl.turns - turns_before;
      if (static_mode) {
        // no events: the chunk's host time, shared by the turns that ran
        const double ms = (now_s() - t_chunk) * 1e3;
        kernel_ms += ms * (double)ran / (double)chunk;
        launch_ms += ms;
      } else if (cfg.flags & HVM_HIP_GRAPHS) {
        // one span for the chunk (its between kernels included)
        float ms = 0;
        if (!ck(hipEventElapsedTime(&ms, tev[0], tev[1]), "chunk time")) return false;
        kernel_ms += ms * (float)ran / (float)chunk;
        launch_ms += ms;
      } else {
        for (u32 i = 0; i < chunk; ++i) {
          float ms = 0;
          if (!ck(hipEventElapsedTime(&ms, tev[2 * i], tev[2 * i + 1]), "turn time")) return false;
          if (i < ran) kernel_ms += ms;
          launch_ms += ms;
        }
      }
      launches += chunk;
      turns += ran;
      turns_before = host_ctl.turns;
      if (host_ctl.err) {
        char msg[160];
        snprintf(msg, sizeof msg, "evaluator error flags 0x%x (turn %u, output bag %u of %u)", host_ctl.err,
                 host_ctl.turns, host_ctl.n_out, bag_cap);
        return fail(msg);
      }
      if (host_ctl.done) break;
      if (chunk < max_chunk) chunk *= 2;
    }
    return true;
  }

  // src/hvm.c:1233-1244 / src/hvm.cu:1871-1881
  bool expand(u32 port, u32& got) {
    u32 old = vars_load(HVM_ROOT_VAL);
    got = peek(port);
    while (hvm_get_tag(got) == HVM_REF) {
      if (!boot(hvm_new_pair(got, HVM_ROOT))) return false;
      if (!normalize()) return false;
      got = peek(vars_load(HVM_ROOT_VAL));
    }
    vars_store(HVM_ROOT_VAL, old);
    return !copy_failed;
  }

  struct Ctr { int32_t tag = -1; u32 args_len = 0; u32 args[16]; };

  // src/run.cu:67-100 (readback_ctr).  Arguments past the 16th are expanded
  // (as the reference does) but not stored.
  bool readback_ctr(u32 port, Ctr& c) {
    u32 lam;
    if (!expand(port, lam)) return false;
    if (hvm_get_tag(lam) != HVM_CON) return true;
    u64 lam_node = node_load(hvm_get_val(lam));
    u32 app;
    if (!expand(hvm_get_fst(lam_node), app)) return false;
    if (hvm_get_tag(app) != HVM_CON) return true;
    u64 app_node = node_load(hvm_get_val(app));
    u32 arg;
    if (!expand(hvm_get_fst(app_node), arg)) return false;
    if (hvm_get_tag(arg) != HVM_NUM) return true;
    c.tag = (int32_t)hvm_get_u24(hvm_get_val(arg));
    for (u32 it = 0; it < (1u << 24); ++it) {
      if (!expand(hvm_get_snd(app_node), app)) return false;
      if (hvm_get_tag(app) != HVM_CON) break;
      app_node = node_load(hvm_get_val(app));
      if (!expand(hvm_get_fst(app_node), arg)) return false;
      if (c.args_len < 16) c.args[c.args_len] = arg;
      c.args_len++;
    }
    return true;
  }

  // -------------------------------------------------------------------------
  // IO driver (src/run.c:1-809; the CUDA runtime's is src/run.cu:67-925).
  // An IO program's normal form is a λ-encoded IO/Call(magic, name, argument,
  // continuation): the host reads the call back, performs it, writes the
  // result into the net as new nodes and normalises again, until the program
  // returns IO/Done.  Every normalisation pass runs on the GPU; reading ports
  // back and writing the results are host memory copies.
  // -------------------------------------------------------------------------

  // Host-created nodes and vars live in [pool_hi, cap) of each buffer, which
  // the device never allocates; the device frees them (zeroes) when it takes
  // them, like any other slot.  The region is read once per IO call.
  std::vector<u64> hnode;
  std::vector<u32> hvar;
  u32 hnode_cur = 0, hvar_cur = 0;

  bool host_region_read() {
    hnode.resize(node_cap - node_pool_hi);
    hvar.resize(vars_cap - vars_pool_hi);
    hnode_cur = hvar_cur = 0;
    return ck(hipMemcpy(hnode.data(), node + node_pool_hi, hnode.size() * 8, hipMemcpyDeviceToHost), "io region") &&
           ck(hipMemcpy(hvar.data(), vars + vars_pool_hi, hvar.size() * 4, hipMemcpyDeviceToHost), "io region");
  }
  // A fresh var (NONE) in the host region; 0 when the region is full.
  u32 host_var() {
    for (; hvar_cur < hvar.size(); ++hvar_cur) {
      if (hvar[hvar_cur] == 0) {
        u32 v = vars_pool_hi + hvar_cur;
        hvar[hvar_cur] = HVM_NONE;
        vars_store(v, HVM_NONE);
        return v;
      }
    }
    return 0;
  }
  // A new node holding `val` in the host region; 0 when the region is full.
  u32 host_node(u64 val) {
    for (; hnode_cur < hnode.size(); ++hnode_cur) {
      if (hnode[hnode_cur] == 0) {
        u32 loc = node_pool_hi + hnode_cur;
        hnode[hnode_cur] = val;
        copy_ck(hipMemcpy(node + loc, &val, 8, hipMemcpyHostToDevice));
        return loc;
      }
    }
    return 0;
  }
  bool io_full = false;   // the host region ran out during this call
  u32 nod(u32 tag, u64 val) {
    u32 loc = host_node(val);
    if (!loc) io_full = true;
    return hvm_new_port(tag, loc);
  }
  static u32 num_u24(u32 v) { return hvm_new_port(HVM_NUM, hvm_new_u24(v)); }
  static u32 num_i24(int32_t v) { return hvm_new_port(HVM_NUM, hvm_new_i24(v)); }

  // λ-encoded constructors, src/run.c:184-375 (tags: list NIL 0 / CONS 1,
  // result OK 0 / ERR 1, IO error TYPE 0 / NAME 1 / INNER 2).
  u32 ctr_var() { u32 v = host_var(); if (!v) io_full = true; return hvm_new_port(HVM_VAR, v); }
  u32 inject_nil() {                      // λt (t NIL)
    u32 x = ctr_var();
    u32 app = nod(HVM_CON, hvm_new_pair(num_u24(0), x));
    return nod(HVM_CON, hvm_new_pair(app, x));
  }
  u32 inject_cons(u32 head, u32 tail) {   // λt (((t CONS) head) tail)
    u32 x = ctr_var();
    u32 a3 = nod(HVM_CON, hvm_new_pair(tail, x));
    u32 a2 = nod(HVM_CON, hvm_new_pair(head, a3));
    u32 a1 = nod(HVM_CON, hvm_new_pair(num_u24(1), a2));
    return nod(HVM_CON, hvm_new_pair(a1, x));
  }
  u32 inject_bytes(const std::string& b) {
    u32 p = inject_nil();
    for (size_t i = b.size(); i-- > 0;) p = inject_cons(num_u24((unsigned char)b[i]), p);
    return p;
  }
  u32 inject_tagged(u32 tag, u32 val) {   // λt ((t TAG) val)
    u32 x = ctr_var();
    u32 a2 = nod(HVM_CON, hvm_new_pair(val, x));
    u32 a1 = nod(HVM_CON, hvm_new_pair(num_u24(tag), a2));
    return nod(HVM_CON, hvm_new_pair(a1, x));
  }
  u32 inject_ok(u32 val) { return inject_tagged(0, val); }
  u32 inject_err(u32 val) { return inject_tagged(1, val); }
  u32 inject_io_err(u32 tag, u32 val) {
    if (tag <= 1) {                       // IOError/Type, IOError/Name: λt (t TAG)
      u32 x = ctr_var();
      u32 app = nod(HVM_CON, hvm_new_pair(num_u24(tag), x));
      return inject_err(nod(HVM_CON, hvm_new_pair(app, x)));
    }
    return inject_err(inject_tagged(2, val));   // IOError/Inner(val)
  }
  u32 io_err_type() { return inject_io_err(0, 0); }
  u32 io_err_name() { return inject_io_err(1, 0); }
  u32 io_err_inner(u32 val) { return inject_io_err(2, val); }
  u32 io_err_errno(int e) { return io_err_inner(num_i24(e)); }
  u32 io_err_str(const std::string& msg) { return io_err_inner(inject_bytes(msg)); }

  // src/run.c:103-117: a right-nested CON tuple of at most `size` elements.
  bool readback_tup(u32 port, u32 size, std::vector<u32>& elems) {
    elems.clear();
    for (u32 it = 0; it < (1u << 24) && hvm_get_tag(port) == HVM_CON && elems.size() + 1 < size; ++it) {
      u64 nd = node_load(hvm_get_val(port));
      u32 e;
      if (!expand(hvm_get_fst(nd), e)) return false;
      elems.push_back(e);
      if (!expand(hvm_get_snd(nd), port)) return false;
    }
    elems.push_back(port);
    return true;
  }
  // src/run.c:123-159: a λ-encoded list of bytes, one normalisation per
  // element (the tail is booted against ROOT and read from there).
  bool readback_bytes(u32 port, std::string& out) {
    out.clear();
    for (u64 it = 0; it < (1ull << 32); ++it) {
      if (!normalize()) return false;
      Ctr c;
      if (!readback_ctr(peek(port), c)) return false;
      if (c.tag != 1 || c.args_len != 2 || hvm_get_tag(c.args[0]) != HVM_NUM) break;
      out.push_back((char)hvm_get_u24(hvm_get_val(c.args[0])));
      if (!boot(hvm_new_pair(c.args[1], HVM_ROOT))) return false;
      port = HVM_ROOT;
    }
    return true;
  }
  // src/run.c:169-180: the bytes as a C string (cut at the first NUL).
  bool readback_str(u32 port, std::string& out) {
    if (!readback_bytes(port, out)) return false;
    out.resize(strlen(out.c_str()));
    return true;
  }

  // Open files by descriptor; 0, 1 and 2 are stdin, stdout and stderr
  // (src/run.c:380-411).
  FILE* files[256] = {nullptr};
  FILE* readback_file(u32 port) {
    if (hvm_get_tag(port) != HVM_NUM) {
      fprintf(stderr, "non-num where file descriptor was expected: %i\n", (int)hvm_get_tag(port));
      return nullptr;
    }
    u32 fd = hvm_get_u24(hvm_get_val(port));
    if (fd == 0) return stdin;
    if (fd == 1) return stdout;
    if (fd == 2) return stderr;
    return fd < 256 ? files[fd] : nullptr;
  }

  // The IO functions of src/run.c:433-720, by name.  DL_OPEN/DL_CALL/DL_CLOSE
  // would call into a dylib with the C runtime's in-memory net, which this
  // evaluator does not have; they return an IO error.
  bool io_call(const std::string& name, u32 argm, u32& ret) {
    std::vector<u32> t;
    if (name == "READ") {
      if (!readback_tup(argm, 2, t)) return false;
      if (t.size() != 2) { ret = io_err_type(); return true; }
      FILE* fp = readback_file(t[0]);
      u32 n = hvm_get_u24(hvm_get_val(t[1]));
      if (!fp) { ret = io_err_errno(EBADF); return true; }
      std::string buf(n, '\0');
      size_t got = fread(&buf[0], 1, n, fp);
      if (got != n && ferror(fp)) { ret = io_err_errno(ferror(fp)); return true; }
      buf.resize(got);
      ret = inject_ok(inject_bytes(buf));
    } else if (name == "OPEN") {
      if (!readback_tup(argm, 2, t)) return false;
      if (t.size() != 2) { ret = io_err_type(); return true; }
      std::string fname, mode;
      if (!readback_str(t[0], fname) || !readback_str(t[1], mode)) return false;
      for (u32 fd = 3; fd < 256; ++fd) {
        if (!files[fd]) {
          files[fd] = fopen(fname.c_str(), mode.c_str());
          ret = files[fd] ? inject_ok(num_u24(fd)) : io_err_errno(errno);
          return true;
        }
      }
      ret = io_err_errno(EMFILE);
    } else if (name == "CLOSE") {
      FILE* fp = readback_file(argm);
      if (!fp) { ret = io_err_errno(EBADF); return true; }
      if (fclose(fp) != 0) { ret = io_err_errno(ferror(fp)); return true; }
      u32 fd = hvm_get_u24(hvm_get_val(argm));
      if (fd < 256) files[fd] = nullptr;
      ret = inject_ok(hvm_new_port(HVM_ERA, 0));
    } else if (name == "WRITE") {
      if (!readback_tup(argm, 2, t)) return false;
      if (t.size() != 2) { ret = io_err_type(); return true; }
      FILE* fp = readback_file(t[0]);
      std::string bytes;
      if (!readback_bytes(t[1], bytes)) return false;
      if (!fp) { ret = io_err_errno(EBADF); return true; }
      if (fwrite(bytes.data(), 1, bytes.size(), fp) != bytes.size()) { ret = io_err_errno(ferror(fp)); return true; }
      ret = inject_ok(hvm_new_port(HVM_ERA, 0));
    } else if (name == "FLUSH") {
      FILE* fp = readback_file(argm);
      if (!fp) { ret = io_err_errno(EBADF); return true; }
      if (fflush(fp) != 0) { ret = io_err_errno(ferror(fp)); return true; }
      ret = inject_ok(hvm_new_port(HVM_ERA, 0));
    } else if (name == "SEEK") {
      if (!readback_tup(argm, 3, t)) return false;
      if (t.size() != 3) { ret = io_err_type(); return true; }
      FILE* fp = readback_file(t[0]);
      int32_t off = hvm_get_i24(hvm_get_val(t[1]));
      u32 whence = (u32)hvm_get_i24(hvm_get_val(t[2]));
      if (!fp) { ret = io_err_errno(EBADF); return true; }
      int w = whence == 0 ? SEEK_SET : whence == 1 ? SEEK_CUR : whence == 2 ? SEEK_END : -1;
      if (w < 0) { ret = io_err_type(); return true; }
      if (fseek(fp, off, w) != 0) { ret = io_err_errno(ferror(fp)); return true; }
      ret = inject_ok(hvm_new_port(HVM_ERA, 0));
    } else if (name == "GET_TIME") {
      u64 ns = (u64)(now_s() * 1e9);
      u32 hi = (u32)(ns >> 24) & 0xFFFFFFF, lo = (u32)(ns & 0xFFFFFFF);
      ret = inject_ok(nod(HVM_CON, hvm_new_pair(num_u24(hi), num_u24(lo))));
    } else if (name == "SLEEP") {
      if (!readback_tup(argm, 2, t)) return false;
      if (t.size() != 2) { ret = io_err_type(); return true; }
      u64 ns = ((u64)hvm_get_u24(hvm_get_val(t[0])) << 24) | hvm_get_u24(hvm_get_val(t[1]));
      struct timespec ts;
      ts.tv_sec = (time_t)(ns / 1000000000ull);
      ts.tv_nsec = (long)(ns % 1000000000ull);
      nanosleep(&ts, nullptr);
      ret = inject_ok(hvm_new_port(HVM_ERA, 0));
    } else if (name == "DL_OPEN" || name == "DL_CALL" || name == "DL_CLOSE") {
      ret = io_err_str("dynamic libraries are not supported by the HIP evaluator");
    } else {
      ret = io_err_name();
    }
    return true;
  }

  // src/run.c:740-809 (do_run_io).  Returns 0, or a status code on failure.
  int run_io() {
    u32 port = HVM_ROOT;
    for (u64 it = 0; it < (1ull << 40); ++it) {
      if (!normalize()) return HVM_HIP_E_EVAL;
      Ctr c;
      if (!readback_ctr(peek(port), c)) return HVM_HIP_E_EVAL;
      if (c.args_len < 1 || hvm_get_tag(c.args[0]) != HVM_CON) break;
      u64 magic = node_load(hvm_get_val(c.args[0]));
      if (copy_failed) return HVM_HIP_E_DEVICE;
      if (hvm_get_val(hvm_get_fst(magic)) != hvm_new_u24(0xD0CA11) ||
          hvm_get_val(hvm_get_snd(magic)) != hvm_new_u24(0xFF1FF1)) break;
      if (c.tag != 1) break;                       // IO/Done (or not an IO call)
      if (c.args_len != 4) {
        fprintf(stderr, "invalid IO_CALL: args_len = %u\n", c.args_len);
        break;
      }
      std::string name;
      if (!readback_str(c.args[1], name)) return HVM_HIP_E_EVAL;
      const u32 argm = c.args[2], cont = c.args[3];
      io_full = false;
      if (!host_region_read()) return HVM_HIP_E_DEVICE;
      u32 ret = 0;
      if (!io_call(name, argm, ret)) return HVM_HIP_E_EVAL;
      u32 app = io_full ? 0 : host_node(hvm_new_pair(ret, HVM_ROOT));
      if (copy_failed) return HVM_HIP_E_DEVICE;
      if (io_full || !app) { fail("IO call results exceed the host region of the net"); return HVM_HIP_E_IO; }
      if (!boot(hvm_new_pair(hvm_new_port(HVM_CON, app), cont))) return HVM_HIP_E_DEVICE;
      port = HVM_ROOT;
    }
    return 0;
  }

  // Reads back everything reachable from `port` and prints it the way
  // pretty_print_port does (src/hvm.c:1618-1699).
  bool print_result(u32 port, std::string& out) {
    u32 cap = static_mode ? s_gather_cap : 1u << 20;
    for (int attempt = 0; attempt < (static_mode ? 1 : 6); ++attempt, cap <<= 2) {
      u32 *d_nidx = nullptr, *d_vidx = nullptr, *d_vval = nullptr, *d_q = nullptr, *d_cnt = nullptr;
      u64* d_nval = nullptr;
      bool ok = true;
      if (static_mode) {
        // the readback lists live in the region: node values first (8-byte aligned)
        d_nval = (u64*)s_gather;
        d_nidx = (u32*)(s_gather + (size_t)cap * 8);
        d_vidx = d_nidx + cap;
        d_vval = d_vidx + cap;
        d_q = d_vval + cap;
        d_cnt = d_q + cap;
      } else {
        ok = ck(hipMalloc(&d_nidx, (size_t)cap * 4), "gather alloc") &&
             ck(hipMalloc(&d_nval, (size_t)cap * 8), "gather alloc") &&
             ck(hipMalloc(&d_vidx, (size_t)cap * 4), "gather alloc") &&
             ck(hipMalloc(&d_vval, (size_t)cap * 4), "gather alloc") &&
             ck(hipMalloc(&d_q, (size_t)cap * 4), "gather alloc") &&
             ck(hipMalloc(&d_cnt, 16), "gather alloc");
      }
      u32 cnt[3] = {0, 0, 0};
      std::vector<u32> nidx, vidx_, vval;
      std::vector<u64> nval;
      if (ok) ok = ck(hipMemsetAsync(d_cnt, 0, 16, stream), "gather counts clear");
      if (ok) {
        HvmGatherArgs g;
        g.node = node; g.vars = vars; g.node_cap = node_cap; g.vars_cap = vars_cap;
        g.root_port = port;
        g.node_idx = d_nidx; g.node_val = d_nval; g.n_node = d_cnt;
        g.var_idx = d_vidx; g.var_val = d_vval; g.n_var = d_cnt + 1;
        g.cap = cap; g.queue = d_q; g.err = d_cnt + 2;
        ok = ck(hvm_launch_gather(&g, stream), "launch gather") &&
             ck(hipMemcpyAsync(cnt, d_cnt, 12, hipMemcpyDeviceToHost, stream), "gather counts") &&
             ck(hipStreamSynchronize(stream), "gather sync");
        if (ok && !(cnt[2] & 2) && cnt[0] <= cap && cnt[1] <= cap) {
          nidx.resize(cnt[0]); nval.resize(cnt[0]); vidx_.resize(cnt[1]); vval.resize(cnt[1]);
          ok = ck(hipMemcpy(nidx.data(), d_nidx, (size_t)cnt[0] * 4, hipMemcpyDeviceToHost), "gather copy") &&
               ck(hipMemcpy(nval.data(), d_nval, (size_t)cnt[0] * 8, hipMemcpyDeviceToHost), "gather copy") &&
               ck(hipMemcpy(vidx_.data(), d_vidx, (size_t)cnt[1] * 4, hipMemcpyDeviceToHost), "gather copy") &&
               ck(hipMemcpy(vval.data(), d_vval, (size_t)cnt[1] * 4, hipMemcpyDeviceToHost), "gather copy");
        }
      }
      // scratch of this readback: nothing to undo if a free fails
      if (!static_mode) {
        (void)hipFree(d_nidx); (void)hipFree(d_nval); (void)hipFree(d_vidx); (void)hipFree(d_vval);
        (void)hipFree(d_q); (void)hipFree(d_cnt);
      }
      if (!ok) return false;
      if (cnt[2] & 1) return fail("readback reached a port outside the net");
      if (cnt[2] & 2 || cnt[0] > cap || cnt[1] > cap) continue;  // grow and retry
      std::unordered_map<u32, u64> nodes;
      std::unordered_map<u32, u32> varv;
      nodes.reserve(nidx.size() * 2);
      for (size_t i = 0; i < nidx.size(); ++i) nodes[nidx[i]] = nval[i];
      for (size_t i = 0; i < vidx_.size(); ++i) varv[vidx_[i]] = vval[i];
      out.clear();
      std::vector<u32> st{port};
      while (!st.empty()) {
        u32 c = st.back();
        st.pop_back();
        switch (hvm_get_tag(c)) {
          case HVM_CON: case HVM_DUP: case HVM_OPR: case HVM_SWI: {
            u64 nd = nodes.count(hvm_get_val(c)) ? nodes[hvm_get_val(c)] : 0;
            hvm_tag t = hvm_get_tag(c);
            out += t == HVM_CON ? "(" : t == HVM_DUP ? "{" : t == HVM_OPR ? "$(" : "?(";
            st.push_back(hvm_new_port(HVM_ERA, t == HVM_DUP ? '}' : ')'));
            st.push_back(hvm_get_snd(nd));
            st.push_back(hvm_new_port(HVM_ERA, ' '));
            st.push_back(hvm_get_fst(nd));
            break;
          }
          case HVM_ERA:
            if (hvm_get_val(c)) out += (char)hvm_get_val(c); else out += "*";
            break;
          case HVM_VAR: {
            auto it = varv.find(hvm_get_val(c));
            u32 got = it == varv.end() ? HVM_NONE : it->second;
            if (got != HVM_NONE) {
              st.push_back(got);
            } else {
              char b[16];
              snprintf(b, sizeof b, "x%x", hvm_get_val(c));
              out += b;
            }
            break;
          }
          case HVM_NUM: out += hvm::c_show_numb(hvm_get_val(c)); break;
          case HVM_REF: {
            u32 fid = hvm_get_val(c) & 0xFFFFFFFu;
            out += "@" + (fid < book.defs.size() ? book.defs[fid].name : std::string("?"));
            break;
          }
        }
      }
      return true;
    }
    return fail(static_mode ? "result too large for the static readback region (" + std::to_string(s_gather_cap) +
                                  " nodes/vars)"
                            : "result too large to read back");
  }

  // The 64-bit-port evaluator: boot, normalise, read back (no IO driver).
  int eval_wide(hvm_hip_out* out) {
    double t0 = now_s();
    const double dl = cfg.time_limit_s ? t0 + cfg.time_limit_s : 0;
    HvmDevStats s;
    memset(&s, 0, sizeof s);
    u32 wturns = 0, wlaunches = 0;
    int rc = wide->normalize(d_defs, d_pairs, ndefs, dl, s, wturns, wlaunches);
    double t_eval = now_s();
    if (out) {
      memset(out, 0, sizeof *out);
      out->itrs = s.itrs;
      for (int k = 0; k < HVM_HIP_KINDS; ++k) out->kind[k] = s.kind[k];
      out->bytes = s.bytes;
      out->steps = s.steps; out->spills = s.spills; out->defers = s.defers; out->retries = s.retries;
      out->eval_ms = (t_eval - t0) * 1e3;
      out->turns = wturns;
      out->launches = wlaunches;
      out->waves = nwaves;
      out->normalizes = 1;
    }
    if (rc != HVM_HIP_OK) {
      fail(wide->err);
      if (out) {
        const size_t at = wide->err.find("flags 0x");
        out->err = at == std::string::npos ? 0u : (u32)strtoul(wide->err.c_str() + at + 8, nullptr, 16);
      }
      return rc;
    }
    if (!(cfg.flags & HVM_HIP_NO_READBACK)) {
      std::vector<std::string> names;
      for (const hvm::Def& d : book.defs) names.push_back(d.name);
      std::string text;
      if (!wide->print_root(names, text)) { fail(wide->err); return HVM_HIP_E_DEVICE; }
      if (out) {
        out->wall_s = now_s() - t0;
        out->result = dup_cstr(text);
        std::string canon;
        try { canon = hvm::canonical_tree_text(text); } catch (std::exception& ex) { canon = "<unparsable: " + std::string(ex.what()) + ">"; }
        out->canonical = dup_cstr(canon);
      }
    } else if (out) {
      out->wall_s = t_eval - t0;
    }
    return HVM_HIP_OK;
  }

  int eval(hvm_hip_out* out) {
    if (!loaded) { fail("no book loaded"); return HVM_HIP_E_ARG; }
    if (wide) return eval_wide(out);
    kernel_ms = 0;
    launch_ms = 0;
    turns = 0;
    launches = 0;
    normalizes = 0;
    eval_started = false;
    if (!reset_net()) return HVM_HIP_E_DEVICE;
    double t0 = now_s();
    deadline = cfg.time_limit_s ? t0 + cfg.time_limit_s : 0;
    if (!boot(hvm_new_pair(hvm_new_port(HVM_REF, 0), HVM_ROOT))) return HVM_HIP_E_DEVICE;
    int rc = run_io();
    double t_eval = now_s();
    HvmDevStats s;
    if (!read_stats(s) && rc == 0) rc = HVM_HIP_E_DEVICE;
    if (out) {
      memset(out, 0, sizeof *out);
      out->itrs = s.itrs;
      for (int k = 0; k < HVM_HIP_KINDS; ++k) out->kind[k] = s.kind[k];
      out->bytes = s.bytes;
      out->steps = s.steps; out->spills = s.spills; out->defers = s.defers;
      out->retries = s.retries; out->probes = s.probes;
      for (int k = 0; k < 6; ++k) out->phase_ms[k] = (double)s.phase[k] / (double)clock_khz;
      out->kernel_ms = kernel_ms;
      out->eval_ms = (t_eval - t0) * 1e3;
      out->turns = turns;
      out->launches = launches;
      out->launch_ms = launch_ms;
      out->waves = nwaves;
      out->leak = s.leaks; out->saves = s.saves; out->local_left = s.local_left;
      out->normalizes = normalizes;
    }
    if (rc != 0) {
      u32 e = 0;
      if (hipMemcpy(&e, &ctl->err, 4, hipMemcpyDeviceToHost) != hipSuccess) e = 0;   // the flags only add detail
      if (out) out->err = e;
      return rc;
    }
    if (!(cfg.flags & HVM_HIP_NO_READBACK)) {
      std::string text;
      if (!print_result(peek(HVM_ROOT), text)) return HVM_HIP_E_EVAL;
      double t1 = now_s();
      if (out) {
        out->wall_s = t1 - t0;
        out->result = dup_cstr(text);
        std::string canon;
        try { canon = hvm::canonical_tree_text(text); } catch (std::exception& ex) { canon = "<unparsable: " + std::string(ex.what()) + ">"; }
        out->canonical = dup_cstr(canon);
      }
    } else if (out) {
      out->wall_s = t_eval - t0;
    }
    return HVM_HIP_OK;
  }

  // Boots a group of nets: net j's entry def ~ its root var (slot 1 + j).
  bool boot_batch(const u32* entries, u32 n) {
    if (!sched_start(n, !eval_started)) return false;
    eval_started = true;
    std::vector<u32> roots(n, HVM_NONE);
    std::vector<u64> rx(n);
    for (u32 j = 0; j < n; ++j)
      rx[j] = hvm_new_pair(hvm_new_port(HVM_REF, entries[j]), hvm_new_port(HVM_VAR, 1 + j));
    if (!ck(hipMemcpyAsync(vars + 1, roots.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream), "boot roots")) return false;
    if (!ck(hipMemcpyAsync(bag[0], rx.data(), (size_t)n * 8, hipMemcpyHostToDevice, stream), "boot redexes")) return false;
    HvmCtl c;
    memset(&c, 0, sizeof c);
    c.n_in = n;
    c.in_head = std::min<u32>(n, nwaves * HVM_WAVE);
    c.fixed_steps = std::min(cfg.turn_steps, 1u << 20);
    c.idle_quit = c.idle_back = turn_args().idle_quit;
    c.max_steps = cfg.turn_steps ? std::min(cfg.turn_steps, 1u << 20) : hvm_turn_policy_host(n, nwaves);
    host_ctl = c;
    if (!ck(hipMemcpyAsync(ctl, &host_ctl, sizeof(HvmCtl), hipMemcpyHostToDevice, stream), "boot ctl")) return false;
    return ck(hipStreamSynchronize(stream), "boot sync");
  }

  bool fill_out(hvm_hip_out* out, double t0, double t_eval) {
    HvmDevStats s;
    const bool good = read_stats(s);
    memset(out, 0, sizeof *out);
    out->itrs = s.itrs;
    for (int k = 0; k < HVM_HIP_KINDS; ++k) out->kind[k] = s.kind[k];
    out->bytes = s.bytes;
    out->steps = s.steps; out->spills = s.spills; out->defers = s.defers;
    out->retries = s.retries; out->probes = s.probes;
    for (int k = 0; k < 6; ++k) out->phase_ms[k] = (double)s.phase[k] / (double)clock_khz;
    out->kernel_ms = kernel_ms;
    out->eval_ms = (t_eval - t0) * 1e3;
    out->turns = turns;
    out->launches = launches;
    out->launch_ms = launch_ms;
    out->waves = nwaves;
    out->leak = s.leaks; out->saves = s.saves; out->local_left = s.local_left;
    out->normalizes = normalizes;
    return good;
  }

  int eval_batch(const u32* entries, u32 n, const uint64_t* est, hvm_hip_out* outs, hvm_hip_out* total) {
    if (wide) { fail("batched evaluation runs on the 32-bit evaluator (no HVM_HIP_WIDE)"); return HVM_HIP_E_ARG; }
    if (!loaded) { fail("no book loaded"); return HVM_HIP_E_ARG; }
    if (!entries || !outs || n == 0) { fail("empty batch"); return HVM_HIP_E_ARG; }
    for (u32 j = 0; j < n; ++j) {
      if (entries[j] >= ndefs) { fail("batch entry is not a def"); return HVM_HIP_E_ARG; }
      memset(&outs[j], 0, sizeof outs[j]);
    }
    kernel_ms = 0;
    launch_ms = 0;
    turns = 0;
    launches = 0;
    normalizes = 0;
    eval_started = false;
    if (!ck(hipMemsetAsync(stats, 0, sizeof(HvmDevStats) * HVM_STAT_STRIPES, stream), "clear stats")) return HVM_HIP_E_DEVICE;
    double t0 = now_s();
    deadline = cfg.time_limit_s ? t0 + cfg.time_limit_s : 0;
    // admission budget: a quarter of the node capacity, leaving headroom for
    // the breadth-first growth of parallel evaluation and for the pool
    const double budget = 0.25 * (double)node_cap;
    u32 start = 0;
    while (start < n) {
      u32 end = start;
      double sum = 0;
      while (end < n && end - start < HVM_HIP_MAX_NETS &&
             (end == start || !est || sum + (double)est[end] <= budget)) {
        sum += est ? (double)est[end] : 0.0;
        ++end;
      }
      if (!reset_heap()) return HVM_HIP_E_DEVICE;
      if (!boot_batch(entries + start, end - start)) return HVM_HIP_E_DEVICE;
      if (!normalize()) return HVM_HIP_E_EVAL;
      // Per-net readback, as the IO driver does for its one root (src/run.cu:857-925).
      for (u32 j = start; j < end; ++j) {
        u32 root = hvm_new_port(HVM_VAR, 1 + (j - start));
        Ctr c;
        if (!readback_ctr(peek(root), c)) return HVM_HIP_E_EVAL;
        if (c.args_len >= 1 && hvm_get_tag(c.args[0]) == HVM_CON && c.tag == 1 && c.args_len == 4) {
          u64 magic = node_load(hvm_get_val(c.args[0]));
          if (copy_failed) return HVM_HIP_E_DEVICE;
          if (hvm_get_val(hvm_get_fst(magic)) == hvm_new_u24(0xD0CA11) &&
              hvm_get_val(hvm_get_snd(magic)) == hvm_new_u24(0xFF1FF1)) {
            outs[j].err = HVM_HIP_E_IO;
            continue;
          }
        }
        if (cfg.flags & HVM_HIP_NO_READBACK) continue;
        std::string text;
        if (!print_result(peek(root), text)) return HVM_HIP_E_EVAL;
        outs[j].result = dup_cstr(text);
        std::string canon;
        try { canon = hvm::canonical_tree_text(text); } catch (std::exception& ex) { canon = "<unparsable>"; }
        outs[j].canonical = dup_cstr(canon);
      }
      start = end;
    }
    if (total) {
      if (!fill_out(total, t0, now_s())) return HVM_HIP_E_DEVICE;
      total->wall_s = now_s() - t0;
    }
    return HVM_HIP_OK;
  }
};

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" hvm_hip_ctx* hvm_hip_create(const hvm_hip_cfg* cfg) {
  hvm_hip_ctx* c = new hvm_hip_ctx();
  if (!c->init(cfg)) {
    g_create_error = c->last_error;
    c->release();
    delete c;
    return nullptr;
  }
  return c;
}

extern "C" int hvm_hip_load(hvm_hip_ctx* ctx, const uint32_t* book_buf, size_t n_words) {
  if (!ctx) return HVM_HIP_E_ARG;
  return ctx->load(book_buf, n_words) ? HVM_HIP_OK : HVM_HIP_E_BOOK;
}

extern "C" int hvm_hip_set_scheduler(hvm_hip_ctx* ctx, uint32_t mode, uint32_t seed, uint32_t explore_pct,
                                     const uint16_t* params, uint32_t trace_cap) {
  if (!ctx) return HVM_HIP_E_ARG;
  if (ctx->wide) {
    // the wide evaluator runs fixed-length turns ended by idle quitting
    ctx->last_error = "scheduler modes apply to the 32-bit evaluator (no HVM_HIP_WIDE)";
    return mode == HVM_HIP_SCHED_PLAIN ? HVM_HIP_OK : HVM_HIP_E_ARG;
  }
  return ctx->set_scheduler(mode, seed, explore_pct, params, trace_cap) ? HVM_HIP_OK : HVM_HIP_E_ARG;
}

extern "C" int hvm_hip_read_trace(hvm_hip_ctx* ctx, float* recs, uint32_t cap, uint32_t* n_recs) {
  if (!ctx || ctx->wide) return HVM_HIP_E_ARG;
  return ctx->read_trace(recs, cap, n_recs);
}

extern "C" int hvm_hip_eval_batch(hvm_hip_ctx* ctx, const uint32_t* entries, uint32_t n_nets,
                                  const uint64_t* est_nodes, hvm_hip_out* outs, hvm_hip_out* total) {
  if (!ctx) return HVM_HIP_E_ARG;
  return ctx->eval_batch(entries, n_nets, est_nodes, outs, total);
}

extern "C" int hvm_hip_eval(hvm_hip_ctx* ctx, hvm_hip_out* out) {
  if (!ctx) return HVM_HIP_E_ARG;
  return ctx->eval(out);
}

extern "C" void hvm_hip_destroy(hvm_hip_ctx* ctx) {
  if (!ctx) return;
  ctx->release();
  delete ctx;
}

extern "C" const char* hvm_hip_last_error(const hvm_hip_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : g_create_error.c_str();
}

extern "C" void hvm_hip_free(void* p) { free(p); }

extern "C" int hvm_hip_run(const uint32_t* book_buf, size_t n_words, const hvm_hip_cfg* cfg, hvm_hip_out* out) {
  if (!book_buf) return HVM_HIP_E_ARG;
  hvm_hip_ctx* c = hvm_hip_create(cfg);
  if (!c) return HVM_HIP_E_DEVICE;
  int rc = c->load(book_buf, n_words) ? HVM_HIP_OK : HVM_HIP_E_BOOK;
  if (rc == HVM_HIP_OK) rc = c->eval(out);
  hvm_hip_destroy(c);
  return rc;
}

// ---------------------------------------------------------------------------
// No-allocation entry (hvm_hip_run_static): the net, bags, control blocks,
// book and readback lists live in `scratch`, a static device array of
// HVM_HIP_STATIC_BYTES that each GPU the library runs on holds until the
// process ends.  Its device address is looked up once per device with that
// device current; host code passes that address, never the array's name.
// Every call on a device shares that device's array: calls that overlap on
// one device race, and each call clears what must start at zero.
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[HVM_HIP_STATIC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};

// The device address of this device's scratch array (device `dev` current).
bool scratch_ptr(int dev, void** out, std::string& err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) { err = "device index out of range"; return false; }
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return false;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return true;
}
}  // namespace

extern "C" int hvm_hip_run_static(const uint32_t* book_buf, size_t n_words, const hvm_hip_cfg* cfg, hvm_hip_out* out,
                                  char* result, size_t result_cap) {
  if (!book_buf || !out) { g_create_error = "hvm_hip_run_static: NULL book or output"; return HVM_HIP_E_ARG; }
  memset(out, 0, sizeof *out);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return HVM_HIP_E_DEVICE; }
  if (hipSetDevice(0) != hipSuccess) { g_create_error = "hipSetDevice failed"; return HVM_HIP_E_DEVICE; }
  void* base = nullptr;
  if (!scratch_ptr(0, &base, g_create_error)) return HVM_HIP_E_DEVICE;
  hvm_hip_cfg c{};
  if (cfg) c = *cfg;
  hvm_hip_ctx ctx;
  int rc = HVM_HIP_OK;
  if (!ctx.init_static(&c, base, HVM_HIP_STATIC_BYTES)) {
    rc = ctx.last_error.rfind("hvm_hip_run_static", 0) == 0 ? HVM_HIP_E_ARG : HVM_HIP_E_DEVICE;
  } else if (!ctx.load(book_buf, n_words)) {
    rc = ctx.last_error.find("static region") != std::string::npos ? HVM_HIP_E_ARG : HVM_HIP_E_BOOK;
  } else {
    rc = ctx.eval(out);
  }
  if (rc == HVM_HIP_OK) {
    // the printed result goes to the caller's buffer
    const char* text = out->result ? out->result : "";
    const size_t n = strlen(text);
    if (result && result_cap > n) {
      memcpy(result, text, n + 1);
    } else if (result || !(c.flags & HVM_HIP_NO_READBACK)) {
      ctx.last_error = "hvm_hip_run_static: result buffer too small (" + std::to_string(n + 1) + " bytes needed)";
      rc = HVM_HIP_E_ARG;
    }
  }
  free(out->result);
  free(out->canonical);
  out->result = out->canonical = nullptr;
  if (rc != HVM_HIP_OK) g_create_error = ctx.last_error;
  ctx.release();
  return rc;
}

extern "C" void hvm_hip(uint32_t* book_buffer) {
  if (!book_buffer) {
    fprintf(stderr, "failed to load book\n");
    return;
  }
  hvm_hip_ctx* c = hvm_hip_create(nullptr);
  if (!c) {
    fprintf(stderr, "Failed to initialise the GPU runtime: %s\n", g_create_error.c_str());
    exit(EXIT_FAILURE);
  }
  if (!c->load(book_buffer, (size_t)-1)) {
    fprintf(stderr, "%s\n", c->last_error.c_str());
    fprintf(stderr, "failed to load book\n");
    hvm_hip_destroy(c);
    return;
  }
  setlinebuf(stdout);
  setlinebuf(stderr);
  hvm_hip_out o;
  int rc = c->eval(&o);
  if (rc != HVM_HIP_OK) {
    fprintf(stderr, "Failed to run the evaluator: %s\n", c->last_error.c_str());
    hvm_hip_destroy(c);
    exit(EXIT_FAILURE);
  }
  printf("Result: %s\n", o.result ? o.result : "");
  printf("- ITRS: %llu\n", (unsigned long long)o.itrs);
  printf("- LEAK: %llu\n", (unsigned long long)o.leak);
  printf("- TIME: %.2fs\n", o.wall_s);
  printf("- MIPS: %.2f\n", o.wall_s > 0 ? (double)o.itrs / o.wall_s / 1e6 : 0.0);
  free(o.result);
  free(o.canonical);
  hvm_hip_destroy(c);
}

// ---------------------------------------------------------------------------
// Front end C ABI
// ---------------------------------------------------------------------------

extern "C" int hvm_front_compile(const char* code, uint32_t** words, size_t* n_words, char** err) {
  if (words) *words = nullptr;
  if (n_words) *n_words = 0;
  if (err) *err = nullptr;
  if (!code || !words || !n_words) return HVM_HIP_E_ARG;
  try {
    hvm::AstBook ast = hvm::parse_book(code);
    hvm::Book b = hvm::build_book(ast);
    std::vector<uint32_t> w = hvm::to_buffer(b);
    uint32_t* p = (uint32_t*)malloc(w.size() * 4 + 4);
    memcpy(p, w.data(), w.size() * 4);
    *words = p;
    *n_words = w.size();
    return 0;
  } catch (std::exception& ex) {
    if (err) *err = dup_cstr(ex.what());
    return 1;
  }
}

extern "C" int hvm_front_specialize(const uint32_t* words, size_t n_words, uint32_t node_budget,
                                    uint32_t** out_words, size_t* out_n, uint32_t** counts, char** err) {
  if (out_words) *out_words = nullptr;
  if (out_n) *out_n = 0;
  if (counts) *counts = nullptr;
  if (err) *err = nullptr;
  if (!words || !out_words || !out_n || !counts || node_budget > HVM_GPU_DEF_MAX) return HVM_HIP_E_ARG;
  hvm::Book b;
  std::string e;
  if (!hvm::load_buffer(words, n_words, b, e)) {
    if (err) *err = dup_cstr(e.c_str());
    return HVM_HIP_E_BOOK;
  }
  std::vector<hvm::SpecCounts> spec;
  hvm::Book sb = hvm::specialize_book(b, node_budget, HVM_GPU_DEF_MAX, spec);
  std::vector<uint32_t> w = hvm::to_buffer(sb);
  uint32_t* p = (uint32_t*)malloc(w.size() * 4 + 4);
  uint32_t* c = (uint32_t*)malloc(spec.size() * 12 + 4);
  if (!p || !c) { free(p); free(c); return HVM_HIP_E_ARG; }
  memcpy(p, w.data(), w.size() * 4);
  for (size_t f = 0; f < spec.size(); ++f) {
    c[3 * f] = spec[f].call;
    c[3 * f + 1] = spec[f].anni;
    c[3 * f + 2] = spec[f].bytes;
  }
  *out_words = p;
  *out_n = w.size();
  *counts = c;
  return HVM_HIP_OK;
}

extern "C" int hvm_front_canonical(const char* printed, char** out, char** err) {
  if (out) *out = nullptr;
  if (err) *err = nullptr;
  if (!printed || !out) return HVM_HIP_E_ARG;
  try {
    *out = dup_cstr(hvm::canonical_tree_text(printed));
    return 0;
  } catch (std::exception& ex) {
    if (err) *err = dup_cstr(ex.what());
    return 1;
  }
}
