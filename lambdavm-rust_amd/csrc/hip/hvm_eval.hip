// hvm_eval.hip — the HVM2 evaluator for gfx950 (MI355X).
//
// One kernel launch is one *turn*.  Every wave (64 lanes) owns:
//   * a redex deque per lane in LDS (HVM_DEQ entries, LIFO pop at the top,
//     sharing from the bottom), fed from the HBM input bag;
//   * a node slice and a var slice of the HBM net, from which it allocates
//     with a wave-cooperative scan (one coalesced 64-slot load per window).
// Each *step* every lane with a redex performs one interaction.  The rule code
// runs as lockstep phases shared by all rules, so the 64 lanes issue their
// memory operations together instead of serialising one rule at a time:
//   P0 peek + classify (rule table, swap)       src/hvm.c:1073-1095
//   P1 allocation-table reservation (wave scan)
//   P2 wave-cooperative node/var allocation     replaces src/hvm.c:713-780
//   P3 node takes (64-bit atomic exchange)      src/hvm.c:909-1069
//   P4 node/var creation (device-scope stores)
//   P5 release: wait for the creation stores (or an agent release fence)
//   P6 links (32-bit exchanges on vars)         src/hvm.c:784-846
//   P7 counting (ITRS += 1 per non-LINK rule)   src/hvm.c:1116-1118
// Between steps, idle lanes take the oldest redex of a lane holding two or
// more (in-wave sharing through LDS).  Redexes that overflow a deque, and all
// redexes left at the end of a turn, go to the HBM output bag, which the host
// hands to the next turn; a turn with an empty output bag ends normalisation.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "hvm_dev.hpp"
#include "hvm_wave.hpp"

typedef unsigned long long u64;
typedef uint32_t u32;

#define AGENT __HIP_MEMORY_SCOPE_AGENT

// Steps between a busy wave's looks at the idle-quit count: a power of two
// of at most 64, so that the looks include every 64th step, where the wave
// also reads the error flag.
#ifndef HVM_QUIT_EVERY
#define HVM_QUIT_EVERY 16u
#endif
static_assert((HVM_QUIT_EVERY & (HVM_QUIT_EVERY - 1u)) == 0 && HVM_QUIT_EVERY <= 64u, "HVM_QUIT_EVERY");

// Mid-turn pool (HBM evaluator, HVM_MID_POOL): at its idle-quit looks a busy
// wave hands the oldest redex of every lane holding three or more to the
// output bag while fewer than HVM_POOL_LOW handed-over redexes wait there, and
// a wave that runs dry first claims up to a wave's worth of the handed-over
// slots (a cursor in HvmCtl, advanced by compare-and-swap to no further than
// the bag's count) and takes each by an exchange with 0.  Bag slots outside
// the live range are kept zero for this (every input-bag read is an exchange
// with 0, and the runtime clears both bags with the heap), so a slot whose
// store has not landed reads 0 and keeps its redex for the next turn, which
// skips the zeros.  No wave waits for another.
#ifndef HVM_POOL_LOW
#define HVM_POOL_LOW 4096u
#endif
// An input-bag slot read by the turn that consumes it (cleared with the mid-turn pool).
__device__ __forceinline__ u64 bag_take(u64* p) {
  if (HVM_MID_POOL) return __hip_atomic_exchange(p, 0ull, __ATOMIC_RELAXED, AGENT);
  return *p;
}
// An output-bag write.  With the mid-turn pool it is a device-scope store:
// a plain store can sit in the writer's L2 and be written back after a
// pull's exchange took the slot, which would hand the redex out twice.
__device__ __forceinline__ void bag_put(u64* p, u64 v) {
  if (HVM_MID_POOL) __hip_atomic_store(p, v, __ATOMIC_RELAXED, AGENT);
  else *p = v;
}

// ---------------------------------------------------------------------------
// Net accessors: every shared node/var access is a device-scope atomic.
// ---------------------------------------------------------------------------

__device__ __forceinline__ u32 vidx(u32 val) { return val == HVM_ROOT_VAL ? 0u : val; }

// Per-turn hash of the scheduler seed (placement rotation, turn-length jitter
// and exploration draws).
__host__ __device__ __forceinline__ u32 sched_hash(u32 seed, u32 turn) {
  u32 h = (seed * 0x9E3779B9u) ^ (turn * 0x85EBCA6Bu);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

// Every access is bounds-checked against the net's capacity: a port that
// points outside the net reaches the spare slot at index cap (each buffer has
// one past its capacity), never memory the net does not own.  The check is a
// branch-free clamp, so that independent accesses stay in flight together;
// hvm_between_kernel raises HVM_ERR_BAD_PORT once a spare slot was written,
// and a take through such a port raises it on the retry path.
__device__ __forceinline__ u32 node_slot(const HvmTurnArgs& A, u32 loc) { return min(loc, A.node_cap); }
__device__ __forceinline__ u32 vars_slot(const HvmTurnArgs& A, u32 val) { return min(vidx(val), A.vars_cap); }
__device__ __forceinline__ u64 node_take(const HvmTurnArgs& A, u32 loc) {
  return __hip_atomic_exchange(&A.node[node_slot(A, loc)], 0ull, __ATOMIC_RELAXED, AGENT);
}
// A take as a load: only the lane holding the redex that points at a node
// ever reaches it, so reading it and zeroing the slot once the interaction
// commits (node_put of 0) is a take; nothing is removed before then, so a
// lane that retries has nothing to put back.
__device__ __forceinline__ u64 node_peek(const HvmTurnArgs& A, u32 loc) {
  return __hip_atomic_load(&A.node[node_slot(A, loc)], __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ void node_put(const HvmTurnArgs& A, u32 loc, u64 v) {
  __hip_atomic_store(&A.node[node_slot(A, loc)], v, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ u32 vars_xchg(const HvmTurnArgs& A, u32 val, u32 x) {
  return __hip_atomic_exchange(&A.vars[vars_slot(A, val)], x, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ void vars_put(const HvmTurnArgs& A, u32 val, u32 x) {
  __hip_atomic_store(&A.vars[vars_slot(A, val)], x, __ATOMIC_RELAXED, AGENT);
}

// ---------------------------------------------------------------------------
// Per-wave LDS state
// ---------------------------------------------------------------------------

// HBM evaluator: node slots that the wave's interactions consumed, and var
// slots that its links freed, are kept on per-wave LDS stacks of up to
// HVM_NODE_RECYCLE and HVM_VAR_RECYCLE entries, not zeroed, and the wave's
// next allocations take them first; what is left is zeroed (freed) at the end
// of the turn.  A consumed node or a freed var is named by nothing else, and
// its slot stays non-zero (a node, or the port or NONE the link's exchange
// left), so no allocator hands it out meanwhile.  0 zeroes every such slot at
// once.  The two stacks share what LDS 8 waves per CU leave (about 1 KiB).
// Measured (median of 5 evaluations, two alternating rounds): 256 node slots
// and no var stack is 1% faster on the sorts, the batch and church_loop than
// 192 + 64, so the var stack is off by default.
#ifndef HVM_NODE_RECYCLE
#define HVM_NODE_RECYCLE 256
#endif
#ifndef HVM_VAR_RECYCLE
#define HVM_VAR_RECYCLE 0
#endif

struct WaveLds {
  u64 deq[HVM_DEQ * HVM_WAVE];  // deque entry j of lane l at [j*64 + l]
  u32 tbl_n[HVM_TBL_NODE + HVM_TBL_SLACK];
  u32 tbl_v[HVM_TBL_VARS + HVM_TBL_SLACK];
  u64 inbox[HVM_WAVE];
  u32 target[HVM_WAVE];
  u32 cursor_n, cursor_v;       // allocator scratch
#if HVM_NODE_RECYCLE
  u32 rec[HVM_NODE_RECYCLE];    // consumed node slots, not yet zeroed
#endif
#if HVM_VAR_RECYCLE
  u32 rec_v[HVM_VAR_RECYCLE];   // var slots freed by links, not yet zeroed
  u32 rec_vn;                   // pushes so far (entries: min(rec_vn, HVM_VAR_RECYCLE))
#endif
};

struct Lane {
  u32 bot, top;                 // deque ring counters
  u64 bytes;
  u32 spills;
  u64 spec;                     // static CALLs (low half) and ANNIs (high half)
                                // of the compiled defs this lane expanded
  u32 fsw0, fswn;               // SWITCHes performed inside fast CALLs
  u32 voids;                    // VOID redexes counted where they were created
};

// Wave-wide interaction counts (wave-uniform, kept in scalar registers).
struct WaveCounts {
  u64 kind[HVM_K_KINDS];
  u64 leaks, saves;             // local-net evaluator only
};

// ---------------------------------------------------------------------------
// Local net (LDS) of one wave, for the local-net evaluator.  Slot i of the
// node (var) buffer is addressed by port value HVM_LOCAL_BASE + i.  A local
// port is meaningful only inside its wave, so no HBM node or var ever holds
// one: a link that would store a local node into an HBM var becomes a LEAK
// interaction, which copies the node to HBM first (src/hvm.cu:1129-1175).
// Free slots are kept on two LDS stacks.
// ---------------------------------------------------------------------------

struct LocalNet {
  u64 node[HVM_LNODE];
  u32 vars[HVM_LVARS];
  uint16_t nfree[HVM_LNODE];
  uint16_t vfree[HVM_LVARS];
  u32 ntop, vtop;               // free slots on each stack
  u32 sp;                       // entries on the wave's private HBM spill stack
  u32 xtn[HVM_TBL_X], xtv[HVM_TBL_X];   // HBM slots of one step's LEAK / SAVE
};
static_assert((HVM_LNODE & (HVM_LNODE - 1)) == 0 && HVM_LNODE <= 65536, "HVM_LNODE: power of two <= 2^16");
static_assert((HVM_LVARS & (HVM_LVARS - 1)) == 0 && HVM_LVARS <= 65536, "HVM_LVARS: power of two <= 2^16");
static_assert(HVM_GPU_DEF_MAX <= HVM_LNODE && HVM_GPU_DEF_MAX <= HVM_LVARS, "a def must fit a local net");

#define WGS __HIP_MEMORY_SCOPE_WORKGROUP

__device__ __forceinline__ bool is_loc_val(u32 v) { return v >= HVM_LOCAL_BASE && v != HVM_ROOT_VAL; }
__device__ __forceinline__ bool is_loc_port(u32 p) {
  u32 t = hvm_get_tag(p);
  return (t == HVM_VAR || t >= HVM_CON) && is_loc_val(hvm_get_val(p));
}
// Local slot of a local port (masked: always inside the LDS buffers).
__device__ __forceinline__ u32 loc_n(const HvmTurnArgs& A, u32 v) { return (v - HVM_LOCAL_BASE) & (HVM_LNODE - 1); }
__device__ __forceinline__ u32 loc_v(const HvmTurnArgs& A, u32 v) { return (v - HVM_LOCAL_BASE) & (HVM_LVARS - 1); }

// Net accessors of the evaluator: local slots in LDS (LOC only), else HBM.
template <bool LOC>
__device__ __forceinline__ u64 net_take(const HvmTurnArgs& A, LocalNet* N, u32 loc) {
  if (LOC && is_loc_val(loc)) return __hip_atomic_exchange(&N->node[loc_n(A, loc)], 0ull, __ATOMIC_RELAXED, WGS);
  return node_take(A, loc);
}
template <bool LOC>
__device__ __forceinline__ void net_put(const HvmTurnArgs& A, LocalNet* N, u32 loc, u64 v) {
  if (LOC && is_loc_val(loc)) { __hip_atomic_store(&N->node[loc_n(A, loc)], v, __ATOMIC_RELAXED, WGS); return; }
  node_put(A, loc, v);
}
template <bool LOC>
__device__ __forceinline__ u32 var_xchg(const HvmTurnArgs& A, LocalNet* N, u32 val, u32 x) {
  if (LOC && is_loc_val(val)) return __hip_atomic_exchange(&N->vars[loc_v(A, val)], x, __ATOMIC_RELAXED, WGS);
  return vars_xchg(A, val, x);
}
// A coherent read of a var (the link loop's ENTER step; see HVM_ENTER_LOAD).
template <bool LOC>
__device__ __forceinline__ u32 var_load(const HvmTurnArgs& A, LocalNet* N, u32 val) {
  if (LOC && is_loc_val(val)) return __hip_atomic_load(&N->vars[loc_v(A, val)], __ATOMIC_RELAXED, WGS);
  return __hip_atomic_load(&A.vars[vars_slot(A, val)], __ATOMIC_RELAXED, AGENT);
}
template <bool LOC>
__device__ __forceinline__ void var_put(const HvmTurnArgs& A, LocalNet* N, u32 val, u32 x) {
  if (LOC && is_loc_val(val)) { __hip_atomic_store(&N->vars[loc_v(A, val)], x, __ATOMIC_RELAXED, WGS); return; }
  vars_put(A, val, x);
}
// Frees a var slot: HBM slots are zeroed, local slots go back on the stack.
template <bool LOC>
__device__ __forceinline__ void var_free(const HvmTurnArgs& A, LocalNet* N, u32 val) {
  if (LOC && is_loc_val(val)) {
    u32 i = loc_v(A, val);
    __hip_atomic_store(&N->vars[i], 0u, __ATOMIC_RELAXED, WGS);
    u32 p = atomicAdd(&N->vtop, 1u);
    N->vfree[p & (HVM_LVARS - 1)] = (uint16_t)i;
    return;
  }
  vars_put(A, val, 0u);
}
// Frees a var slot that a link owns (its content is the port or NONE the
// link's exchange left there): onto the wave's recycling stack, or zeroed.
template <bool LOC>
__device__ __forceinline__ void var_release(WaveLds& L, const HvmTurnArgs& A, LocalNet* N, u32 val) {
#if HVM_VAR_RECYCLE && HVM_EARLY_TAKES && HVM_PEEK_TAKES
  if (!LOC) {
    const u32 i = vidx(val);
    if (i >= A.root_vars && i < A.vars_cap) {
      const u32 p = atomicAdd(&L.rec_vn, 1u);
      if (p < HVM_VAR_RECYCLE) { L.rec_v[p] = i; return; }
    }
  }
#endif
  var_free<LOC>(A, N, val);
}
// Frees a node slot whose content was taken or never written.
template <bool LOC>
__device__ __forceinline__ void node_free(const HvmTurnArgs& A, LocalNet* N, u32 loc) {
  if (LOC && is_loc_val(loc)) {
    u32 i = loc_n(A, loc);
    __hip_atomic_store(&N->node[i], 0ull, __ATOMIC_RELAXED, WGS);
    u32 p = atomicAdd(&N->ntop, 1u);
    N->nfree[p & (HVM_LNODE - 1)] = (uint16_t)i;
    return;
  }
  node_put(A, loc, 0ull);
}

__device__ __forceinline__ u64& deq_at(WaveLds& L, u32 idx) {
  return L.deq[(idx & (HVM_DEQ - 1)) * HVM_WAVE + lane_id()];
}

// Pushes a redex on this lane's deque; when full, the oldest redex goes to
// the HBM output bag.
// A redex holding a local port stays with its wave: with the local net it
// goes to the wave's private HBM stack instead.
template <bool LOC>
__device__ __forceinline__ void push_redex(WaveLds& L, LocalNet* N, Lane& ln, const HvmTurnArgs& A, u64 r) {
  if (ln.top - ln.bot == HVM_DEQ) {
    u64 old = deq_at(L, ln.bot);
    ln.bot++;
    if (LOC && (is_loc_port(hvm_get_fst(old)) || is_loc_port(hvm_get_snd(old)))) {
      u32 pos = atomicAdd(&N->sp, 1u);
      if (pos < HVM_LSPILL) A.lspill[(u64)blockIdx.x * HVM_LSPILL + pos] = old;
      else atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
    } else {
      // one bag reservation for all lanes spilling at this push
      const u64 m = __ballot(1);
      const u32 leader = (u32)__ffsll((long long)m) - 1u;
      u32 base = 0;
      if (__lane_id() == leader) base = atomicAdd(&A.ctl->n_out, (u32)__popcll(m));
      const u32 pos = lane_u32(base, (int)leader) + rank_in(m);
      if (pos < A.bag_cap) bag_put(&A.bag[A.ctl->parity ^ 1u][pos], old);
      else atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
    }
    ln.spills++;
  }
  deq_at(L, ln.top) = r;
  ln.top++;
}

// ---------------------------------------------------------------------------
// Linker (src/hvm.c:796-842)
// ---------------------------------------------------------------------------

#define HVM_LINK_CAP (1u << 16)

template <bool LOC>
__device__ __forceinline__ u32 enter(const HvmTurnArgs& A, LocalNet* N, u32 var) {
  for (u32 it = 0; it < HVM_LINK_CAP; ++it) {
    if (hvm_get_tag(var) != HVM_VAR) return var;
    u32 val = var_xchg<LOC>(A, N, hvm_get_val(var), HVM_NONE);
    if (val == HVM_NONE || val == 0) return var;
    var_free<LOC>(A, N, hvm_get_val(var));
    var = val;
  }
  atomicOr(&A.ctl->err, (u32)HVM_ERR_LINK_LOOP);
  return var;
}

// Links up to HVM_LB pairs of one interaction together.  Each pair runs the
// reference's link loop (src/hvm.c:812-842) as a small state machine; every
// round, each unfinished pair issues its next exchange, and all of a lane's
// exchanges of the round are in flight at once.  Interleaving the links of one
// interaction is the same as running them on different threads, which the
// lock-free linker allows.  A link ends either by publishing its var or with
// one new redex; the redexes are pushed in pair order once every pair is done,
// so the deque receives them in the order a one-at-a-time linker would push
// them (the deque is LIFO, and that order keeps the evaluation depth-first).
#ifndef HVM_LB
#define HVM_LB 4
#endif
enum { LK_DONE = 0, LK_ENTER = 1, LK_PUBLISH = 2, LK_ORIENT = 3 };
// Before publishing A ~> B the reference first follows B while it is a bound
// var (src/hvm.c:796-808).  With 0 a link publishes at once, one exchange
// fewer when B is an unbound var, but var chains are never shortened and grow
// without bound on programs such as examples/stress (the link-loop guard
// fires), so 1 is the only setting for general programs.
#ifndef HVM_LINK_ENTER
#define HVM_LINK_ENTER 1
#endif
// ENTER reads B with a load instead of exchanging NONE into it.  An unbound B
// reads NONE either way (the exchange writes NONE over NONE).  A bound B holds
// the port its other end published, and that end is done with it, so only
// this link touches B again: it takes the port and frees B with a store, as
// after the exchange.  A load that still sees NONE after the other end's
// publish only makes A point to B, a chain a later ENTER follows.  It trades a
// memory-side exchange for a load: measured (median of 5 evaluations, two
// alternating rounds) 4-7% slower on every workload (LSD 2^18 205 -> 216 ms):
// a device-scope load's round trip is longer than an exchange's, and the
// link loop waits for it.  Off by default.
#ifndef HVM_ENTER_LOAD
#define HVM_ENTER_LOAD 0
#endif

// With the local net, a pair about to publish an HBM var A with a local B
// (src/hvm.cu:1061-1072) either swaps (B is a var: the local var points to
// the HBM one) or ends as a LEAK redex A ~ B (B is a node).
template <bool LOC>
__device__ __forceinline__ void link_batch(WaveLds& L, LocalNet* N, Lane& ln, const HvmTurnArgs& A, u32* pa, u32* pb, u32 n) {
  u32 st[HVM_LB];
  u32 got = 0;          // pairs that ended with a redex (pa[i] ~ pb[i])
  bool busy = true;     // some pair is not done
#pragma unroll
  for (int i = 0; i < HVM_LB; ++i) st[i] = (u32)i < n ? (u32)LK_ORIENT : (u32)LK_DONE;
  for (u32 round = 0; round < HVM_LINK_CAP; ++round) {
#pragma unroll
    for (int i = 0; i < HVM_LB; ++i) {
      if (st[i] == LK_ORIENT) {
        // A points to B: make A the var; two non-vars are a new redex
        if (hvm_get_tag(pa[i]) != HVM_VAR && hvm_get_tag(pb[i]) == HVM_VAR) { u32 x = pa[i]; pa[i] = pb[i]; pb[i] = x; }
        if (hvm_get_tag(pa[i]) != HVM_VAR) {
          st[i] = LK_DONE;
          got |= 1u << i;
        } else {
          st[i] = (HVM_LINK_ENTER && hvm_get_tag(pb[i]) == HVM_VAR) ? (u32)LK_ENTER : (u32)LK_PUBLISH;
        }
      }
    }
    if (LOC) {
#pragma unroll
      for (int i = 0; i < HVM_LB; ++i) {
        if (st[i] == LK_PUBLISH && !is_loc_val(hvm_get_val(pa[i])) && is_loc_port(pb[i])) {
          if (hvm_get_tag(pb[i]) == HVM_VAR) {
            u32 x = pa[i]; pa[i] = pb[i]; pb[i] = x;
            st[i] = HVM_LINK_ENTER ? (u32)LK_ENTER : (u32)LK_PUBLISH;   // B is now the HBM var
          } else {
            st[i] = LK_DONE;                        // LEAK redex
            got |= 1u << i;
          }
        }
      }
    }
    // one exchange per pair: extend B (B := NONE) or publish A ~> B (A := B)
    u32 r[HVM_LB];
#pragma unroll
    for (int i = 0; i < HVM_LB; ++i) {
      const bool ent = st[i] == LK_ENTER;
      const u32 v = hvm_get_val(ent ? pb[i] : pa[i]);
      const u32 x = ent ? (u32)HVM_NONE : pb[i];
      r[i] = 0;
      if (HVM_ENTER_LOAD && ent) r[i] = var_load<LOC>(A, N, v);
      else if (ent || st[i] == LK_PUBLISH) r[i] = var_xchg<LOC>(A, N, v, x);
    }
    busy = false;
#pragma unroll
    for (int i = 0; i < HVM_LB; ++i) {
      if (st[i] == LK_ENTER) {
        if (r[i] == HVM_NONE || r[i] == 0) {
          st[i] = LK_PUBLISH;                       // B is an unbound var
        } else {
          var_release<LOC>(L, A, N, hvm_get_val(pb[i]));  // we own both: free B
          pb[i] = r[i];
          st[i] = (HVM_LINK_ENTER && hvm_get_tag(pb[i]) == HVM_VAR) ? (u32)LK_ENTER : (u32)LK_PUBLISH;
        }
      } else if (st[i] == LK_PUBLISH) {
        if (r[i] == HVM_NONE) {
          st[i] = LK_DONE;                          // the other side will find B
        } else {
          var_release<LOC>(L, A, N, hvm_get_val(pa[i]));  // A was bound: free it, link A' ~ B
          pa[i] = r[i];
          st[i] = LK_ORIENT;
        }
      }
      busy |= st[i] != LK_DONE;
    }
    if (!busy) break;
  }
  if (__builtin_expect(busy, 0)) atomicOr(&A.ctl->err, (u32)HVM_ERR_LINK_LOOP);
#pragma unroll
  for (int i = 0; i < HVM_LB; ++i) {
    if (!(got & (1u << i))) continue;
    // Two nullary ports (REF, ERA, NUM) make a VOID redex: the rule does
    // nothing (src/hvm.c:909-911), so it is counted here, where it is
    // created, instead of taking a lane-step later.  ITRS and the kind
    // counts are those of running it.
    if (hvm_get_tag(pa[i]) - 1u < 3u && hvm_get_tag(pb[i]) - 1u < 3u) ln.voids++;
    else push_redex<LOC>(L, N, ln, A, hvm_new_pair(pa[i], pb[i]));
  }
}

// ---------------------------------------------------------------------------
// Wave-cooperative allocator: scans the wave's slice 64 slots at a time,
// writing the first `need` free slot indices into `tbl`.  Only the owning
// wave ever turns a zero slot of its slice non-zero, so a slot read as zero
// is free (src/hvm.c:713-780 uses the same "zero means free" rule).
// `need` and `cur` are wave-uniform.
// ---------------------------------------------------------------------------

// Scan state of one buffer.  A wave first scans its own slice, where only it
// turns zero slots non-zero, so a zero read there is free.  When its slice
// has no room it scans the shared overflow pool, where every wave allocates:
// there a slot is claimed by compare-and-swap from zero.
struct Scan {
  u32 lo, len;        // range scanned (own slice, then the pool)
  u32 cur;            // next window start
  u32 got, need;      // slots found / wanted
  u32 scanned;        // slots looked at in the current range
  bool pool;          // scanning the overflow pool
  u32 slice_next;     // where the next scan of the slice starts, once it moved on to the pool
};

// Slots of its own slice a wave looks at per allocation before it moves on to
// the shared pool.  A dense slice is then scanned a few windows per step,
// continuing where the last scan stopped, instead of in full on every step
// (a full slice of 2^28 * 3/4 / 2048 slots is ~1,500 dependent window loads).
#ifndef HVM_SLICE_SCAN
#define HVM_SLICE_SCAN (8u * HVM_WAVE)
#endif

__device__ __forceinline__ u32 scan_idx(const Scan& s, u32 lane, u32& rel) {
  rel = (s.cur - s.lo + lane) % s.len;
  return s.lo + rel;
}

// Processes one 64-slot window (value v at idx for this lane).  Records free
// slots into tbl[got...]; when the window satisfies the request the lane
// holding the last slot taken writes the next cursor to *cursor_slot.
// A scan covers each slot of its range at most once: the slots it has handed
// out in this call are not created until the creation phase, so they still
// read zero, and a window that wrapped around onto them would hand them out a
// second time (two interactions writing one slot).
template <typename T>
__device__ __forceinline__ void scan_window(T* buf, Scan& s, T v, u32 idx, u32 rel, u32 reserved, T claim,
                                            u32* tbl, u32* cursor_slot) {
  bool cand = (v == 0) && (idx >= reserved) && (s.scanned + lane_id() < s.len);
  if (s.pool && cand) cand = atomicCAS(&buf[idx], (T)0, claim) == (T)0;
  u64 m = __ballot(cand);
  u32 r = rank_in(m);
  u32 nfree = __popcll(m);
  u32 want = s.need - s.got;
  if (cand && r < want) tbl[s.got + r] = idx;
  // claimed pool slots beyond the request go back to the pool
  if (s.pool && cand && r >= want) __hip_atomic_store(&buf[idx], (T)0, __ATOMIC_RELAXED, AGENT);
  if (nfree >= want) {
    if (cand && r == want - 1) *cursor_slot = s.lo + ((rel + 1) % s.len);
    s.got = s.need;
  } else {
    s.got += nfree;
    s.cur = s.lo + ((s.cur - s.lo + HVM_WAVE) % s.len);
    s.scanned += HVM_WAVE;
  }
}

// After a window that left the request open: moves from an exhausted slice
// (every slot looked at once) to the pool; returns false when the pool is
// exhausted too.
__device__ __forceinline__ bool scan_advance(Scan& s, u32 pool_lo, u32 pool_len, u32 pool_cur) {
  if (s.got >= s.need || s.scanned < (s.pool ? s.len : min(s.len, (u32)HVM_SLICE_SCAN))) return true;
  if (s.pool || pool_len < HVM_WAVE) return false;
  s.slice_next = s.cur;   // the next allocation continues the slice from here
  s.pool = true;
  s.lo = pool_lo;
  s.len = pool_len;
  s.cur = pool_cur;
  s.scanned = 0;
  return true;
}

// Allocates need_n node slots and need_v var slots for the whole wave.  Each
// round issues one coalesced probe load per buffer that still needs slots, so
// both buffers are scanned concurrently.  All arguments are wave-uniform.
// With `pre`, pre_n/pre_v are this lane's values of the first window of each
// buffer, loaded earlier at the same cursor.
// With rec_n > 0 the first node slots come from the wave's recycled ones.
__device__ __forceinline__ bool wave_alloc2(WaveLds& L, const HvmTurnArgs& A, u32 n_lo, u32 v_lo,
                                            HvmWaveState& cur, u32 need_n, u32 need_v, u64& probes, u32& oom,
                                            u32* tbl_n, u32* tbl_v, bool pre, u64 pre_n, u32 pre_v,
                                            u32& rec_n) {
  const u32 lane = lane_id();
  u32 got_n = 0;
#if HVM_NODE_RECYCLE
  if (rec_n && need_n) {
    got_n = min(need_n, rec_n);
    __syncthreads();
    for (u32 i = lane; i < got_n; i += HVM_WAVE) tbl_n[i] = L.rec[rec_n - got_n + i];
    rec_n -= got_n;
  }
#endif
  u32 got_v = 0;
#if HVM_VAR_RECYCLE
  if (need_v) {
    __syncthreads();
    const u32 vn = min(L.rec_vn, (u32)HVM_VAR_RECYCLE);
    got_v = min(need_v, vn);
    for (u32 i = lane; i < got_v; i += HVM_WAVE) tbl_v[i] = L.rec_v[vn - got_v + i];
    __syncthreads();
    if (lane == 0) L.rec_vn = vn - got_v;
  }
#endif
  Scan sn = {n_lo, A.node_slice, cur.node_cur, got_n, need_n, 0, false, cur.node_cur};
  Scan sv = {v_lo, A.vars_slice, cur.vars_cur, got_v, need_v, 0, false, cur.vars_cur};
  L.cursor_n = cur.node_cur;
  L.cursor_v = cur.vars_cur;
  while (sn.got < sn.need || sv.got < sv.need) {
    const bool pn = sn.got < sn.need, pv = sv.got < sv.need;
    u32 rn, rv;
    u32 in_ = scan_idx(sn, lane, rn);
    u32 iv = scan_idx(sv, lane, rv);
    u64 vn = !pn ? 1ull : pre ? pre_n : __hip_atomic_load(&A.node[in_], __ATOMIC_RELAXED, AGENT);
    u32 vv = !pv ? 1u : pre ? pre_v : __hip_atomic_load(&A.vars[iv], __ATOMIC_RELAXED, AGENT);
    pre = false;
    if (pn) {
      scan_window(A.node, sn, vn, in_, rn, 1u, ~0ull, tbl_n, &L.cursor_n);
      probes++;
      if (!scan_advance(sn, A.node_pool_lo, A.node_pool_hi - A.node_pool_lo, cur.node_ovf)) { oom = HVM_ERR_OOM_NODE; return false; }
    }
    if (pv) {
      scan_window(A.vars, sv, vv, iv, rv, A.root_vars, (u32)HVM_NONE, tbl_v, &L.cursor_v);
      probes++;
      if (!scan_advance(sv, A.vars_pool_lo, A.vars_pool_hi - A.vars_pool_lo, cur.vars_ovf)) { oom = HVM_ERR_OOM_VARS; return false; }
    }
  }
  __syncthreads();
  // a satisfied request left its cursor just after its last slot
  if (need_n) {
    if (sn.pool) { cur.node_ovf = L.cursor_n; cur.node_cur = sn.slice_next; }
    else cur.node_cur = L.cursor_n;
  }
  if (need_v) {
    if (sv.pool) { cur.vars_ovf = L.cursor_v; cur.vars_cur = sv.slice_next; }
    else cur.vars_cur = L.cursor_v;
  }
  return true;
}

// ---------------------------------------------------------------------------
// The turn kernel
// ---------------------------------------------------------------------------

// Extra rules of the local-net evaluator (device-only codes).
enum {
  HVM_LEAK = 8,   // HBM var ~ local node: copy the node to HBM (src/hvm.cu:1129-1175)
  HVM_SAVE = 9,   // end of turn: a redex with local ports gets HBM vars (src/hvm.cu:1426-1446)
  HVM_EMIT = 10,  // end of turn: a redex of HBM ports goes to the output bag as is
};

// A def template's port in a fresh instance: node and var indices of the
// template become the slots allocated for this CALL (src/hvm.c:351-364).
// Branch-free: one table read per port (slot 0 of the var table for ports
// that are neither), so that the reads of several ports go out together.
__device__ __forceinline__ u32 adjust_port(u32 p, const u32* tn, const u32* tv) {
  const bool nod = hvm_is_nod(p), var = hvm_is_var(p);
  const u32 x = (nod ? tn : tv)[(nod || var) ? hvm_get_val(p) : 0u];
  return (nod || var) ? hvm_new_port(hvm_get_tag(p), x) : p;
}
__device__ __forceinline__ u64 adjust_pair(u64 p, const u32* tn, const u32* tv) {
  return hvm_new_pair(adjust_port(hvm_get_fst(p), tn, tv), adjust_port(hvm_get_snd(p), tn, tv));
}

// Appends the link pair (p1, p2) at position *np of a batch.  The position
// is selected with constant indices so that the batch stays in registers
// (a runtime index would move it to scratch memory).
__device__ __forceinline__ void put_link(u32* pa, u32* pb, u32& np, u32 p1, u32 p2) {
#pragma unroll
  for (int k = 0; k < HVM_LB; ++k)
    if (np == (u32)k) { pa[k] = p1; pb[k] = p2; }
  np++;
}


#ifndef HVM_EARLY_TAKES
#define HVM_EARLY_TAKES 1
#endif
// HBM evaluator: every early take as a coherent load, the slot zeroed at
// commit (node_peek), instead of an exchange (lanes with two takes load in
// either mode).  A node is named by exactly one redex, so no other lane takes
// it concurrently.  Measured, median of 5 evaluations in two alternating
// rounds: LSD radix 2^18 2.7%, bitonic 2^16 3.7%, trie radix 2^18 3%, the
// 256-net batch 3%, church_loop(12, 12) 1.5% faster than exchanges, and
// memory-side atomics per interaction on the LSD sort 1.24 -> 1.13.  0 builds
// the exchanges.
#ifndef HVM_PEEK_TAKES
#define HVM_PEEK_TAKES 1
#endif

// Template pairs loaded together per chunk of a CALL's creation; the
// allocation tables are read up to HVM_XCHUNK - 1 entries past a def's own.
#ifndef HVM_XCHUNK
#define HVM_XCHUNK 8
#endif
static_assert(HVM_TBL_SLACK >= HVM_XCHUNK - 1, "table reads stay inside the tables");

// Steps a wave may spend draining its local net before giving up.
#define HVM_EVAC_CAP (1u << 16)

// FAST (compiled CALLs): a CALL whose caller's port is a CON node takes that
// node (with the other late takes, once the def record shows a CON root) and
// performs the ANNI between it and the def's root node instead of creating
// the root; if the root's first port is
// a SWI node whose first port is a CON node and the caller passes a number,
// the SWITCH and the ANNI of the branch pair happen too, without creating
// either node (src/cmp.rs:106-170 fast SWITCH, 230-262 fast ANNI).  Each is
// counted as the interaction it is.
template <bool LOC, bool FAST>
__device__ __forceinline__ void turn_body(const HvmTurnArgs& A, WaveLds& L, LocalNet* N) {
  const u32 lane = lane_id();
  const u32 w = blockIdx.x;

  // Turn parameters come from the control block (written by the host before
  // the first turn and by hvm_between_kernel after every turn).
  const HvmCtl* C = A.ctl;
  if (C->done) return;
  // A turn runs on the first `act` waves (the scheduler's choice); the others
  // end at once, their slices untouched.
  const u32 act = C->active ? min(C->active, A.nwaves) : A.nwaves;
  if (w >= act) return;
#if HVM_VAR_RECYCLE
  if (lane_id() == 0) L.rec_vn = 0;
  __syncthreads();
#endif
  const u32 n_in = C->n_in;
  const u32 max_steps = C->max_steps;
  const u32 idle_quit = C->idle_quit;
  const u32 idle_back = C->idle_back;
  u64* const bag_in = A.bag[C->parity];
  u64* bag_out = A.bag[C->parity ^ 1u];

  Lane ln;
  ln.bot = ln.top = 0;
  WaveCounts wc;
#pragma unroll
  for (int k = 0; k < HVM_K_KINDS; ++k) wc.kind[k] = 0;
  wc.leaks = wc.saves = 0;
  ln.bytes = 0;
  ln.spills = 0;
  ln.spec = 0;
  ln.fsw0 = ln.fswn = 0;
  ln.voids = 0;
  u64 probes = 0, steps = 0, defers = 0, retries = 0, lane_steps = 0, alloc_slots = 0;
  u32 left_tot = 0;   // redexes this wave leaves to the output bag
  // node slots consumed this turn and not yet handed out again or freed
  // (wave-uniform; HBM evaluator with early takes by load only)
  constexpr bool RECYCLE = !LOC && HVM_EARLY_TAKES && HVM_PEEK_TAKES && HVM_NODE_RECYCLE > 0;
  u32 rec_n = 0;
  constexpr bool POOL = !LOC && HVM_MID_POOL;

  if (LOC) {
    // an empty local net: every slot free
    for (u32 i = lane; i < HVM_LNODE; i += HVM_WAVE) { N->node[i] = 0; N->nfree[i] = (uint16_t)(HVM_LNODE - 1 - i); }
    for (u32 i = lane; i < HVM_LVARS; i += HVM_WAVE) { N->vars[i] = 0; N->vfree[i] = (uint16_t)(HVM_LVARS - 1 - i); }
    if (lane == 0) { N->ntop = HVM_LNODE; N->vtop = HVM_LVARS; N->sp = 0; }
    __syncthreads();
  }

  // Slices start after the reserved slots: node slot 0, var slots [0, root_vars).
  const u32 n_lo = 1 + w * A.node_slice, v_lo = A.root_vars + w * A.vars_slice;
  HvmWaveState ws = A.wave_state[w];
  if (ws.node_cur < n_lo || ws.node_cur >= n_lo + A.node_slice) ws.node_cur = n_lo;
  if (ws.vars_cur < v_lo || ws.vars_cur >= v_lo + A.vars_slice) ws.vars_cur = v_lo;
  // overflow-pool cursors start spread out, one region per wave
  if (ws.node_ovf < A.node_pool_lo || ws.node_ovf >= A.node_pool_hi)
    ws.node_ovf = A.node_pool_lo + (u32)(((u64)(A.node_pool_hi - A.node_pool_lo) * w) / A.nwaves);
  if (ws.vars_ovf < A.vars_pool_lo || ws.vars_ovf >= A.vars_pool_hi)
    ws.vars_ovf = A.vars_pool_lo + (u32)(((u64)(A.vars_pool_hi - A.vars_pool_lo) * w) / A.nwaves);

  // Initial spread: lane l of wave w takes item w' + l*W, so a small bag is
  // spread one redex per wave before any wave takes a second one.  A nonzero
  // scheduler seed rotates the waves (w' = w + a per-turn offset), so that
  // seeds give different placements, hence different schedules; results do
  // not depend on the placement.
  {
    const u32 seed = A.sched->seed;
    const u32 rot = seed ? sched_hash(seed, C->turns) % act : 0u;
    u32 idx = (w + rot) % act + lane * act;
    if (idx < n_in && idx < HVM_WAVE * act) {
      const u64 x = bag_take(&bag_in[idx]);
      if (!HVM_MID_POOL || x) push_redex<LOC>(L, N, ln, A, x);   // 0: taken from the pool
    }
  }
  bool more = n_in > HVM_WAVE * act;
  // With the local net a wave ends its turn by draining the net (evacuation):
  // high-priority redexes run, every other redex leaves through the output bag.
  bool evac = false;
  u32 evac_steps = 0;

  // optional per-phase timing (HVM_TURN_PROFILE), wall clock at 100 MHz
  const bool prof = (A.flags & HVM_TURN_PROFILE) != 0;
  u64 ph[6] = {0, 0, 0, 0, 0, 0};
  u64 t_prev = 0;
#define HVM_PHASE(k)                        \
  if (prof) {                               \
    u64 t_now = wall_clock64();             \
    ph[k] += t_now - t_prev;                \
    t_prev = t_now;                         \
  }

  for (u32 step = 0;; ++step) {
    if (!evac && step >= max_steps) {
      if (!LOC) break;
      evac = true;
    }
    if (evac && ++evac_steps > HVM_EVAC_CAP) {
      if (lane == 0) atomicOr(&A.ctl->err, (u32)HVM_ERR_EVAC);
      break;
    }
    if (prof) t_prev = wall_clock64();
    // ---- refill: the wave's private spill stack first, then the input bag ----
    if (LOC) {
      u32 sp = min(__hip_atomic_load(&N->sp, __ATOMIC_RELAXED, WGS), (u32)HVM_LSPILL);
      if (sp) {
        bool empty = ln.top == ln.bot;
        u64 em = __ballot(empty);
        u32 k = min((u32)__popcll(em), sp);
        u32 r = rank_in(em);
        u64 got = 0;
        if (empty && r < k) got = A.lspill[(u64)w * HVM_LSPILL + (sp - 1 - r)];
        __syncthreads();
        if (lane == 0) __hip_atomic_store(&N->sp, sp - k, __ATOMIC_RELAXED, WGS);
        __syncthreads();
        if (empty && r < k) push_redex<LOC>(L, N, ln, A, got);
      }
    }
    if (more && !evac) {
      bool empty = ln.top == ln.bot;
      u64 em = __ballot(empty);
      if (em) {
        u32 k = __popcll(em);
        u32 base = 0;
        if (lane == 0) base = atomicAdd(&A.ctl->in_head, k);
        base = lane_u32(base, 0);
        if (base >= n_in) more = false;
        u32 r = rank_in(em);
        if (empty && base + r < n_in) {
          const u64 x = bag_take(&bag_in[base + r]);
          if (!HVM_MID_POOL || x) push_redex<LOC>(L, N, ln, A, x);
        }
      }
    }

    // ---- in-wave sharing: idle lanes take the oldest redex of rich lanes ----
    {
      u32 n = ln.top - ln.bot;
      bool needy = n == 0, rich = n >= 2;
      u64 nm = __ballot(needy), rm = __ballot(rich);
      if (nm && rm) {
        u32 k = min((u32)__popcll(nm), (u32)__popcll(rm));
        u32 rn = rank_in(nm), rr = rank_in(rm);
        if (needy && rn < k) L.target[rn] = lane;
        __syncthreads();
        if (rich && rr < k) {
          u32 t = L.target[rr];
          L.inbox[t] = deq_at(L, ln.bot);
          ln.bot++;
        }
        __syncthreads();
        if (needy && rn < k) push_redex<LOC>(L, N, ln, A, L.inbox[lane]);
      }
    }

    bool has = ln.top != ln.bot;
    if (!__ballot(has)) {
      if (more && !evac) continue;
      if (POOL && !evac) {
        // before counting as idle: up to a wave of the redexes busy waves
        // handed to the output bag this turn
        // claim only what waits (a compare-and-swap of the cursor, a few
        // tries), so that the cursor never runs past the redexes handed over
        u32 t = 0, k = 0;
        if (lane == 0) {
          const u32 out_n = min(__hip_atomic_load(&A.ctl->n_out, __ATOMIC_RELAXED, AGENT), A.bag_cap);
          u32 cur = __hip_atomic_load(&A.ctl->out_take, __ATOMIC_RELAXED, AGENT);
          for (int tries = 0; tries < 4 && cur < out_n; ++tries) {
            const u32 want = min((u32)HVM_WAVE, out_n - cur);
            const u32 prev = atomicCAS(&A.ctl->out_take, cur, cur + want);
            if (prev == cur) { t = cur; k = want; break; }
            cur = prev;
          }
        }
        t = lane_u32(t, 0);
        k = lane_u32(k, 0);
        if (k) {
          u64 x = 0;
          if ((u32)lane < k) x = __hip_atomic_exchange(&bag_out[t + lane], 0ull, __ATOMIC_RELAXED, AGENT);
          if (x) push_redex<LOC>(L, N, ln, A, x);
          if (__ballot(x != 0)) continue;
        }
      }
      // out of work: count this wave idle, so that busy waves end the turn
      // early and their redexes get redistributed
      if (lane == 0) atomicAdd(&A.ctl->idle_waves, 1u);
      break;
    }
    if ((step & (HVM_QUIT_EVERY - 1u)) == HVM_QUIT_EVERY - 1u) {
      if (!evac) {
        // A turn that started without idle quit (few redexes, not growing:
        // independent sequential chains) gets it back once some wave holds
        // more redexes than two per lane: work grew inside the turn, and
        // only the end of the turn hands it to the idle waves.
        u32 iq = idle_quit;
        if (!iq && idle_back) {
          u32 mine = wave_sum64(ln.top - ln.bot);
          u32 sur = __hip_atomic_load(&A.ctl->surplus, __ATOMIC_RELAXED, AGENT);
          if (!sur && mine > 2u * HVM_WAVE) {
            if (lane == 0) __hip_atomic_store(&A.ctl->surplus, 1u, __ATOMIC_RELAXED, AGENT);
            sur = 1;
          }
          if (sur) iq = idle_back;
        }
        u32 idle = __hip_atomic_load(&A.ctl->idle_waves, __ATOMIC_RELAXED, AGENT);
        if (iq && idle >= iq) {
          if (!LOC) break;
          evac = true;
        }
        if (POOL) {
          // hand surplus to the pool while it runs low
          const u32 out_n = __hip_atomic_load(&A.ctl->n_out, __ATOMIC_RELAXED, AGENT);
          const u32 took = __hip_atomic_load(&A.ctl->out_take, __ATOMIC_RELAXED, AGENT);
          if (out_n <= took + HVM_POOL_LOW && out_n < A.bag_cap / 2u) {
            const bool give = ln.top - ln.bot >= 3u;
            const u64 gm = __ballot(give);
            if (gm) {
              u32 base = 0;
              if (lane == 0) base = atomicAdd(&A.ctl->n_out, (u32)__popcll(gm));
              const u32 pos = lane_u32(base, 0) + rank_in(gm);
              if (give) {
                if (pos < A.bag_cap) {
                  bag_put(&bag_out[pos], deq_at(L, ln.bot));
                  ln.bot++;
                } else {
                  atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
                }
              }
            }
          }
        }
      }
      if ((step & 63) == 63 && __hip_atomic_load(&A.ctl->err, __ATOMIC_RELAXED, AGENT) != 0) break;
    }
    steps++;
    lane_steps += __popcll(__ballot(has));
    HVM_PHASE(0);

    // ---- P0: peek + classify ----
    u64 rx = has ? deq_at(L, ln.top - 1) : 0ull;
    u32 a = hvm_get_fst(rx), b = hvm_get_snd(rx);
    u32 rule = hvm_get_rule(a, b);
    // A REF against a root var is a boot redex: CALL (src/hvm.c:1088-1090).
    // Var slots [1, root_vars) are the roots of batched nets.
    if (hvm_get_tag(a) == HVM_REF && hvm_get_tag(b) == HVM_VAR &&
        (b == HVM_ROOT || (hvm_get_val(b) != 0 && hvm_get_val(b) < A.root_vars))) rule = HVM_CALL;
    else if (hvm_should_swap(a, b)) { u32 x = a; a = b; b = x; }
    if (LOC) {
      // an HBM var against a local node
      if (rule == HVM_LINK && hvm_get_tag(a) == HVM_VAR && !is_loc_val(hvm_get_val(a)) &&
          hvm_get_tag(b) >= HVM_CON && is_loc_val(hvm_get_val(b))) rule = HVM_LEAK;
      // draining: low-priority redexes leave the wave
      if (evac && (rule == HVM_CALL || rule == HVM_COMM || rule == HVM_OPER || rule == HVM_SWIT))
        rule = (is_loc_port(a) || is_loc_port(b)) ? HVM_SAVE : HVM_EMIT;
    }
    // HBM evaluator: takes are issued before the reservation, so that their
    // round trip is the one of the CALLs' def records and the allocator's
    // first window.  With PEEK a take is a load and the slot is zeroed when
    // the interaction commits, so a lane that retries has removed nothing.
    // The local-net evaluator takes after the reservation (P3).
    constexpr bool EARLY = !LOC && HVM_EARLY_TAKES;
    const bool take_a0 = has && (rule == HVM_ANNI || rule == HVM_COMM);
    const bool take_b0 = has && (rule == HVM_ERAS || rule == HVM_ANNI || rule == HVM_COMM ||
                                 rule == HVM_OPER || rule == HVM_SWIT || (LOC && rule == HVM_LEAK));
    // Lanes with two takes (ANNI, COMM) always read their nodes with loads
    // and zero the slots once the interaction commits.  An early take by
    // exchange frees the slot at once, so the owning wave's allocator may
    // hand it out in this very step; if the lane's other node were then not
    // there, putting the first one back would overwrite the new node.  With
    // loads nothing is removed before the commit, and a lane that retries
    // has nothing to put back (src/hvm.c:921-924 checks both nodes before
    // taking either for the same reason).
    const bool PEEK = EARLY && (HVM_PEEK_TAKES || take_a0);
    auto hbm_take = [](const HvmTurnArgs& A, u32 loc, bool peek) { return peek ? node_peek(A, loc) : node_take(A, loc); };
    // slots from the net's main allocator (local with LOC), known here for
    // every rule but CALL
    u32 nn = 0, nv = 0;
    u32 xn = 0, xv = 0;   // HBM slots of LEAK / SAVE
    if (has) {
      if (rule == HVM_COMM) { nn = 4; nv = 4; }
      else if (rule == HVM_OPER) { nn = 1; }
      else if (rule == HVM_SWIT) { nn = hvm_get_u24(hvm_get_val(a)) == 0 ? 1 : 2; }
      else if (LOC && rule == HVM_LEAK) { xn = 1; xv = 2; }
      else if (LOC && rule == HVM_SAVE) { xv = (u32)is_loc_port(a) + (u32)is_loc_port(b); }
    }
    // HBM evaluator: lanes whose rule is not CALL reserve table space now,
    // and those that fit take their nodes at once, so that the takes share
    // one round trip with the CALLs' def records and the allocator's first
    // window.  CALL lanes reserve after their def record arrives, behind the
    // others.  A lane that waits has taken nothing, so a slot read as free in
    // this step stays free (a taken-and-put-back node could be handed out).
    // The node and var counts are scanned together (16 bits each; a wave's
    // total is at most 64 * HVM_GPU_DEF_MAX < 2^16).
    u32 off_n = 0, off_v = 0;
    bool run = false;
    u64 NA = 0, NB = 0;
    if (EARLY) {
      u32 tot;
      const u32 o = wave_scan(rule != HVM_CALL ? (nn | (nv << 16)) : 0u, &tot);
      off_n = o & 0xFFFFu; off_v = o >> 16;
      run = has && rule != HVM_CALL && off_n + nn <= HVM_TBL_NODE && off_v + nv <= HVM_TBL_VARS;
      NA = (run && take_a0) ? hbm_take(A, hvm_get_val(a), PEEK) : 0ull;
      NB = (run && take_b0) ? hbm_take(A, hvm_get_val(b), PEEK) : 0ull;
    }
    // A CALL's def record with its first template pairs (one 128-byte load).
    bool copy = false;
    HvmDevDefX dx;
    if (has && rule == HVM_CALL) {
      u32 fid = hvm_get_val(a) & 0xFFFFFFFu;
      dx = A.defs[fid < A.ndefs ? fid : 0];
    } else {
      dx.h.rlen = dx.h.nlen = dx.h.vlen = 0; dx.h.rbag_off = dx.h.node_off = 0;
      dx.h.root = 0; dx.h.safe = 0; dx.h.call_bytes = 0;
#pragma unroll
      for (int i = 0; i < HVM_PRE_NODE; ++i) dx.node[i] = 0;
#pragma unroll
      for (int i = 0; i < HVM_PRE_RBAG; ++i) dx.rbag[i] = 0;
    }
    // The allocator's first probe window, loaded in the same round trip
    // (HBM evaluator; used by wave_alloc2 if this step allocates).
    u64 pre_n = 1;
    u32 pre_v = 1;
    if (EARLY) {
      u32 rel;
      Scan s0n = {n_lo, A.node_slice, ws.node_cur, 0, 0, 0, false, ws.node_cur};
      Scan s0v = {v_lo, A.vars_slice, ws.vars_cur, 0, 0, 0, false, ws.vars_cur};
      pre_n = __hip_atomic_load(&A.node[scan_idx(s0n, lane, rel)], __ATOMIC_RELAXED, AGENT);
      pre_v = __hip_atomic_load(&A.vars[scan_idx(s0v, lane, rel)], __ATOMIC_RELAXED, AGENT);
    }
    const HvmDevDef& d = dx.h;
    if (has && rule == HVM_CALL) {
      if ((d.safe & 1u) && hvm_get_tag(b) == HVM_DUP) copy = true;
      else { nn = d.nlen; nv = d.vlen; }
    }

    // ---- P1: reserve allocation-table space (lanes that do not fit wait) ----
    u32 tot_x = 0;
    u32 off_xn = 0, off_xv = 0;
    if (EARLY) {
      // CALL lanes, after the space the other lanes took
      const u32 base_n = wave_max(run ? off_n + nn : 0), base_v = wave_max(run ? off_v + nv : 0);
      u32 tot;
      const u32 o = wave_scan(rule == HVM_CALL ? (nn | (nv << 16)) : 0u, &tot);
      if (has && rule == HVM_CALL) {
        off_n = base_n + (o & 0xFFFFu);
        off_v = base_v + (o >> 16);
        run = off_n + nn <= HVM_TBL_NODE && off_v + nv <= HVM_TBL_VARS;
      }
    } else {
      u32 tot_m;
      u32 off_m = wave_scan(nn | (nv << 16), &tot_m);
      off_n = off_m & 0xFFFFu; off_v = off_m >> 16;
      u32 cap_n = HVM_TBL_NODE, cap_v = HVM_TBL_VARS;
      if (LOC) {
        u32 off_x = wave_scan(xn | (xv << 16), &tot_x);
        off_xn = off_x & 0xFFFFu; off_xv = off_x >> 16;
        cap_n = min(cap_n, __hip_atomic_load(&N->ntop, __ATOMIC_RELAXED, WGS));
        cap_v = min(cap_v, __hip_atomic_load(&N->vtop, __ATOMIC_RELAXED, WGS));
      }
      run = has && off_n + nn <= cap_n && off_v + nv <= cap_v &&
            off_xn + xn <= HVM_TBL_X && off_xv + xv <= HVM_TBL_X;
    }
    if (has && !run) defers++;
    if (LOC && !evac && !__ballot(run)) {
      // the local net is too full for any redex here: drain it
      evac = true;
      continue;
    }
    u32 need_n = wave_max(run ? off_n + nn : 0);
    u32 need_v = wave_max(run ? off_v + nv : 0);
    alloc_slots += need_n + need_v;
    u32 need_xn = 0, need_xv = 0;
    if (LOC && tot_x) {
      need_xn = wave_max(run ? off_xn + xn : 0);
      need_xv = wave_max(run ? off_xv + xv : 0);
    }
    if (run) ln.top--;

    HVM_PHASE(1);
    // ---- P3: takes of the running lanes that have not taken yet (all of
    //      them in the local-net evaluator; a safe def's copy rule) ----
    const bool take_a = run && take_a0;
    // fast CALL: a def whose root is a CON (template node 0) against the
    // caller's CON node takes that node (it is consumed: no put-back, so its
    // slot is never seen free while the node lives)
    const bool fastb = FAST && EARLY && run && rule == HVM_CALL && !copy && hvm_get_tag(b) == HVM_CON &&
                       d.root == hvm_new_port(HVM_CON, 0);
    const bool take_b = run && (take_b0 || (rule == HVM_CALL && copy) || fastb);
    if (!EARLY) {
      NA = take_a ? net_take<LOC>(A, N, hvm_get_val(a)) : 0ull;
      NB = take_b ? net_take<LOC>(A, N, hvm_get_val(b)) : 0ull;
    } else if (run && rule == HVM_CALL && (copy || fastb)) {
      NB = hbm_take(A, hvm_get_val(b), PEEK);
    }
    const u64* tpre = dx.node;
    const u64* rpre = dx.rbag;

    // ---- P2: allocation (wave-cooperative, nodes and vars together) ----
    if (LOC) {
      // local slots come off the free stacks, top first
      if (need_n) {
        u32 t = __hip_atomic_load(&N->ntop, __ATOMIC_RELAXED, WGS);
        for (u32 i = lane; i < need_n; i += HVM_WAVE) L.tbl_n[i] = HVM_LOCAL_BASE + N->nfree[(t - 1 - i) & (HVM_LNODE - 1)];
        __syncthreads();
        if (lane == 0) __hip_atomic_store(&N->ntop, t - need_n, __ATOMIC_RELAXED, WGS);
      }
      if (need_v) {
        u32 t = __hip_atomic_load(&N->vtop, __ATOMIC_RELAXED, WGS);
        for (u32 i = lane; i < need_v; i += HVM_WAVE) L.tbl_v[i] = HVM_LOCAL_BASE + N->vfree[(t - 1 - i) & (HVM_LVARS - 1)];
        __syncthreads();
        if (lane == 0) __hip_atomic_store(&N->vtop, t - need_v, __ATOMIC_RELAXED, WGS);
      }
      __syncthreads();
      if (need_xn || need_xv) {
        u32 oom = 0;
        if (!wave_alloc2(L, A, n_lo, v_lo, ws, need_xn, need_xv, probes, oom, N->xtn, N->xtv, false, 0, 0, rec_n)) {
          atomicOr(&A.ctl->err, oom);
          if (run) ln.top++;
          break;
        }
      }
    } else if (need_n || need_v) {
      u32 oom = 0;
      if (!wave_alloc2(L, A, n_lo, v_lo, ws, need_n, need_v, probes, oom, L.tbl_n, L.tbl_v, true, pre_n, pre_v,
                       rec_n)) {
        // out of slots in the slice and the pool: the evaluation stops with an error
        atomicOr(&A.ctl->err, oom);
        if (run) ln.top++;
        break;
      }
    }
    const u32* tn = L.tbl_n + off_n;
    const u32* tv = L.tbl_v + off_v;
    const u32* gn = LOC ? N->xtn + off_xn : tn;   // HBM slots of LEAK / SAVE
    const u32* gv = LOC ? N->xtv + off_xv : tv;

    if ((take_a && NA == 0) || (take_b && NB == 0)) {
      // A node this redex points to is not visible yet: put everything back
      // and retry the redex on a later step (src/hvm.c:1112-1114).  In the
      // HBM evaluator nothing taken needs to go back: a lane with two takes
      // read its nodes with loads, and a lane with one take found nothing.
      if ((take_a && NA == 0 && !is_loc_val(hvm_get_val(a)) && hvm_get_val(a) >= A.node_cap) ||
          (take_b && NB == 0 && !is_loc_val(hvm_get_val(b)) && hvm_get_val(b) >= A.node_cap))
        atomicOr(&A.ctl->err, (u32)HVM_ERR_BAD_PORT);   // not a node of this net
      if (!PEEK && take_a && NA) net_put<LOC>(A, N, hvm_get_val(a), NA);
      if (!PEEK && take_b && NB) net_put<LOC>(A, N, hvm_get_val(b), NB);
      // slots allocated for this interaction go back (they may be pool claims)
      for (u32 i = 0; i < nn; ++i) node_free<LOC>(A, N, tn[i]);
      for (u32 i = 0; i < nv; ++i) var_free<LOC>(A, N, tv[i]);
      for (u32 i = 0; i < xn; ++i) node_put(A, gn[i], 0ull);
      for (u32 i = 0; i < xv; ++i) vars_put(A, gv[i], 0u);
      ln.top++;
      run = false;
      retries++;
    }
    // nodes read by a peek are taken now: their slots go on the wave's
    // recycling stack (zeroed, i.e. freed, when it is full)
    if (RECYCLE) {
      const bool fa = run && take_a && hvm_get_val(a) < A.node_cap;
      const bool fb = run && take_b && hvm_get_val(b) < A.node_cap;
      u32 tot;
      u32 k = wave_scan((u32)fa + (u32)fb, &tot);
      const u32 room = HVM_NODE_RECYCLE - rec_n;
#if HVM_NODE_RECYCLE
      if (fa) { if (k < room) L.rec[rec_n + k] = hvm_get_val(a); else node_put(A, hvm_get_val(a), 0ull); ++k; }
      if (fb) { if (k < room) L.rec[rec_n + k] = hvm_get_val(b); else node_put(A, hvm_get_val(b), 0ull); }
#endif
      rec_n += min(tot, room);
      // a take through a port outside the net is zeroed as before
      if (run && take_a && !fa) node_put(A, hvm_get_val(a), 0ull);
      if (run && take_b && !fb) node_put(A, hvm_get_val(b), 0ull);
    } else if (PEEK && run) {
      if (take_a) node_put(A, hvm_get_val(a), 0ull);
      if (take_b) node_put(A, hvm_get_val(b), 0ull);
    }
    // taken local nodes are dead now: their slots go back on the stack
    if (LOC && run) {
      if (take_a && is_loc_val(hvm_get_val(a))) node_free<LOC>(A, N, hvm_get_val(a));
      if (take_b && is_loc_val(hvm_get_val(b))) node_free<LOC>(A, N, hvm_get_val(b));
    }

    u32 B1 = hvm_get_fst(NB), B2 = hvm_get_snd(NB);
    // fast CALL: the def's root node (a CON, template node 0) against the
    // caller's CON; with a number there and the root's first port a SWI
    // (node 1) whose first port is a CON (node 2) — the pre-order numbering
    // of a def that switches on its first argument — the SWITCH too.
    bool fa = false, fs = false, shared_k = false;
    const u32 r0 = 0, s0 = 1, c0 = 2;
    u32 skip_m = 0;
    const u64 R0 = dx.node[0], S0 = dx.node[1], C0 = dx.node[2];
    if (FAST && run && fastb) {
      fa = true;
      fs = hvm_get_fst(R0) == hvm_new_port(HVM_SWI, 1) && hvm_get_fst(S0) == hvm_new_port(HVM_CON, 2) &&
           hvm_get_tag(B1) == HVM_NUM;
      shared_k = fs && hvm_get_tag(hvm_get_snd(S0)) == HVM_VAR && hvm_get_snd(S0) == hvm_get_snd(R0);
      skip_m = fa ? (fs ? 7u : 1u) : 0u;
    }
    bool oper1 = false;
    u32 cv = 0;
    if (run && rule == HVM_OPER) {
      // The reference first follows B2's var chain (src/hvm.c:1024).  Here B2
      // is used as it is: OPER-1 links it, and the linker follows the chain
      // itself; OPER-2 stores it in the new node's second port, which no rule
      // inspects (OPER-1 vs OPER-2 reads the first port), and the chain is
      // followed when that port is linked.  The result and ITRS are the same;
      // the step saves a dependent exchange.
      oper1 = hvm_get_tag(B1) == HVM_NUM;
      if (oper1) {
        cv = hvm_operate(hvm_get_val(a), hvm_get_val(B1));
        node_free<LOC>(A, N, tn[0]);  // the node reserved for OPER-2 is not needed
      }
    }

    HVM_PHASE(2);
    // ---- P4: creation ----
    u32 g1 = 0, g2 = 0;   // LEAK: HBM stand-ins for the node's ports
    if (run) {
      if (rule == HVM_CALL && !copy) {
        // Table entries are read unconditionally, a chunk at a time, before
        // the stores that need them (entries past the def's own lie inside
        // the table: offsets are at most HVM_TBL_NODE/VARS, the tables have
        // HVM_TBL_SLACK more entries).
        static_assert(HVM_TBL_SLACK >= 8 && HVM_PRE_NODE % 4 == 0, "table reads stay inside the tables");
        // Template nodes past the prefetched ones come first, eight at a
        // time: a chunk's pair loads are issued together, before any store
        // of the creation.  vmcnt counts loads and stores in issue order, so
        // a wait for a load behind a write-through store also waits for that
        // store; loading one pair per node between the stores cost a store
        // round trip per node.
        for (u32 c = HVM_PRE_NODE; c < d.nlen; c += HVM_XCHUNK) {
          u64 xp[HVM_XCHUNK];
          u32 locs[HVM_XCHUNK];
#pragma unroll
          for (int k = 0; k < HVM_XCHUNK; ++k) xp[k] = c + k < d.nlen ? A.pairs[d.node_off + c + k] : 0ull;
#pragma unroll
          for (int k = 0; k < HVM_XCHUNK; ++k) locs[k] = tn[c + k];
#pragma unroll
          for (int k = 0; k < HVM_XCHUNK; ++k)
            if (c + k < d.nlen) net_put<LOC>(A, N, locs[k], adjust_pair(xp[k], tn, tv));
        }
#pragma unroll
        for (int c = 0; c < 8; c += 4) {
          u32 vs[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) vs[k] = tv[c + k];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if ((u32)(c + k) < d.vlen) var_put<LOC>(A, N, vs[k], HVM_NONE);
        }
        for (u32 i = 8; i < d.vlen; ++i) var_put<LOC>(A, N, tv[i], HVM_NONE);
#pragma unroll
        for (int c = 0; c < HVM_PRE_NODE; c += 4) {
          u32 locs[4];
          u64 vals[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) { locs[k] = tn[c + k]; vals[k] = adjust_pair(tpre[c + k], tn, tv); }
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if ((u32)(c + k) < d.nlen && !((skip_m >> (c + k)) & 1u)) net_put<LOC>(A, N, locs[k], vals[k]);
        }
        if (FAST && fa) {
          // the slots of the nodes not created go back; s0's holds the
          // SWITCH's (n-1, continuation) node when n > 0
          const bool nz = fs && hvm_get_u24(hvm_get_val(B1)) != 0;
          node_free<LOC>(A, N, tn[r0]);
          if (fs) node_free<LOC>(A, N, tn[c0]);
          if (fs && !nz) node_free<LOC>(A, N, tn[s0]);
          // the continuation: with one var shared by the SWI and the root
          // (its only two occurrences) the caller's port replaces it, and
          // the var goes back unused
          if (fs && shared_k) var_free<LOC>(A, N, tv[hvm_get_val(hvm_get_snd(S0))]);
          if (nz)
            net_put<LOC>(A, N, tn[s0], hvm_new_pair(hvm_new_port(HVM_NUM, hvm_new_u24(hvm_get_u24(hvm_get_val(B1)) - 1)),
                                                    shared_k ? B2 : adjust_port(hvm_get_snd(S0), tn, tv)));
        }
      } else if (rule == HVM_COMM) {
        var_put<LOC>(A, N, tv[0], HVM_NONE);
        var_put<LOC>(A, N, tv[1], HVM_NONE);
        var_put<LOC>(A, N, tv[2], HVM_NONE);
        var_put<LOC>(A, N, tv[3], HVM_NONE);
        net_put<LOC>(A, N, tn[0], hvm_new_pair(hvm_new_port(HVM_VAR, tv[0]), hvm_new_port(HVM_VAR, tv[1])));
        net_put<LOC>(A, N, tn[1], hvm_new_pair(hvm_new_port(HVM_VAR, tv[2]), hvm_new_port(HVM_VAR, tv[3])));
        net_put<LOC>(A, N, tn[2], hvm_new_pair(hvm_new_port(HVM_VAR, tv[0]), hvm_new_port(HVM_VAR, tv[2])));
        net_put<LOC>(A, N, tn[3], hvm_new_pair(hvm_new_port(HVM_VAR, tv[1]), hvm_new_port(HVM_VAR, tv[3])));
      } else if (rule == HVM_OPER && !oper1) {
        net_put<LOC>(A, N, tn[0], hvm_new_pair(a, B2));
      } else if (rule == HVM_SWIT) {
        u32 av = hvm_get_u24(hvm_get_val(a));
        if (av == 0) {
          net_put<LOC>(A, N, tn[0], hvm_new_pair(B2, hvm_new_port(HVM_ERA, 0)));
        } else {
          net_put<LOC>(A, N, tn[0], hvm_new_pair(hvm_new_port(HVM_ERA, 0), hvm_new_port(HVM_CON, tn[1])));
          net_put<LOC>(A, N, tn[1], hvm_new_pair(hvm_new_port(HVM_NUM, hvm_new_u24(av - 1)), B2));
        }
      } else if (LOC && rule == HVM_LEAK) {
        // The node's ports, with local ones replaced by fresh HBM vars.  The
        // ports are copied as they are, not entered: whether a port is a var
        // decides later rules (OPER-1 needs a NUM in the fst port itself,
        // src/hvm.c:1023-1028), and ITRS must stay that of run-c.
        if (is_loc_port(B1)) { g1 = hvm_new_port(HVM_VAR, gv[0]); vars_put(A, gv[0], HVM_NONE); }
        else { g1 = B1; vars_put(A, gv[0], 0u); }
        if (is_loc_port(B2)) { g2 = hvm_new_port(HVM_VAR, gv[1]); vars_put(A, gv[1], HVM_NONE); }
        else { g2 = B2; vars_put(A, gv[1], 0u); }
        node_put(A, gn[0], hvm_new_pair(g1, g2));
      } else if (LOC && rule == HVM_SAVE) {
        for (u32 i = 0; i < xv; ++i) vars_put(A, gv[i], HVM_NONE);
      }
    }

    HVM_PHASE(3);
    // ---- P5: publish creations before any link can expose them ----
    if (A.flags & HVM_TURN_WAITCNT_FENCE) {
      // Node/var creations are device-scope (write-through) stores: waiting
      // for their completion orders them before the link exchanges.
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }

    HVM_PHASE(5);
    // ---- P6: links ----
    u64 emit = 0;         // SAVE / EMIT: the redex that leaves through the output bag
    if (run) {
      u32 pa[HVM_LB], pb[HVM_LB], np = 0;
      switch (rule) {
        case HVM_LINK:
          pa[0] = a; pb[0] = b; np = 1;
          break;
        case HVM_CALL:
          if (copy) {
            pa[0] = a; pb[0] = B1; pa[1] = a; pb[1] = B2; np = 2;
          } else {
            // the def's redexes, then its root against the caller's port,
            // HVM_LB links at a time (the first ones were loaded ahead)
            static_assert(HVM_LB >= HVM_PRE_RBAG, "the redexes loaded ahead fit one batch");
#pragma unroll
            for (int k = 0; k < HVM_PRE_RBAG; ++k) {
              if ((u32)k < d.rlen) {
                u64 p = adjust_pair(rpre[k], tn, tv);
                pa[k] = hvm_get_fst(p); pb[k] = hvm_get_snd(p);
              }
            }
            np = min(d.rlen, (u32)HVM_PRE_RBAG);
            auto emit = [&](u32 x, u32 y) {
              if (np == HVM_LB) { link_batch<LOC>(L, N, ln, A, pa, pb, np); np = 0; }
              put_link(pa, pb, np, x, y);
            };
            for (u32 i = np; i < d.rlen; ++i) {
              u64 p = adjust_pair(A.pairs[d.rbag_off + i], tn, tv);
              emit(hvm_get_fst(p), hvm_get_snd(p));
            }
            if (FAST && fs) {
              // SWITCH on n: n = 0 links the zero branch to the continuation
              // and erases the other; n > 0 erases the zero branch and
              // passes (n-1, continuation) to the other
              const u32 z = adjust_port(hvm_get_fst(C0), tn, tv), sc = adjust_port(hvm_get_snd(C0), tn, tv);
              const u32 k = shared_k ? B2 : adjust_port(hvm_get_snd(S0), tn, tv);
              if (hvm_get_u24(hvm_get_val(B1)) == 0) {
                emit(z, k);
                emit(sc, hvm_new_port(HVM_ERA, 0));
              } else {
                emit(z, hvm_new_port(HVM_ERA, 0));
                emit(sc, hvm_new_port(HVM_CON, tn[s0]));
              }
              if (!shared_k) emit(adjust_port(hvm_get_snd(R0), tn, tv), B2);
            } else if (FAST && fa) {
              emit(adjust_port(hvm_get_fst(R0), tn, tv), B1);
              emit(adjust_port(hvm_get_snd(R0), tn, tv), B2);
            } else {
              emit(adjust_port(d.root, tn, tv), b);
            }
          }
          break;
        case HVM_ERAS:
          pa[0] = a; pb[0] = B1; pa[1] = a; pb[1] = B2; np = 2;
          break;
        case HVM_ANNI:
          pa[0] = hvm_get_fst(NA); pb[0] = B1; pa[1] = hvm_get_snd(NA); pb[1] = B2; np = 2;
          break;
        case HVM_COMM:
          pa[0] = hvm_new_port(hvm_get_tag(b), tn[0]); pb[0] = hvm_get_fst(NA);
          pa[1] = hvm_new_port(hvm_get_tag(b), tn[1]); pb[1] = hvm_get_snd(NA);
          pa[2] = hvm_new_port(hvm_get_tag(a), tn[2]); pb[2] = B1;
          pa[3] = hvm_new_port(hvm_get_tag(a), tn[3]); pb[3] = B2;
          np = 4;
          break;
        case HVM_OPER:
          if (oper1) { pa[0] = hvm_new_port(HVM_NUM, cv); pb[0] = B2; }
          else { pa[0] = B1; pb[0] = hvm_new_port(HVM_OPR, tn[0]); }
          np = 1;
          break;
        case HVM_SWIT:
          pa[0] = hvm_new_port(HVM_CON, tn[0]); pb[0] = B1; np = 1;
          break;
        default:
          if (LOC && rule == HVM_LEAK) {
            if (is_loc_port(B1)) put_link(pa, pb, np, g1, B1);
            if (is_loc_port(B2)) put_link(pa, pb, np, g2, B2);
            put_link(pa, pb, np, a, hvm_new_port(hvm_get_tag(b), gn[0]));
          } else if (LOC && rule == HVM_SAVE) {
            u32 x = a, y = b;
            const bool la = is_loc_port(a);
            if (la) { x = hvm_new_port(HVM_VAR, gv[0]); put_link(pa, pb, np, x, a); }
            if (is_loc_port(b)) { y = hvm_new_port(HVM_VAR, gv[la ? 1 : 0]); put_link(pa, pb, np, y, b); }
            emit = hvm_new_pair(x, y);
          } else if (LOC && rule == HVM_EMIT) {
            emit = hvm_new_pair(a, b);
          }
          break;
      }
      if (np) link_batch<LOC>(L, N, ln, A, pa, pb, np);

    }
    // ---- P7: counting (ITRS += 1 per non-LINK rule, src/hvm.c:1116-1118) ----
    {
      u32 kk = HVM_K_KINDS, kb = 0;   // kind slot and canonical bytes of this lane's interaction
      if (run) {
        switch (rule) {
          case HVM_CALL:
            if (copy) { kk = HVM_K_CALL_COPY; kb = 32; }
            else {
              kk = HVM_K_CALL; kb = d.call_bytes;
              // a compiled def's static interactions (zero for plain defs)
              ln.spec += (u64)((d.safe >> 8) & 0xFFu) | ((u64)(d.safe >> 16) << 32);
              if (FAST && fa) {
                // the ANNI with the caller's node; with a SWITCH, that and
                // the branch pair's ANNI
                ln.spec += (fs ? 2ull : 1ull) << 32;
                kb += fs ? 96u : 48u;
                if (fs) {
                  const bool z0 = hvm_get_u24(hvm_get_val(B1)) == 0;
                  ln.fsw0 += z0 ? 1u : 0u;
                  ln.fswn += z0 ? 0u : 1u;
                  kb += z0 ? 32u : 40u;
                }
              }
            }
            break;
          case HVM_VOID: kk = HVM_K_VOID; break;
          case HVM_ERAS: kk = HVM_K_ERAS; kb = 32; break;
          case HVM_ANNI: kk = HVM_K_ANNI; kb = 48; break;
          case HVM_COMM: kk = HVM_K_COMM; kb = 112; break;
          case HVM_OPER:
            kk = oper1 ? HVM_K_OPER1 : HVM_K_OPER2; kb = oper1 ? 32 : 40;
            break;
          case HVM_SWIT:
            if (hvm_get_u24(hvm_get_val(a)) == 0) { kk = HVM_K_SWIT0; kb = 32; }
            else { kk = HVM_K_SWITN; kb = 40; }
            break;
          default: break;
        }
      }
      ln.bytes += kb;
#pragma unroll
      for (int k = 0; k < HVM_K_KINDS; ++k) wc.kind[k] += __popcll(__ballot(kk == (u32)k));
      if (LOC) {
        wc.leaks += __popcll(__ballot(run && rule == HVM_LEAK));
        wc.saves += __popcll(__ballot(run && rule == HVM_SAVE));
      }
    }
    if (LOC && evac) {
      // redexes leaving the wave go to the output bag, one atomic per wave
      bool out1 = emit != 0;
      u64 om = __ballot(out1);
      if (om) {
        u32 k = __popcll(om);
        u32 base = 0;
        if (lane == 0) base = atomicAdd(&A.ctl->n_out, k);
        base = lane_u32(base, 0);
        u32 pos = base + rank_in(om);
        if (out1) {
          if (pos < A.bag_cap) bag_put(&bag_out[pos], emit);
          else atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
        }
      }
    }
    HVM_PHASE(4);
  }
#undef HVM_PHASE

  // ---- end of turn: recycled node and var slots not handed out are freed ----
#if HVM_NODE_RECYCLE
  if (RECYCLE && rec_n) {
    __syncthreads();
    for (u32 i = lane; i < rec_n; i += HVM_WAVE) node_put(A, L.rec[i], 0ull);
    rec_n = 0;
  }
#endif
#if HVM_VAR_RECYCLE
  {
    __syncthreads();
    const u32 vn = min(L.rec_vn, (u32)HVM_VAR_RECYCLE);
    for (u32 i = lane; i < vn; i += HVM_WAVE) vars_put(A, L.rec_v[i], 0u);
  }
#endif

  // ---- end of turn: everything left goes to the output bag ----
  {
    u32 n = ln.top - ln.bot;
    u32 tot;
    u32 off = wave_scan(n, &tot);
    left_tot = tot;
    u32 base = 0;
    if (lane == 0 && tot) base = atomicAdd(&A.ctl->n_out, tot);
    base = lane_u32(base, 0);
    if (base + tot > A.bag_cap) {
      if (lane == 0) atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
    } else {
      for (u32 i = 0; i < n; ++i) bag_put(&bag_out[base + off + i], deq_at(L, ln.bot + i));
    }
  }
  u64 left = 0;
  if (LOC) {
    // a drained wave holds no redex; what the local net still holds is garbage
    // (a cycle no redex reaches) or, after an error, the net's lost part
    u32 sp = __hip_atomic_load(&N->sp, __ATOMIC_RELAXED, WGS);
    left = (u64)(HVM_LNODE - __hip_atomic_load(&N->ntop, __ATOMIC_RELAXED, WGS)) +
           (u64)(HVM_LVARS - __hip_atomic_load(&N->vtop, __ATOMIC_RELAXED, WGS));
    if ((ln.top != ln.bot || sp != 0) && lane == 0) atomicOr(&A.ctl->err, (u32)HVM_ERR_EVAC);
  }

  // ---- input redexes no wave pulled this turn are carried to the output ----
  while (more) {
    u32 base = 0;
    if (lane == 0) base = atomicAdd(&A.ctl->in_head, (u32)HVM_WAVE);
    base = lane_u32(base, 0);
    if (base >= n_in) break;
    u32 cnt = min((u32)HVM_WAVE, n_in - base);
    u32 obase = 0;
    if (lane == 0) obase = atomicAdd(&A.ctl->n_out, cnt);
    obase = lane_u32(obase, 0);
    if (obase + cnt > A.bag_cap) {
      if (lane == 0) atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
      break;
    }
    if (lane < cnt) bag_put(&bag_out[obase + lane], bag_take(&bag_in[base + lane]));
  }

  // ---- statistics: one atomic per counter per wave ----
  {
    u64 s_itrs = 0;
    u64 kinds[HVM_K_KINDS];
#pragma unroll
    for (int k = 0; k < HVM_K_KINDS; ++k) kinds[k] = wc.kind[k];
    // per lane each half stays below 2^32 for turns of up to 2^20 steps
    kinds[HVM_K_CALL] += wave_sum64(ln.spec & 0xFFFFFFFFull);
    kinds[HVM_K_ANNI] += wave_sum64(ln.spec >> 32);
    kinds[HVM_K_VOID] += wave_sum64(ln.voids);
    if (FAST) {
      kinds[HVM_K_SWIT0] += wave_sum64(ln.fsw0);
      kinds[HVM_K_SWITN] += wave_sum64(ln.fswn);
    }
#pragma unroll
    for (int k = 0; k < HVM_K_KINDS; ++k) s_itrs += kinds[k];
    u64 s_bytes = wave_sum64(ln.bytes);
    u64 s_spill = wave_sum64(ln.spills);
    defers = wave_sum64(defers);    // per-lane counts; steps/probes/lane_steps are wave-uniform
    retries = wave_sum64(retries);
    u64 s_leaks = wc.leaks, s_saves = wc.saves;
    if (lane == 0) {
      HvmDevStats* St = A.stats + (w & (HVM_STAT_STRIPES - 1));
      if (s_itrs) atomicAdd(&St->itrs, s_itrs);
      if (s_bytes) atomicAdd(&St->bytes, s_bytes);
      if (s_spill) atomicAdd(&St->spills, s_spill);
#pragma unroll
      for (int k = 0; k < HVM_K_KINDS; ++k) if (kinds[k]) atomicAdd(&St->kind[k], kinds[k]);
      if (steps) atomicAdd(&St->steps, steps);
      if (defers) atomicAdd(&St->defers, defers);
      if (retries) atomicAdd(&St->retries, retries);
      if (probes) atomicAdd(&St->probes, probes);
      if (s_leaks) atomicAdd(&St->leaks, s_leaks);
      if (s_saves) atomicAdd(&St->saves, s_saves);
      if (left) atomicAdd(&St->local_left, left);
      if (prof) {
#pragma unroll
        for (int k = 0; k < 6; ++k) if (ph[k]) atomicAdd(&St->phase[k], ph[k]);
      }
      // for the scheduler's features
      if (lane_steps) atomicAdd(&St->lane_steps, lane_steps);
      const u32 cls = left_tot == 0 ? 0u : left_tot <= HVM_WAVE ? 1u : left_tot <= 2u * HVM_WAVE ? 2u : 3u;
      atomicAdd(&St->wave_left[cls], 1ull);
      A.wave_state[w] = ws;
      if (alloc_slots) atomicAdd(&St->alloc_slots, alloc_slots);
    }
  }
}

__global__ __launch_bounds__(HVM_WAVE, HVM_TURN_OCC) void hvm_turn_kernel(HvmTurnArgs A) {
  __shared__ WaveLds L;
  turn_body<false, false>(A, L, nullptr);
}

// Compiled CALLs (HVM_TURN_FAST): the same turn with the fast CALL paths.
__global__ __launch_bounds__(HVM_WAVE, HVM_TURN_OCC) void hvm_turn_fast_kernel(HvmTurnArgs A) {
  __shared__ WaveLds L;
  turn_body<false, true>(A, L, nullptr);
}

// The local-net evaluator: the same turn with node/var allocation in LDS.
// Its LDS allows one wave per SIMD, so registers are not capped further.
__global__ __launch_bounds__(HVM_WAVE, 1) void hvm_turn_local_kernel(HvmTurnArgs A) {
  __shared__ WaveLds L;
  __shared__ LocalNet N;
  turn_body<true, false>(A, L, &N);
}

// ---------------------------------------------------------------------------
// Readback: level-synchronous traversal from a port, recording every node and
// var it reaches, for the host printer.  One workgroup; bounded by `cap`.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void hvm_gather_kernel(HvmGatherArgs G) {
  __shared__ u32 s_head, s_tail, s_next;
  if (threadIdx.x == 0) {
    G.queue[0] = G.root_port;
    s_head = 0;
    s_tail = 1;
    s_next = 1;
  }
  __syncthreads();
  for (u32 round = 0; round < (1u << 30); ++round) {
    u32 head = s_head, tail = s_tail;
    if (head >= tail) break;
    for (u32 i = head + threadIdx.x; i < tail; i += blockDim.x) {
      u32 p = G.queue[i];
      u32 tag = hvm_get_tag(p);
      if (tag == HVM_VAR) {
        u32 vi = vidx(hvm_get_val(p));
        if (vi >= G.vars_cap) { atomicOr(G.err, 1u); continue; }
        u32 v = G.vars[vi];
        u32 k = atomicAdd(G.n_var, 1u);
        if (k < G.cap) { G.var_idx[k] = hvm_get_val(p); G.var_val[k] = v; }
        if (v != HVM_NONE && v != 0) {
          u32 q = atomicAdd(&s_next, 1u);
          if (q < G.cap) G.queue[q] = v; else atomicOr(G.err, 2u);
        }
      } else if (tag >= HVM_CON) {
        u32 ni = hvm_get_val(p);
        if (ni >= G.node_cap) { atomicOr(G.err, 1u); continue; }
        u64 nd = G.node[ni];
        u32 k = atomicAdd(G.n_node, 1u);
        if (k < G.cap) { G.node_idx[k] = ni; G.node_val[k] = nd; }
        u32 q = atomicAdd(&s_next, 2u);
        if (q + 1 < G.cap) { G.queue[q] = hvm_get_fst(nd); G.queue[q + 1] = hvm_get_snd(nd); }
        else atomicOr(G.err, 2u);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_head = tail;
      s_tail = s_next < G.cap ? s_next : G.cap;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Between turns: the output bag becomes the next input, and the scheduler
// chooses the next turn's length.  One wave; thread 0 does the bookkeeping,
// all 64 lanes evaluate the MLP in learned mode.
// ---------------------------------------------------------------------------

// Scheduler actions: the share of the waves a turn runs on (1/div), its
// idle-quit share (0: the configured one, hvm_hip_cfg.idle_pct) and the
// factor on the redex count below which a turn that did not grow makes the
// next one steady (the steady rule below); turns are at most 4096 steps.
// Action 0 is the plain heuristic's.  Fewer waves put a small frontier's
// redexes on fewer, fuller waves; a larger idle-quit share lets waves run
// longer before the HBM bag redistributes; a 16 times larger steady bound
// (four redexes per wave) lets a frontier that narrows towards a join run
// its chains without turn ends.  (Shorter fixed turns, 512 and 128 steps,
// were slower than idle quitting on every workload measured, by up to 19%.)
#ifndef HVM_SCHED_WDIV
#define HVM_SCHED_WDIV 1, 1, 2, 1
#endif
#ifndef HVM_SCHED_IDLE
#define HVM_SCHED_IDLE 0, 50, 0, 0
#endif
#ifndef HVM_SCHED_STEADY
#define HVM_SCHED_STEADY 1, 1, 1, 16
#endif
#define HVM_TURN_LEN 4096u
__device__ const u32 kWaveDiv[HVM_ACTIONS] = {HVM_SCHED_WDIV};
__device__ const u32 kIdlePct[HVM_ACTIONS] = {HVM_SCHED_IDLE};
__device__ const u32 kSteadyMul[HVM_ACTIONS] = {HVM_SCHED_STEADY};

// The plain heuristic: every wave, the longest turn, ended early once the
// configured idle-quit share of the waves has run out of work, which
// shortens turns by itself when redexes are scarce.
__host__ __device__ __forceinline__ u32 plain_action(u32 n, u32 nwaves) {
  (void)n;
  (void)nwaves;
  return 0;
}

// The flat-frontier rule: after a turn that did not grow its redexes by more
// than GROWTH_PCT% and left fewer of them than the turn's waves, the idle-quit
// share applies to the waves the initial spread gives a redex (the empty ones
// count as idle at once); otherwise the empty waves alone end such a turn at
// the first look (church_loop(12, 12): 360 -> 343 ms, the Church fan 1.02x,
// the other workloads within 0.5%).
#ifndef HVM_FLAT_FIT
#define HVM_FLAT_FIT 1
#endif

// The steady rule: after a turn that ends with at most NUM/DEN redexes per
// wave (times the action's factor) and at most GROWTH_PCT% of those it
// started with, the next turn runs without idle quit (waves run their chains
// until they run dry); idle quit comes back inside the turn once a wave holds
// surplus work.
#ifndef HVM_STEADY_NUM
#define HVM_STEADY_NUM 1u
#endif
#ifndef HVM_STEADY_DEN
#define HVM_STEADY_DEN 4u
#endif
#ifndef HVM_STEADY_GROWTH_PCT
#define HVM_STEADY_GROWTH_PCT 150u
#endif

__device__ __forceinline__ float log2p(float x) { return __log2f(1.0f + x); }

// Scheduler features of the turn that just ended (all in [0, ~1]).
__device__ void sched_features(const HvmCtl* c, const HvmSched* s, const HvmTurnStats& T, u32 n_next, u32 nwaves,
                               float dt_us, float* f) {
  float lanes = (float)nwaves * HVM_WAVE;
  float itrs = (float)T.itrs;
  float ws = (float)(T.wave_steps ? T.wave_steps : 1);
  for (int i = 0; i < HVM_FEAT; ++i) f[i] = 0.0f;
  f[0] = log2p((float)c->n_in) / 24.0f;
  f[1] = log2p((float)n_next) / 24.0f;
  f[2] = fminf(1.0f, (float)n_next / lanes);
  f[3] = (float)T.lane_steps / (HVM_WAVE * ws);
  f[4] = log2p(itrs) / 32.0f;
  f[5] = log2p(dt_us) / 20.0f;
  f[6] = itrs > 0 ? (float)T.spills / itrs : 0.0f;
  f[7] = itrs > 0 ? (float)T.defers / itrs : 0.0f;
  f[8] = fminf(1.0f, (float)T.probes / ws / 4.0f);
  for (int k = 0; k < HVM_K_KINDS; ++k) f[9 + k] = itrs > 0 ? (float)T.kind[k] / itrs : 0.0f;
  // how the turn ended: waves that ran dry, without idle quit (steady), growth
  f[19] = fminf(1.0f, (float)c->idle_waves / (float)nwaves);
  f[20] = c->idle_quit == 0 ? 1.0f : 0.0f;
  f[21] = fminf(1.0f, fmaxf(0.0f, 0.5f + 0.0625f * (log2p((float)n_next) - log2p((float)c->n_in))));
  f[22] = c->surplus ? 1.0f : 0.0f;
  f[23] = log2p((float)c->turns) / 16.0f;
  f[24] = fminf(1.0f, ws / ((float)nwaves * (float)(c->max_steps ? c->max_steps : 1)));
  f[25] = 1.0f;
  // per-wave occupancy at the end of the turn, free-slot density of the
  // allocator's windows, and whether some wave held surplus work
  for (int k = 0; k < 4; ++k) f[26 + k] = (float)T.wave_left[k] / (float)nwaves;
  f[30] = T.probes ? fminf(1.0f, (float)T.alloc_slots / (64.0f * (float)T.probes)) : 1.0f;
  f[31] = log2p(itrs / fmaxf(dt_us, 1.0f)) / 16.0f;
}

__device__ __forceinline__ float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }

__global__ __launch_bounds__(HVM_WAVE) void hvm_between_kernel(HvmTurnArgs A, u32 clock_khz) {
  __shared__ float f[HVM_FEAT], h1[HVM_HID], h2[HVM_HID], y[HVM_ACTIONS];
  __shared__ u32 s_go, s_n, s_flat;
  __shared__ HvmTurnStats tsum;
  HvmCtl* c = A.ctl;
  HvmSched* s = A.sched;
  const u32 t = threadIdx.x;
  // the turn's counters: the evaluator's counters summed over their copies
  // (lane t sums copies t, t + 64, ..., every load in flight at once), minus
  // their sums at the previous decision
  static_assert(HVM_STAT_STRIPES % HVM_WAVE == 0, "copies of the counters: a multiple of the lanes");
  static_assert(sizeof(HvmTurnStats) % 8 == 0 && sizeof(HvmTurnStats) / 8 <= HVM_WAVE, "turn counters: u64, one per lane");
  {
    constexpr u32 NF = sizeof(HvmTurnStats) / 8;
    unsigned long long v[NF];
#pragma unroll
    for (u32 k = 0; k < NF; ++k) v[k] = 0;
    for (u32 j = t; j < HVM_STAT_STRIPES; j += HVM_WAVE) {
      HvmTurnStats x;
      hvm_turn_view(A.stats[j], x);
      const unsigned long long* xv = (const unsigned long long*)&x;
#pragma unroll
      for (u32 k = 0; k < NF; ++k) v[k] += xv[k];
    }
    unsigned long long* sum = (unsigned long long*)&tsum;
    unsigned long long* prev = (unsigned long long*)&s->prev;
#pragma unroll
    for (u32 k = 0; k < NF; ++k) {
      const unsigned long long x = wave_sum64(v[k]);
      if (t == k) {
        sum[k] = x - prev[k];
        prev[k] = x;
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    s_go = 0;
    s_flat = 0;
    if (!c->done) {
      u32 n = c->n_out;
      c->turns += 1;
      // the spare slots past the capacities are written only through ports
      // that point outside the net
      if (__hip_atomic_load(&A.node[A.node_cap], __ATOMIC_RELAXED, AGENT) != 0 ||
          __hip_atomic_load(&A.vars[A.vars_cap], __ATOMIC_RELAXED, AGENT) != 0)
        c->err |= HVM_ERR_BAD_PORT;
      if (c->err || n > A.bag_cap) {
        if (n > A.bag_cap) c->err |= HVM_ERR_BAG_OVERFLOW;
        c->done = 1;
      } else {
        // reward of the previous decision: log2 of this turn's interactions per µs
        unsigned long long now = wall_clock64();
        float dt_us = s->last_clock ? (float)(now - s->last_clock) * 1000.0f / (float)clock_khz : 0.0f;
        if (s->trace && s->last_clock && s->trace_len > 0 && s->trace_len <= s->trace_cap) {
          float* rec = s->trace + (size_t)(s->trace_len - 1) * HVM_TRACE_REC;
          rec[HVM_FEAT + 1] = __log2f((1.0f + (float)tsum.itrs) / fmaxf(dt_us, 1.0f));
          rec[HVM_FEAT + 2] = (float)tsum.itrs;
          rec[HVM_FEAT + 3] = dt_us;
        }
        sched_features(c, s, tsum, n, A.nwaves, dt_us, f);
        // not growing: with few redexes (checked once the action is known),
        // independent sequential chains, which a redistribution cannot
        // spread; the next turn then runs without idle quit
        s_flat = (u64)n * 100u <= (u64)c->n_in * HVM_STEADY_GROWTH_PCT;
        c->n_in = n;
        c->n_out = 0;
        c->out_take = 0;
        c->idle_waves = 0;
        c->surplus = 0;
        c->parity ^= 1u;
        s_n = n;
        if (n == 0) c->done = 1;
        else s_go = 1;
        s->last_clock = now;
      }
    }
  }
  __syncthreads();
  if (!s_go) return;
  const u32 n = s_n;
  u32 action = plain_action(n, A.nwaves);
  if (s->mode == HVM_SCHED_LEARNED && s->params) {
    // 32 -> 64 -> 64 -> 4 MLP, fp16 weights, fp32 accumulation
    const unsigned short* P = s->params;
    const unsigned short* W1 = P;
    const unsigned short* b1 = W1 + HVM_FEAT * HVM_HID;
    const unsigned short* W2 = b1 + HVM_HID;
    const unsigned short* b2 = W2 + HVM_HID * HVM_HID;
    const unsigned short* W3 = b2 + HVM_HID;
    const unsigned short* b3 = W3 + HVM_HID * HVM_ACTIONS;
    float acc = h2f(b1[t]);
    for (int i = 0; i < HVM_FEAT; ++i) acc += f[i] * h2f(W1[i * HVM_HID + t]);
    h1[t] = fmaxf(acc, 0.0f);
    __syncthreads();
    acc = h2f(b2[t]);
    for (int i = 0; i < HVM_HID; ++i) acc += h1[i] * h2f(W2[i * HVM_HID + t]);
    h2[t] = fmaxf(acc, 0.0f);
    __syncthreads();
    if (t < HVM_ACTIONS) {
      acc = h2f(b3[t]);
      for (int i = 0; i < HVM_HID; ++i) acc += h2[i] * h2f(W3[i * HVM_ACTIONS + t]);
      // The decision uses the outputs averaged over the turns of this
      // normalisation (weight 1/8 for the newest turn): the network is
      // trained on evaluations with one action held throughout, and a choice
      // that follows single turns mixes actions in ways no training run saw
      // (measured: 15% slower than any fixed action on a Church fan).
      const float q = s->qn ? s->qavg[t] + 0.125f * (acc - s->qavg[t]) : acc;
      s->qavg[t] = q;
      y[t] = q;
    }
    __syncthreads();
    if (t == 0) s->qn += 1;
    // The plain heuristic is the fallback: the network moves off the action
    // in force (the plain one when a normalisation starts) only to an action
    // whose predicted whole-evaluation speedup beats it by more than the
    // margin (explore_pct / 100, in log2 units), so that near-ties do not
    // alternate actions from turn to turn.
    const u32 cur = s->last_action < HVM_ACTIONS ? s->last_action : action;
    u32 best = 0;
    for (u32 k = 1; k < HVM_ACTIONS; ++k) if (y[k] > y[best]) best = k;
    action = y[best] > y[cur] + 0.01f * (float)s->explore_pct ? best : cur;
  } else if (s->mode == HVM_SCHED_EXPLORE) {
    const u32 h = sched_hash(s->seed, c->turns);
    if (h % 100u < s->explore_pct) action = (h >> 8) & 3u;
  } else if (s->mode == HVM_SCHED_FIXED) {
    action = s->explore_pct % HVM_ACTIONS;
  }
  if (t == 0) {
    // a nonzero seed shortens each turn by up to an eighth (share jitter)
    const u32 len = HVM_TURN_LEN;
    const u32 jit = s->seed ? (sched_hash(s->seed ^ 0xA5A5A5A5u, c->turns) >> 4) % (len / 8u + 1u) : 0u;
    c->max_steps = c->fixed_steps ? c->fixed_steps : len - jit;
    const u32 act = max(1u, A.nwaves / kWaveDiv[action]);
    // the configured idle-quit count, scaled to the waves that run
    const u32 back = A.idle_quit ? max(1u, (u32)(((u64)A.idle_quit * act) / A.nwaves)) : 0u;
    c->active = act;
    c->in_head = min(n, act * (u32)HVM_WAVE);
    c->idle_back = back;
    c->idle_quit = kIdlePct[action] ? max(1u, (act * kIdlePct[action]) / 100u) : back;
    if (HVM_FLAT_FIT && s_flat && n < act && c->idle_quit) {
      // a frontier that did not grow, smaller than the waves: the waves the
      // spread leaves empty count as idle at once, so the share applies to
      // the n waves that got a redex
      c->idle_quit = (act - n) + max(1u, (u32)(((u64)c->idle_quit * n) / act));
    }
    if (s_flat && (u64)n * HVM_STEADY_DEN <= (u64)A.nwaves * HVM_STEADY_NUM * kSteadyMul[action])
      c->idle_quit = 0;
    s->last_action = action;
    if (s->trace && s->trace_len < s->trace_cap) {
      float* rec = s->trace + (size_t)s->trace_len * HVM_TRACE_REC;
      for (int i = 0; i < HVM_FEAT; ++i) rec[i] = f[i];
      rec[HVM_FEAT] = (float)action;
      rec[HVM_FEAT + 1] = 0.0f;
      rec[HVM_FEAT + 2] = 0.0f;
      rec[HVM_FEAT + 3] = 0.0f;
      s->trace_len += 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Launch helpers (host)
// ---------------------------------------------------------------------------

extern "C" hipError_t hvm_launch_turn(const HvmTurnArgs* a, hipStream_t s) {
  if (a->flags & HVM_TURN_LOCAL)
    hipLaunchKernelGGL(hvm_turn_local_kernel, dim3(a->nwaves), dim3(HVM_WAVE), 0, s, *a);
  else if (a->flags & HVM_TURN_FAST)
    hipLaunchKernelGGL(hvm_turn_fast_kernel, dim3(a->nwaves), dim3(HVM_WAVE), 0, s, *a);
  else
    hipLaunchKernelGGL(hvm_turn_kernel, dim3(a->nwaves), dim3(HVM_WAVE), 0, s, *a);
  return hipGetLastError();
}

extern "C" hipError_t hvm_launch_between(const HvmTurnArgs* a, uint32_t clock_khz, hipStream_t s) {
  hipLaunchKernelGGL(hvm_between_kernel, dim3(1), dim3(HVM_WAVE), 0, s, *a, clock_khz);
  return hipGetLastError();
}

// Turn length of the first turn (the host boots it; later turns are chosen
// on the device).
extern "C" uint32_t hvm_turn_policy_host(uint32_t n, uint32_t nwaves) {
  (void)n;
  (void)nwaves;
  return HVM_TURN_LEN;
}

extern "C" uint32_t hvm_plain_action_host(uint32_t n, uint32_t nwaves) { return plain_action(n, nwaves); }

// Whether this build's turn kernel runs the mid-turn pool (the runtime then
// keeps the bags cleared; variant builds recompile only this file).
extern "C" uint32_t hvm_turn_mid_pool(void) { return HVM_MID_POOL ? 1u : 0u; }

extern "C" hipError_t hvm_launch_gather(const HvmGatherArgs* g, hipStream_t s) {
  hipLaunchKernelGGL(hvm_gather_kernel, dim3(1), dim3(256), 0, s, *g);
  return hipGetLastError();
}

extern "C" size_t hvm_turn_lds_bytes(int local) { return sizeof(WaveLds) + (local ? sizeof(LocalNet) : 0); }

// Turn-kernel workgroups (one wave each) that can be resident on one CU at
// once, given the kernel's registers and LDS; 0 if the runtime cannot tell.
extern "C" int hvm_turn_waves_per_cu(int local) {
  int n = 0;
  hipError_t e = local ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, hvm_turn_local_kernel, HVM_WAVE, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, hvm_turn_kernel, HVM_WAVE, 0);
  return e == hipSuccess ? n : 0;
}
