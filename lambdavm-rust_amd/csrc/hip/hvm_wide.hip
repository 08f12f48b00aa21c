// hvm_wide.hip — the 64-bit-port evaluator (HVM_HIP_WIDE).
//
// Same interaction rules, counting and turn structure as the 32-bit
// evaluator (src/hvm.cu:1197-1420 for the rules and the dispatch,
// src/hvm.c:796-842 for the linker), with 64-bit ports and 128-bit nodes so
// that a net is no longer capped at 2^29 slots (layout and node protocol in
// hvm_wide.hpp).  One redex per lane per step: pop, reserve slots (one
// wave-wide table of up to 256 nodes and 256 vars), take, create, publish
// (wait for the creation stores), link.  It is the capacity path, not the
// tuned one: no early takes, no batched link rounds, no local nets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../../include/hvm_hip.h"
#include "../front/hvm_front.hpp"
#include "hvm_wave.hpp"
#include "hvm_wide.hpp"

namespace {

typedef unsigned long long u64;
typedef uint32_t u32;

#define AGENT __HIP_MEMORY_SCOPE_AGENT

constexpr u64 W_NONE = ~0ull;
constexpr u64 W_ROOT_VAL = (1ull << 61) - 1;
constexpr u64 W_ROOT = (W_ROOT_VAL << 3) | HVM_VAR;
constexpr u32 W_DEQ = 16;        // per-lane redex deque (power of two)
constexpr u32 W_TBL = HVM_GPU_DEF_MAX;
constexpr u32 W_LINK_CAP = 1u << 16;
constexpr int W_XCHUNK = 8;      // template pairs loaded together at a CALL

__host__ __device__ inline u32 wtag(u64 p) { return (u32)(p & 7u); }
__host__ __device__ inline u64 wval(u64 p) { return p >> 3; }
__host__ __device__ inline u64 wport(u32 t, u64 v) { return (v << 3) | t; }
// A number port: the 32-bit evaluator keeps the low 29 bits of a Numb (its
// port is `numb << 3` in 32 bits, src/hvm.c:197-222); so does this one.
__device__ inline u64 wnum(u32 numb) { return wport(HVM_NUM, numb & 0x1FFFFFFFu); }

struct WRedex {
  u64 a, b;
};

struct WLds {
  WRedex deq[W_DEQ][HVM_WAVE];
  u32 tbl_n[W_TBL], tbl_v[W_TBL];
  WRedex inbox[HVM_WAVE];
  u32 target[HVM_WAVE];
};

struct WLane {
  u32 bot, top;
  u64 bytes, spec;
  u32 spills, voids;
};

// ---- net accessors (every shared access is a device-scope atomic) ----------

__device__ inline u32 nslot(const HvmWideArgs& A, u64 v) { return v < A.node_cap ? (u32)v : A.node_cap; }
__device__ inline u32 vslot(const HvmWideArgs& A, u64 v) {
  return v == W_ROOT_VAL ? 0u : (v < A.vars_cap ? (u32)v : A.vars_cap);
}
__device__ inline u64 ld(const u64* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT); }
__device__ inline void st(u64* p, u64 x) { __hip_atomic_store(p, x, __ATOMIC_RELAXED, AGENT); }
__device__ inline u64 xchg(u64* p, u64 x) { return __hip_atomic_exchange(p, x, __ATOMIC_RELAXED, AGENT); }
__device__ inline void node_put(const HvmWideArgs& A, u32 s, u64 f, u64 g) {
  st(&A.node[2ull * s], f);
  st(&A.node[2ull * s + 1], g);
}
// The creation stores are complete (visible device-wide) before the links
// that publish the new nodes and vars.
__device__ inline void release_creations() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ---- redex deque -------------------------------------------------------------

__device__ inline void spill(const HvmWideArgs& A, u64 a, u64 b) {
  const u64 m = __ballot(1);
  const u32 leader = (u32)__ffsll((long long)m) - 1u;
  u32 base = 0;
  if (lane_id() == leader) base = atomicAdd(&A.ctl->n_out, (u32)__popcll(m));
  const u32 pos = lane_u32(base, (int)leader) + rank_in(m);
  if (pos < A.bag_cap) {
    u64* out = A.bag[A.ctl->parity ^ 1u];
    out[2ull * pos] = a;
    out[2ull * pos + 1] = b;
  } else {
    atomicOr(&A.ctl->err, (u32)HVM_ERR_BAG_OVERFLOW);
  }
}

__device__ inline void push(WLds& L, WLane& ln, const HvmWideArgs& A, u64 a, u64 b) {
  if (ln.top - ln.bot == W_DEQ) {
    WRedex old = L.deq[ln.bot % W_DEQ][lane_id()];
    ln.bot++;
    spill(A, old.a, old.b);
    ln.spills++;
  }
  L.deq[ln.top % W_DEQ][lane_id()] = WRedex{a, b};
  ln.top++;
}

// ---- linker (src/hvm.c:796-842) ----------------------------------------------

__device__ u64 enter(const HvmWideArgs& A, u64 b) {
  for (u32 it = 0; it < W_LINK_CAP; ++it) {
    if (wtag(b) != HVM_VAR) return b;
    u64* slot = &A.vars[vslot(A, wval(b))];
    u64 v = xchg(slot, W_NONE);
    if (v == W_NONE || v == 0) return b;
    st(slot, 0);
    b = v;
  }
  atomicOr(&A.ctl->err, (u32)HVM_ERR_LINK_LOOP);
  return b;
}

__device__ void link(WLds& L, WLane& ln, const HvmWideArgs& A, u64 a, u64 b) {
  for (u32 it = 0; it < W_LINK_CAP; ++it) {
    if (wtag(a) != HVM_VAR && wtag(b) == HVM_VAR) { u64 t = a; a = b; b = t; }
    if (wtag(a) != HVM_VAR) {
      // two nullary ports make a VOID redex, counted here instead of run
      if (wtag(a) - 1u < 3u && wtag(b) - 1u < 3u) ln.voids++;
      else push(L, ln, A, a, b);
      return;
    }
    b = enter(A, b);
    u64* slot = &A.vars[vslot(A, wval(a))];
    u64 a2 = xchg(slot, b);
    if (a2 == W_NONE) return;
    if (a2 == 0) { atomicOr(&A.ctl->err, (u32)HVM_ERR_BAD_PORT); return; }
    st(slot, 0);
    a = a2;
  }
  atomicOr(&A.ctl->err, (u32)HVM_ERR_LINK_LOOP);
}

// ---- allocation ----------------------------------------------------------------

// Fills tbl[0, need) with free slots (wave-uniform `need`).  The wave's own
// slice [lo, lo + len) is scanned 64 slots at a time from `cur` (only this
// wave allocates there); then the shared pool [plo, cap) with a
// compare-and-swap claim per slot.  `words` is 2 for nodes (the first half
// tells free), 1 for vars.  Returns false if the net is full.
__device__ bool wave_alloc(const HvmWideArgs& A, u64* buf, u32 words, u32 lo, u32 len, u32& cur, u32 plo,
                           u32 cap, u32& pcur, u32 need, u32* tbl) {
  const u32 lane = lane_id();
  u32 got = 0;
  for (u32 scanned = 0; got < need && scanned < len; scanned += HVM_WAVE) {
    const u32 s = lo + ((cur + lane) % len);
    const bool fr = ld(&buf[(u64)words * s]) == 0;
    const u64 m = __ballot(fr);
    const u32 r = rank_in(m);
    if (fr && got + r < need) tbl[got + r] = s;
    got = min(need, got + (u32)__popcll(m));
    cur = (cur + HVM_WAVE) % len;
  }
  const u32 plen = cap > plo ? ((cap - plo) / HVM_WAVE) * HVM_WAVE : 0;
  for (u32 scanned = 0; got < need && scanned < plen; scanned += HVM_WAVE) {
    const u32 s = plo + ((pcur + lane) % plen);
    bool fr = ld(&buf[(u64)words * s]) == 0;
    const u32 r0 = rank_in(__ballot(fr));
    bool mine = false;
    if (fr && got + r0 < need) {
      unsigned long long expect = 0;
      mine = __hip_atomic_compare_exchange_strong(&buf[(u64)words * s], &expect, W_NONE, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, AGENT);
    }
    const u64 m = __ballot(mine);
    const u32 r = rank_in(m);
    if (mine) tbl[got + r] = s;
    got += (u32)__popcll(m);
    pcur = (pcur + HVM_WAVE) % plen;
  }
  return got >= need;
}

// A reserved slot that the interaction did not use goes back: pool slots
// were claimed by writing their first word.
__device__ inline void unreserve(const HvmWideArgs& A, u64* buf, u32 words, u32 plo, u32 s) {
  if (s >= plo) st(&buf[(u64)words * s], 0);
}

// ---- the turn --------------------------------------------------------------------

__global__ __launch_bounds__(HVM_WAVE) void hvm_wide_turn_kernel(HvmWideArgs A) {
  __shared__ WLds L;
  HvmWideCtl* C = A.ctl;
  if (C->done) return;
  const u32 lane = lane_id();
  const u32 w = blockIdx.x;
  const u32 n_in = C->n_in, par = C->parity, max_steps = C->max_steps, idle_quit = C->idle_quit;
  const u64* bin = A.bag[par];
  HvmWaveState ws = A.wave_state[w];
  const u32 n_lo = 1 + w * A.node_slice, v_lo = 1 + w * A.vars_slice;
  if (ws.node_cur >= A.node_slice) ws.node_cur = 0;
  if (ws.vars_cur >= A.vars_slice) ws.vars_cur = 0;
  WLane ln{0, 0, 0, 0, 0, 0};
  u64 kinds[HVM_K_KINDS];
#pragma unroll
  for (int k = 0; k < HVM_K_KINDS; ++k) kinds[k] = 0;
  u64 steps = 0, defers = 0, retries = 0;

  {
    const u32 idx = w + lane * A.nwaves;
    if (idx < n_in && idx < HVM_WAVE * A.nwaves) push(L, ln, A, bin[2ull * idx], bin[2ull * idx + 1]);
  }
  bool more = n_in > HVM_WAVE * A.nwaves;

  for (u32 step = 0; step < max_steps; ++step) {
    // refill from the input bag
    if (more) {
      const bool empty = ln.top == ln.bot;
      const u64 em = __ballot(empty);
      if (em) {
        const u32 k = (u32)__popcll(em);
        u32 base = 0;
        if (lane == 0) base = atomicAdd(&C->in_head, k) + HVM_WAVE * A.nwaves;
        base = lane_u32(base, 0);
        if (base >= n_in) more = false;
        const u32 r = rank_in(em);
        if (empty && base + r < n_in) push(L, ln, A, bin[2ull * (base + r)], bin[2ull * (base + r) + 1]);
      }
    }
    // idle lanes take the oldest redex of rich lanes
    {
      const u32 n = ln.top - ln.bot;
      const bool needy = n == 0, rich = n >= 2;
      const u64 nm = __ballot(needy), rm = __ballot(rich);
      if (nm && rm) {
        const u32 k = min((u32)__popcll(nm), (u32)__popcll(rm));
        const u32 rn = rank_in(nm), rr = rank_in(rm);
        if (needy && rn < k) L.target[rn] = lane;
        __syncthreads();
        if (rich && rr < k) {
          L.inbox[L.target[rr]] = L.deq[ln.bot % W_DEQ][lane];
          ln.bot++;
        }
        __syncthreads();
        if (needy && rn < k) push(L, ln, A, L.inbox[lane].a, L.inbox[lane].b);
      }
    }
    const bool has = ln.top != ln.bot;
    if (!__ballot(has)) {
      if (more) continue;
      if (lane == 0) atomicAdd(&C->idle_waves, 1u);
      break;
    }
    if ((step & 15) == 15) {
      // a turn that started without idle quit (few redexes, not growing) gets
      // it back once some wave holds more than two redexes per lane: work
      // grew inside the turn, and only the end of the turn spreads it
      u32 iq = idle_quit;
      if (!iq && C->idle_base) {
        u32 sur = __hip_atomic_load(&C->surplus, __ATOMIC_RELAXED, AGENT);
        if (!sur && wave_sum64(ln.top - ln.bot) > 2u * HVM_WAVE) {
          if (lane == 0) __hip_atomic_store(&C->surplus, 1u, __ATOMIC_RELAXED, AGENT);
          sur = 1;
        }
        if (sur) iq = C->idle_base;
      }
      const u32 idle = __hip_atomic_load(&C->idle_waves, __ATOMIC_RELAXED, AGENT);
      if (iq && idle >= iq) break;
      if (__hip_atomic_load(&C->err, __ATOMIC_RELAXED, AGENT) != 0) break;
    }
    steps++;

    // classify
    const WRedex rx = has ? L.deq[(ln.top - 1) % W_DEQ][lane] : WRedex{0, 0};
    u64 a = rx.a, b = rx.b;
    u32 rule = hvm_get_rule((u32)a, (u32)b);
    if (wtag(a) == HVM_REF && b == W_ROOT) rule = HVM_CALL;
    else if (wtag(b) < wtag(a)) { u64 t = a; a = b; b = t; }
    HvmDevDef d;
    d.rbag_off = d.node_off = d.rlen = d.nlen = d.vlen = d.root = d.safe = d.call_bytes = 0;
    bool copy = false;
    u32 nn = 0, nv = 0;
    if (has) {
      switch (rule) {
        case HVM_CALL: {
          const u32 fid = (u32)wval(a) & 0xFFFFFFFu;
          d = A.defs[fid < A.ndefs ? fid : 0].h;
          if ((d.safe & 1u) && wtag(b) == HVM_DUP) copy = true;
          else { nn = d.nlen; nv = d.vlen; }
          break;
        }
        case HVM_COMM: nn = 4; nv = 4; break;
        case HVM_OPER: nn = 1; break;
        case HVM_SWIT: nn = 2; break;
        default: break;
      }
    }
    u32 tot;
    const u32 o = wave_scan(has ? (nn | (nv << 16)) : 0u, &tot);
    const u32 off_n = o & 0xFFFFu, off_v = o >> 16;
    const bool run = has && off_n + nn <= W_TBL && off_v + nv <= W_TBL;
    if (has && !run) defers++;
    const u32 need_n = wave_max(run ? off_n + nn : 0u);
    const u32 need_v = wave_max(run ? off_v + nv : 0u);
    if (need_n && !wave_alloc(A, A.node, 2, n_lo, A.node_slice, ws.node_cur, A.node_pool_lo, A.node_cap,
                              ws.node_ovf, need_n, L.tbl_n)) {
      if (lane == 0) atomicOr(&C->err, (u32)HVM_ERR_OOM_NODE);
      break;
    }
    if (need_v && !wave_alloc(A, A.vars, 1, v_lo, A.vars_slice, ws.vars_cur, A.vars_pool_lo, A.vars_cap,
                              ws.vars_ovf, need_v, L.tbl_v)) {
      if (lane == 0) atomicOr(&C->err, (u32)HVM_ERR_OOM_VARS);
      break;
    }
    __syncthreads();

    // takes: two coherent loads per node; the slot is freed once both returned
    const bool take_a = run && (rule == HVM_ANNI || rule == HVM_COMM);
    const bool take_b = run && (rule == HVM_ERAS || rule == HVM_ANNI || rule == HVM_COMM || rule == HVM_OPER ||
                                rule == HVM_SWIT || (rule == HVM_CALL && copy));
    const u32 sa = nslot(A, wval(a)), sb = nslot(A, wval(b));
    u64 A1 = 0, A2 = 0, B1 = 0, B2 = 0;
    if (take_a) { A1 = ld(&A.node[2ull * sa]); A2 = ld(&A.node[2ull * sa + 1]); }
    if (take_b) { B1 = ld(&A.node[2ull * sb]); B2 = ld(&A.node[2ull * sb + 1]); }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // a node whose creation is not visible yet: try again next step
    const bool retry = (take_a && A1 == 0) || (take_b && B1 == 0);
    if (retry) retries++;
    const bool go = run && !retry;
    if (go) {
      ln.top--;
      if (take_a) st(&A.node[2ull * sa], 0);
      if (take_b) st(&A.node[2ull * sb], 0);
    }
    if (run && retry) {
      for (u32 i = 0; i < nn; ++i) unreserve(A, A.node, 2, A.node_pool_lo, L.tbl_n[off_n + i]);
      for (u32 i = 0; i < nv; ++i) unreserve(A, A.vars, 1, A.vars_pool_lo, L.tbl_v[off_v + i]);
    }

    // interactions (src/hvm.cu:1197-1364)
    u32 kk = HVM_K_KINDS, kb = 0;
    if (go) {
      switch (rule) {
        case HVM_LINK:
          link(L, ln, A, a, b);
          break;
        case HVM_VOID:
          kk = HVM_K_VOID;
          break;
        case HVM_ERAS:
          link(L, ln, A, a, B1);
          link(L, ln, A, a, B2);
          kk = HVM_K_ERAS; kb = 32;
          break;
        case HVM_ANNI:
          link(L, ln, A, A1, B1);
          link(L, ln, A, A2, B2);
          kk = HVM_K_ANNI; kb = 48;
          break;
        case HVM_COMM: {
          const u32 v0 = L.tbl_v[off_v], v1 = L.tbl_v[off_v + 1], v2 = L.tbl_v[off_v + 2], v3 = L.tbl_v[off_v + 3];
          const u32 l0 = L.tbl_n[off_n], l1 = L.tbl_n[off_n + 1], l2 = L.tbl_n[off_n + 2], l3 = L.tbl_n[off_n + 3];
          st(&A.vars[v0], W_NONE); st(&A.vars[v1], W_NONE); st(&A.vars[v2], W_NONE); st(&A.vars[v3], W_NONE);
          node_put(A, l0, wport(HVM_VAR, v0), wport(HVM_VAR, v1));
          node_put(A, l1, wport(HVM_VAR, v2), wport(HVM_VAR, v3));
          node_put(A, l2, wport(HVM_VAR, v0), wport(HVM_VAR, v2));
          node_put(A, l3, wport(HVM_VAR, v1), wport(HVM_VAR, v3));
          release_creations();
          link(L, ln, A, wport(wtag(b), l0), A1);
          link(L, ln, A, wport(wtag(b), l1), A2);
          link(L, ln, A, wport(wtag(a), l2), B1);
          link(L, ln, A, wport(wtag(a), l3), B2);
          kk = HVM_K_COMM; kb = 112;
          break;
        }
        case HVM_OPER: {
          const u32 l0 = L.tbl_n[off_n];
          const u64 b2 = enter(A, B2);
          if (wtag(B1) == HVM_NUM) {
            unreserve(A, A.node, 2, A.node_pool_lo, l0);
            link(L, ln, A, wnum(hvm_operate((u32)wval(a), (u32)wval(B1))), b2);
            kk = HVM_K_OPER1; kb = 32;
          } else {
            node_put(A, l0, a, b2);
            release_creations();
            link(L, ln, A, B1, wport(HVM_OPR, l0));
            kk = HVM_K_OPER2; kb = 40;
          }
          break;
        }
        case HVM_SWIT: {
          const u32 l0 = L.tbl_n[off_n], l1 = L.tbl_n[off_n + 1];
          const u32 av = hvm_get_u24((u32)wval(a));
          if (av == 0) {
            unreserve(A, A.node, 2, A.node_pool_lo, l1);
            node_put(A, l0, B2, wport(HVM_ERA, 0));
            release_creations();
            link(L, ln, A, wport(HVM_CON, l0), B1);
            kk = HVM_K_SWIT0; kb = 32;
          } else {
            node_put(A, l0, wport(HVM_ERA, 0), wport(HVM_CON, l1));
            node_put(A, l1, wnum(hvm_new_u24(av - 1)), B2);
            release_creations();
            link(L, ln, A, wport(HVM_CON, l0), B1);
            kk = HVM_K_SWITN; kb = 40;
          }
          break;
        }
        case HVM_CALL: {
          if (copy) {
            link(L, ln, A, a, B1);
            link(L, ln, A, a, B2);
            kk = HVM_K_CALL_COPY; kb = 32;
            break;
          }
          // template ports: nodes and vars by table, the rest widened as is
          auto adj = [&](u32 p) -> u64 {
            const u32 t = p & 7u, v = p >> 3;
            if (t >= HVM_CON) return wport(t, L.tbl_n[off_n + v]);
            if (t == HVM_VAR) return wport(HVM_VAR, L.tbl_v[off_v + v]);
            return (u64)p;
          };
          // template pairs eight at a time, each chunk's loads issued
          // together before its stores, and the var stores last: a wait for
          // a load also waits for every store issued before it, so loading
          // one pair per node between the stores cost a store round trip
          // per node
          for (u32 c = 0; c < d.nlen; c += W_XCHUNK) {
            u64 xp[W_XCHUNK];
#pragma unroll
            for (int k = 0; k < W_XCHUNK; ++k) xp[k] = c + k < d.nlen ? A.pairs[d.node_off + c + k] : 0ull;
#pragma unroll
            for (int k = 0; k < W_XCHUNK; ++k)
              if (c + k < d.nlen) node_put(A, L.tbl_n[off_n + c + k], adj(hvm_get_fst(xp[k])), adj(hvm_get_snd(xp[k])));
          }
          for (u32 i = 0; i < d.vlen; ++i) st(&A.vars[L.tbl_v[off_v + i]], W_NONE);
          // the first redex pairs load while the creation stores complete
          u64 xr[W_XCHUNK];
#pragma unroll
          for (int k = 0; k < W_XCHUNK; ++k) xr[k] = (u32)k < d.rlen ? A.pairs[d.rbag_off + k] : 0ull;
          release_creations();
          for (u32 c = 0; c < d.rlen; c += W_XCHUNK) {
            if (c) {
#pragma unroll
              for (int k = 0; k < W_XCHUNK; ++k) xr[k] = c + k < d.rlen ? A.pairs[d.rbag_off + c + k] : 0ull;
            }
            const u32 m = min(d.rlen - c, (u32)W_XCHUNK);
            for (u32 k = 0; k < m; ++k) {
              link(L, ln, A, adj(hvm_get_fst(xr[0])), adj(hvm_get_snd(xr[0])));
#pragma unroll
              for (int j = 0; j + 1 < W_XCHUNK; ++j) xr[j] = xr[j + 1];
            }
          }
          link(L, ln, A, adj(d.root), b);
          kk = HVM_K_CALL; kb = d.call_bytes;
          // a compiled def's static interactions (zero for plain defs)
          ln.spec += (u64)((d.safe >> 8) & 0xFFu) | ((u64)(d.safe >> 16) << 32);
          break;
        }
        default:
          break;
      }
    }
    ln.bytes += kb;
#pragma unroll
    for (int k = 0; k < HVM_K_KINDS; ++k) kinds[k] += (u64)__popcll(__ballot(kk == (u32)k));
    __syncthreads();   // the tables are rewritten next step
  }

  // input redexes no wave pulled (a turn cut short) go to the output bag
  while (more) {
    u32 base = 0;
    if (lane == 0) base = atomicAdd(&C->in_head, (u32)HVM_WAVE) + HVM_WAVE * A.nwaves;
    base = lane_u32(base, 0);
    if (base >= n_in) break;
    if (base + lane < n_in) spill(A, bin[2ull * (base + lane)], bin[2ull * (base + lane) + 1]);
  }
  // what the lanes still hold goes to the output bag
  for (;;) {
    const bool h = ln.top != ln.bot;
    if (!__ballot(h)) break;
    if (h) {
      const WRedex r = L.deq[ln.bot % W_DEQ][lane];
      ln.bot++;
      spill(A, r.a, r.b);
    }
  }

  kinds[HVM_K_CALL] += wave_sum64(ln.spec & 0xFFFFFFFFull);
  kinds[HVM_K_ANNI] += wave_sum64(ln.spec >> 32);
  kinds[HVM_K_VOID] += wave_sum64(ln.voids);
  u64 s_itrs = 0;
#pragma unroll
  for (int k = 0; k < HVM_K_KINDS; ++k) s_itrs += kinds[k];
  const u64 s_bytes = wave_sum64(ln.bytes), s_spill = wave_sum64(ln.spills);
  defers = wave_sum64(defers);
  if (lane == 0) {
    HvmDevStats* St = A.stats + (w & (HVM_STAT_STRIPES - 1));   // one of the counters' copies
    if (s_itrs) atomicAdd(&St->itrs, s_itrs);
    if (s_bytes) atomicAdd(&St->bytes, s_bytes);
    if (s_spill) atomicAdd(&St->spills, s_spill);
#pragma unroll
    for (int k = 0; k < HVM_K_KINDS; ++k) if (kinds[k]) atomicAdd(&St->kind[k], kinds[k]);
    if (steps) atomicAdd(&St->steps, steps);
    if (defers) atomicAdd(&St->defers, defers);
    if (retries) atomicAdd(&St->retries, retries);
    A.wave_state[w] = ws;
  }
}

// Between two turns: the output bag becomes the input; a spare slot that was
// written means a port pointed outside the net.
__global__ void hvm_wide_between_kernel(HvmWideArgs A) {
  HvmWideCtl* C = A.ctl;
  if (C->done) return;
  if (A.node[2ull * A.node_cap] || A.node[2ull * A.node_cap + 1] || A.vars[A.vars_cap])
    C->err |= HVM_ERR_BAD_PORT;
  const u32 n = min(C->n_out, A.bag_cap);
  // few redexes, not growing: sequential chains that a redistribution cannot
  // spread; the next turn runs without idle quit
  C->idle_quit = (n <= A.nwaves / 4 && 2 * n <= 3 * C->n_in) ? 0u : C->idle_base;
  C->n_in = n;
  C->n_out = 0;
  C->in_head = 0;
  C->parity ^= 1u;
  C->turns += 1;
  C->idle_waves = 0;
  C->surplus = 0;
  if (C->n_in == 0 || C->err) C->done = 1;
}

// Readback: the values of listed node and var slots.
__global__ void hvm_wide_fetch_kernel(const u64* node, const u64* vars, const u32* nidx, u32 nn, const u32* vidx,
                                      u32 nv, u64* out_n, u64* out_v) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) {
    out_n[2ull * i] = node[2ull * nidx[i]];
    out_n[2ull * i + 1] = node[2ull * nidx[i] + 1];
  }
  if (i < nv) out_v[i] = vars[vidx[i]];
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

bool HvmWide::create(uint64_t node_cap, uint64_t vars_cap, uint32_t bag_cap, uint32_t waves, int cus,
                     hipStream_t s) {
  stream = s;
  if (node_cap < 4096 || vars_cap < 4096 || node_cap > 0xFFFFFFFEull || vars_cap > 0xFFFFFFFEull) {
    err = "wide net capacities must be within [2^12, 2^32 - 2] slots";
    return false;
  }
  a.node_cap = (u32)node_cap;
  a.vars_cap = (u32)vars_cap;
  a.bag_cap = bag_cap;
  // slices of at least 256 slots, multiples of 64, over three quarters of the
  // net; the rest is the shared pool
  u32 nw = waves ? waves : (u32)std::max(1, cus) * 8u;
  const u64 room = std::min((node_cap - 1) * 3 / 4, (vars_cap - 1) * 3 / 4);
  nw = (u32)std::max<u64>(1, std::min<u64>(nw, room / 256));
  a.nwaves = nw;
  a.node_slice = (u32)((((node_cap - 1) * 3 / 4) / nw) / HVM_WAVE * HVM_WAVE);
  a.vars_slice = (u32)((((vars_cap - 1) * 3 / 4) / nw) / HVM_WAVE * HVM_WAVE);
  a.node_pool_lo = 1 + nw * a.node_slice;
  a.vars_pool_lo = 1 + nw * a.vars_slice;
  auto ok = [&](hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  };
  return ok(hipMalloc(&a.node, (node_cap + 1) * 16), "hipMalloc(wide node)") &&
         ok(hipMalloc(&a.vars, (vars_cap + 1) * 8), "hipMalloc(wide vars)") &&
         ok(hipMalloc(&a.bag[0], (size_t)bag_cap * 16), "hipMalloc(wide bag)") &&
         ok(hipMalloc(&a.bag[1], (size_t)bag_cap * 16), "hipMalloc(wide bag)") &&
         ok(hipMalloc(&a.ctl, sizeof(HvmWideCtl)), "hipMalloc(wide ctl)") &&
         ok(hipMalloc(&a.wave_state, nw * sizeof(HvmWaveState)), "hipMalloc(wide wave state)") &&
         ok(hipMalloc(&a.stats, sizeof(HvmDevStats) * HVM_STAT_STRIPES), "hipMalloc(wide stats)");
}

void HvmWide::destroy() {
  // teardown: a failing free has nothing left to undo
  (void)hipFree(a.node); (void)hipFree(a.vars); (void)hipFree(a.bag[0]); (void)hipFree(a.bag[1]);
  (void)hipFree(a.ctl); (void)hipFree(a.wave_state); (void)hipFree(a.stats);
  a = HvmWideArgs{};
}

bool HvmWide::ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  err = std::string(what) + ": " + hipGetErrorString(e);
  return false;
}

bool HvmWide::boot(unsigned long long x, unsigned long long y) {
  const u64 root_none = W_NONE;
  const u64 redex[2] = {x, y};
  HvmWideCtl c;
  memset(&c, 0, sizeof c);
  c.n_in = 1;
  c.max_steps = turn_steps ? std::min<u32>(turn_steps, 1u << 20) : 4096u;
  c.idle_quit = idle_pct > 100 ? 0u : std::max<u32>(1u, (u32)(((u64)a.nwaves * idle_pct) / 100));
  c.idle_base = c.idle_quit;
  return ok(hipMemcpyAsync(a.vars, &root_none, 8, hipMemcpyHostToDevice, stream), "boot") &&
         ok(hipMemcpyAsync(a.bag[0], redex, 16, hipMemcpyHostToDevice, stream), "boot") &&
         ok(hipMemcpyAsync(a.ctl, &c, sizeof c, hipMemcpyHostToDevice, stream), "boot") &&
         ok(hipStreamSynchronize(stream), "boot");
}

int HvmWide::run() {
  HvmWideCtl c;
  for (u32 chunk = 2;; chunk = std::min<u32>(chunk * 2, 64)) {
    for (u32 i = 0; i < chunk; ++i) {
      hipLaunchKernelGGL(hvm_wide_turn_kernel, dim3(a.nwaves), dim3(HVM_WAVE), 0, stream, a);
      hipLaunchKernelGGL(hvm_wide_between_kernel, dim3(1), dim3(1), 0, stream, a);
      launches++;
    }
    if (!ok(hipGetLastError(), "launch wide turn") ||
        !ok(hipMemcpyAsync(&c, a.ctl, sizeof c, hipMemcpyDeviceToHost, stream), "read wide ctl") ||
        !ok(hipStreamSynchronize(stream), "wide turn"))
      return HVM_HIP_E_DEVICE;
    if (c.done) break;
    if (deadline > 0 && now_s() > deadline) {
      err = "time limit reached";
      return HVM_HIP_E_EVAL;
    }
  }
  turns += c.turns;
  if (c.err) {
    char b[96];
    snprintf(b, sizeof b, "evaluator error flags 0x%x (turn %u)", c.err, turns);
    err = b;
    return HVM_HIP_E_EVAL;
  }
  return HVM_HIP_OK;
}

// A failed copy is sticky: later reads return 0 and expand fails.
unsigned long long HvmWide::vars_load(unsigned long long val) {
  const u64 s = val == W_ROOT_VAL ? 0 : std::min<u64>(val, a.vars_cap);
  u64 v = 0;
  if (!copy_failed) copy_failed = !ok(hipMemcpy(&v, a.vars + s, 8, hipMemcpyDeviceToHost), "readback copy");
  return v;
}

void HvmWide::node_load(unsigned long long loc, unsigned long long& f, unsigned long long& g) {
  u64 v[2] = {0, 0};
  if (!copy_failed)
    copy_failed = !ok(hipMemcpy(v, a.node + 2 * std::min<u64>(loc, a.node_cap), 16, hipMemcpyDeviceToHost), "readback copy");
  f = v[0];
  g = v[1];
}

unsigned long long HvmWide::peek(unsigned long long p) {
  for (int it = 0; it < (1 << 24) && wtag(p) == HVM_VAR && !copy_failed; ++it) {
    const u64 v = vars_load(wval(p));
    if (v == W_NONE || v == 0) break;
    p = v;
  }
  return p;
}

// src/hvm.c:1233-1244: a REF met at readback is called, its net normalised.
int HvmWide::expand(unsigned long long p, unsigned long long& got) {
  const u64 old = vars_load(W_ROOT_VAL);
  got = peek(p);
  while (wtag(got) == HVM_REF) {
    if (!boot(got, W_ROOT)) return HVM_HIP_E_DEVICE;
    int rc = run();
    if (rc) return rc;
    got = peek(vars_load(W_ROOT_VAL));
  }
  if (copy_failed) return HVM_HIP_E_DEVICE;
  return ok(hipMemcpy(a.vars, &old, 8, hipMemcpyHostToDevice), "expand") ? HVM_HIP_OK : HVM_HIP_E_DEVICE;
}

int HvmWide::normalize(const HvmDevDefX* defs, const unsigned long long* pairs, uint32_t ndefs, double dl,
                       HvmDevStats& out, uint32_t& n_turns, uint32_t& n_launches) {
  a.defs = defs;
  a.pairs = pairs;
  a.ndefs = ndefs;
  deadline = dl;
  turns = launches = 0;
  copy_failed = false;
  if (!ok(hipMemsetAsync(a.node, 0, ((size_t)a.node_cap + 1) * 16, stream), "clear wide net") ||
      !ok(hipMemsetAsync(a.vars, 0, ((size_t)a.vars_cap + 1) * 8, stream), "clear wide net") ||
      !ok(hipMemsetAsync(a.wave_state, 0, a.nwaves * sizeof(HvmWaveState), stream), "clear wide net") ||
      !ok(hipMemsetAsync(a.stats, 0, sizeof(HvmDevStats) * HVM_STAT_STRIPES, stream), "clear wide stats"))
    return HVM_HIP_E_DEVICE;
  if (!boot(wport(HVM_REF, 0), W_ROOT)) return HVM_HIP_E_DEVICE;
  int rc = run();
  // The IO driver's first readback (src/run.c:746-760): the root read as a
  // lambda-encoded constructor, expanding the REFs on the way (their
  // interactions count).  An IO call is not driven here.
  if (rc == HVM_HIP_OK) {
    u64 lam = 0, app = 0, arg = 0, f = 0, g = 0;
    rc = expand(peek(W_ROOT), lam);
    std::vector<u64> args;
    int32_t tag = -1;
    if (rc == HVM_HIP_OK && wtag(lam) == HVM_CON) {
      node_load(wval(lam), f, g);
      rc = expand(f, app);
      if (rc == HVM_HIP_OK && wtag(app) == HVM_CON) {
        u64 af = 0, ag = 0;
        node_load(wval(app), af, ag);
        rc = expand(af, arg);
        if (rc == HVM_HIP_OK && wtag(arg) == HVM_NUM) {
          tag = (int32_t)hvm_get_u24((u32)wval(arg));
          for (u32 it = 0; rc == HVM_HIP_OK && it < (1u << 24); ++it) {
            rc = expand(ag, app);
            if (rc || wtag(app) != HVM_CON) break;
            node_load(wval(app), af, ag);
            rc = expand(af, arg);
            args.push_back(arg);
          }
        }
      }
    }
    if (rc == HVM_HIP_OK && !args.empty() && wtag(args[0]) == HVM_CON && tag == 1) {
      u64 m0 = 0, m1 = 0;
      node_load(wval(args[0]), m0, m1);
      if (copy_failed) rc = HVM_HIP_E_DEVICE;
      else if ((u32)wval(m0) == hvm_new_u24(0xD0CA11) && (u32)wval(m1) == hvm_new_u24(0xFF1FF1)) {
        err = "IO calls run on the 32-bit evaluator (no HVM_HIP_WIDE)";
        rc = HVM_HIP_E_IO;
      }
    }
  }
  n_turns = turns;
  n_launches = launches;
  {
    // the counters' copies, summed
    std::vector<HvmDevStats> all(HVM_STAT_STRIPES);
    if (!ok(hipMemcpy(all.data(), a.stats, sizeof(HvmDevStats) * HVM_STAT_STRIPES, hipMemcpyDeviceToHost),
            "wide stats"))
      return HVM_HIP_E_DEVICE;
    memset(&out, 0, sizeof out);
    unsigned long long* sum = (unsigned long long*)&out;
    for (const HvmDevStats& x : all) {
      const unsigned long long* v = (const unsigned long long*)&x;
      for (size_t k = 0; k < sizeof(HvmDevStats) / 8; ++k) sum[k] += v[k];
    }
  }
  return rc;
}

bool HvmWide::print_root(const std::vector<std::string>& names, std::string& out) {
  // Fetch everything reachable from ROOT, one level per round.
  std::unordered_map<u32, std::pair<u64, u64>> nodes;
  std::unordered_map<u32, u64> vars;
  std::vector<u32> want_n, want_v{0};
  vars.reserve(1024);
  u32 *d_idx = nullptr;
  u64* d_out = nullptr;
  size_t cap = 0;
  bool good = true;
  auto want_port = [&](u64 p) {
    const u32 t = wtag(p);
    if (t >= HVM_CON) {
      const u64 v = wval(p);
      if (v < a.node_cap && !nodes.count((u32)v)) { nodes[(u32)v] = {0, 0}; want_n.push_back((u32)v); }
    } else if (t == HVM_VAR) {
      const u64 v = wval(p);
      const u32 s = v == W_ROOT_VAL ? 0u : (v < a.vars_cap ? (u32)v : a.vars_cap);
      if (!vars.count(s)) { vars[s] = W_NONE; want_v.push_back(s); }
    }
  };
  vars[0] = W_NONE;
  while (good && (!want_n.empty() || !want_v.empty())) {
    const size_t nn = want_n.size(), nv = want_v.size(), need = std::max<size_t>(nn + nv, 1);
    if (need > cap) {
      (void)hipFree(d_idx); (void)hipFree(d_out);
      cap = need * 2;
      good = hipMalloc(&d_idx, cap * 4) == hipSuccess && hipMalloc(&d_out, cap * 16) == hipSuccess;
      if (!good) break;
    }
    std::vector<u32> idx(want_n);
    idx.insert(idx.end(), want_v.begin(), want_v.end());
    std::vector<u64> res(2 * nn + nv);
    good = hipMemcpy(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (!good) break;
    const u32 m = (u32)std::max(nn, nv);
    hipLaunchKernelGGL(hvm_wide_fetch_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, a.node, a.vars, d_idx,
                       (u32)nn, d_idx + nn, (u32)nv, d_out, d_out + 2 * nn);
    good = hipStreamSynchronize(stream) == hipSuccess &&
           hipMemcpy(res.data(), d_out, res.size() * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (!good) break;
    std::vector<u32> got_n(want_n), got_v(want_v);
    want_n.clear();
    want_v.clear();
    for (size_t i = 0; i < nn; ++i) {
      nodes[got_n[i]] = {res[2 * i], res[2 * i + 1]};
      want_port(res[2 * i]);
      want_port(res[2 * i + 1]);
    }
    for (size_t i = 0; i < nv; ++i) {
      vars[got_v[i]] = res[2 * nn + i];
      if (res[2 * nn + i] != W_NONE && res[2 * nn + i] != 0) want_port(res[2 * nn + i]);
    }
  }
  (void)hipFree(d_idx);   // scratch of this readback: nothing to undo if a free fails
  (void)hipFree(d_out);
  if (!good) { err = "wide readback failed"; return false; }
  // print (src/hvm.c:1618-1699)
  out.clear();
  std::vector<u64> stack{W_ROOT};
  u64 budget = 1ull << 32;
  while (!stack.empty() && budget--) {
    const u64 p = stack.back();
    stack.pop_back();
    const u32 t = wtag(p);
    switch (t) {
      case HVM_CON: case HVM_DUP: case HVM_OPR: case HVM_SWI: {
        auto it = nodes.find((u32)wval(p));
        const std::pair<u64, u64> nd = it == nodes.end() ? std::pair<u64, u64>{0, 0} : it->second;
        out += t == HVM_CON ? "(" : t == HVM_DUP ? "{" : t == HVM_OPR ? "$(" : "?(";
        stack.push_back(wport(HVM_ERA, t == HVM_DUP ? '}' : ')'));
        stack.push_back(nd.second);
        stack.push_back(wport(HVM_ERA, ' '));
        stack.push_back(nd.first);
        break;
      }
      case HVM_ERA:
        if (wval(p)) out += (char)wval(p); else out += "*";
        break;
      case HVM_VAR: {
        const u64 v = wval(p);
        const u32 s = v == W_ROOT_VAL ? 0u : (v < a.vars_cap ? (u32)v : a.vars_cap);
        auto it = vars.find(s);
        const u64 got = it == vars.end() ? W_NONE : it->second;
        if (got != W_NONE && got != 0) {
          stack.push_back(got);
        } else {
          char b[32];
          snprintf(b, sizeof b, "x%llx", (unsigned long long)v);
          out += b;
        }
        break;
      }
      case HVM_NUM: out += hvm::c_show_numb((u32)wval(p)); break;
      case HVM_REF: {
        const u32 fid = (u32)wval(p) & 0xFFFFFFFu;
        out += "@" + (fid < names.size() ? names[fid] : std::string("?"));
        break;
      }
    }
  }
  return true;
}
