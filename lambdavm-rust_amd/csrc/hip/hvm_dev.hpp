// hvm_dev.hpp — device-side layout shared by the evaluator kernels and the
// host runtime.  Everything here is plain data: pointers and sizes.
#pragma once
#include <stdint.h>

#include "../hvm_core.h"

// ---------------------------------------------------------------------------
// Tunables that shape the LDS footprint of one evaluator wave.
// ---------------------------------------------------------------------------

#define HVM_WAVE        64     // gfx950 wavefront width
#ifndef HVM_DEQ
#define HVM_DEQ         32     // per-lane redex deque entries (power of two)
#endif
// resident waves per SIMD the HBM turn kernel is compiled for (its registers
// are capped so that this many fit; LDS allows 8 waves per CU)
#ifndef HVM_TURN_OCC
#define HVM_TURN_OCC    2
#endif
#define HVM_TBL_NODE    256    // per-wave node-slot table (one step's allocations)
#define HVM_TBL_VARS    256    // per-wave var-slot table
#ifndef HVM_TBL_SLACK
#define HVM_TBL_SLACK   8      // entries past the tables, read ahead and never used
#endif
#define HVM_TBL_X       128    // tables of the local-net evaluator's HBM slots
                               // (LEAK and SAVE steps)

// Local net (the local-net evaluator): each wave owns HVM_LNODE node slots and
// HVM_LVARS var slots in LDS.  A node or var port whose value is
// HVM_LOCAL_BASE + i (and not ROOT) addresses slot i of the owning wave's
// local net; HBM slots are below HVM_LOCAL_BASE, so the net's capacities are
// at most 2^28 slots with it.
#ifndef HVM_LNODE
#define HVM_LNODE       1024
#endif
#ifndef HVM_LVARS
#define HVM_LVARS       1024
#endif
#define HVM_LOCAL_BASE  0x10000000u
// per-wave HBM stack for deque overflow that holds local ports
#define HVM_LSPILL      1024

// Largest def the evaluator instantiates in one interaction.  Defs are
// checked at load time.
#define HVM_GPU_DEF_MAX 256

// ROOT is var 0x1FFFFFFF; the evaluator stores it in var slot 0, which the
// allocators never hand out (src/hvm.c:717: slot 0 is skipped).
#define HVM_ROOT_VAL 0x1FFFFFFFu

// ---------------------------------------------------------------------------
// Book as the device reads it: one record per fid plus a flat pair array.
// ---------------------------------------------------------------------------

struct HvmDevDef {
  uint32_t rbag_off;   // index into pairs[] of the def's redexes
  uint32_t node_off;   // index into pairs[] of the def's nodes
  uint32_t rlen, nlen, vlen;
  uint32_t root;
  uint32_t safe;
  uint32_t call_bytes; // 8N + 4V + 8(R+1), the canonical CALL byte count
};

// The def record the evaluator loads for a CALL: the header and the first
// template pairs in one 128-byte record, so that one round trip brings them.
// Further pairs are read from the flat pair array.
#ifndef HVM_PRE_NODE
#define HVM_PRE_NODE 8
#endif
#define HVM_PRE_RBAG 4
struct HvmDevDefX {
  HvmDevDef h;
  unsigned long long node[HVM_PRE_NODE];  // node_buf[0, 8), zero past nlen
  unsigned long long rbag[HVM_PRE_RBAG];  // rbag_buf[0, 4), zero past rlen
};

// The evaluator's counters (HvmDevStats) are kept in HVM_STAT_STRIPES
// copies: wave w adds its counters to copy w % HVM_STAT_STRIPES at the end of
// its turn, and the readers sum the copies (the host for the totals; the
// between kernel every turn, for the scheduler, which takes the difference
// from the previous turn's sums).  With one copy, and a second set of
// per-turn counters, the 2048 waves of a turn ended with ~34 atomics each on
// the same ~34 addresses, which the memory side serialises (measured on
// church_loop(12, 12), 1,941 turns: 426 ms; 370 ms without the atomics,
// 395 ms with 64 copies).  256 copies cost the between kernel more than they
// save (421 ms).
#ifndef HVM_STAT_STRIPES
#define HVM_STAT_STRIPES 64
#endif

// Mid-turn pool of redexes in the output bag (turn kernel); 0: off.
#ifndef HVM_MID_POOL
#define HVM_MID_POOL 0
#endif

// Statistic slots accumulated by the evaluator (same order as the oracle).
enum {
  HVM_K_CALL = 0, HVM_K_CALL_COPY, HVM_K_VOID, HVM_K_ERAS, HVM_K_ANNI, HVM_K_COMM,
  HVM_K_OPER1, HVM_K_OPER2, HVM_K_SWIT0, HVM_K_SWITN, HVM_K_KINDS
};

struct HvmDevStats {
  unsigned long long itrs;
  unsigned long long bytes;
  unsigned long long kind[HVM_K_KINDS];
  unsigned long long steps;       // wave-steps executed
  unsigned long long spills;      // redexes spilled from LDS deques to the HBM bag
  unsigned long long defers;      // lane-steps deferred for allocation-table space
  unsigned long long retries;     // interactions re-queued because a node was not yet visible
  unsigned long long probes;      // 64-slot allocator windows scanned
  // wall-clock ticks (100 MHz) spent per step phase, summed over waves, when
  // the evaluator runs with HVM_TURN_PROFILE: 0 refill+share, 1 classify+
  // reserve, 2 takes+allocation, 3 creation, 4 links+counting, 5 release wait
  unsigned long long phase[6];
  unsigned long long leaks;       // LEAK interactions (local node copied to HBM), not in itrs
  unsigned long long saves;       // redexes moved out of a local net at the end of a turn
  unsigned long long local_left;  // local slots still in use after a wave's evacuation
  // for the scheduler's features
  unsigned long long lane_steps;   // lanes with work, summed over the wave-steps
  unsigned long long wave_left[4]; // waves that ended a turn leaving 0, 1-64, 65-128, >128 redexes
  unsigned long long alloc_slots;  // slots the allocator handed out
};

// Error flags (bitwise) reported by the evaluator.
enum {
  HVM_ERR_OOM_NODE = 1, HVM_ERR_OOM_VARS = 2, HVM_ERR_BAG_OVERFLOW = 4,
  HVM_ERR_LINK_LOOP = 8, HVM_ERR_BAD_DEF = 16, HVM_ERR_BAD_PORT = 32,
  HVM_ERR_EVAC = 64,    // a local net did not drain
};

// Per-wave persistent allocator state (cursor into the wave's slices).
struct HvmWaveState {
  uint32_t node_cur;      // scan cursor in the wave's own slices
  uint32_t vars_cur;
  uint32_t node_ovf;      // scan cursor in the shared overflow pools
  uint32_t vars_ovf;
};

// Turn control block, in device memory.  The turn kernel reads it at start;
// the between kernel advances it, so a chain of turns needs no host round trip.
struct HvmCtl {
  uint32_t n_in;        // redexes in the input bag of the current turn
  uint32_t n_out;       // output-bag counter
  uint32_t in_head;     // dynamic-pull counter into the input bag
  uint32_t err;         // HVM_ERR_* flags
  uint32_t parity;      // input bag index (0/1)
  uint32_t max_steps;   // steps of the current turn
  uint32_t turns;       // turns completed
  uint32_t done;        // set once a turn leaves no redex (or on error)
  uint32_t fixed_steps; // if nonzero, every turn runs this many steps
  uint32_t idle_waves;  // waves of the current turn that ran out of work
  uint32_t idle_quit;   // the current turn ends once this many waves are idle (0: never early)
  uint32_t surplus;     // some wave of the current turn held more redexes than two per lane
  uint32_t active;      // waves that run the current turn (0: all); the others end at once
  uint32_t idle_back;   // idle quit a turn started without gets back on surplus (0: never)
  uint32_t out_take;    // mid-turn pull cursor into the output bag (HVM_MID_POOL)
  uint32_t pad[1];
};

// ---------------------------------------------------------------------------
// Scheduler: picks the length of the next turn (how long waves run on their
// own redexes before the HBM bag redistributes work).  It never changes what
// an interaction does, only when redexes are shared.
// ---------------------------------------------------------------------------

#define HVM_ACTIONS     4      // (wave share, idle-quit share, steady bound) triples the scheduler chooses from
#define HVM_FEAT        32     // MLP input features
#define HVM_HID         64     // MLP hidden width
// packed fp16 parameters: W1[32][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]
#define HVM_MLP_PARAMS  (HVM_FEAT * HVM_HID + HVM_HID + HVM_HID * HVM_HID + HVM_HID + HVM_HID * HVM_ACTIONS + HVM_ACTIONS)
// trace record: features, action, reward (log2 interactions per µs), itrs, dt µs
#define HVM_TRACE_REC   (HVM_FEAT + 4)

enum { HVM_SCHED_PLAIN = 0, HVM_SCHED_LEARNED = 1, HVM_SCHED_EXPLORE = 2, HVM_SCHED_FIXED = 3 };

// Per-turn counters (the scheduler's view of a turn): the difference of the
// evaluator's summed counters between two decisions.
struct HvmTurnStats {
  unsigned long long itrs;
  unsigned long long wave_steps;   // steps in which a wave had work
  unsigned long long lane_steps;   // lanes with work, summed over those steps
  unsigned long long spills;
  unsigned long long defers;
  unsigned long long probes;
  unsigned long long kind[HVM_K_KINDS];
  // per-wave occupancy at the end of the turn: waves that left 0, 1-64,
  // 65-128 and more than 128 redexes (two per lane) to the output bag
  unsigned long long wave_left[4];
  unsigned long long alloc_slots;  // slots the allocator handed out (over probes * 64: free density)
};

struct HvmSched {
  uint32_t mode;                 // HVM_SCHED_*
  uint32_t seed;                 // exploration jitter
  uint32_t explore_pct;          // % of turns with a random action (explore mode), the margin
                                 // (learned mode) or the action (fixed mode)
  uint32_t trace_cap;            // records the trace buffer holds
  uint32_t trace_len;            // records written
  uint32_t last_action;
  unsigned long long last_clock; // wall clock (100 MHz) at the previous decision
  const unsigned short* params;  // HVM_MLP_PARAMS fp16 values (learned mode)
  float* trace;                  // trace_cap * HVM_TRACE_REC floats, or null
  float qavg[HVM_ACTIONS];       // learned mode: the network's outputs averaged over the turns
  uint32_t qn;                   // of this normalisation (exponentially), and their count
  HvmTurnStats prev;             // the evaluator's counters, summed, at the previous decision
};

// The scheduler's view of summed counters (host and device).
__host__ __device__ inline void hvm_turn_view(const HvmDevStats& x, HvmTurnStats& v) {
  v.itrs = x.itrs;
  v.wave_steps = x.steps;
  v.lane_steps = x.lane_steps;
  v.spills = x.spills;
  v.defers = x.defers;
  v.probes = x.probes;
  for (int k = 0; k < HVM_K_KINDS; ++k) v.kind[k] = x.kind[k];
  for (int k = 0; k < 4; ++k) v.wave_left[k] = x.wave_left[k];
  v.alloc_slots = x.alloc_slots;
}

struct HvmTurnArgs {
  // net
  unsigned long long* node;
  uint32_t* vars;
  uint32_t node_cap, vars_cap;     // slots (slot 0 never allocated)
  uint32_t node_slice, vars_slice; // slots per wave
  uint32_t node_pool_lo, vars_pool_lo; // shared overflow pools: [pool_lo, pool_hi)
  uint32_t node_pool_hi, vars_pool_hi; // slots from pool_hi to cap belong to the host (IO)
  // book
  const HvmDevDefX* defs;
  const unsigned long long* pairs;
  uint32_t ndefs;
  // bags: bag[ctl->parity] is the input, the other one the output
  unsigned long long* bag[2];
  uint32_t bag_cap;
  // control
  HvmCtl* ctl;
  uint32_t nwaves;
  HvmWaveState* wave_state;
  HvmDevStats* stats;
  uint32_t root_vars;              // var slots [0, root_vars) are roots, never allocated
  HvmSched* sched;
  uint32_t flags;                  // HVM_TURN_* behaviour switches
  uint32_t idle_quit;              // the plain scheduler's idle-quit wave count (0: never early)
  unsigned long long* lspill;      // local-net evaluator: HVM_LSPILL entries per wave
};

// Release before the links.  With this flag (the runtime sets it unless
// HVM_HIP_FENCE_RELEASE is given) the turn kernel waits for the completion of
// its creation stores (s_waitcnt vmcnt(0)); without it, it issues an
// agent-scope release fence (L2 writeback, then the wait).  The wait alone
// orders the stores before the link exchanges only because device-scope
// stores on gfx950 are written through to the point where the memory-side
// atomics execute: a property of the hardware, not of the HIP memory model.
// Both modes run the parity tests.
#define HVM_TURN_WAITCNT_FENCE 1u
// Accumulate per-phase wall-clock time into HvmDevStats::phase.
#define HVM_TURN_PROFILE 2u
// Run the local-net evaluator (node/var allocation in LDS).
#define HVM_TURN_LOCAL 4u
// Fast CALL paths of compiled CALLs (HBM evaluator only).
#define HVM_TURN_FAST 8u

struct HvmGatherArgs {
  const unsigned long long* node;
  const uint32_t* vars;
  uint32_t node_cap, vars_cap;
  uint32_t root_port;
  // outputs: visited node indices/values and var indices/values
  uint32_t* node_idx; unsigned long long* node_val; uint32_t* n_node;
  uint32_t* var_idx; uint32_t* var_val; uint32_t* n_var;
  uint32_t cap;                    // capacity of each output list
  uint32_t* queue;                 // work queue of ports (cap entries)
  uint32_t* err;
};
