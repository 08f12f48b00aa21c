/* hvm_hip.h — C ABI of the MI355X HVM2 evaluator (libhvm_hip.so).
 *
 * The reference's host calls its GPU runtime through one entry,
 *     extern "C" void hvm_cu(u32* book_buffer);        src/hvm.cu:2320
 * declared on the Rust side at src/main.rs:17-20 and called at
 * src/main.rs:93-96 with the bytes of Book::to_buffer (src/hvm.rs:957-1003).
 * `hvm_hip` below is the drop-in for it.  The remaining entries expose the
 * same evaluation without stdout parsing (status codes, counters, result
 * text), a persistent context for repeated runs, and the front end that the
 * reference keeps in Rust (src/ast.rs), which this build provides in C++.
 *
 * All pointers are plain host pointers; no torch or HIP types cross the ABI.
 */
#ifndef HVM_HIP_H
#define HVM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Drop-in entry                                                            */
/* ------------------------------------------------------------------------ */

/* Replaces `hvm_cu` (src/hvm.cu:2320-2397) and `hvm_c` (src/hvm.c:1743-1791).
 * Loads the book (src/hvm.cu:1898-1947 semantics; prints "failed to load
 * book" to stderr and returns on failure), evaluates @main on the GPU under
 * the IO driver (src/run.c:740-809: READ/OPEN/CLOSE/FLUSH/WRITE/SEEK/
 * GET_TIME/SLEEP run on the host; DL_OPEN/DL_CALL/DL_CLOSE answer with an IO
 * error) and prints, exactly as run-cu does:
 *     Result: <tree>\n- ITRS: <n>\n- LEAK: <n>\n- TIME: <s>s\n- MIPS: <m>\n
 * TIME is wall clock from boot to the result being read back.  A NULL
 * buffer is rejected (the reference would run an uninitialised book). */
void hvm_hip(uint32_t* book_buffer);

/* ------------------------------------------------------------------------ */
/* Structured entry                                                         */
/* ------------------------------------------------------------------------ */

typedef struct hvm_hip_cfg {
  uint64_t node_cap;    /* node slots in the net (0: 2^28; at most 2^29, or
                           2^32 - 2 with HVM_HIP_WIDE)                        */
  uint64_t vars_cap;    /* var slots in the net (0: 2^28; limits as node_cap) */
  uint32_t bag_cap;     /* redex-bag entries per bag (0: node_cap + 4097, as
                           many as the net can hold redexes, so the bags
                           never overflow; 2^24 with HVM_HIP_WIDE)           */
  uint32_t waves;       /* evaluator waves (0: fill every CU)                 */
  uint32_t turn_steps;  /* steps per turn (0: adaptive; at most 2^20)         */
  uint32_t flags;       /* HVM_HIP_NO_READBACK: skip the result text          */
  uint32_t time_limit_s;/* abandon evaluation after this many seconds (0: none) */
  uint32_t idle_pct;    /* end a turn early once this % of waves are idle
                           (0: 25; above 100: never end early)                */
} hvm_hip_cfg;

#define HVM_HIP_NO_READBACK 1u
/* Node/var creations are ordered before the links that publish them by
 * waiting for the creation stores (device-scope, write-through) to complete.
 * This flag uses a full agent-scope release fence (with L2 writeback)
 * instead. */
#define HVM_HIP_FENCE_RELEASE 2u
/* Time the phases of every evaluator step (reported in phase_ms). */
#define HVM_HIP_PROFILE 4u
/* Local-net evaluator: every wave allocates nodes and vars in its own LDS net
 * and copies a node to HBM only when an HBM var must point to it (a LEAK,
 * counted in `leak`; src/hvm.cu:1129-1187); at the end of a turn each wave
 * moves what its local net still holds to HBM (src/hvm.cu:1426-1482).  Net
 * capacities are limited to 2^28 slots with it. */
#define HVM_HIP_LOCAL_NET 8u
/* Compiled CALLs (src/cmp.rs:10-351, the reference's compiled defs; its gen-cu
 * leaves them out, src/main.rs:140-145): at load every def's template is
 * pre-reduced — CALLs of REFs against the template's own nodes inlined, ANNIs
 * between template nodes performed — within HVM_HIP_SPEC_NODES live nodes,
 * and each expansion adds the def's static CALL and ANNI counts to ITRS and
 * kind[], so ITRS, kind[] and the normal form equal run-c's.  Fewer nodes are
 * created and taken per interaction counted. */
#define HVM_HIP_COMPILED_CALLS 16u
/* 64-bit ports (paper/HVM2.typst:1212-1222, the reference's planned
 * extension): ports of 64 bits, nodes of two 64-bit halves, slots indexed by
 * 32-bit integers, so node_cap and vars_cap may reach 2^32 - 2 (64 GiB of
 * nodes, 32 GiB of vars) instead of 2^29.  Numbers keep the reference's
 * 29-bit encoding: results, ITRS and kind[] equal run-c's.  A net of its own
 * with its own kernel (one redex per lane per step); no IO driver, no batched
 * nets, no local nets, plain scheduling only. */
#define HVM_HIP_WIDE 32u
/* Launch each chunk of turns (turn and between kernels with their timing
 * events) as a HIP graph captured once per chunk length and replayed, instead
 * of one launch call per kernel.  The 32-bit evaluators only. */
#define HVM_HIP_GRAPHS 64u
#define HVM_HIP_SPEC_NODES 64u

/* Interaction kinds, in the order of hvm_hip_out.kind[]. */
enum {
  HVM_HIP_K_CALL = 0, HVM_HIP_K_CALL_COPY = 1, HVM_HIP_K_VOID = 2, HVM_HIP_K_ERAS = 3,
  HVM_HIP_K_ANNI = 4, HVM_HIP_K_COMM = 5, HVM_HIP_K_OPER1 = 6, HVM_HIP_K_OPER2 = 7,
  HVM_HIP_K_SWIT0 = 8, HVM_HIP_K_SWITN = 9, HVM_HIP_KINDS = 10
};

typedef struct hvm_hip_out {
  uint64_t itrs;                 /* ITRS, counted as src/hvm.c:1116-1118      */
  uint64_t kind[HVM_HIP_KINDS];  /* per-kind interaction counts               */
  uint64_t bytes;                /* algorithmic bytes of all interactions     */
  uint64_t leak;                 /* LEAK line value: LEAK interactions (0
                                    without HVM_HIP_LOCAL_NET)                */
  uint64_t steps, spills, defers, retries, probes;  /* scheduler diagnostics  */
  double kernel_ms;              /* sum of evaluator kernel durations         */
  double eval_ms;                /* boot → normal form, incl. readback expansions */
  double wall_s;                 /* boot → result read back (reference window) */
  uint32_t turns;                /* evaluator turns that ran                  */
  uint32_t normalizes;           /* normalize passes (1 + readback expansions) */
  uint32_t err;                  /* evaluator error flags (0 = none)          */
  uint32_t launches;             /* turn-kernel dispatches, including those that
                                    found the net already normal and returned */
  double launch_ms;              /* summed duration of all `launches`         */
  uint32_t waves;                /* evaluator waves (workgroups) per turn     */
  uint64_t saves;                /* HVM_HIP_LOCAL_NET: redexes moved to HBM at turn ends */
  uint64_t local_left;           /* HVM_HIP_LOCAL_NET: local slots left in use after
                                    turn ends (unreachable garbage)           */
  double phase_ms[6];            /* with HVM_HIP_PROFILE: wave-time per step phase
                                    (refill+share, classify, takes+alloc,
                                    creation, links+count, release wait)     */
  char* result;                  /* "Result:" text as the reference prints it */
  char* canonical;               /* result with canonical var names (tests/run.rs:112-146) */
} hvm_hip_out;

/* Status codes */
#define HVM_HIP_OK          0
#define HVM_HIP_E_BOOK      1   /* book failed to load / exceeds limits       */
#define HVM_HIP_E_IO        2   /* an IO call's result did not fit the net's host
                                   region, or a batched net asked for IO      */
#define HVM_HIP_E_DEVICE    3   /* HIP error or no GPU                        */
#define HVM_HIP_E_EVAL      4   /* evaluator error (see out->err)             */
#define HVM_HIP_E_ARG       5   /* bad argument                               */

/* One-shot: create a context, load, evaluate, read back, destroy. */
int hvm_hip_run(const uint32_t* book_buf, size_t n_words, const hvm_hip_cfg* cfg, hvm_hip_out* out);

/* No counterpart in the reference.  hvm_hip_run for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the net, bags, control blocks, book and
 * readback lists live in a static device array of HVM_HIP_STATIC_BYTES that
 * the library holds on each GPU until the process ends, so the call makes no
 * hipMalloc/hipFree and creates no stream, event or graph (it runs on the
 * default stream and times turns with the host clock).  Limits: node_cap and
 * vars_cap at most HVM_HIP_STATIC_SLOTS (0 means that), bag_cap at most
 * HVM_HIP_STATIC_BAG, books up to 4 MiB, results up to 2^18 nodes/vars; no
 * HVM_HIP_WIDE, HVM_HIP_LOCAL_NET or HVM_HIP_GRAPHS.  Larger requests return
 * HVM_HIP_E_ARG with the reason in hvm_hip_last_error(NULL).  The printed
 * result goes to `result` (result_cap bytes, NUL-terminated; may be NULL with
 * HVM_HIP_NO_READBACK); out->result and out->canonical are left NULL.  Calls
 * on one GPU share its array: calls that overlap in time on one GPU race. */
#define HVM_HIP_STATIC_BYTES (64u << 20)
#define HVM_HIP_STATIC_SLOTS (1u << 21)
#define HVM_HIP_STATIC_BAG   (1u << 20)
int hvm_hip_run_static(const uint32_t* book_buf, size_t n_words, const hvm_hip_cfg* cfg, hvm_hip_out* out,
                       char* result, size_t result_cap);

/* Frees result/canonical strings and anything else this library returns. */
void hvm_hip_free(void* p);

/* ------------------------------------------------------------------------ */
/* Persistent context: the net, bags and book stay resident in HBM          */
/* ------------------------------------------------------------------------ */

typedef struct hvm_hip_ctx hvm_hip_ctx;

hvm_hip_ctx* hvm_hip_create(const hvm_hip_cfg* cfg);          /* NULL on failure */
int  hvm_hip_load(hvm_hip_ctx* ctx, const uint32_t* book_buf, size_t n_words);
/* Clears the net, boots @main, normalises (with readback expansions) and,
 * unless HVM_HIP_NO_READBACK, reads the result back. */
int  hvm_hip_eval(hvm_hip_ctx* ctx, hvm_hip_out* out);
void hvm_hip_destroy(hvm_hip_ctx* ctx);
/* Last error message of a context (or of context creation when ctx is NULL). */
const char* hvm_hip_last_error(const hvm_hip_ctx* ctx);

/* Batched evaluation of independent nets (no counterpart in the reference,
 * which evaluates one net per call; the per-net semantics are those of
 * hvm_hip_eval).  The loaded book holds every program; entries[i] is the fid
 * of net i's entry def.  Nets are admitted in order while the sum of their
 * estimated peak node counts (est_nodes, or NULL for "admit all") fits the
 * net; each admitted group is booted into one heap, each net with its own
 * root var, and normalised by one chain of turns.  outs[i] receives net i's
 * result text (itrs of a single net is not separable and is left 0); *total
 * receives the summed counters.  At most HVM_HIP_MAX_NETS nets per group. */
#define HVM_HIP_MAX_NETS 4096u
int hvm_hip_eval_batch(hvm_hip_ctx* ctx, const uint32_t* entries, uint32_t n_nets,
                       const uint64_t* est_nodes, hvm_hip_out* outs, hvm_hip_out* total);

/* ------------------------------------------------------------------------ */
/* Scheduler (no counterpart in the reference: its CUDA runtime switches    */
/* between fixed SEED/GROW/WORK modes, src/hvm.cu:1509-1548)                */
/* ------------------------------------------------------------------------ */

/* After every turn the scheduler picks how the next turn shares work: one of
 * 4 actions (wave share, idle-quit share, steady bound) = (all, idle_pct, 1),
 * (all, 50%, 1), (1/2, idle_pct, 1), (all, idle_pct, 16).  A turn runs on
 * that share of the waves (the others end at once; fewer, fuller waves for a
 * small frontier) and ends once its idle-quit share of them has run out of
 * work, when the HBM bag redistributes the redexes; turns are at most 4096
 * steps.  idle_pct is hvm_hip_cfg.idle_pct (25% by default).  A turn that
 * follows one that did not grow the redexes by more than half, and left at
 * most (steady bound)/4 of them per wave, runs without idle quit until some
 * wave holds more than two redexes per lane (steady chains); one that
 * follows a turn that did not grow them and left fewer than there are waves
 * counts its idle-quit share among the waves that got a redex (the empty
 * ones count as idle at once).  A turn on part of the waves
 * allocates from those waves' slices and the shared pool only.  The choice
 * never changes what an interaction does, so normal form and ITRS do not
 * depend on it.
 *   HVM_HIP_SCHED_PLAIN   action 0, every turn
 *   HVM_HIP_SCHED_LEARNED argmax of a 32->64->64->4 MLP (fp16 params,
 *                         HVM_HIP_MLP_PARAMS values) evaluated on the device
 *                         from the finished turn's features (bag sizes, lane
 *                         use, interaction mix, how the turn ended, per-wave
 *                         occupancy at its end, free-slot density, rate);
 *                         it moves off the action in force (action 0
 *                         when a normalisation starts) only to an action
 *                         whose predicted value is higher by more than
 *                         explore_pct/100 (the plain choice stays the
 *                         fallback; near-ties keep the action in force)
 *   HVM_HIP_SCHED_EXPLORE action 0 with explore_pct% random choices
 *                         (seeded), for collecting training traces
 *   HVM_HIP_SCHED_FIXED   action explore_pct % 4, every turn
 * A nonzero seed rotates the initial placement of redexes over the waves and
 * shortens each turn by up to an eighth, in every mode.
 * trace_cap > 0 records one HVM_HIP_TRACE_REC-float record per decision:
 * 32 features, action, log2 interactions per µs of the turn the action
 * governed, its interactions, its µs. */
#define HVM_HIP_SCHED_PLAIN   0u
#define HVM_HIP_SCHED_LEARNED 1u
#define HVM_HIP_SCHED_EXPLORE 2u
#define HVM_HIP_SCHED_FIXED   3u
#define HVM_HIP_MLP_FEAT   32
#define HVM_HIP_MLP_PARAMS (32 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4)
#define HVM_HIP_TRACE_REC  36
int hvm_hip_set_scheduler(hvm_hip_ctx* ctx, uint32_t mode, uint32_t seed, uint32_t explore_pct,
                          const uint16_t* params, uint32_t trace_cap);
int hvm_hip_read_trace(hvm_hip_ctx* ctx, float* recs, uint32_t cap, uint32_t* n_recs);

/* The scheduler MLP's training kernels (HIP, fp16 weights, fp32 math).
 * Parameters are packed W1[32][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4].
 * Loss per sample: 0.5 * (y[action] - reward)^2 (only the taken action).
 *   hvm_mlp_forward: y[n][4] for inputs x[n][32]
 *   hvm_mlp_grad:    mean gradient (fp32, packed like params) and mean loss
 *   hvm_mlp_train:   `epochs` full-batch Adam steps from params_in (fp32
 *                    master copy), writes fp16 params_out and the final loss */
int hvm_mlp_forward(const uint16_t* params, const float* x, uint32_t n, float* y);
int hvm_mlp_grad(const uint16_t* params, const float* x, const uint32_t* actions, const float* rewards,
                 uint32_t n, float* grad, float* loss);
int hvm_mlp_train(const float* params_in, const float* x, const uint32_t* actions, const float* rewards,
                  uint32_t n, uint32_t epochs, float lr, uint16_t* params_out, float* params_fp32_out,
                  float* loss);

/* ------------------------------------------------------------------------ */
/* Measurement (no counterpart in the reference)                            */
/* ------------------------------------------------------------------------ */

/* The atomic roofline: device-scope exchanges per second on pseudo-random
 * slots of a fresh buffer of `slots` (a power of two, 2^10..2^31) elements of
 * `bytes` (4: var-sized, 8: node-sized), every lane keeping 8 exchanges in
 * flight.  The evaluator's takes and links are such exchanges. */
int hvm_hip_atomic_rate(uint64_t slots, uint32_t bytes, double* ops_per_s);

/* The request rooflines: the same measurement for op 0 (exchanges, as
 * hvm_hip_atomic_rate), 1 (device-scope stores) or 2 (device-scope loads) —
 * the three kinds of memory-side request the evaluator issues (links and
 * pool claims; creations and frees; takes and the allocator's windows). */
int hvm_hip_request_rate(uint64_t slots, uint32_t bytes, uint32_t op, double* ops_per_s);

/* The attainable HBM bandwidth: bytes read plus bytes written per second by
 * the best of 5 device-to-device copies of `bytes` (1 MiB..8 GiB). */
int hvm_hip_copy_rate(uint64_t bytes, double* bytes_per_s);

/* Self-test of the evaluator's cross-lane primitives (wave prefix sum, max,
 * 64-bit sum) on n_waves * 64 inputs; per lane 7 words: exclusive prefix
 * sum, wave sum, wave max, the same prefix sum and wave sum over LDS
 * permutes, and the 64-bit wave sum (low, high). */
int hvm_hip_wave_selftest(uint32_t n_waves, const uint32_t* in32, const uint64_t* in64, uint32_t* out);

/* ------------------------------------------------------------------------ */
/* Front end (the reference's Rust host, restated in C++)                   */
/* ------------------------------------------------------------------------ */

/* Parses `.hvm` text and builds the wire buffer, as Book::parse + build +
 * to_buffer do (src/ast.rs:528-559, src/hvm.rs:957-1003).  On success returns
 * 0 and a malloc'd word array; on failure returns nonzero and a malloc'd
 * message in *err (the reference panics with the same text). */
int hvm_front_compile(const char* code, uint32_t** words, size_t* n_words, char** err);

/* Canonical form of a printed tree: parse, rename vars x0, x1, … in
 * first-occurrence order, re-print (tests/run.rs:112-146). */
int hvm_front_canonical(const char* printed, char** out, char** err);

/* The specialisation HVM_HIP_COMPILED_CALLS applies at load, on a wire buffer
 * (no counterpart in the reference, whose compiled defs are C text,
 * src/cmp.rs:10-34): returns 0, the specialised book as a malloc'd wire
 * buffer, and a malloc'd array of 3 words per def: the static CALLs and ANNIs
 * an expansion of the def adds to the counters, and the canonical bytes of one
 * expansion (the unspecialised def's own plus the static interactions'),
 * which replace the specialised template's own 8N+4V+8(R+1); all zero for a
 * def left as it was.
 * `node_budget` ≤ 256 bounds a def's live template nodes. */
int hvm_front_specialize(const uint32_t* words, size_t n_words, uint32_t node_budget,
                         uint32_t** out_words, size_t* out_n, uint32_t** counts, char** err);

#ifdef __cplusplus
}
#endif
#endif
