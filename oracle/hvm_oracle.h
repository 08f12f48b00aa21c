/* hvm_oracle.h — CPU oracle for the HVM2 GPU path.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference C runtime's
 * evaluation path (`hvm run-c`), used as the checker for the HIP evaluator by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  It is
 * never linked into, loaded by, or called from the product library.
 *
 * What it restates (reference file:line):
 *   redex bags            src/hvm.c:589-624
 *   node/var accessors    src/hvm.c:652-694
 *   allocators            src/hvm.c:713-780
 *   peek/enter/link       src/hvm.c:784-846
 *   interaction rules     src/hvm.c:852-1069
 *   interact + counting   src/hvm.c:1073-1121
 *   boot/expand           src/hvm.c:1204-1244
 *   IO driver readback    src/run.c:67-100 (readback_ctr), src/run.c:740-810 (do_run_io)
 *   printer               src/hvm.c:1535-1699
 *   entry/output          src/hvm.c:1743-1791
 * with TPC = 1 (one evaluator thread), so the schedule is deterministic.
 *
 * Pinning (what checks each part; see tests/test_oracle.py):
 *   - normal forms: the reference's own snapshots (tests/snapshots/ files) for
 *     every non-IO program, and closed forms of the synthetic workloads;
 *   - interaction counts: the compiled reference run-c (oracle/_ref) on the
 *     same book buffers, recorded in tests/golden/ fixtures;
 *   - per-rule histogram: its total equals ITRS (reference-pinned) and its
 *     per-rule split matches closed-form counts on hand-sized nets.
 */
#ifndef HVM_ORACLE_H
#define HVM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Histogram slots (interaction kinds; LINK is not an interaction). */
enum {
  HVM_ORC_CALL = 0,      /* CALL that expands a def                        */
  HVM_ORC_CALL_COPY = 1, /* CALL of a safe def against DUP: REF is copied   */
  HVM_ORC_VOID = 2,
  HVM_ORC_ERAS = 3,
  HVM_ORC_ANNI = 4,
  HVM_ORC_COMM = 5,
  HVM_ORC_OPER1 = 6,     /* OPER whose first operand is already a number   */
  HVM_ORC_OPER2 = 7,     /* OPER that re-creates the OPR node              */
  HVM_ORC_SWIT0 = 8,
  HVM_ORC_SWITN = 9,
  HVM_ORC_KINDS = 10,
};

typedef struct {
  uint64_t itrs;                    /* ITRS exactly as run-c counts it        */
  uint64_t kind[HVM_ORC_KINDS];     /* per-kind interaction counts            */
  uint64_t bytes;                   /* algorithmic bytes (canonical model)    */
  uint64_t normalizes;              /* normalize() calls incl. expansions     */
  uint64_t peak_nodes;              /* max live nodes                         */
  double seconds;                   /* wall clock, boot → result printed      */
} hvm_oracle_stats;

/* Runs a book buffer (Book::to_buffer format, src/hvm.rs:957-1003).
 * On success returns 0, stores a malloc'd NUL-terminated copy of the text the
 * reference prints after "Result: " in *result (free with hvm_oracle_free),
 * and fills *st.  Returns 1 if the book fails to load, 2 if the program
 * performs IO (out of scope), 3 on allocation failure. */
int hvm_oracle_run(const uint32_t* book_buf, size_t n_words, char** result, hvm_oracle_stats* st);

/* As hvm_oracle_run, also counting the CALLs that expand each def: calls[fid]
 * for fid < ncalls (a checker for def-level transformations of a book). */
int hvm_oracle_run_calls(const uint32_t* book_buf, size_t n_words, char** result, hvm_oracle_stats* st,
                         uint64_t* calls, uint32_t ncalls);

void hvm_oracle_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
