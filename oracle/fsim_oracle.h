/*
 * CPU oracle for stuck-at fault simulation: a plain-C restatement of the
 * reference simulator fsim::Simulator (fsim/simulation.{hpp,cpp}).
 *
 * TEST INFRASTRUCTURE ONLY. Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this code, and only as the checker.
 * The product library never links or calls it.
 */
#ifndef FSIM_ORACLE_H
#define FSIM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* Parse the three texts (fsim/simulation.cpp:37-63). Returns NULL on an
 * input the reference cannot process; err receives the reason. */
oracle_sim *oracle_read(const char *ckt, size_t ckt_len, const char *flst, size_t flst_len,
                        const char *ptn, size_t ptn_len, char *err, size_t err_len);
void oracle_free(oracle_sim *s);

enum {
  ORACLE_NUM_PIS = 0,
  ORACLE_NUM_POS,
  ORACLE_NUM_GATES,
  ORACLE_NUM_WIRES,
  ORACLE_NUM_ROUNDS,
  ORACLE_NUM_FAULTS,
  ORACLE_NUM_NODES,
  ORACLE_NUM_LEVELS, /* size of the gates-per-level table */
  ORACLE_NUM_PATTERNS,
};
int64_t oracle_count(const oracle_sim *s, int which);

/* Reference accessors (fsim/simulation.hpp:64-89). */
void oracle_order_map(const oracle_sim *s, uint64_t *out);       /* N: file id -> position */
void oracle_gates_per_level(const oracle_sim *s, uint64_t *out); /* NUM_LEVELS */
void oracle_levels(const oracle_sim *s, uint64_t *out);          /* N: per position */
void oracle_faults(const oracle_sim *s, uint64_t *site, uint64_t *val); /* F each; val 0 or ~0 */
void oracle_patterns(const oracle_sim *s, uint64_t *out);        /* R x P */

/* The reference algorithm (fsim/simulation.cpp:104-228): for every round, a
 * good-machine pass, then for every fault a whole-circuit faulty pass and a
 * masked PO compare. Faults are split over `threads` OpenMP threads; every
 * fault's result is independent of the split.
 *   fault_idx/num_sel: optional subset of faults to simulate (NULL = all);
 *   table: 2*F words {found, round} with the last round overwriting
 *          (fsim/simulation.cpp:142-144), written for the selected faults;
 *   det:   optional F x R bytes, det[f*R+r] = d(f, r), selected faults only;
 *   g_pi/g_gate/g_po: optional masked good values of the last round.
 * Returns 0. */
int oracle_run_literal(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel,
                       uint64_t *table, uint8_t *det, uint64_t *g_pi, uint64_t *g_gate, uint64_t *g_po);

/* Event-driven single-fault restatement of the same detection function:
 * per (fault, round) only the nodes downstream of the fault site whose fanins
 * changed are re-evaluated. Same outputs as oracle_run_literal for `det`
 * (F x R bytes, rows of unselected faults untouched). Runs through
 * oracle_run_event_range over all rounds. Returns 0, -1 on allocation failure. */
int oracle_run_event(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel,
                     uint8_t *det);

/* The same over rounds [r0, r1), for large circuits: the good machine is
 * evaluated `chunk` rounds at a time (N x chunk words of memory, not N x R),
 * rounds of a chunk in parallel, then the faults of the chunk in parallel.
 * det holds num_sel x (r1 - r0) bytes, det[k * (r1 - r0) + r - r0] =
 * d(fault_idx[k], r) (fault_idx NULL: all F faults). With stop_at_first a
 * fault's rounds after its first detection are not simulated and stay 0, so
 * the first detecting round of each fault costs only the rounds up to it.
 * Returns 0, -1 on a bad range or allocation failure. */
int oracle_run_event_range(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel,
                           int64_t r0, int64_t r1, int64_t chunk, int stop_at_first, uint8_t *det);

/* Masked good values of round rd alone (fsim/simulation.cpp:149-184): a
 * round's good machine reads only that round's pattern words, so this equals
 * the last-round values of oracle_run_literal for rd = R - 1. Returns 0, -1
 * when rd is out of range. */
int oracle_good_round(const oracle_sim *s, int64_t rd, uint64_t *g_pi, uint64_t *g_gate, uint64_t *g_po);

#ifdef __cplusplus
}
#endif
#endif
