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
 * changed are re-evaluated. Same outputs as oracle_run_literal for `det`. */
int oracle_run_event(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel,
                     uint8_t *det);

#ifdef __cplusplus
}
#endif
#endif
