/*
 * lsim: MI355X (gfx950) stuck-at fault simulator - C ABI.
 *
 * Drop-in for the reference simulation path fsim::Simulator
 * (fsim/simulation.hpp:39-188): the same .ckt/.flst/.ptn inputs, the same
 * read / run / accessor sequence and the same detected-fault report
 * (found_fault_to_pattern, fsim/simulation.hpp:87-89). Every entry point
 * returns 0 on success or a negative LSIM_E_* code; lsim_last_error() gives
 * the message. No C++ exception crosses this boundary. Output buffers are
 * caller-allocated and their element counts are checked.
 *
 * One handle is one simulator; calls on one handle must not overlap.
 */
#ifndef LSIM_H
#define LSIM_H
#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define LSIM_API __attribute__((visibility("default")))
#else
#define LSIM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define LSIM_OK 0
#define LSIM_E_ARG (-1)    /* bad argument or buffer size */
#define LSIM_E_INPUT (-2)  /* malformed or unsupported circuit/fault/pattern text */
#define LSIM_E_IO (-3)     /* a file could not be opened */
#define LSIM_E_STATE (-4)  /* call out of order (e.g. run before read) */
#define LSIM_E_DEVICE (-5) /* HIP runtime error or no usable GPU */
#define LSIM_E_NOMEM (-6)  /* device memory exhausted */

typedef struct lsim_sim lsim_sim;

/* Modes of lsim_run. All run on the GPU and give identical results.
 * The reference's modes (fsim::Simulator::Mode, fsim/simulation.hpp:45)
 * SEQUENTIAL(0), PARALLEL_TF(1) and PARALLEL_OMP(2) all run the persistent
 * engine. Two explicit engine choices have values of their own:
 *  PLAIN:        per-level kernel launches (good machine and fault sweeps).
 *  PARTITIONED:  one persistent launch per sweep; each workgroup owns a tile
 *                and walks the levels with workgroup barriers. */
#define LSIM_MODE_REF_SEQUENTIAL 0
#define LSIM_MODE_REF_PARALLEL_TF 1
#define LSIM_MODE_REF_PARALLEL_OMP 2
#define LSIM_MODE_PLAIN 16
#define LSIM_MODE_PARTITIONED 17

/* Create a simulator bound to HIP device `device`.
 * Replaces constructing fsim::Simulator (fsim/simulation.hpp:39). */
LSIM_API int lsim_create(int device, lsim_sim **out);
/* Replaces fsim::Simulator's destructor. */
LSIM_API void lsim_destroy(lsim_sim *s);
/* Message for the last failed call on `s` (or of lsim_create when s is NULL). */
LSIM_API const char *lsim_last_error(const lsim_sim *s);

/* Replaces Simulator::read(const std::string&, const std::string&, const
 * std::string&) (fsim/simulation.cpp:12-35): LSIM_E_IO when a file cannot be
 * opened, with the reference's message text. */
LSIM_API int lsim_read_files(lsim_sim *s, const char *ckt_path, const char *flst_path, const char *ptn_path);
/* Replaces Simulator::read(std::istream&, std::istream&, std::istream&)
 * (fsim/simulation.cpp:37-63) for in-memory text. */
LSIM_API int lsim_read_buffers(lsim_sim *s, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len,
                      const char *ptn, size_t ptn_len);

/* Options (no reference counterpart; the reference has compile-time flags).
 *  "drop"               0/1  fault dropping: a detected fault is not simulated
 *                            in later rounds (default 0, as the reference).
 *  "report_last_round"  0/1  with dropping on, also evaluate every fault in the
 *                            last round so found_fault_to_pattern stays exact
 *                            (default 1).
 *  "first_words"        with dropping: words (64 patterns each) of the first
 *                            step, 1..64 (default 1).
 *  "word_growth"        with dropping: factor of words per step after each
 *                            step, 1..64 (default 2).
 *  "max_step_words"     with dropping: the growth stops at this many words per
 *                            step, 1..4096 (default 128).
 *  "scratch_mb"         MiB of device memory the fault sweeps may use for
 *                            per-tile state (default 0 = automatic).
 *  "direct"             0/1  inject every fault directly instead of stem
 *                            tracing (default 0; forced on for circuits with MUX).
 *  "sweep_kernel"       persistent sweep: 0 flattened (node, lane) pairs (default),
 *                            1 one node per wave.
 *  "pair_threads"       pair sweep workgroup: 0 by step (default), 256 or 512.
 *  "pair_dense"         pair sweep: a level pass whose staged nodes average at
 *                            least this many active lanes evaluates one node per
 *                            wave (default 24; 0 = always flattened pairs).
 *  "pair_nopend"        pair sweep without pending flags (every node of a walked
 *                            level is a candidate): 0 never, 1 dense steps (the
 *                            first step of a dropping job, full-matrix steps;
 *                            default), 2 every step.
 *  "pair_nopend_pct"    0..100: with pair_nopend 1, a step after one that staged
 *                            more than this percentage of the nodes its tiles
 *                            walked also runs without pending flags (default 65).
 *  "pair_fo_pre"        0/1  pair sweep with pending flags: a staged node's first
 *                            live fanout is loaded beside its fanin records, so a
 *                            node with one live fanout marks it pending without a
 *                            further dependent load (default 1).
 *  "pool_cap"           pair sweep: at most this many pool words per slot
 *                            (default 0 = as memory allows); a tile that needs
 *                            more is swept again in a slot that holds any tile.
 *  "block_waves"        waves per workgroup of the one-node-per-wave sweep (default 8).
 *  "autotune"           0/1  successive runs of a job set each measure one sweep
 *                            configuration, chosen by coordinate descent: the word
 *                            schedule (first step 1 or 2 words, growth x2 or x4),
 *                            then the pair sweep's workgroup size (by step, 256,
 *                            512), its dense-pass threshold (24, 16, 32, never),
 *                            the step cap (128, 64, 256 words) and the sweep kernel,
 *                            each with the others at the best so far; later runs
 *                            use the fastest.
 *  "good_graph"         0/1  a lone job replays its circuit's captured
 *                            good-machine graph (default 1).
 *  "good_form"          good machine of a lone job: 0 one launch per level,
 *                            1 LDS-staged level windows (the partitioned form,
 *                            fsim/cuda/simulation.cu:487-733; default).
 *  "win_levels"         level windows: most levels per window, 1..8 (default 4);
 *                            a window whose deepest cone outgrows a block is
 *                            cut shorter.
 *  "win_cap"            level windows: nodes per block, 0..4096 (default 0 = half
 *                            the CU's LDS at 544 B per node: 150 on gfx950).
 *  "device_steps"       0/1  pair sweep: each word step is planned by kernels
 *                            (fault filtering, units, lanes, tiles) with no host
 *                            round trip between steps (default 1).
 *  "step_graphs"        0/1  device-planned steps: a step whose launch signature
 *                            recurs replays its captured hipGraphs (default 1).
 *  "slot_pct"           1..100: the pair sweep uses at most this percentage of
 *                            the device's resident workgroup slots (default 100). */
LSIM_API int lsim_set_option(lsim_sim *s, const char *key, int64_t value);

/* Replaces Simulator::run(Mode, size_t) (fsim/simulation.cpp:65-92).
 * num_threads is accepted for signature parity and ignored. */
LSIM_API int lsim_run(lsim_sim *s, int mode, size_t num_threads);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). lsim_run with the same modes and results (read them with the
 * same accessors), but every device buffer the run needs - the circuit's
 * device copy, the job's good values, fault state and results, the sweeps'
 * tile state - is carved from a scratch array of LSIM_NOALLOC_SCRATCH_BYTES
 * (64 MiB) that the library holds on each GPU it runs on until the process
 * ends; host staging is pageable. The call makes no hipMalloc, hipFree,
 * page-locked or other allocating device call; its run stats report
 * device_allocs 0. A run whose footprint exceeds the scratch fails with
 * LSIM_E_NOMEM (circuit nodes x rounds x 8 bytes of good values, plus the
 * fault state and the tile state, must fit). Every call on a GPU shares that
 * GPU's scratch and leaves its contents behind: calls that overlap on one GPU
 * (from any handles or threads) race, so callers serialise them. The handle
 * must not run lsim_run at the same time either (the two share its stream). */
#define LSIM_NOALLOC_SCRATCH_BYTES (64ll << 20)
LSIM_API int lsim_run_noalloc(lsim_sim *s, int mode, size_t num_threads);

/* Counts: Simulator::num_PIs/num_POs/num_gates/num_wires/num_rounds
 * (fsim/simulation.hpp:75-80), faults().size() (:70) and
 * num_gates_per_level().size() (:84). */
LSIM_API int64_t lsim_num_pis(const lsim_sim *s);
LSIM_API int64_t lsim_num_pos(const lsim_sim *s);
LSIM_API int64_t lsim_num_gates(const lsim_sim *s);
LSIM_API int64_t lsim_num_wires(const lsim_sim *s);
LSIM_API int64_t lsim_num_rounds(const lsim_sim *s);
LSIM_API int64_t lsim_num_faults(const lsim_sim *s);
LSIM_API int64_t lsim_num_levels(const lsim_sim *s);
LSIM_API int64_t lsim_num_patterns(const lsim_sim *s);

/* Simulator::gate_idx_to_tpg_order_idx() (fsim/simulation.hpp:81-83); n = P+O+G. */
LSIM_API int lsim_gate_idx_to_tpg_order_idx(const lsim_sim *s, uint64_t *out, size_t n);
/* Simulator::num_gates_per_level() (fsim/simulation.hpp:84-86); n = lsim_num_levels. */
LSIM_API int lsim_num_gates_per_level(const lsim_sim *s, uint64_t *out, size_t n);
/* Levels of pi(i), gate(i), po(i) in topological order (ElementBase::level(),
 * fsim/base_ppg.hpp:49); n = P+G+O. */
LSIM_API int lsim_levels(const lsim_sim *s, uint64_t *out, size_t n);
/* Simulator::visualization_graph(path) (fsim/simulation.hpp:93,
 * fsim/simulation.cpp:1157-1225): writes the circuit as a Graphviz digraph,
 * one edge per fanin of every gate and PO in topological order. LSIM_E_IO
 * when the file cannot be opened (the reference prints the message and
 * returns). lsim_visualization_text returns the same text in a buffer the
 * caller frees with lsim_free_text. */
LSIM_API int lsim_visualization_graph(const lsim_sim *s, const char *path);
LSIM_API int lsim_visualization_text(const lsim_sim *s, char **text, size_t *len);
/* Structural diagnostics per level, over groups of `wave` consecutive nodes of
 * the level in topological order (the idea of the reference's
 * fsim/cuda_MA_partition/cuda_MA_partition.cu:227-346, at wave64 by default):
 * nodes[l]; avg_dist[l] = mean over the level's nodes with fanins of the mean
 * level distance to their fanins; max_span[l] = largest such distance;
 * avg_range[l] = mean over groups of (largest - smallest fanin position);
 * avg_unique[l] = mean over groups of the number of distinct fanin positions.
 * Arrays hold n_levels = max level + 1 entries. Needs no device. */
LSIM_API int lsim_level_diagnostics(const lsim_sim *s, int wave, int64_t *nodes, double *avg_dist, int64_t *max_span,
                                    double *avg_range, double *avg_unique, size_t n_levels);
/* Simulator::faults() (fsim/simulation.hpp:70): site position and stuck value
 * as the reference stores it (0 or all ones); n = number of faults. */
LSIM_API int lsim_faults(const lsim_sim *s, uint64_t *site, uint64_t *value, size_t n);
/* Simulator::patterns() (fsim/simulation.hpp:71): rounds x PIs words. */
LSIM_API int lsim_patterns(const lsim_sim *s, uint64_t *out, size_t n);

/* Simulator::g_pi_results/g_gate_results/g_po_results (fsim/simulation.hpp:58-60):
 * good-machine values of the last round, masked to its live bits. */
LSIM_API int lsim_good_results(const lsim_sim *s, uint64_t *pi, size_t n_pi, uint64_t *gate, size_t n_gate, uint64_t *po,
                      size_t n_po);
/* Simulator::found_fault_to_pattern() (fsim/simulation.hpp:87-89): 2*F words,
 * {found in the last round, last round index}. LSIM_E_STATE when dropping was
 * on and report_last_round off. */
LSIM_API int lsim_found_fault_to_pattern(const lsim_sim *s, uint64_t *out, size_t n);
/* No reference counterpart: detected in any round (1/0) per fault. */
LSIM_API int lsim_detected(const lsim_sim *s, uint8_t *out, size_t n);
/* No reference counterpart: first detecting round per fault, -1 if never. */
LSIM_API int lsim_first_detect_round(const lsim_sim *s, int64_t *out, size_t n);
/* No reference counterpart: d(f, r) for every fault and round, n = F*R bytes,
 * row-major by fault. LSIM_E_STATE when the run used fault dropping. */
LSIM_API int lsim_detect_matrix(const lsim_sim *s, uint8_t *out, size_t n);

/* Timing and traffic of the last run (no reference counterpart). */
typedef struct {
  double t_total_ms;       /* run() wall time after the circuit is resident */
  double t_good_ms;        /* good-machine kernels (device time) */
  double t_sweep_ms;       /* fault sweep kernels (device time) */
  double t_detect_ms;      /* stem-tracing / detection kernels (device time) */
  int64_t sweep_launches;  /* number of sweep kernel launches */
  int64_t sweep_tiles;     /* tiles processed by the sweeps */
  int64_t sweep_node_steps;/* (tile, node) pairs visited by the sweeps */
  int64_t sweep_bytes;     /* bytes of node operands touched by the sweeps */
  int64_t good_bytes;      /* algorithmic bytes of the good-machine kernels */
  int64_t live_faults;     /* faults whose site can reach a PO */
  int64_t stems;           /* stem sweeps injected (summed over steps) */
  int64_t steps;           /* word steps executed */
  int64_t slots;           /* concurrent sweep tiles (scratch slots) */
  int64_t live_nodes;
  int64_t sweep_reruns;    /* tiles swept again after outgrowing their slot's pool */
  int64_t tune_trials;     /* autotuner candidates measured so far for this job set (0: off) */
  int64_t tuned;           /* configuration used: sweep kernel | first step words << 8 |
                              word growth << 16 | pair workgroup / 256 << 24 (0: by
                              step) | dense-pass threshold << 32 | step cap words << 40 */
  int64_t engine;          /* sweep kernel that ran: 0 k_sweep_level (per-level form),
                              1 k_sweep_pairs, 2 k_sweep_events (persistent forms) */
  int64_t tune_candidates; /* configurations the autotuner measures for this job set (0: off) */
  int64_t good_replayed;   /* 1 when the good machine replayed a captured hipGraph */
  int64_t device_allocs;   /* device and page-locked host allocations the library made
                              during the call (lsim_run, lsim_run_noalloc; counted over
                              all threads of the process) */
  int64_t steps_replayed;  /* word steps whose launches replayed captured hipGraphs */
} lsim_run_stats;
LSIM_API int lsim_last_run_stats(const lsim_sim *s, lsim_run_stats *out);

/* ------------------------------------------------------------------------
 * Batching (no reference counterpart): many circuits of uneven size, each
 * with its own fault list and pattern block, simulated together; every
 * iteration sweeps the tiles of all unfinished jobs in one persistent launch.
 * Calls on one batch must not overlap. */
typedef struct lsim_batch lsim_batch;
LSIM_API int lsim_batch_create(int device, lsim_batch **out);
LSIM_API void lsim_batch_destroy(lsim_batch *b);
LSIM_API const char *lsim_batch_last_error(const lsim_batch *b);
/* Options as lsim_set_option (engine options: "scratch_mb", "block_waves", "sweep_kernel", "pair_threads",
 * "pair_dense", "pair_nopend", "pair_nopend_pct", "pair_fo_pre", "max_step_words", "pool_cap", "autotune", "good_graph",
 * "good_form", "win_levels", "win_cap", "device_steps", "step_graphs", "slot_pct") and "mode" (0 plain,
 * 1 persistent). */
LSIM_API int lsim_batch_set_option(lsim_batch *b, const char *key, int64_t value);
/* Makes a circuit (.ckt text) resident; returns its id. */
LSIM_API int lsim_batch_add_circuit(lsim_batch *b, const char *ckt, size_t ckt_len, int64_t *circuit_id);
/* Adds a job (.flst and .ptn text against a resident circuit); returns its id. */
LSIM_API int lsim_batch_add_job(lsim_batch *b, int64_t circuit_id, const char *flst, size_t flst_len,
                                const char *ptn, size_t ptn_len, int drop, int64_t *job_id);
/* Runs every job added since the last run, together, to completion. */
LSIM_API int lsim_batch_run(lsim_batch *b);
LSIM_API int64_t lsim_batch_job_faults(const lsim_batch *b, int64_t job_id);
LSIM_API int64_t lsim_batch_job_rounds(const lsim_batch *b, int64_t job_id);
LSIM_API int lsim_batch_first_detect_round(const lsim_batch *b, int64_t job_id, int64_t *out, size_t n);
LSIM_API int lsim_batch_detect_matrix(const lsim_batch *b, int64_t job_id, uint8_t *out, size_t n);
LSIM_API int lsim_batch_found_fault_to_pattern(const lsim_batch *b, int64_t job_id, uint64_t *out, size_t n);
LSIM_API int lsim_batch_last_run_stats(const lsim_batch *b, lsim_run_stats *out);
LSIM_API int lsim_batch_release_job(lsim_batch *b, int64_t job_id);

/* ------------------------------------------------------------------------
 * Serving (no reference counterpart): circuits stay resident; jobs are
 * submitted from any thread and admitted in arrival order while their HBM
 * footprint fits the budget. Two lanes, each an engine with its own stream
 * and worker thread, run concurrently on the device: small jobs are batched
 * continuously (every admitted job advances one word step per iteration, so
 * a job that arrives while others run joins at the next one); large jobs run
 * one at a time in arrival order. While both lanes have work the small-job
 * lane's sweeps take a quarter of the workgroup slots and the other lane the
 * rest. */
typedef struct lsim_server lsim_server;
/* job_budget_mb: HBM for the resident circuits (one device copy serves both
 * lanes) and the admitted jobs (good values, fault state, results); a job that
 * cannot fit beside the circuits fails at submit with LSIM_E_NOMEM.
 * scratch_mb: HBM for sweep tiles (0 = 90 % of free HBM less the job budget and 4 GiB).
 * When an iteration fails, each of its jobs is run again alone from its
 * start, so only a job that fails by itself reports the error. */
LSIM_API int lsim_server_create(int device, int64_t job_budget_mb, int64_t scratch_mb, lsim_server **out);
LSIM_API void lsim_server_destroy(lsim_server *s);
LSIM_API const char *lsim_server_last_error(const lsim_server *s);
LSIM_API int lsim_server_add_circuit(lsim_server *s, const char *ckt, size_t ckt_len, int64_t *circuit_id);
LSIM_API int lsim_server_submit(lsim_server *s, int64_t circuit_id, const char *flst, size_t flst_len,
                                const char *ptn, size_t ptn_len, int drop, int64_t *job_id);
/* Blocks until the job finished (0), failed (its error code) or timeout_s
 * passed (LSIM_E_STATE); timeout_s < 0 waits forever. */
LSIM_API int lsim_server_wait(lsim_server *s, int64_t job_id, double timeout_s);
LSIM_API int lsim_server_first_detect_round(lsim_server *s, int64_t job_id, int64_t *out, size_t n);
LSIM_API int64_t lsim_server_job_faults(lsim_server *s, int64_t job_id);
/* times[0..2] = submit, admit, finish in seconds since the server started. */
LSIM_API int lsim_server_job_times(lsim_server *s, int64_t job_id, double *times);
LSIM_API int lsim_server_release_job(lsim_server *s, int64_t job_id);
/* counters[0..3] = iterations, jobs finished, jobs failed, peak admitted jobs. */
LSIM_API int lsim_server_counters(lsim_server *s, int64_t *counters);
/* Engine options as lsim_set_option; the word schedule of jobs submitted
 * later: "first_words" (default 1) and "word_growth" (default 4); and
 * "big_evals" and "big_node_words": jobs with at least that many fault x
 * pattern evaluations (default 4e9), or circuit nodes x 64-pattern rounds
 * (default 5e7), run in the large-job lane, one at a time in arrival order
 * ("big_evals" 0 puts every job in the batching lane); "small_share_pct"
 * (1..99, default 25): the batching lane's share of the sweep's workgroup
 * slots while both lanes have work (the large-job lane takes the rest). */
LSIM_API int lsim_server_set_option(lsim_server *s, const char *key, int64_t value);

/* Seeded synthetic workload in the reference formats (no reference
 * counterpart; the reference ships fsim/circuit_generator.py and
 * fsim/generate_fault_patterns.py as separate producers). Buffers are
 * allocated by the library and released with lsim_free_text. */
typedef struct {
  uint64_t seed;
  int64_t num_pis, num_pos, num_gates, num_levels;
  int32_t fanout_cap;    /* 0 = uniform source weights, no cap */
  double power_exp;      /* source-weight power law exponent (2.2) */
  double locality;       /* P(second fanin from the last `window` levels) (0.7) */
  int32_t window;        /* 3 */
  double unary_frac;     /* INV/CLKBUF share (0.1) */
  int64_t num_patterns;
  double fault_fraction; /* (0,1]: seeded subset of the full SA0/SA1 list */
} lsim_gen_spec;
LSIM_API int lsim_generate(const lsim_gen_spec *spec, char **ckt, size_t *ckt_len, char **flst, size_t *flst_len, char **ptn,
                  size_t *ptn_len);
LSIM_API void lsim_free_text(char *p);

/* Loop removal by traversal from the PIs, the rule the reference's netlist
 * producer applies before it writes a .ckt (fsim/parser.py:163-211, its
 * 'bfs' and 'dfs' modes): nodes are taken from a queue seeded with the PIs,
 * from the front (BFS) or the back (DFS); a node taken before is skipped;
 * otherwise its wires into nodes taken before are dropped and its other
 * destinations are queued. A node's wires are its first wire record. Returns
 * the new .ckt text (free with lsim_free_text) and the number of dropped
 * (destination, pin) entries. */
#define LSIM_ORDER_BFS 0
#define LSIM_ORDER_DFS 1
LSIM_API int lsim_break_loops(const char *ckt, size_t ckt_len, int mode, char **out, size_t *out_len,
                              int64_t *removed);
/* OpenTimer timing-graph dump -> .ckt (fsim/parser.py:1-227): sections by their
 * marker lines, cell types to gate codes by name prefix (AOI/OAI/DFF become
 * NAND/NOR), pins through the converter's per-cell table, node ids PIs, POs,
 * cells in file order, loops removed by a BFS or DFS traversal from the PIs
 * (mode LSIM_ORDER_BFS / LSIM_ORDER_DFS). A source's destinations keep the
 * dump's order (the converter's Python set has no fixed order). Returns the
 * .ckt text (free with lsim_free_text); `stats` may be null. */
typedef struct lsim_ingest_stats {
  int64_t pis, pos, cells, wires;
  int64_t connections;   /* "a -> b" lines read */
  int64_t records;       /* wire records written */
  int64_t removed;       /* destinations dropped by the loop removal */
  int64_t unknown_cells; /* cell types mapped to AND */
  int64_t unknown_pins;  /* pins mapped to pin 0 */
  int64_t unreached;     /* nodes reached that drive nothing (POs excluded) */
} lsim_ingest_stats;
LSIM_API int lsim_convert_timing_dump(const char *dump, size_t dump_len, int mode, char **ckt, size_t *ckt_len,
                                      lsim_ingest_stats *stats);
/* Producers with the conventions of fsim/generate_fault_patterns.py (its
 * random draws replaced by the library's seeded stream):
 *  - lsim_produce_faults: F distinct (wire id, value) pairs, ids uniform over
 *    [0, N) less the PO ids [P, P+O), values 0/1 (:7-30); LSIM_E_ARG when
 *    F > 2 (N - O).
 *  - lsim_pack_patterns: explicit patterns (bits[k*P + i] = PI i of pattern k)
 *    in the "new" writer's layout (:56-94): rounds of bits_per_round patterns,
 *    one decimal word per PI with the round's first pattern most significant,
 *    then the same chunks in binary. bits_per_round 64 is what the simulator
 *    reads; 32 is the layout of the reference's 32-bit GPU path.
 *  - lsim_produce_patterns: n distinct random patterns (n <= 2^P), packed so. */
LSIM_API int lsim_produce_faults(uint64_t seed, int64_t num_nodes, int64_t num_faults, int64_t num_pis,
                                 int64_t num_pos, char **flst, size_t *flst_len);
LSIM_API int lsim_pack_patterns(int64_t num_pis, int64_t num_patterns, int bits_per_round, const uint8_t *bits,
                                char **ptn, size_t *ptn_len);
LSIM_API int lsim_produce_patterns(uint64_t seed, int64_t num_pis, int64_t num_patterns, int bits_per_round,
                                   char **ptn, size_t *ptn_len);
/* The fault list or the patterns alone (same seeds and rules as lsim_generate):
 * SA0/SA1 over inner gate ids [P+O, P+O+G) thinned to `fraction`, or
 * ceil(n/64) rounds of P random words. */
LSIM_API int lsim_generate_faults(uint64_t seed, int64_t num_pis, int64_t num_pos, int64_t num_gates,
                                  double fraction, char **flst, size_t *flst_len);
LSIM_API int lsim_generate_patterns(uint64_t seed, int64_t num_pis, int64_t num_patterns, char **ptn,
                                    size_t *ptn_len);

#ifdef __cplusplus
}
#endif
#endif
