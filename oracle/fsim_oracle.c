/*
 * CPU oracle: plain-C restatement of the reference stuck-at fault simulator
 * fsim::Simulator (fsim/simulation.{hpp,cpp}).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by the tests, the smoke check and the CPU
 * baseline leg of the benchmark, always as the checker. Nothing in the product
 * library links or calls it.
 *
 * Where the reference's behaviour is defined this file reproduces it,
 * including the quirks (MUX uses the word-level `!s`, non-dense input pins
 * read node 0 through the zero-initialised pin table, the last round
 * overwrites the report). Where the reference indexes out of bounds
 * (fsim/simulation.cpp:685-711,856-877,885,967-969) the input is rejected.
 */
#include "fsim_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { T_INV = 0, T_AND, T_OR, T_XOR, T_NAND, T_NOR, T_XNOR, T_MUX, T_CLKBUF, T_PI, T_PO };

struct oracle_sim {
  int64_t P, O, G, W, N;
  int64_t *order_idx; /* file id -> position (_gate_idx_to_tpg_order_idx) */
  int *type;          /* per position */
  int64_t *in_off;    /* per position, fanin list offsets (positions) */
  int64_t *in;
  int64_t *level;     /* per position */
  int64_t num_levels;
  int64_t *per_level; /* _num_gates_per_level */
  int64_t F;
  int64_t *fault_site; /* position */
  uint64_t *fault_val; /* 0 or ~0 */
  int64_t num_patterns, R;
  uint64_t *patterns;  /* R x P, round-major */
  /* fanout lists for the event-driven restatement */
  int64_t *out_off, *out;
};

/* ---- text scanning: whitespace separated unsigned decimals ---- */
typedef struct {
  const char *p, *end;
  int bad;
} scan_t;

static int scan_u64(scan_t *s, uint64_t *v) {
  while (s->p < s->end && (*s->p == ' ' || *s->p == '\n' || *s->p == '\t' || *s->p == '\r' || *s->p == '\f' || *s->p == '\v')) s->p++;
  if (s->p >= s->end) return 0;
  if (*s->p < '0' || *s->p > '9') {
    s->bad = 1;
    return 0;
  }
  uint64_t x = 0;
  while (s->p < s->end && *s->p >= '0' && *s->p <= '9') {
    uint64_t d = (uint64_t)(*s->p - '0');
    if (x > (UINT64_MAX - d) / 10) {
      s->bad = 1;
      return 0;
    }
    x = x * 10 + d;
    s->p++;
  }
  *v = x;
  return 1;
}

static void set_err(char *err, size_t n, const char *msg) {
  if (err && n) {
    strncpy(err, msg, n - 1);
    err[n - 1] = 0;
  }
}

void oracle_free(oracle_sim *s) {
  if (!s) return;
  free(s->order_idx);
  free(s->type);
  free(s->in_off);
  free(s->in);
  free(s->level);
  free(s->per_level);
  free(s->fault_site);
  free(s->fault_val);
  free(s->patterns);
  free(s->out_off);
  free(s->out);
  free(s);
}

#define FAIL(msg)                  \
  do {                             \
    set_err(err, err_len, (msg));  \
    goto fail;                     \
  } while (0)
#define NEXT(sc, var, msg)               \
  do {                                   \
    if (!scan_u64((sc), &(var))) FAIL(msg); \
  } while (0)

oracle_sim *oracle_read(const char *ckt, size_t ckt_len, const char *flst, size_t flst_len, const char *ptn,
                        size_t ptn_len, char *err, size_t err_len) {
  oracle_sim *s = (oracle_sim *)calloc(1, sizeof(oracle_sim));
  int64_t *adj_cnt = NULL, *adj_off = NULL, *adj = NULL, *indeg = NULL, *order = NULL, *queue = NULL;
  int64_t *wsrc = NULL, *wdst = NULL;
  uint64_t(*table)[5] = NULL; /* gate_input_order_table[N][5] (fsim/simulation.cpp:676-680) */
  int *type_id = NULL;
  int64_t nedge = 0, cap = 0;
  scan_t c = {ckt, ckt + ckt_len, 0};
  uint64_t P, O, G, W;
  NEXT(&c, P, "circuit header");
  NEXT(&c, O, "circuit header");
  NEXT(&c, G, "circuit header");
  NEXT(&c, W, "circuit header");
  s->P = (int64_t)P;
  s->O = (int64_t)O;
  s->G = (int64_t)G;
  s->W = (int64_t)W;
  s->N = s->P + s->O + s->G;
  const int64_t N = s->N;
  if (s->P < 1) FAIL("no PI");
  table = calloc((size_t)N, sizeof(*table));
  indeg = calloc((size_t)N, sizeof(int64_t));
  adj_cnt = calloc((size_t)N + 1, sizeof(int64_t));
  /* wires (fsim/simulation.cpp:684-711): keep the edge list in file order */
  for (uint64_t w = 0; w < W; w++) {
    uint64_t src, pin_y, k;
    NEXT(&c, src, "wire");
    NEXT(&c, pin_y, "wire");
    NEXT(&c, k, "wire");
    (void)pin_y;
    if (src >= (uint64_t)N) FAIL("wire source out of range");
    for (uint64_t j = 0; j < k; j++) {
      uint64_t dst, pin;
      NEXT(&c, dst, "wire");
      NEXT(&c, pin, "wire");
      if (dst >= (uint64_t)N) FAIL("wire destination out of range");
      if (pin < 1 || pin > 5) FAIL("pin id out of range");
      if (pin == 5) pin = 3; /* MUX select (fsim/simulation.cpp:703) */
      if (nedge == cap) {
        cap = cap ? cap * 2 : 1024;
        wsrc = realloc(wsrc, (size_t)cap * sizeof(int64_t));
        wdst = realloc(wdst, (size_t)cap * sizeof(int64_t));
      }
      wsrc[nedge] = (int64_t)src;
      wdst[nedge] = (int64_t)dst;
      nedge++;
      indeg[dst]++;
      table[dst][pin] = src;
      table[dst][0]++;
      if (table[dst][0] > 4) FAIL("more than 4 inputs");
    }
  }
  /* adjacency in insertion order */
  for (int64_t e = 0; e < nedge; e++) adj_cnt[wsrc[e] + 1]++;
  adj_off = calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t i = 0; i < N; i++) adj_off[i + 1] = adj_off[i] + adj_cnt[i + 1];
  adj = malloc((size_t)(nedge ? nedge : 1) * sizeof(int64_t));
  {
    int64_t *fill = calloc((size_t)N, sizeof(int64_t));
    for (int64_t e = 0; e < nedge; e++) adj[adj_off[wsrc[e]] + fill[wsrc[e]]++] = wdst[e];
    free(fill);
  }
  /* Kahn with a FIFO queue seeded by PIs 0..P-1 (fsim/simulation.cpp:719-738) */
  order = malloc((size_t)(N + s->P + 1) * sizeof(int64_t));
  queue = malloc((size_t)(N + s->P + 1) * sizeof(int64_t));
  int64_t qh = 0, qt = 0, no = 0;
  for (int64_t i = 0; i < s->P; i++) queue[qt++] = i;
  while (qh < qt) {
    int64_t u = queue[qh++];
    if (no >= N) FAIL("node visited twice (fanin into a PI)");
    order[no++] = u;
    for (int64_t e = adj_off[u]; e < adj_off[u + 1]; e++) {
      int64_t v = adj[e];
      if (--indeg[v] == 0) {
        if (qt >= N + s->P) FAIL("node visited twice (fanin into a PI)");
        queue[qt++] = v;
      }
    }
  }
  if (no != N) FAIL("unreachable node or cycle");
  /* drop POs and append them in id order (fsim/simulation.cpp:742-751) */
  {
    int64_t k = 0;
    for (int64_t i = 0; i < no; i++)
      if (!(order[i] >= s->P && order[i] < s->P + s->O)) order[k++] = order[i];
    if (k != s->P + s->G) FAIL("order size mismatch");
    for (int64_t i = 0; i < s->O; i++) order[k++] = s->P + i;
  }
  s->order_idx = malloc((size_t)N * sizeof(int64_t));
  for (int64_t i = 0; i < N; i++) s->order_idx[order[i]] = i;
  /* types (fsim/simulation.cpp:760-772) */
  type_id = malloc((size_t)N * sizeof(int));
  for (int64_t i = 0; i < s->P; i++) type_id[i] = T_PI;
  for (int64_t i = s->P; i < s->P + s->O; i++) type_id[i] = T_PO;
  for (int64_t i = s->P + s->O; i < N; i++) {
    uint64_t t;
    NEXT(&c, t, "gate types");
    if (t == T_PI || t > T_PO) FAIL("gate type without defined evaluation");
    type_id[i] = (int)t;
  }
  s->type = malloc((size_t)N * sizeof(int));
  for (int64_t p = 0; p < N; p++) s->type[p] = type_id[order[p]];
  /* fanins in pin order (fsim/simulation.cpp:851-877); positions */
  s->in_off = calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t p = 0; p < N; p++) {
    int64_t id = order[p];
    int64_t cnt = (p >= s->P) ? (int64_t)table[id][0] : 0;
    s->in_off[p + 1] = s->in_off[p] + cnt;
  }
  s->in = malloc((size_t)(s->in_off[N] ? s->in_off[N] : 1) * sizeof(int64_t));
  for (int64_t p = s->P; p < N; p++) {
    int64_t id = order[p];
    int64_t cnt = (int64_t)table[id][0];
    int is_po_dst = id >= s->P && id < s->P + s->O;
    if (cnt == 0) FAIL("node without inputs");
    for (int64_t j = 1; j <= cnt; j++) {
      int64_t from = (int64_t)table[id][j];
      int from_pi = from < s->P;
      int from_po = from >= s->P && from < s->P + s->O;
      /* the reference resolves PI->gate, gate->PO and gate->gate only */
      if (from_pi && is_po_dst) FAIL("PI drives a PO");
      if (from_po) FAIL("PO drives a node");
      s->in[s->in_off[p] + j - 1] = s->order_idx[from];
    }
    if (s->type[p] == T_MUX && cnt < 3) FAIL("MUX with fewer than 3 inputs");
  }
  /* levels (fsim/simulation.cpp:882-916) */
  s->level = calloc((size_t)N, sizeof(int64_t));
  int64_t maxl = 0;
  for (int64_t p = s->P; p < s->P + s->G; p++) {
    int64_t l = s->level[s->in[s->in_off[p]]] + 1;
    for (int64_t e = s->in_off[p] + 1; e < s->in_off[p + 1]; e++)
      if (s->level[s->in[e]] + 1 > l) l = s->level[s->in[e]] + 1;
    s->level[p] = l;
    if (l > maxl) maxl = l;
  }
  for (int64_t p = s->P + s->G; p < N; p++) s->level[p] = maxl + 1;
  maxl++;
  s->num_levels = maxl + 1;
  s->per_level = calloc((size_t)s->num_levels, sizeof(int64_t));
  for (int64_t p = 0; p < N; p++) s->per_level[s->level[p]]++;
  /* faults (fsim/simulation.cpp:943-959) */
  {
    scan_t f = {flst, flst + flst_len, 0};
    uint64_t F;
    NEXT(&f, F, "fault header");
    s->F = (int64_t)F;
    s->fault_site = malloc((size_t)(F ? F : 1) * sizeof(int64_t));
    s->fault_val = malloc((size_t)(F ? F : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < F; i++) {
      uint64_t id, v;
      NEXT(&f, id, "fault");
      NEXT(&f, v, "fault");
      if (id >= (uint64_t)N) FAIL("fault id out of range");
      s->fault_site[i] = s->order_idx[id];
      s->fault_val[i] = v ? ~(uint64_t)0 : 0;
    }
  }
  /* patterns (fsim/simulation.cpp:961-976): round i lands in slot order_idx[i] */
  {
    scan_t t = {ptn, ptn + ptn_len, 0};
    uint64_t n;
    NEXT(&t, n, "pattern header");
    s->num_patterns = (int64_t)n;
    s->R = (s->num_patterns + 63) / 64;
    for (int64_t i = 0; i < s->R; i++)
      if (i >= N || s->order_idx[i] != i) FAIL("more rounds than PIs");
    s->patterns = malloc((size_t)((s->R * s->P) ? s->R * s->P : 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < s->R; i++)
      for (int64_t pi = 0; pi < s->P; pi++) NEXT(&t, s->patterns[s->order_idx[i] * s->P + pi], "pattern word");
  }
  /* fanout lists for the event-driven restatement */
  s->out_off = calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t p = 0; p < N; p++)
    for (int64_t e = s->in_off[p]; e < s->in_off[p + 1]; e++) s->out_off[s->in[e] + 1]++;
  for (int64_t p = 0; p < N; p++) s->out_off[p + 1] += s->out_off[p];
  s->out = malloc((size_t)(s->out_off[N] ? s->out_off[N] : 1) * sizeof(int64_t));
  {
    int64_t *fill = calloc((size_t)N, sizeof(int64_t));
    for (int64_t p = 0; p < N; p++)
      for (int64_t e = s->in_off[p]; e < s->in_off[p + 1]; e++) s->out[s->out_off[s->in[e]] + fill[s->in[e]]++] = p;
    free(fill);
  }
  free(adj_cnt);
  free(adj_off);
  free(adj);
  free(indeg);
  free(order);
  free(queue);
  free(wsrc);
  free(wdst);
  free(table);
  free(type_id);
  return s;
fail:
  free(adj_cnt);
  free(adj_off);
  free(adj);
  free(indeg);
  free(order);
  free(queue);
  free(wsrc);
  free(wdst);
  free(table);
  free(type_id);
  oracle_free(s);
  return NULL;
}

int64_t oracle_count(const oracle_sim *s, int which) {
  switch (which) {
    case ORACLE_NUM_PIS: return s->P;
    case ORACLE_NUM_POS: return s->O;
    case ORACLE_NUM_GATES: return s->G;
    case ORACLE_NUM_WIRES: return s->W;
    case ORACLE_NUM_ROUNDS: return s->R;
    case ORACLE_NUM_FAULTS: return s->F;
    case ORACLE_NUM_NODES: return s->N;
    case ORACLE_NUM_LEVELS: return s->num_levels;
    case ORACLE_NUM_PATTERNS: return s->num_patterns;
  }
  return -1;
}

void oracle_order_map(const oracle_sim *s, uint64_t *out) {
  for (int64_t i = 0; i < s->N; i++) out[i] = (uint64_t)s->order_idx[i];
}
void oracle_gates_per_level(const oracle_sim *s, uint64_t *out) {
  for (int64_t i = 0; i < s->num_levels; i++) out[i] = (uint64_t)s->per_level[i];
}
void oracle_levels(const oracle_sim *s, uint64_t *out) {
  for (int64_t i = 0; i < s->N; i++) out[i] = (uint64_t)s->level[i];
}
void oracle_faults(const oracle_sim *s, uint64_t *site, uint64_t *val) {
  for (int64_t i = 0; i < s->F; i++) {
    site[i] = (uint64_t)s->fault_site[i];
    val[i] = s->fault_val[i];
  }
}
void oracle_patterns(const oracle_sim *s, uint64_t *out) {
  memcpy(out, s->patterns, (size_t)(s->R * s->P) * sizeof(uint64_t));
}

/* One node evaluation (fsim/simulation.cpp:979-1124) reading fanin values
 * from v[]. PI index == position for PIs. */
static uint64_t eval_node(const oracle_sim *s, int64_t p, const uint64_t *v, const uint64_t *pattern) {
  const int64_t a = s->in_off[p], b = s->in_off[p + 1];
  uint64_t r;
  switch (s->type[p]) {
    case T_INV: return ~v[s->in[a]];
    case T_AND:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r &= v[s->in[e]];
      return r;
    case T_OR:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r |= v[s->in[e]];
      return r;
    case T_XOR:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r ^= v[s->in[e]];
      return r;
    case T_NAND:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r &= v[s->in[e]];
      return ~r;
    case T_NOR:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r |= v[s->in[e]];
      return ~r;
    case T_XNOR:
      r = v[s->in[a]];
      for (int64_t e = a + 1; e < b; e++) r ^= v[s->in[e]];
      return ~r;
    case T_MUX: {
      uint64_t x = v[s->in[a]], y = v[s->in[a + 1]], sel = v[s->in[a + 2]];
      return (sel & y) | (x & (uint64_t)(!sel)); /* word-level !s (fsim/simulation.cpp:1056) */
    }
    case T_CLKBUF:
    case T_PO: return v[s->in[a]];
    case T_PI: return pattern[p];
  }
  return 0;
}

static uint64_t round_mask(const oracle_sim *s, int64_t rd) {
  /* fsim/simulation.cpp:119-122 and the shift at :94-102 */
  int64_t bits = (s->num_patterns / (64 * (rd + 1))) ? 64 : (s->num_patterns % 64);
  return bits < 64 ? ((((uint64_t)1) << bits) - 1) : ~(uint64_t)0;
}

int oracle_run_literal(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel, uint64_t *table,
                       uint8_t *det, uint64_t *g_pi, uint64_t *g_gate, uint64_t *g_po) {
  const int64_t N = s->N, R = s->R, F = s->F;
  const int64_t nsel = fault_idx ? num_sel : F;
  uint64_t *good = calloc((size_t)N, sizeof(uint64_t));
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  if (table)
    for (int64_t k = 0; k < nsel; k++) {
      int64_t f = fault_idx ? fault_idx[k] : k;
      table[2 * f] = 0;
      table[2 * f + 1] = 0;
    }
  for (int64_t rd = 0; rd < R; rd++) {
    const uint64_t *pat = s->patterns + rd * s->P;
    const uint64_t mask = round_mask(s, rd);
    for (int64_t p = 0; p < N; p++) good[p] = eval_node(s, p, good, pat);
    if (rd == R - 1) {
      if (g_pi)
        for (int64_t i = 0; i < s->P; i++) g_pi[i] = good[i] & mask;
      if (g_gate)
        for (int64_t i = 0; i < s->G; i++) g_gate[i] = good[s->P + i] & mask;
      if (g_po)
        for (int64_t i = 0; i < s->O; i++) g_po[i] = good[s->P + s->G + i] & mask;
    }
#pragma omp parallel
    {
      uint64_t *bad = malloc((size_t)N * sizeof(uint64_t));
#pragma omp for schedule(dynamic, 16)
      for (int64_t k = 0; k < nsel; k++) {
        const int64_t f = fault_idx ? fault_idx[k] : k;
        const int64_t site = s->fault_site[f];
        /* whole-circuit faulty pass with the stuck-at override first
         * (fsim/simulation.cpp:186-205,1083-1086) */
        for (int64_t p = 0; p < N; p++) bad[p] = (p == site) ? s->fault_val[f] : eval_node(s, p, bad, pat);
        uint64_t found = 0;
        for (int64_t i = 0; i < s->O; i++) {
          const int64_t p = s->P + s->G + i;
          if ((good[p] & mask) != (bad[p] & mask)) {
            found = 1;
            break;
          }
        }
        if (table) {
          table[2 * f] = found; /* the last round overwrites (fsim/simulation.cpp:142-144) */
          table[2 * f + 1] = (uint64_t)rd;
        }
        if (det) det[f * R + rd] = (uint8_t)found;
      }
      free(bad);
    }
  }
  if (R == 0) {
    if (g_pi) memset(g_pi, 0, (size_t)s->P * sizeof(uint64_t));
    if (g_gate) memset(g_gate, 0, (size_t)s->G * sizeof(uint64_t));
    if (g_po) memset(g_po, 0, (size_t)s->O * sizeof(uint64_t));
  }
  free(good);
  return 0;
}

/* min-heap of positions */
typedef struct {
  int64_t *a;
  int64_t n;
} heap_t;
static void heap_push(heap_t *h, int64_t x) {
  int64_t i = h->n++;
  h->a[i] = x;
  while (i > 0) {
    int64_t par = (i - 1) / 2;
    if (h->a[par] <= h->a[i]) break;
    int64_t t = h->a[par];
    h->a[par] = h->a[i];
    h->a[i] = t;
    i = par;
  }
}
static int64_t heap_pop(heap_t *h) {
  int64_t top = h->a[0];
  h->a[0] = h->a[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->n && h->a[l] < h->a[m]) m = l;
    if (r < h->n && h->a[r] < h->a[m]) m = r;
    if (m == i) break;
    int64_t t = h->a[m];
    h->a[m] = h->a[i];
    h->a[i] = t;
    i = m;
  }
  return top;
}

/* Good machine of one round (fsim/simulation.cpp:149-184): the values of a
 * round depend on that round's pattern words only, so one pass gives them. */
static void good_round(const oracle_sim *s, int64_t rd, uint64_t *g) {
  const uint64_t *pat = s->patterns + rd * s->P;
  for (int64_t p = 0; p < s->N; p++) g[p] = eval_node(s, p, g, pat);
}

int oracle_good_round(const oracle_sim *s, int64_t rd, uint64_t *g_pi, uint64_t *g_gate, uint64_t *g_po) {
  if (rd < 0 || rd >= s->R) return -1;
  uint64_t *g = malloc((size_t)(s->N ? s->N : 1) * sizeof(uint64_t));
  if (!g) return -1;
  good_round(s, rd, g);
  const uint64_t mask = round_mask(s, rd);
  if (g_pi)
    for (int64_t i = 0; i < s->P; i++) g_pi[i] = g[i] & mask;
  if (g_gate)
    for (int64_t i = 0; i < s->G; i++) g_gate[i] = g[s->P + i] & mask;
  if (g_po)
    for (int64_t i = 0; i < s->O; i++) g_po[i] = g[s->P + s->G + i] & mask;
  free(g);
  return 0;
}

/* One (fault, round) of the event-driven restatement: the faulty machine
 * differs from the good one only downstream of the site, and a node whose
 * fanins all hold good values keeps its good value. Nodes are evaluated in
 * position (topological) order from a min-heap. Returns d(f, rd). */
typedef struct {
  uint64_t *val;
  int64_t *stamp, *queued;
  heap_t h;
  int64_t epoch;
} event_ws;

static int event_one(const oracle_sim *s, event_ws *w, int64_t f, const uint64_t *g, uint64_t mask) {
  const int64_t po0 = s->P + s->G;
  const int64_t site = s->fault_site[f];
  const int64_t ep = w->epoch++;
  int found = 0;
  w->h.n = 0;
  w->val[site] = s->fault_val[f];
  w->stamp[site] = ep;
  if (w->val[site] != g[site]) {
    if (site >= po0 && ((w->val[site] ^ g[site]) & mask)) found = 1;
    for (int64_t e = s->out_off[site]; e < s->out_off[site + 1]; e++)
      if (w->queued[s->out[e]] != ep) {
        w->queued[s->out[e]] = ep;
        heap_push(&w->h, s->out[e]);
      }
  }
  while (w->h.n) {
    const int64_t p = heap_pop(&w->h);
    if (p == site) continue;
    /* fanin values: faulty where stamped, good otherwise */
    const int64_t a = s->in_off[p], b = s->in_off[p + 1];
    uint64_t loc[4];
    for (int64_t e = a; e < b; e++) loc[e - a] = (w->stamp[s->in[e]] == ep) ? w->val[s->in[e]] : g[s->in[e]];
    uint64_t r = loc[0];
    switch (s->type[p]) {
      case T_INV: r = ~r; break;
      case T_AND: for (int64_t j = 1; j < b - a; j++) r &= loc[j]; break;
      case T_OR: for (int64_t j = 1; j < b - a; j++) r |= loc[j]; break;
      case T_XOR: for (int64_t j = 1; j < b - a; j++) r ^= loc[j]; break;
      case T_NAND: for (int64_t j = 1; j < b - a; j++) r &= loc[j]; r = ~r; break;
      case T_NOR: for (int64_t j = 1; j < b - a; j++) r |= loc[j]; r = ~r; break;
      case T_XNOR: for (int64_t j = 1; j < b - a; j++) r ^= loc[j]; r = ~r; break;
      case T_MUX: r = (loc[2] & loc[1]) | (loc[0] & (uint64_t)(!loc[2])); break;
      default: break; /* CLKBUF / PO copy input 0 */
    }
    if (r == g[p]) continue;
    w->val[p] = r;
    w->stamp[p] = ep;
    if (p >= po0 && ((r ^ g[p]) & mask)) found = 1;
    for (int64_t e = s->out_off[p]; e < s->out_off[p + 1]; e++)
      if (w->queued[s->out[e]] != ep) {
        w->queued[s->out[e]] = ep;
        heap_push(&w->h, s->out[e]);
      }
  }
  return found;
}

int oracle_run_event_range(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel, int64_t r0,
                           int64_t r1, int64_t chunk, int stop_at_first, uint8_t *det) {
  const int64_t N = s->N, F = s->F;
  const int64_t nsel = fault_idx ? num_sel : F;
  if (r0 < 0 || r1 > s->R || r0 > r1) return -1;
  const int64_t nr = r1 - r0;
  if (chunk < 1) chunk = 1;
  if (chunk > nr) chunk = nr > 0 ? nr : 1;
  memset(det, 0, (size_t)(nsel * nr));
  if (nsel == 0 || nr == 0) return 0;
  uint64_t *good = malloc((size_t)(N ? N : 1) * (size_t)chunk * sizeof(uint64_t)); /* [chunk][N] */
  uint8_t *done = calloc((size_t)nsel, 1);
  if (!good || !done) {
    free(good);
    free(done);
    return -1;
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (int64_t c0 = r0; c0 < r1; c0 += chunk) {
    const int64_t c1 = c0 + chunk < r1 ? c0 + chunk : r1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t rd = c0; rd < c1; rd++) good_round(s, rd, good + (rd - c0) * N);
#pragma omp parallel
    {
      event_ws w;
      w.val = malloc((size_t)N * sizeof(uint64_t));
      w.stamp = malloc((size_t)N * sizeof(int64_t));
      w.queued = malloc((size_t)N * sizeof(int64_t));
      w.h.a = malloc((size_t)(s->out_off[N] + 2) * sizeof(int64_t));
      w.epoch = 0;
      for (int64_t p = 0; p < N; p++) w.stamp[p] = w.queued[p] = -1;
#pragma omp for schedule(dynamic, 1)
      for (int64_t k = 0; k < nsel; k++) {
        if (done[k]) continue;
        const int64_t f = fault_idx ? fault_idx[k] : k;
        for (int64_t rd = c0; rd < c1; rd++) {
          const int d = event_one(s, &w, f, good + (rd - c0) * N, round_mask(s, rd));
          det[k * nr + (rd - r0)] = (uint8_t)d;
          if (d && stop_at_first) {
            done[k] = 1;
            break;
          }
        }
      }
      free(w.val);
      free(w.stamp);
      free(w.queued);
      free(w.h.a);
    }
  }
  free(good);
  free(done);
  return 0;
}

int oracle_run_event(const oracle_sim *s, int threads, const int64_t *fault_idx, int64_t num_sel, uint8_t *det) {
  /* all rounds in chunks of 64; rows of the selected faults land at their
   * fault index */
  const int64_t R = s->R, F = s->F;
  const int64_t nsel = fault_idx ? num_sel : F;
  uint8_t *rows = malloc((size_t)(nsel * R ? nsel * R : 1));
  if (!rows) return -1;
  int rc = oracle_run_event_range(s, threads, fault_idx, num_sel, 0, R, 64, 0, rows);
  if (rc == 0)
    for (int64_t k = 0; k < nsel; k++) {
      const int64_t f = fault_idx ? fault_idx[k] : k;
      memcpy(det + f * R, rows + k * R, (size_t)R);
    }
  free(rows);
  return rc;
}
