// GPU stuck-at fault simulation for gfx950 (MI355X).
//
// What is computed is the reference's detection function d(f, r): fault f is
// detected in 64-pattern round r iff some PO's faulty word differs from its
// good word in the round's live bits (fsim/simulation.cpp:118-146,186-228).
// The reference re-simulates the whole circuit per (fault, round); here:
//
//  * Good machine: one pass over all rounds, node-major layout good[node][R]
//    (lane = round word), one batched launch per level over all started jobs.
//  * Faults whose site cannot reach a PO are never detected and are not
//    simulated.
//  * Stem tracing (circuits without MUX): inside a fanout-free region a flip
//    of node x reaches the region's head h exactly on the patterns where
//    every gate on the path is sensitised, so
//        d(x/v) = L_x & D_h,   L_x = (good_x ^ v) & sens(x->h) & live bits,
//    with D_h the patterns on which flipping h changes a PO: all ones for a
//    PO head, an explicit forward sweep for a stem head. The sweep of stem h
//    only needs the bits U_h = OR of L over the step's faults in its region,
//    so it flips h on U_h alone and (stem, word) pairs with U_h = 0 are not
//    swept at all. Gates are bitwise per pattern, so this is exact. The
//    reference's MUX uses the word-level `!s` (fsim/simulation.cpp:1056),
//    which couples the 64 patterns of a word through the select only: a
//    fault whose cone reaches a MUX select is injected directly (full
//    word), skipping words where the site's good value already equals the
//    stuck value; every other fault keeps stem tracing.
//  * Sweeps: a tile holds up to 64 lanes of one job, each lane one (injection
//    site, word, injected bits).
//      - plain form (k_sweep_level): one launch per level, every live node
//        of the level visited by every tile of the batch, a wave per node;
//      - persistent forms: one launch over the tiles of all active jobs; a
//        workgroup owns a tile at a time, walks the levels from its lowest
//        injection to the deepest pending node and visits only nodes with a
//        differing fanin. k_sweep_events evaluates a node per wave;
//        k_sweep_pairs (the default) stages a level's nodes in LDS and
//        evaluates their (node, active lane) pairs, with each differing
//        node's faulty words packed into a per-tile pool.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>

#include "engine.hpp"

namespace lsim {

#define LSIM_HIP(x)                                                                                  \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess)                                                                            \
      throw DeviceError(e_ == hipErrorOutOfMemory ? LSIM_E_NOMEM : LSIM_E_DEVICE,                    \
                        std::string(#x) + ": " + hipGetErrorString(e_));                             \
  } while (0)

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

namespace {

enum : uint32_t { K_SA0 = 0, K_SA1 = 1, K_FLIP = 2 };

// One sweep lane: inject at `site` in word `word`; for K_FLIP the good value
// is complemented on the bits of `mask`, for K_SA0/K_SA1 the whole word is
// forced. The OR over POs of (faulty ^ good) goes to out[out].
struct Lane {
  unsigned long long mask;
  int32_t site;
  int32_t word;
  int32_t out;
  uint32_t kind;
  int32_t level;  // level of `site`
  int32_t pad;
};

struct Tile {
  int32_t lane0;        // first lane record of the tile
  int32_t nlanes;       // 1..64
  int32_t start_level;  // level of the lowest injection site
  int32_t end_level;    // level of the highest injection site
  int32_t job;          // index into the iteration's job table
};

// Per-tile record of a node that differs from the good machine in some lane:
// the lane mask, valid only when `epoch` equals the tile's epoch.
struct MaskRec {
  unsigned long long mask;
  uint32_t epoch;
  uint32_t pad;
};

// Everything a kernel needs about one job and its circuit in this iteration.
struct JobDev {
  // circuit (device positions: live nodes first, then dead nodes)
  const uint32_t *info;      // type | fanin count << 4 | deepest live fanout level << 8
  const int4 *fin;           // fanin positions (PI: PI index in .x)
  const int4 *nrec;          // narrow circuits: per live node {info, fanin 0, fanin 1, first live fanout}, NL + 1 entries
  const int32_t *live_lvl;   // first live position of each level, L+2 entries
  const int32_t *dead_lvl;   // first dead position of each level, L+2 entries
  const int32_t *fo_off;     // live fanouts of live nodes
  const int2 *fo;            // {position, level}, ascending
  const int32_t *parent;     // fanout-free-region parent, -1 at heads
  const uint8_t *is_po;      // per live position
  int max_level;
  int NL;
  // job
  uint64_t *good;            // [N][RW]
  const uint64_t *pi_words;  // [P][RW]
  int64_t RW;
  int64_t num_patterns;
  const int32_t *site;       // per fault, live position (dead faults never listed)
  const int32_t *head;       // per fault, region head
  const uint8_t *sa;
  int32_t *head_slot;        // per live position: the stem's unit index in this step, else -1
  int32_t *first_round;      // atomicMin, or null
  uint64_t *det_bits;        // [F][RB], or null
  int64_t RB;
  uint8_t *last_flag;        // d(f, R-1), or null
  int64_t last_word;
  // this step
  const int32_t *flist;      // faults of the step (direct: in unit order)
  int64_t nf;
  int64_t w0;
  int Wn;
  int direct;
  int64_t item0;             // first detection item of the job
  int64_t word0;             // first need/dout word of the job's units
  int64_t lane_unit0;        // first entry of the job's units in the step's unit list
  const int32_t *units;      // the step's units (positions), in order
};

// Device memory pointers in the global address space: loads through them are
// global loads, which the LDS wait counter does not track. A pointer read out
// of the job table is otherwise a generic one, loaded with flat instructions
// that count against both wait counters, so every wait for an LDS read or a
// cross-lane shuffle would also wait for them.
#define LSIM_GLOBAL __attribute__((address_space(1)))
// int4 / int2 / 2 x uint64 as the compiler's own vector types, which (unlike
// HIP's int4 / int2 classes) load from any address space.
typedef int gi4 __attribute__((ext_vector_type(4)));
typedef int gi2 __attribute__((ext_vector_type(2)));
typedef unsigned long long gu2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int4 to_int4(gi4 v) { return make_int4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ int2 to_int2(gi2 v) { return make_int2(v.x, v.y); }
// A live-fanout entry is {position, level | pins << 24}: `pins` has bit k set
// when the fanout reads the node on its input pin k.
__device__ __forceinline__ int fo_level(int2 f) { return f.y & 0xffffff; }
__device__ __forceinline__ uint32_t fo_pins(int2 f) { return uint32_t(f.y) >> 24; }

// First live positions of levels l and l + 1, read with a scalar load: the
// table is read-only while a kernel runs, and a scalar load is waited for on
// the LDS/scalar counter, so it does not wait for the vector stores still in
// flight (a vector load would, its counter being in order with them).
__device__ __forceinline__ int2 level_range(const LSIM_GLOBAL int32_t *lvl, int l) {
  const uint64_t p = reinterpret_cast<uint64_t>(lvl);
  // readfirstlane returns a signed int: each half goes through uint32_t so
  // that the low half is not sign-extended into the high one.
  const uint64_t q = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(p))))) |
                     (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(p >> 32))))) << 32);
  const __attribute__((address_space(4))) int32_t *c = reinterpret_cast<const __attribute__((address_space(4))) int32_t *>(q);
  const int i = __builtin_amdgcn_readfirstlane(l);
  return make_int2(c[i], c[i + 1]);
}

// The fields a sweep reads, copied out of the job table once per tile so
// they stay in registers (stores to sweep state cannot alias them).
struct SweepView {
  const LSIM_GLOBAL uint32_t *__restrict__ info;
  const LSIM_GLOBAL gi4 *__restrict__ fin;
  const LSIM_GLOBAL uint64_t *__restrict__ good;
  const LSIM_GLOBAL int32_t *__restrict__ live_lvl;
  const LSIM_GLOBAL int32_t *__restrict__ fo_off;
  const LSIM_GLOBAL gi2 *__restrict__ fo;
  const LSIM_GLOBAL gi4 *__restrict__ nrec;
  int64_t RW;
};

__device__ __forceinline__ SweepView sweep_view(const JobDev *jobs, int j) {
  SweepView v;
  v.info = (const LSIM_GLOBAL uint32_t *)jobs[j].info;
  v.fin = (const LSIM_GLOBAL gi4 *)jobs[j].fin;
  v.good = (const LSIM_GLOBAL uint64_t *)jobs[j].good;
  v.live_lvl = (const LSIM_GLOBAL int32_t *)jobs[j].live_lvl;
  v.fo_off = (const LSIM_GLOBAL int32_t *)jobs[j].fo_off;
  v.fo = (const LSIM_GLOBAL gi2 *)jobs[j].fo;
  v.nrec = (const LSIM_GLOBAL gi4 *)jobs[j].nrec;
  v.RW = jobs[j].RW;
  return v;
}

__device__ __forceinline__ uint64_t eval_gate(uint32_t type, uint32_t nin, uint64_t v0, uint64_t v1, uint64_t v2,
                                              uint64_t v3) {
  // Same semantics as the reference _apply_* (fsim/simulation.cpp:979-1077).
  switch (type) {
    case G_INV: return ~v0;
    case G_AND:
    case G_NAND: {
      uint64_t r = v0;
      if (nin > 1) r &= v1;
      if (nin > 2) r &= v2;
      if (nin > 3) r &= v3;
      return type == G_AND ? r : ~r;
    }
    case G_OR:
    case G_NOR: {
      uint64_t r = v0;
      if (nin > 1) r |= v1;
      if (nin > 2) r |= v2;
      if (nin > 3) r |= v3;
      return type == G_OR ? r : ~r;
    }
    case G_XOR:
    case G_XNOR: {
      uint64_t r = v0;
      if (nin > 1) r ^= v1;
      if (nin > 2) r ^= v2;
      if (nin > 3) r ^= v3;
      return type == G_XOR ? r : ~r;
    }
    case G_MUX: return (v2 & v1) | (v0 & (uint64_t)(v2 == 0));
    default: return v0;  // CLKBUF, PO
  }
}

// eval_gate without branches: the type only selects among results computed
// for every family, so lanes holding different gate types (the pair sweep's
// flattened passes) or a type under a lane-dependent condition cost selects,
// not a divergent switch. Missing fanins take the fold's identity. Same
// semantics as eval_gate; the 2-input form is for gates of at most two fanins.
__device__ __forceinline__ uint64_t eval_gate_sel2(uint32_t type, uint32_t nin, uint64_t v0, uint64_t v1) {
  const bool f_and = type == G_AND || type == G_NAND, f_or = type == G_OR || type == G_NOR;
  const bool f_xor = type == G_XOR || type == G_XNOR;
  const bool inv = type == G_INV || type == G_NAND || type == G_NOR || type == G_XNOR;
  const uint64_t b = nin > 1 ? v1 : (f_and ? ~uint64_t(0) : uint64_t(0));
  const uint64_t r = f_and ? (v0 & b) : f_or ? (v0 | b) : f_xor ? (v0 ^ b) : v0;
  return inv ? ~r : r;
}

__device__ __forceinline__ uint64_t eval_gate_sel(uint32_t type, uint32_t nin, uint64_t v0, uint64_t v1, uint64_t v2,
                                                  uint64_t v3) {
  const bool f_and = type == G_AND || type == G_NAND, f_or = type == G_OR || type == G_NOR;
  const bool f_xor = type == G_XOR || type == G_XNOR;
  const bool inv = type == G_INV || type == G_NAND || type == G_NOR || type == G_XNOR;
  const uint64_t id = f_and ? ~uint64_t(0) : uint64_t(0);
  const uint64_t b1 = nin > 1 ? v1 : id, b2 = nin > 2 ? v2 : id, b3 = nin > 3 ? v3 : id;
  uint64_t r = f_and ? (v0 & b1 & b2 & b3) : f_or ? (v0 | b1 | b2 | b3) : f_xor ? (v0 ^ b1 ^ b2 ^ b3) : v0;
  if (type == G_MUX) r = (v2 & v1) | (v0 & (uint64_t)(v2 == 0));
  return inv ? ~r : r;
}

// Gates of one or two inputs as 4-bit truth tables: bit 2a + b is the output
// for bit a on pin 0 and bit b on pin 1 (a one-input gate's table does not
// depend on b). Four bits per type code, for two inputs and for one; the same
// functions as eval_gate.
__host__ __device__ constexpr uint64_t tt_entry(int type, uint32_t table) { return uint64_t(table) << (4 * type); }
constexpr uint64_t kGateTT2 = tt_entry(G_INV, 0x3) | tt_entry(G_AND, 0x8) | tt_entry(G_OR, 0xe) |
                              tt_entry(G_XOR, 0x6) | tt_entry(G_NAND, 0x7) | tt_entry(G_NOR, 0x1) |
                              tt_entry(G_XNOR, 0x9) | tt_entry(G_CLKBUF, 0xc) | tt_entry(G_PO, 0xc);
constexpr uint64_t kGateTT1 = tt_entry(G_INV, 0x3) | tt_entry(G_AND, 0xc) | tt_entry(G_OR, 0xc) |
                              tt_entry(G_XOR, 0xc) | tt_entry(G_NAND, 0x3) | tt_entry(G_NOR, 0x3) |
                              tt_entry(G_XNOR, 0x3) | tt_entry(G_CLKBUF, 0xc) | tt_entry(G_PO, 0xc);

// Truth table of a gate of at most two inputs (not MUX, not PI).
__device__ __forceinline__ uint32_t gate_tt(uint32_t type, uint32_t nin) {
  return uint32_t(((nin > 1 ? kGateTT2 : kGateTT1) >> (4 * type)) & 15u);
}

// The table applied bitwise: three bit selects, no branch, whether the table
// is uniform over the wave (scalar masks) or not.
__device__ __forceinline__ uint64_t eval_tt(uint32_t tt, uint64_t a, uint64_t b) {
  // 32-bit masks, applied to each half.
  const uint32_t t0 = 0u - (tt & 1u), t1 = 0u - ((tt >> 1) & 1u), t2 = 0u - ((tt >> 2) & 1u), t3 = 0u - (tt >> 3);
  auto half = [&](uint32_t x, uint32_t y) {
    const uint32_t hi = (y & t3) | (~y & t2), lo = (y & t1) | (~y & t0);
    return (x & hi) | (~x & lo);
  };
  return uint64_t(half(uint32_t(a), uint32_t(b))) | (uint64_t(half(uint32_t(a >> 32), uint32_t(b >> 32))) << 32);
}

__device__ __forceinline__ uint64_t round_mask(int64_t num_patterns, int64_t w) {
  const int64_t bits = (num_patterns / (64 * (w + 1))) ? 64 : (num_patterns % 64);
  return bits < 64 ? ((uint64_t(1) << bits) - 1) : ~uint64_t(0);
}

// Index of the job whose item range [off[j], off[j+1]) holds t.
__device__ __forceinline__ int find_job(const int64_t *off, int njobs, int64_t t) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Runs fn(j) on every active lane of the wave with j the job of the lane's
// item t, j uniform over the lanes of each call: one binary search (scalar
// loads) per distinct job among the wave's items. Items ascend with the lane,
// so a wave usually holds one job.
template <class Fn>
__device__ __forceinline__ void for_wave_jobs(const int64_t *off, int njobs, int64_t t, Fn &&fn) {
  bool done = false;
  for (;;) {
    const uint64_t left = __ballot(!done);
    if (!left) break;
    const int src = int(__builtin_ctzll(left));
    const int64_t ts = int64_t((uint64_t(uint32_t(__builtin_amdgcn_readlane(int(t >> 32), src))) << 32) |
                               uint64_t(uint32_t(__builtin_amdgcn_readlane(int(t), src))));
    const int j = __builtin_amdgcn_readfirstlane(find_job(off, njobs, ts));
    if (!done && t < off[j + 1]) {
      fn(j);
      done = true;
    }
  }
}

// lt / d for a job whose items number `span`: 32-bit division when they fit.
__device__ __forceinline__ int64_t div_items(int64_t lt, int64_t d, int64_t span) {
  return span <= int64_t(UINT32_MAX) ? int64_t(uint32_t(lt) / uint32_t(d)) : lt / d;
}

// ---------------------------------------------------------------- good machine

// Level `level` of every job: threads cover (job, node, word); word is the
// fastest index so a wave reads a node's consecutive words in one coalesced
// access. off = per-job item prefix for this level.
__device__ __forceinline__ void good_item(const JobDev &J, int level, int64_t lt, int64_t span) {
  const int64_t RW = J.RW;
  const int i = int(div_items(lt, RW, span));
  const int64_t w = lt - int64_t(i) * RW;
  const int a0 = J.live_lvl[level], n0 = J.live_lvl[level + 1] - a0;
  const int u = i < n0 ? a0 + i : J.dead_lvl[level] + (i - n0);
  const uint32_t inf = J.info[u];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = J.fin[u];
  const auto good = J.good;
  uint64_t r;
  if (type == G_PI) {
    r = J.pi_words[int64_t(f.x) * RW + w];
  } else {
    const uint64_t v0 = good[int64_t(f.x) * RW + w];
    const uint64_t v1 = nin > 1 ? good[int64_t(f.y) * RW + w] : 0;
    const uint64_t v2 = nin > 2 ? good[int64_t(f.z) * RW + w] : 0;
    const uint64_t v3 = nin > 3 ? good[int64_t(f.w) * RW + w] : 0;
    // A wave spans several nodes when a job has fewer than 64 words per node.
    r = eval_gate_sel(type, nin, v0, v1, v2, v3);
  }
  J.good[int64_t(u) * RW + w] = r;
}

__global__ void k_good_level(const JobDev *__restrict__ jobs, int njobs, const int64_t *__restrict__ off,
                             int level) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= off[njobs]) return;
  for_wave_jobs(off, njobs, t, [&](int j) { good_item(jobs[j], level, t - off[j], off[j + 1] - off[j]); });
}

// Masked good values of one round, gathered per device position.
// Level-window form of the good machine (the partitioned form): the levels
// are cut into windows of up to `win_levels` consecutive levels; a window's
// nodes are split into blocks, each closed under fanins inside the window (the
// backward cones of the window's outputs, shared nodes repeated in every
// block that needs them, each written by one owner). A block holds at most
// `cap` nodes, sized from the device's LDS: a workgroup evaluates one block for
// 64 words level by level, keeping the block's values (cap x 64 words of 8 B)
// and node records (32 B each) in LDS and reading only fanins of earlier
// windows from HBM. On gfx950 (160 KB of LDS per CU, two blocks resident per
// CU) that is 150 nodes in 80 KB. A window whose deepest cone does not fit is
// cut to fewer levels.
constexpr int kWinLevelsMax = 8;

struct WinNode {
  int32_t u;      // device position
  uint32_t info;  // type | fanin count << 4 | owner << 8
  int4 fin;       // fanin position, or ~(block index) for a fanin inside the block; PI index for a PI
};

struct WinBlock {
  int32_t node0;                   // first WinNode of the block
  int32_t seg[kWinLevelsMax + 1];  // seg[k]: first block node on the window's k-th level; seg[depth]: count
};

constexpr int kWinGroup = 4;  // nodes per wave with their loads in flight together

// LDS of one window block: values [cap][64] words, then the block's node records.
__host__ __device__ constexpr size_t win_lds_bytes(int cap) {
  return size_t(cap) * 64 * sizeof(uint64_t) + size_t(cap) * sizeof(WinNode);
}

__global__ void __launch_bounds__(512) k_good_window(const JobDev *__restrict__ jd, const WinBlock *__restrict__ blocks,
                                                     const WinNode *__restrict__ nodes, int block0, int wgroups,
                                                     int depth, int cap) {
  extern __shared__ uint64_t s_val[];  // [cap][64], then WinNode[cap]
  WinNode *s_node = reinterpret_cast<WinNode *>(s_val + int64_t(cap) * 64);
  const JobDev &J = jd[0];
  const WinBlock B = blocks[block0 + int(blockIdx.x) / wgroups];
  const int g = int(blockIdx.x) % wgroups;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int64_t RW = J.RW;
  const int64_t w = int64_t(g) * 64 + lane;
  const bool wok = w < RW;
  const auto good = J.good;
  // The block's node records, once, in one coalesced pass.
  const int nb = B.seg[depth];
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s_node[i] = nodes[B.node0 + i];
  __syncthreads();
  for (int k = 0; k < depth; ++k) {
    const int a = B.seg[k], b = B.seg[k + 1];
    // Nodes of one level do not read each other: a wave issues the fanin
    // loads of kWinGroup nodes before it evaluates any of them.
    for (int i0 = a + wave; i0 < b; i0 += kWinGroup * nwaves) {
      uint64_t v[kWinGroup][4];
      uint32_t info[kWinGroup];
#pragma unroll
      for (int t = 0; t < kWinGroup; ++t) {
        const int i = i0 + t * nwaves;
        info[t] = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[t][q] = 0;
        if (i < b && wok) {
          const WinNode n = s_node[i];
          info[t] = n.info;
          const uint32_t type = n.info & 15u, nin = (n.info >> 4) & 7u;
          if (type == G_PI) {
            v[t][0] = J.pi_words[int64_t(n.fin.x) * RW + w];
          } else {
            const int fi[4] = {n.fin.x, n.fin.y, n.fin.z, n.fin.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (q < int(nin))
                v[t][q] = fi[q] < 0 ? s_val[int64_t(~fi[q]) * 64 + lane] : good[int64_t(fi[q]) * RW + w];
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kWinGroup; ++t) {
        const int i = i0 + t * nwaves;
        if (i >= b) continue;
        uint64_t r = 0;
        if (wok) {
          const uint32_t type = info[t] & 15u, nin = (info[t] >> 4) & 7u;
          r = type == G_PI ? v[t][0] : eval_gate(type, nin, v[t][0], v[t][1], v[t][2], v[t][3]);
          if ((info[t] >> 8) & 1u) J.good[int64_t(s_node[i].u) * RW + w] = r;
        }
        s_val[int64_t(i) * 64 + lane] = r;
      }
    }
    __syncthreads();
  }
}

__global__ void k_gather_round(const uint64_t *__restrict__ good, int64_t RW, int64_t w, int N, uint64_t mask,
                               uint64_t *__restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < N) out[u] = good[int64_t(u) * RW + w] & mask;
}

// ---------------------------------------------------------------- sweeps

struct SweepArgs {
  const JobDev *__restrict__ jobs;
  const Tile *__restrict__ tiles;
  int ntiles;
  int lvl_entries;   // LDS per-level counters (max level over jobs + 1)
  const Lane *__restrict__ lanes;
  uint64_t *rows;    // per slot: [NLmax][64] faulty words
  uint64_t *masks;   // per slot: [NLmax] lane masks (plain form)
  int64_t rows_stride, masks_stride;
  uint64_t *out;     // lane results
  int *queue;        // event form: next tile
  unsigned long long *bytes;  // operand bytes touched
  MaskRec *mrec;     // per slot: [NLmax] (event form)
  uint32_t *pend;    // per slot: [NLmax], == epoch when the node must be evaluated (one-node-per-wave form)
  int64_t rec_stride, pend_stride;
  uint32_t epoch_base;
  // Optional tile statistics (event form), null unless traced:
  // [0] tiles, [1] levels with pending nodes, [2] level scan passes,
  // [3] staged nodes, [4] nodes written (some lane differs), [5] max nodes
  // written by one tile, [6] levels walked, [7] (node, lane) evaluations.
  unsigned long long *tstats;
};

struct LaneSetup {
  unsigned long long mask;
  int64_t w;
  int site;
  int out;
  uint32_t kind;
  int level;
  int active;
};

__device__ __forceinline__ LaneSetup lane_setup(const SweepArgs &a, const Tile &T, int lane) {
  LaneSetup s;
  s.active = T.nlanes;
  if (lane < T.nlanes) {
    const Lane L = a.lanes[T.lane0 + lane];
    s.mask = L.mask;
    s.w = L.word;
    s.site = L.site;
    s.out = L.out;
    s.kind = L.kind;
    s.level = L.level;
  } else {
    const Lane L = a.lanes[T.lane0];  // idle lane: a valid word, never injected
    s.mask = 0;
    s.w = L.word;
    s.site = -1;
    s.out = -1;
    s.kind = K_SA0;
    s.level = -1;
  }
  return s;
}

__device__ __forceinline__ uint64_t inject(const LaneSetup &s, uint64_t gu) {
  return s.kind == K_FLIP ? (gu ^ s.mask) : (s.kind == K_SA1 ? ~uint64_t(0) : uint64_t(0));
}

// Plain form, one (tile, node) pair per wave. `base` is the first live
// position of the tile's start level: positions below it lie outside every
// injection's fanout and hold good values in all lanes. Every node at or after
// `base` is visited and writes its lane mask, so masks need no clearing.
__device__ __forceinline__ void sweep_node(const SweepView &J, int u, int base, int lane, const LaneSetup &s,
                                           uint64_t *rows, uint64_t *masks, uint64_t &acc, uint64_t &bytes) {
  u = __builtin_amdgcn_readfirstlane(u);
  const uint32_t inf = J.info[u];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = to_int4(J.fin[u]);
  uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  if (type != G_PI) {
    if (f.x >= base) m0 = masks[f.x];
    if (nin > 1 && f.y >= base) m1 = masks[f.y];
    if (nin > 2 && f.z >= base) m2 = masks[f.z];
    if (nin > 3 && f.w >= base) m3 = masks[f.w];
  }
  const bool inj = s.site == u;
  const uint64_t anyinj = __ballot(inj);
  bytes += 28 + 8 * nin;
  if (((m0 | m1 | m2 | m3) == 0) && anyinj == 0) {
    if (lane == 0) masks[u] = 0;
    return;
  }
  const int64_t RW = J.RW;
  const auto good = J.good;
  // Lanes without a differing fanin and without an injection keep the good
  // value and load nothing.
  const bool act = (((m0 | m1 | m2 | m3) >> lane) & 1) || inj;
  const uint64_t gu = act ? good[int64_t(u) * RW + s.w] : 0;
  uint64_t r = gu;
  if (act && type != G_PI) {
    const uint64_t v0 = ((m0 >> lane) & 1) ? rows[int64_t(f.x) * 64 + lane] : good[int64_t(f.x) * RW + s.w];
    uint64_t v1 = 0, v2 = 0, v3 = 0;
    if (nin > 1) v1 = ((m1 >> lane) & 1) ? rows[int64_t(f.y) * 64 + lane] : good[int64_t(f.y) * RW + s.w];
    if (nin > 2) v2 = ((m2 >> lane) & 1) ? rows[int64_t(f.z) * 64 + lane] : good[int64_t(f.z) * RW + s.w];
    if (nin > 3) v3 = ((m3 >> lane) & 1) ? rows[int64_t(f.w) * 64 + lane] : good[int64_t(f.w) * RW + s.w];
    r = eval_gate(type, nin, v0, v1, v2, v3);
  }
  if (inj) r = inject(s, gu);
  const uint64_t d = r ^ gu;
  const uint64_t nm = __ballot(d != 0);
  if (d) rows[int64_t(u) * 64 + lane] = r;
  if (lane == 0) masks[u] = nm;
  if (type == G_PO) acc |= d;
  bytes += uint64_t(__popcll(__ballot(act))) * 8 * (type == G_PI ? 1 : nin + 1) + 8 * uint64_t(__popcll(nm));
}

// Plain form: one launch per level; a wave handles one (tile, node) pair. All
// tiles of the launch belong to one job.
__global__ void k_sweep_level(SweepArgs a, int level, int tile0, int nbatch) {
  const int lane = threadIdx.x & 63;
  const int64_t gw = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const SweepView J = sweep_view(a.jobs, a.tiles[tile0].job);
  const int lo = J.live_lvl[level], hi = J.live_lvl[level + 1];
  const int n = hi - lo;
  if (n <= 0) return;
  const int tb = int(gw / n);
  if (tb >= nbatch) return;
  const int u = lo + int(gw - int64_t(tb) * n);
  const Tile T = a.tiles[tile0 + tb];
  if (level < T.start_level) return;
  const LaneSetup s = lane_setup(a, T, lane);
  uint64_t *rows = a.rows + int64_t(tb) * a.rows_stride;
  uint64_t *masks = a.masks + int64_t(tb) * a.masks_stride;
  uint64_t acc = 0, bytes = 0;
  sweep_node(J, u, J.live_lvl[T.start_level], lane, s, rows, masks, acc, bytes);
  if (acc && lane < s.active) atomicOr((unsigned long long *)&a.out[s.out], (unsigned long long)acc);
  if (lane == 0) atomicAdd(a.bytes, (unsigned long long)bytes);
}

// Event form. A level's pending nodes are staged in LDS by the scan that finds
// them: node record, live-fanout range and the lane masks of its fanins (the
// fanins sit on lower levels, so their masks are final). A MaskRec stamped
// with another epoch means "no lane differs".
struct NodeStage {
  uint64_t m[4];  // fanin lane masks
  int4 fin;
  int u;
  uint32_t info;
  int fo0, fo1;   // live fanouts fo[fo0, fo1)
};

__device__ __forceinline__ uint64_t stamped_mask(const MaskRec *mrec, int f, uint32_t epoch) {
  const MaskRec r = mrec[f];
  return r.epoch == epoch ? r.mask : 0;
}

__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  return (uint64_t(hi) << 32) | lo;
}

// Evaluation of up to NB staged nodes for the 64 lanes of one tile, with
// their fanin loads interleaved. When some lane differs, the node's row and
// mask record are written and its live fanouts marked pending (counted per
// level in LDS); the deepest pending level is raised.
template <int NB, bool kStats>
__device__ __forceinline__ void event_nodes(const SweepView &J, const NodeStage *st, int cnt, int lane,
                                            const LaneSetup &s, uint32_t epoch, uint64_t *rows, MaskRec *mrec,
                                            uint32_t *pend, uint64_t &acc, uint64_t &bytes, int *s_maxl,
                                            int *s_lvlcnt, uint32_t &n_written, uint64_t &n_active) {
  int u[NB];
  uint32_t type[NB], nin[NB];
  int4 f[NB];
  uint64_t m[NB][4], gu[NB], r[NB];
  bool act[NB];
  const int64_t RW = J.RW;
  const auto good = J.good;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const NodeStage &n = st[k < cnt ? k : 0];
    u[k] = __builtin_amdgcn_readfirstlane(n.u);
    const uint32_t inf = __builtin_amdgcn_readfirstlane(n.info);
    type[k] = inf & 15u;
    nin[k] = (inf >> 4) & 7u;
    f[k].x = __builtin_amdgcn_readfirstlane(n.fin.x);
    f[k].y = __builtin_amdgcn_readfirstlane(n.fin.y);
    f[k].z = __builtin_amdgcn_readfirstlane(n.fin.z);
    f[k].w = __builtin_amdgcn_readfirstlane(n.fin.w);
#pragma unroll
    for (int q = 0; q < 4; ++q) m[k][q] = uniform64(n.m[q]);
    // A lane takes part iff one of its fanins differs or it injects here;
    // in every other lane the node keeps its good value and loads nothing.
    const uint64_t any = m[k][0] | m[k][1] | m[k][2] | m[k][3];
    act[k] = k < cnt && (((any >> lane) & 1) || s.site == u[k]);
    gu[k] = act[k] ? good[int64_t(u[k]) * RW + s.w] : 0;
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    r[k] = gu[k];
    if (act[k] && type[k] != G_PI) {
      const uint32_t n = nin[k];
      const uint64_t v0 =
          ((m[k][0] >> lane) & 1) ? rows[int64_t(f[k].x) * 64 + lane] : good[int64_t(f[k].x) * RW + s.w];
      uint64_t v1 = 0, v2 = 0, v3 = 0;
      if (n > 1)
        v1 = ((m[k][1] >> lane) & 1) ? rows[int64_t(f[k].y) * 64 + lane] : good[int64_t(f[k].y) * RW + s.w];
      if (n > 2)
        v2 = ((m[k][2] >> lane) & 1) ? rows[int64_t(f[k].z) * 64 + lane] : good[int64_t(f[k].z) * RW + s.w];
      if (n > 3)
        v3 = ((m[k][3] >> lane) & 1) ? rows[int64_t(f[k].w) * 64 + lane] : good[int64_t(f[k].w) * RW + s.w];
      r[k] = eval_gate(type[k], n, v0, v1, v2, v3);
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k >= cnt) break;
    if (s.site == u[k]) r[k] = inject(s, gu[k]);
    const uint64_t d = r[k] ^ gu[k];
    const uint64_t nm = __ballot(d != 0);
    const uint64_t na = __ballot(act[k]);
    if (kStats) n_active += __popcll(na);
    bytes += uint64_t(__popcll(na)) * 8 * (type[k] == G_PI ? 1 : nin[k] + 1);
    if (nm) {
      if (kStats) ++n_written;
      if (d) rows[int64_t(u[k]) * 64 + lane] = r[k];
      if (lane == 0) {
        MaskRec rec;
        rec.mask = nm;
        rec.epoch = epoch;
        rec.pad = 0;
        mrec[u[k]] = rec;
      }
      const int e0 = __builtin_amdgcn_readfirstlane(st[k].fo0), e1 = __builtin_amdgcn_readfirstlane(st[k].fo1);
      for (int e = e0 + lane; e < e1; e += 64) {
        const int2 v = to_int2(J.fo[e]);
        pend[v.x] = epoch;
        atomicAdd(&s_lvlcnt[fo_level(v)], 1);
      }
      if (e1 > e0 && lane == 0) atomicMax(s_maxl, fo_level(to_int2(J.fo[e1 - 1])));
      bytes += 8 * uint64_t(__popcll(nm)) + 16 + 12 * uint64_t(e1 - e0);
    }
    if (type[k] == G_PO) acc |= d;
  }
}

// Event form, persistent: a workgroup pulls tiles (of any job) from a queue.
// For each tile it walks the levels from the lowest injection, skipping levels
// with no pending node (per-level counts in LDS). At a level with pending
// nodes, one thread per node checks the pending flag; those that find one
// append the node to the LDS stage with its record and fanin masks; then the
// waves evaluate the staged nodes, two per wave at a time. The walk ends after
// the deepest level any pending node sits on. Pending flags and mask records
// carry the tile's epoch, so nothing in global memory is cleared between
// tiles; each level's LDS count is zeroed once the level is done. Every
// workgroup leaves once the queue passes the last tile.
constexpr int kChunk = 512;
constexpr int kNodesPerWave = 2;
template <bool kStats>
__global__ void __launch_bounds__(kChunk) __attribute__((amdgpu_waves_per_eu(6))) k_sweep_events(SweepArgs a) {
  __shared__ unsigned long long acc_lds[64];
  __shared__ int s_tile, s_cnt, s_maxl;
  __shared__ NodeStage s_stage[kChunk];
  extern __shared__ int s_lvlcnt[];  // a.lvl_entries entries
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6, nwaves = blockDim.x >> 6;
  uint64_t *rows = a.rows + int64_t(blockIdx.x) * a.rows_stride;
  MaskRec *mrec = a.mrec + int64_t(blockIdx.x) * a.rec_stride;
  uint32_t *pend = a.pend + int64_t(blockIdx.x) * a.pend_stride;
  for (int i = tid; i < a.lvl_entries; i += blockDim.x) s_lvlcnt[i] = 0;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(a.queue, 1);
    if (tid < 64) acc_lds[tid] = 0;
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_tile);
    if (t >= a.ntiles) return;
    const Tile T = a.tiles[t];
    const SweepView J = sweep_view(a.jobs, __builtin_amdgcn_readfirstlane(T.job));
    const LaneSetup s = lane_setup(a, T, lane);
    const uint32_t epoch = a.epoch_base + uint32_t(t) + 1u;
    if (tid < 64 && s.site >= 0) {
      pend[s.site] = epoch;
      atomicAdd(&s_lvlcnt[s.level], 1);
    }
    if (tid == 0) s_maxl = T.end_level;
    uint64_t acc = 0, bytes = 0, stage_bytes = 0;  // bytes: wave-uniform; stage_bytes: per lane
    uint32_t n_lv = 0, n_pass = 0, n_staged = 0, n_written = 0;  // tile statistics
    uint64_t n_active = 0;
    __syncthreads();
    int l_end = T.start_level;
    for (int l = T.start_level; l <= s_maxl; ++l) {
      l_end = l;
      if (s_lvlcnt[l] == 0) continue;
      ++n_lv;
      const int lo = J.live_lvl[l], hi = J.live_lvl[l + 1];
      for (int c0 = lo; c0 < hi; c0 += blockDim.x) {
        ++n_pass;
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        const int u = c0 + tid;
        const bool p = u < hi && pend[u] == epoch;
        const uint64_t b = __ballot(p);
        bytes += 4 * uint64_t(__popcll(__ballot(u < hi)));
        if (b) {
          int base = 0;
          if (lane == 0) base = atomicAdd(&s_cnt, __popcll(b));
          base = __shfl(base, 0);
          if (p) {
            NodeStage &n = s_stage[base + __popcll(b & ((1ull << lane) - 1ull))];
            const uint32_t inf = J.info[u];
            const int4 f = to_int4(J.fin[u]);
            const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
            n.u = u;
            n.info = inf;
            n.fin = f;
            n.fo0 = J.fo_off[u];
            n.fo1 = J.fo_off[u + 1];
            const bool gate = type != G_PI;
            n.m[0] = gate ? stamped_mask(mrec, f.x, epoch) : 0;
            n.m[1] = gate && nin > 1 ? stamped_mask(mrec, f.y, epoch) : 0;
            n.m[2] = gate && nin > 2 ? stamped_mask(mrec, f.z, epoch) : 0;
            n.m[3] = gate && nin > 3 ? stamped_mask(mrec, f.w, epoch) : 0;
            stage_bytes += 28 + (gate ? 16 * nin : 0);
          }
        }
        __syncthreads();
        const int n = s_cnt;
        n_staged += n;
        for (int i = wave * kNodesPerWave; i < n; i += nwaves * kNodesPerWave)
          event_nodes<kNodesPerWave, kStats>(J, &s_stage[i], min(kNodesPerWave, n - i), lane, s, epoch, rows, mrec,
                                             pend, acc, bytes, &s_maxl, s_lvlcnt, n_written, n_active);
        __syncthreads();
      }
      if (tid == 0) s_lvlcnt[l] = 0;
    }
    if (kStats) {
      // n_written is per wave: sum it over the workgroup through LDS.
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      if (lane == 0) atomicAdd(&s_cnt, int(n_written));
      __syncthreads();
      if (tid == 0) {
        atomicAdd(&a.tstats[0], 1ull);
        atomicAdd(&a.tstats[1], (unsigned long long)n_lv);
        atomicAdd(&a.tstats[2], (unsigned long long)n_pass);
        atomicAdd(&a.tstats[3], (unsigned long long)n_staged);
        atomicAdd(&a.tstats[4], (unsigned long long)s_cnt);
        atomicMax(&a.tstats[5], (unsigned long long)s_cnt);
        atomicAdd(&a.tstats[6], (unsigned long long)(l_end - T.start_level + 1));
      }
      if (lane == 0) atomicAdd(&a.tstats[7], (unsigned long long)n_active);
    }
    for (int o = 32; o > 0; o >>= 1) stage_bytes += __shfl_xor(stage_bytes, o);
    bytes += stage_bytes;
    if (lane == 0 && bytes) atomicAdd(a.bytes, (unsigned long long)bytes);
    if (acc) atomicOr(&acc_lds[lane], (unsigned long long)acc);
    __syncthreads();
    if (tid < s.active) a.out[s.out] = acc_lds[tid];
    __syncthreads();
  }
}

// Event form with flattened work. Per pass over a level chunk the workgroup
//  A. scans the chunk's pending flags, loading each node's static data with
//     its flag, and stages each pending node (fanin masks, active lanes =
//     lanes with a differing fanin or an injection, fanout range), numbering
//     nodes and their (node, active lane) pairs in one LDS counter;
//  B. evaluates every (node, active lane) pair, two per thread: lanes where
//     no fanin differs keep the good value and are never touched;
//  C. for the nodes that differ somewhere, writes the mask record;
//  D. marks their live fanouts pending, each wave those of its own nodes.
// A level pass thus costs four dependent memory round trips (flags, fanin
// records, words, fanout lists) and three barriers, however many nodes and
// lanes it has. A node's injecting lanes are found by a binary search
// over the tile's lane sites in LDS. Pending flags and mask records carry
// the tile's epoch, so nothing is cleared between tiles.
constexpr int kTileStats = 13;  // sweep statistics counters (LSIM_TRACE)
constexpr int kPairsPerThread = 2;
constexpr int kFanoutsPerLane = 2;  // pair sweep: fanouts per lane marked at a time
constexpr int kDenseNodes = 3;      // pair sweep dense passes: nodes per wave in flight
constexpr int kDenseNodesOcc = 4;   // the same at more than four waves per SIMD
constexpr int kPairsPerThreadOcc = 2;

// Set bits of m below the calling lane (popcount of m & (2^lane - 1)).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

// An empty use that makes every path wait here for the load that produced x.
// The vector memory counter is in order over loads and stores: a load left
// pending on some path (its value unused there) makes the compiler wait for
// it before the next write of its register, at the head of the next loop
// iteration, and that wait also waits for every store issued in between.
// Settled after their batch of loads, the registers are free again and the
// next batch of loads goes out while the stores are still in flight.
__device__ __forceinline__ void settle(uint64_t x) { asm volatile("" ::"v"(x)); }

// Marks the fanout of live-fanout entry f pending on the pins that read the
// written node: a byte store of the tile's tag per pin (pins != 0).
__device__ __forceinline__ void mark_pins(uint32_t *pend, int2 f, uint32_t tag) {
  LSIM_GLOBAL uint8_t *b = (LSIM_GLOBAL uint8_t *)(pend + f.x);
  uint32_t pins = fo_pins(f);
  do {
    b[__builtin_ctz(pins)] = uint8_t(tag);
    pins &= pins - 1u;
  } while (pins);
}

// Position of the r-th set bit (r < popcount) of m.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
  int pos = 0;
#pragma unroll
  for (int half = 32; half > 0; half >>= 1) {
    const uint64_t low = m & ((uint64_t(1) << half) - 1);
    const int c = __popcll(low);
    if (r >= c) {
      r -= c;
      m >>= half;
      pos += half;
    } else {
      m = low;
    }
  }
  return pos;
}

// Largest k in [0, n) with pref[k] <= q (pref ascending, pref[0] == 0 <= q).
__device__ __forceinline__ int find_range(const int *pref, int n, int q) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pref[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Wave-wide exclusive prefix of a packed (count << 32 | amount) value; the
// last lane adds the wave's total to *counter and the base is added back.
__device__ __forceinline__ uint64_t wave_claim(uint64_t x, int lane, unsigned long long *counter) {
  uint64_t inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  uint64_t base = 0;
  if (lane == 63 && inc) base = atomicAdd(counter, (unsigned long long)inc);
  base = __shfl(base, 63);
  return base + inc - x;
}

// Lanes of the tile that inject at position u. Lane sites are non-decreasing
// over the tile (idle lanes hold INT_MAX), so they form one run.
__device__ __forceinline__ uint64_t inj_lanes(const int *s_lsite, int u) {
  int lo = 0, hi = 64;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_lsite[mid] < u) lo = mid + 1; else hi = mid;
  }
  const int a = lo;
  hi = 64;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_lsite[mid] <= u) lo = mid + 1; else hi = mid;
  }
  const int n = lo - a;
  if (n <= 0) return 0;
  return n == 64 ? ~uint64_t(0) : (((uint64_t(1) << n) - 1) << a);
}

// Record of a node that differs from the good machine in some lane of the
// tile (pair sweep): the differing lanes, the evaluated (active) lanes, the
// tile's epoch and the node's first pool word. The node's faulty word of
// active lane ln is pool[base + popcount(act & (2^ln - 1))]; it is written
// only where the lane differs, which is all a fanout reads.
struct alignas(32) PairRec {  // 32 B: a record never straddles a memory sector
  unsigned long long nm;
  unsigned long long act;
  uint32_t epoch;
  uint32_t base;
  unsigned long long pad;
};

// Sweep state of a step planned on the device (written by k_ds_tiles).
struct SweepDyn {
  int ntiles;           // tiles of the step
  int nopend;           // sweep without pending flags
  int form;             // workgroup size of the pair sweep that runs the step (256 or 512)
  uint32_t epoch_base;  // the step's tiles take epochs epoch_base + 1 ...
  double share;         // staged share of the nodes walked by the last sweep that walked any
};

struct PairArgs {
  const JobDev *__restrict__ jobs;
  const Tile *__restrict__ tiles;
  const int *__restrict__ tile_list;  // tiles to sweep (queue index -> tile), or null for all
  int ntiles;                         // entries of the queue
  int lvl_entries;
  const Lane *__restrict__ lanes;
  uint64_t *out;
  int *queue;
  unsigned long long *bytes;
  PairRec *rec;      // per slot [rec_stride]
  uint32_t *pend;    // per slot [rec_stride], == epoch when pending
  int64_t rec_stride;
  uint64_t *pool;    // per slot [pool_words]
  int64_t pool_words;
  int *ovf_count;    // tiles that outgrew their pool, listed in ovf_list
  int *ovf_list;
  uint32_t epoch_base;
  int dense_lanes;             // a pass whose nodes average this many active lanes goes one node per wave (0: never)
  int nopend;                  // no pending flags: every node of a walked level is a candidate
  int fo_pre;                  // with pending flags: a staged node's first live fanout is loaded in the scan
  unsigned long long *tstats;  // as SweepArgs::tstats, or null
  // A step planned on the device: the tile count, pending-flag mode,
  // workgroup size and epochs come from *dyn (the launch of the other
  // workgroup size returns at once). With ntiles_p set, the launch sweeps the
  // tiles the first launch gave up: their count is *ntiles_p and their epochs
  // follow the first launch's.
  const SweepDyn *dyn;
  const int *ntiles_p;
};

// NIN is the most fanins of any gate of the launch's circuits (2, or 4 as the
// format allows): circuits of 1- and 2-input gates stage half the fanin data,
// which leaves LDS and registers for more tiles in flight per CU (WAVES is
// the waves per SIMD the kernel is compiled for).
template <bool kStats, int CHUNK, int NIN, int WAVES>
__global__ void __launch_bounds__(CHUNK) __attribute__((amdgpu_waves_per_eu(WAVES))) k_sweep_pairs(PairArgs a) {
  static_assert(NIN == 2 || NIN == 4, "fanin slots per node: 2 or 4");
  // Nodes per wave in flight in dense passes and pairs per thread in
  // flattened passes, with constants of their own for more than four waves
  // per SIMD (where the registers are scarcer).
  constexpr int DN = WAVES > 4 ? kDenseNodesOcc : (NIN == 2 ? 4 : kDenseNodes);
  constexpr int PPT = WAVES > 4 ? kPairsPerThreadOcc : kPairsPerThread;
  // Loads of pair passes without a branch around them (measured faster for
  // the 512-thread narrow form, slower for the five-wave 256-thread one).
  constexpr bool kUncond = NIN == 2 && WAVES == 4;
  __shared__ unsigned long long s_acc[64], s_lmask[64];
  __shared__ int s_lsite[64], s_lword[64], s_lkind[64];
  __shared__ int s_u[CHUNK];
  __shared__ uint32_t s_info[CHUNK];
  __shared__ int s_fin[CHUNK][NIN];
  __shared__ unsigned long long s_m[CHUNK][NIN], s_fa[CHUNK][NIN];  // fanin differing / active lanes
  __shared__ uint32_t s_fb[CHUNK][NIN];                              // fanin pool bases
  __shared__ unsigned long long s_act[CHUNK], s_nm[CHUNK];
  __shared__ int s_pref[CHUNK];
  __shared__ int s_fo0[CHUNK], s_deg[CHUNK];  // live fanout range
  __shared__ int2 s_fop[CHUNK];               // first live fanout {position, level} (fo_pre)
  __shared__ int s_tile, s_maxl;
  __shared__ uint32_t s_pool;  // next free word of the slot's pool for this tile
  __shared__ unsigned long long s_cnt2;  // nodes << 32 | pairs
  __shared__ unsigned long long s_nwt;   // statistics: nodes the tile wrote
  extern __shared__ uint8_t s_lvl[];  // per level: some node of the level is pending
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  int ntiles = a.ntiles, nopend = a.nopend;
  uint32_t epoch_base = a.epoch_base;
  if (a.dyn) {
    const SweepDyn D = *a.dyn;
    if (a.ntiles_p) {
      ntiles = *a.ntiles_p;
      epoch_base = D.epoch_base + uint32_t(D.ntiles);
    } else {
      if (D.form != CHUNK) return;
      ntiles = D.ntiles;
      epoch_base = D.epoch_base;
    }
    nopend = D.nopend;
  }
  PairRec *rec = a.rec + int64_t(blockIdx.x) * a.rec_stride;
  uint32_t *pend = a.pend + int64_t(blockIdx.x) * a.rec_stride;
  LSIM_GLOBAL uint64_t *pool = (LSIM_GLOBAL uint64_t *)(a.pool + int64_t(blockIdx.x) * a.pool_words);
  for (int i = tid; i < a.lvl_entries; i += CHUNK) s_lvl[i] = 0;
  if (tid == 0) {
    s_cnt2 = 0;
    s_nwt = 0;
  }
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(s_tile);
    if (q >= ntiles) return;
    const int t = a.tile_list ? a.tile_list[q] : q;
    const Tile T = a.tiles[t];
    const SweepView J = sweep_view(a.jobs, __builtin_amdgcn_readfirstlane(T.job));
    const uint32_t epoch = epoch_base + uint32_t(q) + 1u;
    // Pending marks are per input pin: byte k of pend[v] holds the tag of the
    // last tile whose walk wrote v's fanin on pin k (tags 1..255, so cleared
    // memory never matches). An older tile's tag can match by chance; the
    // fanin record's full epoch decides, so that costs one record read.
    const uint32_t tag = epoch % 255u + 1u, tag4 = tag * 0x01010101u;
    const int64_t RW = J.RW;
    const uint32_t RW32 = uint32_t(RW);  // words per node (< 2^32): 32 x 32-bit products
    const auto good = J.good;
    uint32_t bytes = 0;  // per thread and tile
    uint32_t n_lv = 0, n_pass = 0, n_staged = 0, n_written = 0, n_pairs = 0;
    if (tid < 64) {
      const bool valid = lane < T.nlanes;
      const Lane L = a.lanes[T.lane0 + (valid ? lane : 0)];
      const int site = valid ? L.site : INT_MAX;
      s_lsite[lane] = site;
      s_lword[lane] = L.word;
      s_lkind[lane] = int(L.kind);
      s_lmask[lane] = L.mask;
      s_acc[lane] = 0;
      if (valid) {
        // An injection site is pending on every pin: all its fanins' records
        // are read.
        if (!nopend) pend[site] = tag4;
        s_lvl[L.level] = 1;
      }
    }
    if (tid == 0) {
      s_maxl = T.end_level;
      s_pool = 0;
    }
    __syncthreads();
    int l_end = T.start_level;
    bool aborted = false;
    // Statistics: cycles of the tile's walk and of each pass phase (thread 0,
    // which waits at every barrier): A scan and stage, B pairs (flattened or
    // dense), C-D records and fanout marking.
    uint64_t cy0 = kStats ? clock64() : 0, cy_a = 0, cy_bp = 0, cy_bd = 0, cy_cd = 0, cy_t = 0;
    bool dense_pass = false;
    for (int l = T.start_level; l <= s_maxl && !aborted; ++l) {
      l_end = l;
      if (!nopend && s_lvl[l] == 0) continue;
      ++n_lv;
      const int2 lr = level_range(J.live_lvl, l);
      const int lo = lr.x, hi = lr.y;
      for (int c0 = lo; c0 < hi; c0 += CHUNK) {
        ++n_pass;
        if (kStats) cy_t = clock64();
        // A. scan and stage. The node's static data (type, fanins, fanout
        // range) is loaded together with its pending marks, so a pending node
        // costs two dependent loads: these, then the records of the fanins on
        // its marked pins. With fo_pre its first live fanout is loaded beside
        // the records, so that a written node with one live fanout marks it
        // without a further load.
        const bool pre = a.fo_pre && !nopend;
        const int u = c0 + tid;
        const bool in = u < hi;
        // Threads past the level's end load node lo's data and ignore it, so
        // the loads are unconditional.
        const int uc = in ? u : lo;
        const uint32_t pf = nopend ? 0u : pend[uc];
        uint32_t inf;
        int4 f;
        int fo0, fo1;
        if constexpr (NIN == 2) {
          // Narrow circuits: the node's record in one 16-byte load (the next
          // record's fanout offset ends its range, mostly in the same lines).
          const gi4 r = J.nrec[uc];
          inf = uint32_t(r.x);
          f = make_int4(r.y, r.z, 0, 0);
          fo0 = r.w;
          fo1 = J.nrec[uc + 1].w;
        } else {
          inf = J.info[uc];
          f = to_int4(J.fin[uc]);
          fo0 = J.fo_off[uc];
          fo1 = J.fo_off[uc + 1];
        }
        // Pins whose mark carries this tile's tag; without pending flags every
        // node of the level is a candidate, all its fanins' records are read,
        // and it is staged when some fanin differs in some lane or it injects.
        uint32_t rd = 0xfu;
        if (!nopend) {
          const uint32_t x = pf ^ tag4;
          rd = ((x & 0xffu) == 0 ? 1u : 0u) | ((x & 0xff00u) == 0 ? 2u : 0u) | ((x & 0xff0000u) == 0 ? 4u : 0u) |
               ((x & 0xff000000u) == 0 ? 8u : 0u);
        }
        const bool cand = in && rd != 0;
        uint64_t act = 0;
        uint64_t m[NIN], fa[NIN];
        uint32_t fb[NIN];
#pragma unroll
        for (int k = 0; k < NIN; ++k) {
          m[k] = fa[k] = 0;
          fb[k] = 0;
        }
        int2 fp = make_int2(0, 0);
        if (cand) {
          if (pre && fo1 > fo0) {
            fp = to_int2(J.fo[fo0]);
            bytes += 8;
          }
          const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
          // The records of the fanins on the read pins are loaded together,
          // unconditionally for the first two pins (record 0 stands in for a
          // missing or unread fanin: one line for the whole wave), and masked
          // afterwards by fanin count, pin and epoch, so a node's records cost
          // one round trip. (Under a per-fanin or per-epoch branch the
          // compiler loads them one fanin after another.)
          const int nf = type != G_PI ? int(nin) : 0;
          const int fi[4] = {f.x, f.y, f.z, f.w};
          gu2 ma[NIN];
          gi2 eb[NIN];
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            const PairRec *rp = rec + (k < nf && ((rd >> k) & 1u) ? fi[k] : 0);
            ma[k] = *reinterpret_cast<const gu2 *>(rp);
            eb[k] = *reinterpret_cast<const gi2 *>(&rp->epoch);
          }
          if constexpr (NIN > 2) {
            ma[2] = ma[3] = gu2{0, 0};
            eb[2] = eb[3] = gi2{0, 0};
            if (nf > 2) {
#pragma unroll
              for (int k = 2; k < 4; ++k) {
                const PairRec *rp = rec + (k < nf && ((rd >> k) & 1u) ? fi[k] : 0);
                ma[k] = *reinterpret_cast<const gu2 *>(rp);
                eb[k] = *reinterpret_cast<const gi2 *>(&rp->epoch);
              }
            }
          }
#pragma unroll
          for (int k = 0; k < NIN; ++k) {
            const uint64_t keep =
                (k < nf && ((rd >> k) & 1u) && uint32_t(eb[k].x) == epoch) ? ~uint64_t(0) : uint64_t(0);
            m[k] = ma[k].x & keep;
            fa[k] = ma[k].y & keep;
            fb[k] = uint32_t(eb[k].y) & uint32_t(keep);
            act |= m[k];
          }
          act |= inj_lanes(s_lsite, u);
          bytes += 20 + (type != G_PI ? 24 * uint32_t(__popc(rd & ((1u << nin) - 1u))) : 0);
        }
        if (!nopend) bytes += 4;
        // A pending node always has a differing fanin or injects (act != 0);
        // a mark matched only by an older tile's tag leaves act == 0.
        const bool p = cand && act != 0;
        const uint64_t slot = wave_claim(p ? ((uint64_t(1) << 32) | uint64_t(__popcll(act))) : 0, lane, &s_cnt2);
        if (p) {
          const int k = int(slot >> 32);
          s_u[k] = u;
          s_info[k] = inf;
          const int fj[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
          for (int j = 0; j < NIN; ++j) {
            s_fin[k][j] = fj[j];
            s_m[k][j] = m[j];
            s_fa[k][j] = fa[j];
            s_fb[k][j] = fb[j];
          }
          s_act[k] = act;
          s_nm[k] = 0;
          s_pref[k] = int(uint32_t(slot));
          s_fo0[k] = fo0;
          s_deg[k] = fo1 - fo0;
          s_fop[k] = fp;
        }
        __syncthreads();
        if (kStats) {
          const uint64_t c = clock64();
          cy_a += c - cy_t;
          cy_t = c;
        }
        const uint64_t c2 = s_cnt2;
        const int n = int(c2 >> 32), P = int(uint32_t(c2));
        const uint32_t pb = s_pool;
        if (uint64_t(pb) + uint64_t(P) > uint64_t(a.pool_words)) {
          // The pass would outgrow the slot's pool: the tile is given up and
          // swept again in a slot that holds any tile (same value for every
          // thread, so the whole workgroup leaves together).
          aborted = true;
          break;
        }
        dense_pass = a.dense_lanes > 0 && int64_t(P) >= int64_t(a.dense_lanes) * n;
        if (dense_pass) {
          // B'. Dense pass: one node per wave at a time, DN in flight, the
          // wave's lanes being the tile's lanes. A lane's pool word within a
          // node is the count of the node's active lanes below it (mbcnt), so
          // no pair has to be located by search.
          const int wave = tid >> 6;
          const int64_t w_l = s_lword[lane];
          const int site_l = s_lsite[lane];
          const uint32_t kind_l = uint32_t(s_lkind[lane]);
          const uint64_t mask_l = s_lmask[lane];
          // Node k (position v) of the pass evaluated to r in this lane, gu
          // being its good word: the lane's injection replaces r at the
          // lane's site; the differing lanes are written, packed.
          auto dense_out = [&](int k, int v, uint32_t type, uint32_t nin, bool on, uint64_t gu, uint64_t r) {
            if (on && site_l == v)
              r = kind_l == K_FLIP ? (gu ^ mask_l) : (kind_l == K_SA1 ? ~uint64_t(0) : uint64_t(0));
            const uint64_t d = on ? (r ^ gu) : 0;
            const uint64_t nm = __ballot(d != 0);
            if (on) {
              bytes += 8 * (type == G_PI ? 1 : nin + 1);
              if (kStats) ++n_pairs;
            }
            if (d) {
              // Dense passes write only the differing lanes, packed: the
              // node's record then indexes its block by the differing lanes
              // (its `act` is nm), so readers touch only written words.
              pool[pb + uint32_t(s_pref[k]) + lanes_below(nm)] = r;
              if (type == G_PO) atomicOr(&s_acc[lane], (unsigned long long)d);
              bytes += 8;
            }
            if (lane == 0) s_nm[k] = nm;
          };
#pragma unroll 1
          for (int k0 = wave; k0 < n; k0 += DN * (CHUNK / 64)) {
            int kk[DN], v[DN];
            bool has[DN];
            uint32_t type[DN], nin[DN];
            bool wide = false;
#pragma unroll
            for (int i = 0; i < DN; ++i) {
              const int ki = k0 + i * (CHUNK / 64);
              has[i] = ki < n;
              kk[i] = has[i] ? ki : k0;
              v[i] = __builtin_amdgcn_readfirstlane(s_u[kk[i]]);
              const uint32_t vi = __builtin_amdgcn_readfirstlane(s_info[kk[i]]);
              type[i] = vi & 15u;
              nin[i] = (vi >> 4) & 7u;
              if (NIN > 2 && has[i] && nin[i] > 2) wide = true;
            }
            if (!wide) {
              // Gates of at most two inputs (the group's choice is uniform):
              // every load of the group's nodes first, then the evaluations.
              // Every load is issued unconditionally, with no branch around it:
              // an inactive lane, a PI (its fanin slot holds the PI index) or
              // a missing fanin (slot 0) reads a valid good word and ignores
              // it, so the group's loads share one wait.
              bool on[DN];
              uint64_t gu[DN], vals[DN][2];
#pragma unroll
              for (int i = 0; i < DN; ++i) {
                on[i] = has[i] && ((s_act[kk[i]] >> lane) & 1);
                gu[i] = 0;
                if (kUncond || on[i]) gu[i] = good[uint64_t(uint32_t(v[i])) * RW32 + w_l];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                  // Both addresses first, then one load from the selected one
                  // (a load in each arm of a branch makes the compiler wait).
                  const uint32_t fj = uint32_t(__builtin_amdgcn_readfirstlane(s_fin[kk[i]][j]));
                  const auto pa = &pool[s_fb[kk[i]][j] + lanes_below(s_fa[kk[i]][j])];
                  const auto ga = &good[uint64_t(fj) * RW32 + w_l];
                  const auto src = ((s_m[kk[i]][j] >> lane) & 1) ? pa : ga;
                  vals[i][j] = 0;
                  if (kUncond || (on[i] && type[i] != G_PI && j < int(nin[i]))) vals[i][j] = *src;
                }
              }
              // All of the group's words, on every path (a group short of
              // DN nodes skips some evaluations), before its stores.
#pragma unroll
              for (int i = 0; i < DN; ++i) {
                settle(gu[i]);
                settle(vals[i][0]);
                settle(vals[i][1]);
              }
#pragma unroll
              for (int i = 0; i < DN; ++i) {
                if (!has[i]) continue;
                // The type is uniform: its truth table is computed on the
                // scalar unit and applied with bit selects, no branch.
                const uint64_t e = eval_tt(gate_tt(type[i], nin[i]), vals[i][0], vals[i][1]);
                const uint64_t r = (on[i] && type[i] != G_PI) ? e : gu[i];
                dense_out(kk[i], v[i], type[i], nin[i], on[i], gu[i], r);
              }
            } else if constexpr (NIN > 2) {
              // The group holds a gate of more than two inputs: its nodes go
              // one at a time, with four fanin slots.
              const int k_end = min(n, k0 + DN * (CHUNK / 64));
#pragma unroll 1
              for (int k1 = k0; k1 < k_end; k1 += CHUNK / 64) {
                const int v1 = __builtin_amdgcn_readfirstlane(s_u[k1]);
                const uint32_t vi = __builtin_amdgcn_readfirstlane(s_info[k1]);
                const uint32_t ty = vi & 15u, ni = (vi >> 4) & 7u;
                const bool on1 = (s_act[k1] >> lane) & 1;
                uint64_t gu1 = 0;
                if (on1) gu1 = good[int64_t(v1) * RW + w_l];
                const int fi[4] = {s_fin[k1][0], s_fin[k1][1], s_fin[k1][2], s_fin[k1][3]};
                uint64_t v4[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  const auto pa = &pool[s_fb[k1][j] + lanes_below(s_fa[k1][j])];
                  const auto ga = &good[int64_t(__builtin_amdgcn_readfirstlane(fi[j])) * RW + w_l];
                  const auto src = ((s_m[k1][j] >> lane) & 1) ? pa : ga;
                  v4[j] = 0;
                  if (on1 && ty != G_PI && j < int(ni)) v4[j] = *src;
                }
                const uint64_t e = eval_gate(ty, ni, v4[0], v4[1], v4[2], v4[3]);  // uniform type
                const uint64_t r = (on1 && ty != G_PI) ? e : gu1;
                dense_out(k1, v1, ty, ni, on1, gu1, r);
              }
            }
          }
        } else {
          // B. (node, active lane) pairs, PPT per thread at a time
          // with their loads interleaved: the pairs' good words and fanin
          // words are all loaded before any is used. Every pair writes its
          // pool word (readers use only the lanes that differ): the pass's
          // words then fill one contiguous block, written in whole lines.
  #pragma unroll 1
          for (int q0 = tid; q0 < P; q0 += PPT * CHUNK) {
            int k[PPT], ln[PPT], rk[PPT];
            uint32_t vi[PPT];
            uint64_t gu[PPT], vals[PPT][NIN];
            bool ok[PPT];
  #pragma unroll
            for (int i = 0; i < PPT; ++i) {
              const int qq = q0 + i * CHUNK;
              ok[i] = qq < P;
              const int qe = ok[i] ? qq : q0;
              k[i] = find_range(s_pref, n, qe);
              rk[i] = qe - s_pref[k[i]];
              ln[i] = nth_set_bit(s_act[k[i]], rk[i]);
              vi[i] = s_info[k[i]];
              const int w = s_lword[ln[i]];
              const int b = ln[i];
              const uint64_t below = (uint64_t(1) << b) - 1;
              // With kUncond the loads have no branch around them: a pair past
              // the end repeats pair q0, and PIs and missing fanins read a
              // valid word and ignore it.
              const uint32_t type = vi[i] & 15u, nin = (vi[i] >> 4) & 7u;
              const bool use = kUncond || ok[i];
              gu[i] = 0;
              if (use) gu[i] = good[uint64_t(uint32_t(s_u[k[i]])) * RW32 + uint32_t(w)];
  #pragma unroll
              for (int j = 0; j < NIN; ++j) {
                // Both addresses first, then one load from the selected one
                // (a load in each arm of a branch makes the compiler wait).
                const auto pa = &pool[s_fb[k[i]][j] + uint32_t(__popcll(s_fa[k[i]][j] & below))];
                const auto ga = &good[uint64_t(uint32_t(s_fin[k[i]][j])) * RW32 + uint32_t(w)];
                const auto src = ((s_m[k[i]][j] >> b) & 1) ? pa : ga;
                vals[i][j] = 0;
                if (kUncond || (ok[i] && type != G_PI && j < int(nin))) vals[i][j] = *src;
              }
            }
  #pragma unroll
            for (int i = 0; i < PPT; ++i) {
              if (!ok[i]) continue;
              const uint32_t type = vi[i] & 15u, nin = (vi[i] >> 4) & 7u;
              const int b = ln[i];
              const int v = s_u[k[i]];
              uint64_t e;
              if constexpr (NIN > 2)
                e = eval_gate_sel(type, nin, vals[i][0], vals[i][1], vals[i][2], vals[i][3]);
              else if constexpr (WAVES > 4)
                e = eval_gate_sel2(type, nin, vals[i][0], vals[i][1]);  // fewer registers at five waves
              else
                e = eval_tt(gate_tt(type, nin), vals[i][0], vals[i][1]);
              uint64_t r = type != G_PI ? e : gu[i];
              if (s_lsite[b] == v) {
                const uint32_t kind = uint32_t(s_lkind[b]);
                r = kind == K_FLIP ? (gu[i] ^ s_lmask[b]) : (kind == K_SA1 ? ~uint64_t(0) : uint64_t(0));
              }
              const uint64_t d = r ^ gu[i];
              bytes += 8 * (type == G_PI ? 1 : nin + 1);
              pool[pb + uint32_t(s_pref[k[i]] + rk[i])] = r;
              if (d) {
                atomicOr(&s_nm[k[i]], 1ull << b);
                if (type == G_PO) atomicOr(&s_acc[b], (unsigned long long)d);
                bytes += 8;
              }
              if (kStats) ++n_pairs;
            }
          }
        }
        __syncthreads();
        if (kStats) {
          const uint64_t c = clock64();
          (dense_pass ? cy_bd : cy_bp) += c - cy_t;
          cy_t = c;
        }
        // C. nodes that differ somewhere write their record; each wave then
        // marks the live fanouts of its own written nodes pending.
        if (tid == 0) s_cnt2 = 0;
        int e0 = 0, deg = 0;
        if (tid < n) {
          const uint64_t nm = s_nm[tid];
          if (nm) {
            const int v = s_u[tid];
            PairRec r;
            r.nm = nm;
            // The mask that indexes the node's pool block: its active lanes
            // (flattened passes write every active lane), or the differing
            // lanes where a dense pass packed only those.
            r.act = dense_pass ? nm : s_act[tid];
            r.epoch = epoch;
            r.base = pb + uint32_t(s_pref[tid]);
            r.pad = 0;
            rec[v] = r;
            e0 = s_fo0[tid];
            deg = s_deg[tid];
            if (deg) atomicMax(&s_maxl, int(s_info[tid] >> 8));
            bytes += 40;
            if (kStats) ++n_written;
            if (pre && deg) {
              const int2 fv = s_fop[tid];
              mark_pins(pend, fv, tag);
              s_lvl[fo_level(fv)] = 1;
              bytes += 4;
              ++e0;
              --deg;
            }
          }
        }
        int incl = nopend ? 0 : deg;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(incl, o);
          if (lane >= o) incl += y;
        }
        const int excl = incl - (nopend ? 0 : deg);
        const int Ew = __shfl(incl, 63);  // 0 without pending flags: no fanout is marked
        // D. kFanoutsPerLane fanouts per lane at a time, loads interleaved.
        // The fanout e belongs to the last lane whose prefix is <= e.
        if (tid == 0) s_pool = pb + uint32_t(P);
#pragma unroll 1
        for (int eb = 0; eb < Ew; eb += kFanoutsPerLane * 64) {
          int2 fo[kFanoutsPerLane];
          bool ok[kFanoutsPerLane];
#pragma unroll
          for (int i = 0; i < kFanoutsPerLane; ++i) {
            const int e = eb + i * 64 + lane;
            int k = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
              if (__shfl(excl, k + step) <= e) k += step;
            const int src0 = __shfl(e0, k), srcx = __shfl(excl, k);
            ok[i] = e < Ew;
            // Loaded whatever ok is (entry 0 exists: Ew > 0), so the loads of
            // the lane's fanouts are in flight together.
            fo[i] = to_int2(J.fo[ok[i] ? src0 + (e - srcx) : 0]);
          }
#pragma unroll
          for (int i = 0; i < kFanoutsPerLane; ++i)
            if (ok[i]) {
              mark_pins(pend, fo[i], tag);
              s_lvl[fo_level(fo[i])] = 1;
              bytes += 12;
            }
        }
        n_staged += n;
        __syncthreads();
        if (kStats) cy_cd += clock64() - cy_t;
      }
      if (tid == 0 && !aborted) s_lvl[l] = 0;
    }
    if (aborted) {
      __syncthreads();  // every thread has read the pass counters
      for (int i = tid; i < a.lvl_entries; i += CHUNK) s_lvl[i] = 0;
      if (tid == 0) {
        s_cnt2 = 0;
        a.ovf_list[atomicAdd(a.ovf_count, 1)] = t;
      }
    } else {
      // The index goes through an empty asm so that the LDS address is formed
      // here, once per tile, rather than kept live in a register (or a spill
      // slot) across the whole walk.
      int t_out = tid;
      asm volatile("" : "+v"(t_out));
      if (t_out < T.nlanes) a.out[a.lanes[T.lane0 + t_out].out] = s_acc[t_out];
      if (tid == 0) {
        atomicAdd(&a.bytes[1], (unsigned long long)n_staged);
        atomicAdd(&a.bytes[2], (unsigned long long)(J.live_lvl[l_end + 1] - J.live_lvl[T.start_level]));
      }
    }
    unsigned long long wbytes = bytes;
    for (int o = 32; o > 0; o >>= 1) wbytes += __shfl_xor(wbytes, o);
    if (lane == 0 && wbytes) atomicAdd(a.bytes, wbytes);
    if (kStats) {
      unsigned long long np = n_pairs, nw = n_written;
      for (int o = 32; o > 0; o >>= 1) {
        np += __shfl_xor(np, o);
        nw += __shfl_xor(nw, o);
      }
      if (lane == 0) {
        atomicAdd(&a.tstats[7], np);
        atomicAdd(&s_nwt, nw);
      }
      __syncthreads();
      if (tid == 0) {
        atomicAdd(&a.tstats[0], 1ull);
        atomicAdd(&a.tstats[1], (unsigned long long)n_lv);
        atomicAdd(&a.tstats[2], (unsigned long long)n_pass);
        atomicAdd(&a.tstats[3], (unsigned long long)n_staged);
        atomicAdd(&a.tstats[4], s_nwt);
        atomicMax(&a.tstats[5], s_nwt);
        atomicAdd(&a.tstats[6], (unsigned long long)(l_end - T.start_level + 1));
        atomicAdd(&a.tstats[8], (unsigned long long)(clock64() - cy0));
        atomicAdd(&a.tstats[9], (unsigned long long)cy_a);
        atomicAdd(&a.tstats[10], (unsigned long long)cy_bp);
        atomicAdd(&a.tstats[11], (unsigned long long)cy_bd);
        atomicAdd(&a.tstats[12], (unsigned long long)cy_cd);
        s_nwt = 0;
      }
    }
    __syncthreads();
  }
}

using PairKernel = void (*)(PairArgs);

// The pair sweep instance for a workgroup of `threads` (512 or 256).
// The pair sweep for a workgroup size, with statistics or not, for circuits
// whose gates all have at most two fanins (narrow) or any circuit.
// The 256-thread form of narrow circuits runs five waves per SIMD (five
// workgroups per CU where four fit otherwise). A 384-thread form at five waves
// per SIMD (three workgroups per CU) was slower than the 512-thread one on the
// 2M-gate workload: 576 against 503 ms of sweep per run.
PairKernel pairs_kernel(int form, bool stats, bool narrow) {
  if (narrow) {
    if (form == 256) return stats ? k_sweep_pairs<true, 256, 2, 5> : k_sweep_pairs<false, 256, 2, 5>;
    return stats ? k_sweep_pairs<true, 512, 2, 4> : k_sweep_pairs<false, 512, 2, 4>;
  }
  if (form == 256) return stats ? k_sweep_pairs<true, 256, 4, 4> : k_sweep_pairs<false, 256, 4, 4>;
  return stats ? k_sweep_pairs<true, 512, 4, 4> : k_sweep_pairs<false, 512, 4, 4>;
}

// ---------------------------------------------------------------- detection

__device__ __forceinline__ void record(const JobDev &J, int f, int64_t w, uint64_t det) {
  if (!det) return;
  if (J.first_round) atomicMin(&J.first_round[f], int32_t(w));
  if (J.det_bits) atomicOr((unsigned long long *)&J.det_bits[int64_t(f) * J.RB + (w >> 6)], 1ull << (w & 63));
  if (J.last_flag && w == J.last_word) J.last_flag[f] = 1;
}

// Flip sensitivity of gate g to its input node x at word w: the patterns where
// complementing every fanin slot fed by x changes g's output.
__device__ __forceinline__ uint64_t sens(const JobDev &J, int g, int x, int64_t w) {
  const uint32_t inf = J.info[g];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = J.fin[g];
  const auto good = J.good;
  // The second fanin slot holds node 0 on a one-input gate: loaded anyway,
  // so the two loads have no branch between them. Lanes hold different
  // faults, hence different gate types: the branch-free evaluation.
  const uint64_t v0 = good[int64_t(f.x) * J.RW + w];
  const uint64_t v1 = good[int64_t(f.y) * J.RW + w];
  const uint64_t v2 = nin > 2 ? good[int64_t(f.z) * J.RW + w] : 0;
  const uint64_t v3 = nin > 3 ? good[int64_t(f.w) * J.RW + w] : 0;
  const uint64_t r0 = eval_gate_sel(type, nin, v0, v1, v2, v3);
  const uint64_t r1 = eval_gate_sel(type, nin, f.x == x ? ~v0 : v0, f.y == x ? ~v1 : v1, f.z == x ? ~v2 : v2,
                                    f.w == x ? ~v3 : v3);
  return r0 ^ r1;
}

// L = activation & path sensitisation to the region head & live bits.
__device__ __forceinline__ uint64_t local_term(const JobDev &J, int f, int64_t w) {
  const int x0 = J.site[f];
  uint64_t D = (J.good[int64_t(x0) * J.RW + w] ^ (J.sa[f] ? ~uint64_t(0) : uint64_t(0))) &
               round_mask(J.num_patterns, w);
  for (int x = x0, g = J.parent[x]; g >= 0 && D; x = g, g = J.parent[x]) D &= sens(J, g, x, w);
  return D;
}

struct ItemArgs {
  const JobDev *__restrict__ jobs;
  int njobs;
  const int64_t *__restrict__ item_off;  // njobs + 1
  uint64_t *need;                         // [word] bits a unit's sweep must cover
  const uint64_t *__restrict__ dout;      // [word] sweep results
};

// Phase 1, thread = (job, fault, word). Stem tracing: a fault whose region
// ends at a PO is decided here, otherwise its bits join its stem's need.
// Direct injection: a word needs a lane iff the stuck value differs from the
// site's good value somewhere in the full word.
__device__ __forceinline__ void local_item(const ItemArgs &a, const JobDev &J, int64_t lt, int64_t span) {
  const int fi = int(div_items(lt, J.Wn, span));
  const int j = int(lt - int64_t(fi) * J.Wn);
  const int f = J.flist[fi];
  const int64_t w = J.w0 + j;
  if (J.direct) {
    a.need[J.word0 + lt] = J.good[int64_t(J.site[f]) * J.RW + w] ^ (J.sa[f] ? ~uint64_t(0) : uint64_t(0));
    return;
  }
  const uint64_t L = local_term(J, f, w);
  if (!L) return;
  const int h = J.head[f];
  if (J.is_po[h]) {
    record(J, f, w, L);
  } else {
    atomicOr((unsigned long long *)&a.need[J.word0 + int64_t(J.head_slot[h]) * J.Wn + j], (unsigned long long)L);
  }
}

// Item t of the step (job lookup once per distinct job of the wave).
__device__ __forceinline__ void local_at(const ItemArgs &a, int64_t t) {
  for_wave_jobs(a.item_off, a.njobs, t, [&](int jb) {
    local_item(a, a.jobs[jb], t - a.item_off[jb], a.item_off[jb + 1] - a.item_off[jb]);
  });
}

__global__ void k_local(ItemArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < a.item_off[a.njobs]) local_at(a, t);
}

// Phase 2: stem tracing d = L & D_stem; direct injection d = PO difference.
__device__ __forceinline__ void detect_item(const ItemArgs &a, const JobDev &J, int64_t lt, int64_t span) {
  const int fi = int(div_items(lt, J.Wn, span));
  const int j = int(lt - int64_t(fi) * J.Wn);
  const int f = J.flist[fi];
  const int64_t w = J.w0 + j;
  if (J.direct) {
    record(J, f, w, a.dout[J.word0 + lt] & round_mask(J.num_patterns, w));
    return;
  }
  const int h = J.head[f];
  if (J.is_po[h]) return;
  const uint64_t D = a.dout[J.word0 + int64_t(J.head_slot[h]) * J.Wn + j];
  if (!D) return;
  record(J, f, w, local_term(J, f, w) & D);
}

__device__ __forceinline__ void detect_at(const ItemArgs &a, int64_t t) {
  for_wave_jobs(a.item_off, a.njobs, t, [&](int jb) {
    detect_item(a, a.jobs[jb], t - a.item_off[jb], a.item_off[jb + 1] - a.item_off[jb]);
  });
}

__global__ void k_detect(ItemArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < a.item_off[a.njobs]) detect_at(a, t);
}

// The same over a fixed grid: the item count is read from the device.
__global__ void k_local_gs(ItemArgs a) {
  const int64_t n = a.item_off[a.njobs], stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) local_at(a, t);
}

__global__ void k_detect_gs(ItemArgs a) {
  const int64_t n = a.item_off[a.njobs], stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) detect_at(a, t);
}

// First detecting rounds of the step's faults, packed per dropping job at
// foff (jobs without a first_round table own an empty range).
__global__ void k_gather_first(const JobDev *__restrict__ jobs, int njobs, const int64_t *__restrict__ foff,
                               int32_t *__restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= foff[njobs]) return;
  const int jb = find_job(foff, njobs, t);
  const JobDev &J = jobs[jb];
  out[t] = J.first_round[J.flist[t - foff[jb]]];
}

// units[i] = {job, stem position, unit index}: head_slot of that job maps the
// stem to its unit index within the job (or back to -1).
// One device array of a finished job, copied to byte offset dst of the
// result block.
struct CopySeg {
  const uint8_t *src;
  int64_t dst, bytes;
};

// Gathers the segments (dst ascending, non-overlapping, each starting at a
// multiple of 4 bytes from a 4-byte aligned source) into one block, so that
// the host copies it back once. One thread per 4-byte word; a segment's last,
// partial word goes byte by byte. Bytes between segments stay untouched.
__global__ void k_gather_words(const CopySeg *__restrict__ segs, int nseg, int64_t end, uint8_t *__restrict__ out) {
  const int64_t t = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  if (t >= end) return;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].dst <= t) lo = mid; else hi = mid - 1;
  }
  const CopySeg s = segs[lo];
  const int64_t o = t - s.dst;
  if (o < 0 || o >= s.bytes) return;
  if (o + 4 <= s.bytes) {
    *reinterpret_cast<uint32_t *>(out + t) = *reinterpret_cast<const uint32_t *>(s.src + o);
  } else {
    for (int64_t k = 0; o + k < s.bytes; ++k) out[t + k] = s.src[o + k];
  }
}

// Job of entry t of the step's unit list: the largest i with lane_unit0 <= t.
__device__ __forceinline__ int job_of_unit(const JobDev *__restrict__ jobs, int nj, int64_t t) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].lane_unit0 <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// unit_pos[lane_unit0 + k] = position of unit k of a job. For stem-tracing
// jobs, the stem's slot becomes k (or -1 again with clear).
__global__ void k_set_slots(const JobDev *__restrict__ jobs, int nj, const int32_t *__restrict__ unit_pos, int64_t n,
                            int clear) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const JobDev &J = jobs[job_of_unit(jobs, nj, t)];
  if (J.direct) return;
  J.head_slot[unit_pos[t]] = clear ? -1 : int32_t(t - J.lane_unit0);
}

// ---------------------------------------------------------------- lanes and tiles on the device
//
// Word o of a step's need array belongs to (job, unit k, word jw); it becomes
// sweep lane pos[o] (pos = exclusive prefix count of words with needed bits),
// so lanes keep the (job, unit, word) order, and a job's lanes are cut into
// tiles of 64.

struct NeedNonzero {
  __host__ __device__ int operator()(const uint64_t &x) const { return x != 0 ? 1 : 0; }
};

// Largest job index i with jobs[i].word0 <= o (the job holding word o).
__device__ __forceinline__ int job_of_word(const JobDev *__restrict__ jobs, int nj, int64_t o) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].word0 <= o) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Level of live position u: the largest l with live_lvl[l] <= u.
__device__ __forceinline__ int level_of(const int32_t *__restrict__ live_lvl, int L, int u) {
  int lo = 0, hi = L;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (live_lvl[mid] <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// lane_off[i] = first lane of job i (i < nj), lane_off[nj] = all lanes.
__global__ void k_lane_bounds(const JobDev *__restrict__ jobs, int nj, const uint64_t *__restrict__ need,
                              const int32_t *__restrict__ pos, int64_t words, int64_t *lane_off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nj) return;
  if (words == 0) {
    lane_off[t] = 0;
    return;
  }
  const int64_t o = t < nj ? jobs[t].word0 : words;
  lane_off[t] = o < words ? int64_t(pos[o]) : int64_t(pos[words - 1]) + (need[words - 1] != 0 ? 1 : 0);
}

// units[k] = position of unit k of a job. Its injection is a flip for stem
// tracing, else the stuck value of step fault k (a direct job's step faults
// are in unit order).
__device__ __forceinline__ void lane_fill_word(const JobDev *__restrict__ jobs, int nj, const uint64_t *__restrict__ need,
                                               const int32_t *__restrict__ pos, int64_t o, Lane *__restrict__ lanes) {
  const uint64_t m = need[o];
  if (!m) return;
  const JobDev &J = jobs[job_of_word(jobs, nj, o)];
  const int64_t rel = o - J.word0;
  const int64_t k = rel / J.Wn;
  const int jw = int(rel - k * J.Wn);
  const int32_t site = J.units[k];
  Lane ln;
  ln.mask = m;
  ln.site = site;
  ln.word = int32_t(J.w0 + jw);
  ln.out = int32_t(o);
  ln.kind = uint32_t(J.direct ? (J.sa[J.flist[k]] ? K_SA1 : K_SA0) : K_FLIP);
  ln.level = level_of(J.live_lvl, J.max_level, site);
  ln.pad = 0;
  lanes[pos[o]] = ln;
}

__global__ void k_lane_fill(const JobDev *__restrict__ jobs, int nj, const uint64_t *__restrict__ need,
                            const int32_t *__restrict__ pos, int64_t words, Lane *__restrict__ lanes) {
  const int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (o < words) lane_fill_word(jobs, nj, need, pos, o, lanes);
}

// Tile t of the step: job by tile_off (largest i with tile_off[i] <= t),
// lanes [lane_off[i] + 64 (t - tile_off[i]), +64), and its sort key (levels
// the tile may walk).
__global__ void k_make_tiles(const JobDev *__restrict__ jobs, int nj, const Lane *__restrict__ lanes,
                             const int64_t *__restrict__ lane_off, const int64_t *__restrict__ tile_off, int64_t ntiles,
                             Tile *__restrict__ tiles, int *__restrict__ keys, int *__restrict__ idx) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_off[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const int64_t l0 = lane_off[lo] + 64 * (t - tile_off[lo]);
  const int n = int(min(int64_t(64), lane_off[lo + 1] - l0));
  Tile T;
  T.lane0 = int32_t(l0);
  T.nlanes = n;
  T.start_level = lanes[l0].level;
  T.end_level = lanes[l0 + n - 1].level;
  T.job = lo;
  tiles[t] = T;
  keys[t] = jobs[lo].max_level - T.start_level;
  idx[t] = int(t);
}

__global__ void k_gather_tiles(const Tile *__restrict__ in, const int *__restrict__ idx, int64_t n,
                               Tile *__restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < n) out[t] = in[idx[t]];
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- word steps planned on the device
//
// A step's fault lists, units, detection items, lanes and tiles are counted
// and laid out by kernels: prefix sums over fixed grids that read their
// lengths from device memory, one-workgroup kernels for the per-job offsets,
// and a counting sort of the tiles by depth. The host enqueues a step
// without waiting for the one before it, and every launch of a step has a
// fixed shape.

// One job of a device-planned step.
struct StepJob {
  const int32_t *src;    // candidate faults: the job's previous step list, or its live list
  const int32_t *src_n;  // their count on the device, or null (then src_count)
  int32_t src_count;
  int32_t filter;        // keep only candidates not detected yet (first_round == INT32_MAX)
  int32_t *dst;          // the step's fault list (unit order)
  int32_t *dst_n;        // its count (written by k_ds_plan)
  int32_t *units;        // the step's units
  int64_t fbase;         // first entry of the job in the step's fault space
  int32_t wide;          // the circuit's levels average more than 256 live nodes
  int32_t pad;
};

// Lengths of a device-planned step.
struct StepLen {
  int64_t items;  // detection items
  int64_t words;  // need / dout words
  int64_t lanes;
  int64_t tiles;
};

constexpr int kScanBlocks = 256;   // workgroups of a prefix sum's first and last phase
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;      // consecutive entries per thread in the last phase
constexpr int kGsBlocks = 1024;    // workgroups of the fixed-grid kernels (256 threads)
constexpr int kItemBlocks = 8192;  // at most, for the detection-item kernels
constexpr int kPlanThreads = 1024; // the one-workgroup planning kernels
constexpr int kDsStat = 8;         // step figures before the per-job ones

// Exclusive prefix sum of v over the BS threads of a workgroup; total = the
// sum. s_w holds BS / 64 entries of LDS. Every thread calls it.
template <int BS>
__device__ __forceinline__ int64_t block_excl_sum(int64_t v, int64_t &total, int64_t *s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  int64_t before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) {
    const int64_t t = s_w[i];
    before += i < w ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return before + int64_t(x) - v;
}

// Flag sources of the prefix sums.
struct FlagI32 {
  const int32_t *p;
  __device__ int operator()(int64_t i) const { return p[i]; }
};
struct FlagWord {
  const uint64_t *p;
  __device__ int operator()(int64_t i) const { return p[i] != 0 ? 1 : 0; }
};

__device__ __forceinline__ int64_t scan_len(const int64_t *n_p, int64_t n) { return n_p ? *n_p : n; }

// Entries of each workgroup's range: a multiple of one last-phase tile.
__device__ __forceinline__ int64_t scan_chunk(int64_t n) {
  constexpr int64_t tile = int64_t(kScanThreads) * kScanItems;
  const int64_t per = (n + kScanBlocks - 1) / kScanBlocks;
  return (per + tile - 1) / tile * tile;
}

// Prefix sum, phase 1: the flag count of each workgroup's range.
template <class Flag>
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(Flag flag, const int64_t *n_p, int64_t n_h, int64_t *bsum) {
  __shared__ int64_t s_w[kScanThreads / 64];
  const int64_t n = scan_len(n_p, n_h), chunk = scan_chunk(n);
  const int64_t lo = min(n, int64_t(blockIdx.x) * chunk), hi = min(n, lo + chunk);
  int64_t v = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kScanThreads) v += flag(i);
  int64_t tot;
  (void)block_excl_sum<kScanThreads>(v, tot, s_w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Phase 2 (one workgroup of kScanBlocks threads): range counts to range
// offsets, in place; out[n] = the total.
__global__ void __launch_bounds__(kScanBlocks) k_scan_offsets(int64_t *bsum, const int64_t *n_p, int64_t n_h,
                                                              int32_t *out) {
  __shared__ int64_t s_w[kScanBlocks / 64];
  const int64_t v = bsum[threadIdx.x];
  int64_t tot;
  const int64_t ex = block_excl_sum<kScanBlocks>(v, tot, s_w);
  bsum[threadIdx.x] = ex;
  if (threadIdx.x == 0) out[scan_len(n_p, n_h)] = int32_t(tot);
}

// Phase 3: out[i] = the flags before entry i.
template <class Flag>
__global__ void __launch_bounds__(kScanThreads) k_scan_write(Flag flag, const int64_t *n_p, int64_t n_h,
                                                             const int64_t *boff, int32_t *out) {
  __shared__ int64_t s_w[kScanThreads / 64];
  const int64_t n = scan_len(n_p, n_h), chunk = scan_chunk(n);
  const int64_t lo = min(n, int64_t(blockIdx.x) * chunk), hi = min(n, lo + chunk);
  int64_t run = boff[blockIdx.x];
  for (int64_t b = lo; b < hi; b += int64_t(kScanThreads) * kScanItems) {
    const int64_t i0 = b + int64_t(threadIdx.x) * kScanItems;
    int f[kScanItems];
    int64_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      f[k] = i0 + k < hi ? flag(i0 + k) : 0;
      v += f[k];
    }
    int64_t tot;
    int64_t ex = run + block_excl_sum<kScanThreads>(v, tot, s_w);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (i0 + k < hi) out[i0 + k] = int32_t(ex);
      ex += f[k];
    }
    run += tot;
  }
}

// Job holding entry t of the step's fault space: the largest i with fbase <= t.
__device__ __forceinline__ int job_of_entry(const StepJob *__restrict__ sj, int nj, int64_t t) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sj[mid].fbase <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// End of job j's range in a fault space of n entries.
__device__ __forceinline__ int64_t entry_end(const StepJob *__restrict__ sj, int nj, int j, int64_t n) {
  return j + 1 < nj ? sj[j + 1].fbase : n;
}

// flag[t] = 1 when entry t of the step's fault space is a fault of the
// step: candidate k = t - fbase of its job exists and, when the job filters,
// is not detected yet.
__global__ void k_ds_keep(const StepJob *__restrict__ sj, const JobDev *__restrict__ jobs, int nj, int64_t n,
                          int32_t *__restrict__ flag) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int j = job_of_entry(sj, nj, t);
    const StepJob &S = sj[j];
    const int64_t k = t - S.fbase;
    const int64_t cnt = S.src_n ? int64_t(*S.src_n) : int64_t(S.src_count);
    int keep = 0;
    if (k < cnt) keep = (!S.filter || jobs[j].first_round[S.src[k]] == INT32_MAX) ? 1 : 0;
    flag[t] = keep;
  }
}

// The kept candidates, in their order, into each job's step list.
__global__ void k_ds_compact(const StepJob *__restrict__ sj, int nj, int64_t n, const int32_t *__restrict__ flag,
                             const int32_t *__restrict__ pos) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    if (!flag[t]) continue;
    const StepJob &S = sj[job_of_entry(sj, nj, t)];
    S.dst[pos[t] - pos[S.fbase]] = S.src[t - S.fbase];
  }
}

// flag[t] = 1 when step fault k = t - fbase opens a unit: every fault under
// direct injection; under stem tracing the first fault of a region whose
// head is not a PO (a job's faults are ordered by region head).
__global__ void k_ds_unit_flag(const StepJob *__restrict__ sj, const JobDev *__restrict__ jobs, int nj, int64_t n,
                               const int32_t *__restrict__ pos, int32_t *__restrict__ flag) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int j = job_of_entry(sj, nj, t);
    const StepJob &S = sj[j];
    const JobDev &J = jobs[j];
    const int64_t k = t - S.fbase;
    const int64_t nf = int64_t(pos[entry_end(sj, nj, j, n)]) - int64_t(pos[S.fbase]);
    int f = 0;
    if (k < nf) {
      if (J.direct) {
        f = 1;
      } else {
        const int32_t h = J.head[S.dst[k]];
        f = (!J.is_po[h] && (k == 0 || J.head[S.dst[k - 1]] != h)) ? 1 : 0;
      }
    }
    flag[t] = f;
  }
}

// Unit positions: the region head, or the fault site under direct injection.
__global__ void k_ds_units(const StepJob *__restrict__ sj, const JobDev *__restrict__ jobs, int nj, int64_t n,
                           const int32_t *__restrict__ flag, const int32_t *__restrict__ upos) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    if (!flag[t]) continue;
    const int j = job_of_entry(sj, nj, t);
    const StepJob &S = sj[j];
    const JobDev &J = jobs[j];
    const int32_t x = S.dst[t - S.fbase];
    S.units[upos[t] - upos[S.fbase]] = J.direct ? J.site[x] : J.head[x];
  }
}

// Per job: the step's faults and units, its first detection item and need
// word (written into the job table), and the step's lengths; stat gets
// {faults, units} per job. One workgroup.
__global__ void __launch_bounds__(kPlanThreads) k_ds_plan(const StepJob *__restrict__ sj, JobDev *__restrict__ jobs,
                                                          int nj, int64_t n, const int32_t *__restrict__ pos,
                                                          const int32_t *__restrict__ upos, int64_t *__restrict__ item_off,
                                                          StepLen *__restrict__ len, int64_t *__restrict__ stat) {
  __shared__ int64_t s_w[kPlanThreads / 64];
  int64_t ci = 0, cw = 0;
  for (int j0 = 0; j0 < nj; j0 += kPlanThreads) {
    const int j = j0 + int(threadIdx.x);
    int64_t nf = 0, nu = 0, items = 0, words = 0;
    if (j < nj) {
      const int64_t b = sj[j].fbase, e = entry_end(sj, nj, j, n);
      nf = int64_t(pos[e]) - int64_t(pos[b]);
      nu = int64_t(upos[e]) - int64_t(upos[b]);
      items = nf * jobs[j].Wn;
      words = nu * jobs[j].Wn;
    }
    int64_t ti, tw;
    const int64_t ei = block_excl_sum<kPlanThreads>(items, ti, s_w);
    const int64_t ew = block_excl_sum<kPlanThreads>(words, tw, s_w);
    if (j < nj) {
      JobDev &J = jobs[j];
      J.nf = nf;
      J.item0 = ci + ei;
      J.word0 = cw + ew;
      item_off[j] = ci + ei;
      *sj[j].dst_n = int32_t(nf);
      stat[2 * j] = nf;
      stat[2 * j + 1] = nu;
    }
    ci += ti;
    cw += tw;
  }
  if (threadIdx.x == 0) {
    item_off[nj] = ci;
    len->items = ci;
    len->words = cw;
  }
}

__global__ void k_ds_zero(const StepLen *__restrict__ len, uint64_t *__restrict__ need, uint64_t *__restrict__ dout) {
  const int64_t n = len->words, stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    need[t] = 0;
    dout[t] = 0;
  }
}

// Stem tracing: head_slot of each unit of the step = its index in the job's
// units (clear: back to -1).
__global__ void k_ds_slots(const StepJob *__restrict__ sj, const JobDev *__restrict__ jobs, int nj, int64_t n,
                           const int32_t *__restrict__ upos, int clear) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int j = job_of_entry(sj, nj, t);
    const JobDev &J = jobs[j];
    if (J.direct) continue;
    const StepJob &S = sj[j];
    const int64_t k = t - S.fbase;
    const int64_t nu = int64_t(upos[entry_end(sj, nj, j, n)]) - int64_t(upos[S.fbase]);
    if (k < nu) J.head_slot[S.units[k]] = clear ? -1 : int32_t(k);
  }
}

__global__ void k_lane_fill_gs(const JobDev *__restrict__ jobs, int nj, const uint64_t *__restrict__ need,
                               const int32_t *__restrict__ pos, const StepLen *__restrict__ len,
                               Lane *__restrict__ lanes) {
  const int64_t n = len->words, stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < n; o += stride)
    lane_fill_word(jobs, nj, need, pos, o, lanes);
}

struct TilePlanArgs {
  const StepJob *sj;
  const JobDev *jobs;
  int nj;
  const int32_t *pos;           // lane position of each need word (words + 1 entries)
  StepLen *len;
  int64_t *lane_off, *tile_off; // nj + 1 entries each
  int32_t *hist;                // depth histogram of nkeys entries, cleared here
  int nkeys;
  SweepDyn *dyn;
  unsigned long long *bytes;    // sweep counters: bytes, staged nodes, walked nodes
  int *counters;                // tile queue of the first launch and of the rerun, tiles each gave up
  const uint32_t *epoch;        // the step's first tile epoch
  int dense;                    // a dropping job's first step or a full-matrix step
  int threads;                  // workgroup size asked for (0: chosen here)
  int allow256;                 // the 256-thread form can be resident
  int nopend_mode, nopend_pct;  // as EngineOptions::pair_nopend, pair_nopend_pct
};

// Per job: lanes and tiles of 64 lanes; the step's sweep configuration,
// chosen as the host chooses it for a step it plans itself: 512-thread
// workgroups when most tiles walk levels wider than 256 nodes, in dense
// steps and outside a middle staged share (0.2-0.75) of the last sweep;
// no pending flags in dense steps and after a staged share above
// nopend_pct percent. One workgroup.
__global__ void __launch_bounds__(kPlanThreads) k_ds_tiles(TilePlanArgs a) {
  __shared__ int64_t s_w[kPlanThreads / 64];
  const int64_t words = a.len->words;
  int64_t ct = 0, cwide = 0;
  for (int j0 = 0; j0 < a.nj; j0 += kPlanThreads) {
    const int j = j0 + int(threadIdx.x);
    int64_t lanes = 0, tiles = 0, lo = 0;
    if (j < a.nj) {
      const int64_t w0 = a.jobs[j].word0, w1 = j + 1 < a.nj ? a.jobs[j + 1].word0 : words;
      lo = a.pos[w0];
      lanes = int64_t(a.pos[w1]) - lo;
      tiles = (lanes + 63) / 64;
    }
    int64_t tt, tw;
    const int64_t et = block_excl_sum<kPlanThreads>(tiles, tt, s_w);
    (void)block_excl_sum<kPlanThreads>((j < a.nj && a.sj[j].wide) ? tiles : 0, tw, s_w);
    if (j < a.nj) {
      a.lane_off[j] = lo;
      a.tile_off[j] = ct + et;
    }
    ct += tt;
    cwide += tw;
  }
  for (int i = int(threadIdx.x); i < a.nkeys; i += kPlanThreads) a.hist[i] = 0;
  if (threadIdx.x == 0) {
    const int64_t lanes = a.pos[words];
    a.lane_off[a.nj] = lanes;
    a.tile_off[a.nj] = ct;
    a.len->lanes = lanes;
    a.len->tiles = ct;
    SweepDyn D = *a.dyn;
    if (a.bytes[2]) D.share = double(a.bytes[1]) / double(a.bytes[2]);
    const bool mid = D.share >= 0.2 && D.share <= 0.75;
    const bool big = 2 * cwide > ct && (a.dense || !mid);
    int threads = a.threads ? a.threads : (big ? 512 : 256);
    if (threads == 256 && !a.allow256) threads = 512;
    D.form = threads;
    D.nopend = (a.nopend_mode == 2 || (a.nopend_mode == 1 && (a.dense || 100.0 * D.share > double(a.nopend_pct)))) ? 1 : 0;
    D.ntiles = int(ct);
    D.epoch_base = *a.epoch;
    *a.dyn = D;
    a.bytes[0] = 0;
    a.bytes[1] = 0;
    a.bytes[2] = 0;
    for (int i = 0; i < 4; ++i) a.counters[i] = 0;
  }
}

struct TileSortArgs {
  const JobDev *jobs;
  int nj;
  const Lane *lanes;
  const int64_t *lane_off, *tile_off;
  const StepLen *len;
  Tile *raw;
  int32_t *keys;
  int32_t *hist;
  Tile *out;
};

// Tile t of the step (its job by tile_off), with its depth key (the levels it
// may walk) counted in the histogram.
__global__ void k_ds_tile_hist(TileSortArgs a) {
  const int64_t n = a.len->tiles, stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride) {
    int lo = 0, hi = a.nj - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.tile_off[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int64_t l0 = a.lane_off[lo] + 64 * (t - a.tile_off[lo]);
    const int nl = int(min(int64_t(64), a.lane_off[lo + 1] - l0));
    Tile T;
    T.lane0 = int32_t(l0);
    T.nlanes = nl;
    T.start_level = a.lanes[l0].level;
    T.end_level = a.lanes[l0 + nl - 1].level;
    T.job = lo;
    const int key = a.jobs[lo].max_level - T.start_level;
    a.raw[t] = T;
    a.keys[t] = key;
    atomicAdd(&a.hist[key], 1);
  }
}

// Depth buckets, deepest first: hist[key] becomes the first queue position of
// the key's tiles. One workgroup.
__global__ void __launch_bounds__(kPlanThreads) k_ds_hist_scan(int32_t *hist, int nkeys) {
  __shared__ int64_t s_w[kPlanThreads / 64];
  int64_t carry = 0;
  for (int i0 = 0; i0 < nkeys; i0 += kPlanThreads) {
    const int i = i0 + int(threadIdx.x);
    const int key = nkeys - 1 - i;
    const int64_t v = i < nkeys ? hist[key] : 0;
    int64_t tot;
    const int64_t ex = block_excl_sum<kPlanThreads>(v, tot, s_w);
    if (i < nkeys) hist[key] = int32_t(carry + ex);
    carry += tot;
  }
}

// Tiles into the queue, deepest first (in any order within a depth).
__global__ void k_ds_tile_scatter(TileSortArgs a) {
  const int64_t n = a.len->tiles, stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += stride)
    a.out[atomicAdd(&a.hist[a.keys[t]], 1)] = a.raw[t];
}

// The step's figures for the host: sweep bytes, lanes, tiles, tiles the
// first sweep launch gave up and tiles the rerun gave up, items, words.
__global__ void k_ds_stats(const StepLen *__restrict__ len, const unsigned long long *__restrict__ bytes,
                           const int *__restrict__ counters, int64_t *__restrict__ stat) {
  stat[0] = int64_t(bytes[0]);
  stat[1] = len->lanes;
  stat[2] = len->tiles;
  stat[3] = counters[2];
  stat[4] = counters[3];
  stat[5] = len->items;
  stat[6] = len->words;
  stat[7] = 0;
}

// Stream on which device buffers allocated by the calling thread are
// ordered: an engine's stream while one of its calls runs. Stream-ordered
// allocation and release never wait for other streams (a plain hipFree waits
// for the whole device, so one engine's release would wait for another
// engine's kernels).
thread_local hipStream_t tl_alloc_stream = nullptr;

// Device memory of an engine that must not allocate: one region the caller
// provides, carved front to back in 256-byte aligned pieces and never freed
// piecewise (the whole engine is dropped at once).
struct DeviceArena {
  uint8_t *base = nullptr;
  size_t cap = 0, used = 0;
  void *take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    if (bytes > cap || at > cap - bytes)
      throw DeviceError(LSIM_E_NOMEM, "the run needs more device memory than the " + std::to_string(cap >> 20) +
                                          " MiB no-allocation scratch holds");
    used = at + bytes;
    return base + at;
  }
  size_t left() const { return cap - std::min(cap, (used + 255) & ~size_t(255)); }
};
thread_local DeviceArena *tl_arena = nullptr;
// Device and page-locked host allocations the library has made (all threads).
std::atomic<int64_t> g_allocs{0};

struct AllocStream {
  hipStream_t prev;
  DeviceArena *prev_arena;
  explicit AllocStream(hipStream_t s, DeviceArena *a = nullptr) : prev(tl_alloc_stream), prev_arena(tl_arena) {
    tl_alloc_stream = s;
    tl_arena = a;
  }
  ~AllocStream() {
    tl_alloc_stream = prev;
    tl_arena = prev_arena;
  }
  AllocStream(const AllocStream &) = delete;
  AllocStream &operator=(const AllocStream &) = delete;
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;  // stream of a stream-ordered allocation, null for a plain one
  bool own = true;          // false: a view into memory another buffer owns
  bool carved = false;      // carved from the thread's DeviceArena (reused while large enough, never freed)
  // Under a DeviceArena (tl_arena) the memory is carved from it and no
  // allocation is made.
  bool from_arena(size_t count) {
    if (!tl_arena) return false;
    p = static_cast<T *>(tl_arena->take(count * sizeof(T)));
    n = count;
    s = nullptr;
    own = false;
    carved = true;
    return true;
  }
  void alloc(size_t count) {
    if (count <= n && p && (own || carved)) return;
    release();
    if (count == 0) count = 1;
    if (from_arena(count)) return;
    if (tl_alloc_stream) {
      LSIM_HIP(hipMallocAsync(reinterpret_cast<void **>(&p), count * sizeof(T), tl_alloc_stream));
      s = tl_alloc_stream;
    } else {
      LSIM_HIP(hipMalloc(&p, count * sizeof(T)));
      s = nullptr;
    }
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    n = count;
  }
  // Plain allocation, for the large sweep scratch that lives as long as the engine.
  void alloc_direct(size_t count) {
    if (count <= n && p && (own || carved)) return;
    release();
    if (count == 0) count = 1;
    if (from_arena(count)) return;
    LSIM_HIP(hipMalloc(&p, count * sizeof(T)));
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    s = nullptr;
    n = count;
  }
  void release() {
    if (p && own) {
      if (s) (void)hipFreeAsync(p, s);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    s = nullptr;
    own = true;
    carved = false;
  }
  // Points at `count` elements of memory owned elsewhere (never freed here).
  void view(T *ptr, size_t count) {
    release();
    p = ptr;
    n = count;
    own = false;
  }
  size_t bytes() const { return n * sizeof(T); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
};

// Page-locked host staging buffer (grows, never shrinks). Under a
// DeviceArena it is ordinary pageable memory instead: an engine that must
// not allocate makes no page-locked allocation either.
template <typename T>
struct HostBuf {
  T *p = nullptr;
  size_t n = 0;
  bool pinned = false;
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    if (tl_arena) {
      p = static_cast<T *>(std::malloc(count * sizeof(T)));
      if (!p) throw std::bad_alloc();
      pinned = false;
    } else {
      LSIM_HIP(hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault));
      g_allocs.fetch_add(1, std::memory_order_relaxed);
      pinned = true;
    }
    n = count;
  }
  void release() {
    if (p) {
      if (pinned) (void)hipHostFree(p);
      else std::free(p);
    }
    p = nullptr;
    n = 0;
    pinned = false;
  }
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() { release(); }
};

// LSIM_TRACE set to anything but "" or "0": per-iteration timings on stderr.
inline bool trace_on() {
  static const bool on = [] {
    const char *v = std::getenv("LSIM_TRACE");
    return v && *v && std::strcmp(v, "0") != 0;
  }();
  return on;
}

inline unsigned blocks_for(int64_t threads, int bs) { return unsigned((threads + bs - 1) / bs); }

// Host worker threads kept for the process's life, so that a parallel region
// costs a wake-up instead of starting and joining threads. One region runs at
// a time; the caller works too. A region is finished only when every worker
// has taken part in it (each worker sees each region once), so no worker can
// read a finished region's function.
thread_local bool tl_in_region = false;  // a pool worker, or a caller inside its region

class HostPool {
 public:
  explicit HostPool(int workers) : workers_(workers) {
    for (int i = 0; i < workers_; ++i) th_.emplace_back([this] { loop(); });
  }
  int workers() const { return workers_; }
  // Runs f(k) for k in [0, n) on the caller and every worker; false (nothing
  // run) when another region holds the pool or the caller is inside one.
  bool run(int n, const std::function<void(int)> &f) {
    if (tl_in_region) return false;
    std::unique_lock<std::mutex> region(busy_, std::try_to_lock);
    if (!region.owns_lock()) return false;
    struct InRegion {
      InRegion() { tl_in_region = true; }
      ~InRegion() { tl_in_region = false; }
    } in_region;
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &f;
      n_ = n;
      next_.store(0);
      finished_ = 0;
      failed_.store(false);
      err_ = nullptr;
      ++gen_;
    }
    cv_work_.notify_all();
    drain(f);
    std::exception_ptr err;
    {
      std::unique_lock<std::mutex> lk(m_);
      cv_done_.wait(lk, [&] { return finished_ == workers_; });
      fn_ = nullptr;
      err = err_;
    }
    if (err) std::rethrow_exception(err);
    return true;
  }

 private:
  void loop() {
    tl_in_region = true;
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)> *f;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_work_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        f = fn_;
      }
      drain(*f);
      std::lock_guard<std::mutex> lk(m_);
      if (++finished_ == workers_) cv_done_.notify_all();
    }
  }
  void drain(const std::function<void(int)> &f) {
    for (;;) {
      const int k = next_.fetch_add(1);
      if (k >= n_ || failed_.load()) return;
      try {
        f(k);
      } catch (...) {
        if (!failed_.exchange(true)) {
          std::lock_guard<std::mutex> lk(m_);
          err_ = std::current_exception();
        }
      }
    }
  }
  const int workers_;
  std::vector<std::thread> th_;
  std::mutex busy_, m_;
  std::condition_variable cv_work_, cv_done_;
  const std::function<void(int)> *fn_ = nullptr;
  int n_ = 0, finished_ = 0;
  uint64_t gen_ = 0;
  std::atomic<int> next_{0};
  std::atomic<bool> failed_{false};
  std::exception_ptr err_;
};

// The process's pool: 15 workers at most (16 threads with the caller), never
// destroyed (its workers wait on a condition variable until the process ends).
HostPool &host_pool() {
  static HostPool *pool = new HostPool(
      int(std::max(1u, std::min(15u, std::max(1u, std::thread::hardware_concurrency()) - 1u))));
  return *pool;
}

// Runs fn(order[k]) for k in [0, n) on up to 16 host threads (items are
// handed out in order, so put the largest first); rethrows the first error.
// The pool runs it when free; a nested or concurrent region starts threads.
template <typename Fn>
void parallel_for(const std::vector<int> &order, Fn &&fn) {
  const int n = int(order.size());
  const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
  const int T = std::min(n, std::min(16, hw));
  if (T <= 1) {
    for (int k = 0; k < n; ++k) fn(order[k]);
    return;
  }
  const std::function<void(int)> body_k = [&](int k) { fn(order[size_t(k)]); };
  if (host_pool().run(n, body_k)) return;
  std::atomic<int> next{0};
  std::exception_ptr err;
  std::atomic<bool> failed{false};
  auto body = [&] {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= n || failed.load()) break;
      try {
        fn(order[k]);
      } catch (...) {
        if (!failed.exchange(true)) err = std::current_exception();
      }
    }
  };
  std::vector<std::thread> th;
  th.reserve(size_t(T - 1));
  for (int t = 1; t < T; ++t) th.emplace_back(body);
  body();
  for (auto &x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// Chunks for splitting [0, n) over host threads: 1 below 128k items, else
// one per 64k items, at most 16.
int chunk_count(int64_t n, bool allow) {
  if (!allow) return 1;
  const int hw = int(std::max(1u, std::thread::hardware_concurrency()));
  return int(std::max<int64_t>(1, std::min<int64_t>(std::min(16, hw), n / 65536)));
}

// fn(chunk, begin, end) over `chunks` contiguous pieces of [0, n), in
// parallel (the same pieces for the same n and chunk count).
template <typename Fn>
void run_chunks(int chunks, int64_t n, Fn &&fn) {
  std::vector<int> ids(size_t(std::max(1, chunks)));
  std::iota(ids.begin(), ids.end(), 0);
  const int64_t C = int64_t(ids.size());
  parallel_for(ids, [&](int c) { fn(c, n * c / C, n * (c + 1) / C); });
}

}  // namespace

// ---------------------------------------------------------------- host state

// The good-machine launch sequence of one circuit at one round count,
// captured once: kernels read the job from `job` (rewritten before each
// replay) and per-level item offsets from `off`. Only kernel nodes.
struct GoodGraph {
  int64_t R = 0;
  int form = 0;  // 0: one launch per level, 1: level windows
  int win_levels = 0, win_cap = 0;  // window plan of form 1
  DevBuf<JobDev> job;
  DevBuf<int64_t> off;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GoodGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// The batched good machine of a job set (one launch per level index over all
// jobs), captured once the same set of (circuit, rounds) recurs and replayed
// with the jobs' records copied in first.
struct BatchGoodGraph {
  std::string key;
  std::vector<const void *> circuits;
  DevBuf<JobDev> jobs;
  DevBuf<int64_t> off;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchGoodGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// Level-window blocks of one circuit (built on first use).
struct WindowPlan {
  int levels = 0, cap = 0;         // window depth asked for, block capacity in nodes
  std::vector<int32_t> win_block;  // window k holds blocks [win_block[k], win_block[k+1])
  std::vector<int32_t> win_depth;  // levels of window k (cut below `levels` where a cone outgrows `cap`)
  DevBuf<WinBlock> blocks;
  DevBuf<WinNode> nodes;
  int64_t repeated = 0;  // block nodes beyond one per circuit node
};

struct CircuitRes {
  int64_t N = 0, NL = 0, P = 0, G = 0, O = 0, E = 0;
  int L = 0;
  std::vector<uint32_t> h_info;  // per device position (kept for the window partitioner)
  std::vector<int4> h_fin;
  std::unique_ptr<WindowPlan> win;
  std::vector<std::unique_ptr<GoodGraph>> graphs;  // most recent last, at most kMaxGraphs
  bool wide = false;  // mean live nodes per level above 256 (the pair sweep then uses 512-thread groups)
  bool narrow = false;  // no gate has more than two fanins (the pair sweep then stages two per node)
  bool has_mux = false;
  std::vector<int32_t> pos2dp, dp2pos, level_dp, live_lvl, dead_lvl;
  std::vector<int32_t> head_dp;     // per position: region head (device position), -1 if dead
  std::vector<uint8_t> head_is_po;  // per live device position
  DevBuf<uint32_t> info;
  DevBuf<int4> fin;
  DevBuf<int4> nrec;  // narrow circuits: the pair sweep's per-node record (see JobDev::nrec)
  DevBuf<int32_t> parent, d_live_lvl, d_dead_lvl, fo_off;
  DevBuf<int2> fo;
  DevBuf<uint8_t> is_po;
  int64_t bytes() const {
    return int64_t(info.bytes() + fin.bytes() + nrec.bytes() + parent.bytes() + d_live_lvl.bytes() +
                   d_dead_lvl.bytes() + fo_off.bytes() + fo.bytes() + is_po.bytes());
  }
};

struct JobState {
  CircuitRes *c = nullptr;
  JobOptions opt;
  bool direct = false;
  int64_t F = 0, R = 0, num_patterns = 0, RB = 0;
  std::vector<int32_t> site, head;  // per fault: device position, region head (-1: dead site)
  std::vector<uint8_t> sa;
  // Faults with a live site, ordered by injection unit (stable, so by fault id
  // within a unit): by site for direct injection, by region head otherwise.
  // Dropping filters keep that order, so a step's units come from one pass.
  std::vector<int32_t> live, alive;
  bool live_by_site = false;  // the order of `live`
  DevBuf<uint8_t> arena;  // one allocation holding every per-job device array below
  DevBuf<uint64_t> good, pi_words, detbits;
  DevBuf<int32_t> d_site, d_head, first, head_slot;
  DevBuf<uint8_t> d_sa, last;
  bool started = false, done = false, last_done = false;
  int64_t w0 = 0;
  int Wt = 1;                          // words of the next dropping step
  int Wg = 2;                          // growth of Wt per step
  // current step
  enum Kind { MATRIX, DROP, LAST } kind = MATRIX;
  std::vector<int32_t> step_faults, units;
  int64_t step_w0 = 0;
  int step_Wn = 0;
  RunResults res;
  // Steps planned on the device: the live list, two step lists (a step's list
  // is filtered from the one before), the step's units, and the counts of the
  // two step lists.
  bool dev = false;         // this run's steps are planned on the device
  int dev_steps = 0;        // device steps taken (the parity picks the step list)
  int64_t dev_alive = -1;   // faults of the last step whose figures came back (-1: none yet)
  int64_t stems_cap = 0;    // regions with a non-PO head among the live faults
  DevBuf<int32_t> d_live, d_fl0, d_fl1, d_units, d_cnt;
};

// Owns the engine's stream; declared first in Engine::Impl so that it is
// destroyed last, after every buffer ordered on the stream is released.
struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

// Device memory of released jobs, kept for the engine's later jobs (up to
// kCap bytes): a stream-ordered free or allocation costs a fraction of a
// millisecond, a reuse nothing. Blocks are used on the engine's one stream
// only, so a reuse is ordered after the released job's work.
struct ArenaCache {
  struct Block {
    uint8_t *p;
    size_t n;
    hipStream_t s;
  };
  static constexpr size_t kCap = size_t(16) << 30;
  std::vector<Block> blocks;
  size_t bytes = 0;
  // The smallest block of at least `want` bytes and at most twice that.
  bool take(size_t want, DevBuf<uint8_t> &out) {
    int best = -1;
    for (int i = 0; i < int(blocks.size()); ++i)
      if (blocks[i].n >= want && blocks[i].n <= 2 * want && (best < 0 || blocks[i].n < blocks[size_t(best)].n))
        best = i;
    if (best < 0) return false;
    const Block b = blocks[size_t(best)];
    blocks.erase(blocks.begin() + best);
    bytes -= b.n;
    out.release();
    out.p = b.p;
    out.n = b.n;
    out.s = b.s;
    out.own = true;
    return true;
  }
  // Takes the buffer's memory (a stream-ordered allocation) into the cache.
  void put(DevBuf<uint8_t> &b) {
    if (!b.p || !b.own || !b.s || b.n > kCap) {
      b.release();
      return;
    }
    blocks.push_back({b.p, b.n, b.s});
    bytes += b.n;
    b.p = nullptr;
    b.n = 0;
    b.s = nullptr;
    while (bytes > kCap) {
      (void)hipFreeAsync(blocks.front().p, blocks.front().s);
      bytes -= blocks.front().n;
      blocks.erase(blocks.begin());
    }
  }
  void clear() {
    for (const Block &b : blocks) (void)hipFreeAsync(b.p, b.s);
    blocks.clear();
    bytes = 0;
  }
  ArenaCache() = default;
  ArenaCache(const ArenaCache &) = delete;
  ArenaCache &operator=(const ArenaCache &) = delete;
  ~ArenaCache() { clear(); }
};

struct Engine::Impl {
  StreamHolder holder;
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  ArenaCache arenas;  // destroyed after the jobs, before the stream
  std::vector<std::unique_ptr<CircuitRes>> circuits;
  std::vector<std::unique_ptr<JobState>> jobs;
  // per iteration
  DevBuf<JobDev> d_jobs;
  DevBuf<int64_t> d_item_off, d_good_off, d_fault_off;
  DevBuf<int32_t> d_flist, d_first_out;
  HostBuf<int32_t> h_flist, h_first_out;
  HostBuf<uint64_t> h_need;
  HostBuf<Lane> h_lanes;
  // device-built lanes and tiles (persistent sweeps)
  HostBuf<int32_t> h_unit_pos;  // the step's units (positions), job after job
  DevBuf<int32_t> d_unit_pos;
  DevBuf<int32_t> d_pos;
  DevBuf<int64_t> d_lane_off, d_tile_off;
  HostBuf<int64_t> h_lane_off;
  DevBuf<Tile> d_tiles_raw;
  DevBuf<int> d_tile_key, d_tile_idx;
  DevBuf<uint8_t> d_cub;
  HostBuf<unsigned long long> h_bytes;  // sweep bytes; pair sweep staged and walked nodes
  double staged_share = 0;              // of the last pair sweep launch
  // Steps after one that staged a middle share of the nodes its tiles walked
  // run in 256-thread groups even on wide levels: twice the tiles in flight
  // outweigh the second scan pass per level there (on the 2M-gate circuit,
  // 9-16 ms less in each of the four word steps after shares 0.69, 0.53,
  // 0.37 and 0.25; more at 0.83 and at 0.11, the same at 0.17).
  static constexpr double kMidShareLo = 0.2, kMidShareHi = 0.75;
  DevBuf<uint64_t> d_need, d_dout, d_gather;
  DevBuf<Lane> d_lanes;
  DevBuf<Tile> d_tiles;
  DevBuf<int> d_queue;
  DevBuf<unsigned long long> d_bytes, d_tstats;
  // sweep slots
  DevBuf<uint64_t> d_rows, d_masks;
  DevBuf<MaskRec> d_mrec;
  DevBuf<uint32_t> d_pend;
  // pair sweep scratch: per-slot records and pending flags, one pool carved per launch
  DevBuf<PairRec> d_prec;
  DevBuf<uint32_t> d_ppend;
  DevBuf<uint64_t> d_pool;
  DevBuf<int> d_ovf_count, d_ovf_list, d_ovf_list2;
  int64_t pair_slots = 0, pair_req = 0, pair_nl = 0;
  int64_t slots = 0, slot_nl = 0;
  uint32_t epoch_base = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};

  // An engine that must not allocate: every device buffer is carved from
  // arena_mem, and the stream and events are borrowed from another engine.
  DeviceArena arena_mem;
  DeviceArena *arena = nullptr;
  bool borrowed = false;

  ~Impl() {
    // In-flight steps may still copy into the slots' page-locked buffers.
    if (stream) (void)hipStreamSynchronize(stream);
    jobs.clear();
    circuits.clear();
    for (DsSlot &sl : ds)
      for (hipEvent_t &e : sl.ev)
        if (e) (void)hipEventDestroy(e);
    if (!borrowed) {
      if (ev[0]) (void)hipEventDestroy(ev[0]);
      if (ev[1]) (void)hipEventDestroy(ev[1]);
      if (fin_ev) (void)hipEventDestroy(fin_ev);
    }
    // The stream itself goes with `holder` (null when borrowed), after the
    // member buffers.
  }
  void time_begin() { LSIM_HIP(hipEventRecord(ev[0], stream)); }
  double time_end() {
    LSIM_HIP(hipEventRecord(ev[1], stream));
    LSIM_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    LSIM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    return ms;
  }
  void h2d(void *dst, const void *src, size_t bytes) {
    if (bytes) LSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  }
  void d2h(void *dst, const void *src, size_t bytes) {
    if (bytes) LSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
  }
  JobDev job_dev(JobState &j) const;
  GoodGraph *good_graph(CircuitRes &c, int64_t R, int form, const EngineOptions &eo);
  void build_window_plan(CircuitRes &c, int levels, int cap);
  int win_cap(const EngineOptions &eo) const;
  int lds_cu = 65536, lds_block = 65536;  // LDS bytes per CU and per workgroup
  void finish_start(const std::vector<JobState *> &js, double ms);
  int64_t iterate_jobs(const std::vector<JobState *> &jobs, const EngineOptions &eo);
  // Autotuner: sweep kernel and dropping word schedule. Per job-set
  // signature (resident circuits, rounds, dropping, injection mode) the first
  // runs each use one untried candidate and time their word steps; later
  // runs use the fastest. Exploration runs are ordinary runs (same results).
  // Autotuner: one configuration of the sweep per run of a job set, chosen
  // by coordinate descent over the dimensions below (the first value of each
  // is the default): each dimension's values are tried with the others held
  // at the best found so far, then the best value is kept and the next
  // dimension is explored; after the last one every run uses the best.
  struct TuneChoice {
    int kernel = 0, first = 1, growth = 2, threads = 0, dense = 24, max_words = 128;
    bool operator==(const TuneChoice &o) const {
      return kernel == o.kernel && first == o.first && growth == o.growth && threads == o.threads &&
             dense == o.dense && max_words == o.max_words;
    }
    int64_t code() const {
      return int64_t(kernel) | (int64_t(first) << 8) | (int64_t(growth) << 16) | (int64_t(threads / 256) << 24) |
             (int64_t(dense) << 32) | (int64_t(max_words) << 40);
    }
  };
  struct TuneState {
    std::vector<std::vector<TuneChoice>> dims;  // per dimension: the base with that dimension varied (filled lazily)
    std::vector<int> dim_kind;                  // 0 schedule, 1 threads, 2 dense, 3 max words, 4 kernel
    size_t dim = 0;                             // dimension being explored
    std::vector<TuneChoice> cands;              // configurations measured or in flight, in order
    std::vector<double> ms;                     // their word-step time, -1 until measured
    TuneChoice base;                            // the defaults, then the best so far
    int planned = 1;                            // distinct configurations the descent will measure
    int best() const {
      int b = -1;
      for (size_t i = 0; i < ms.size(); ++i)
        if (ms[i] >= 0 && (b < 0 || ms[i] < ms[size_t(b)])) b = int(i);
      return b;
    }
    // The variants of dimension `kind` around `b`.
    static std::vector<TuneChoice> variants(int kind, const TuneChoice &b) {
      std::vector<TuneChoice> v;
      auto add = [&](TuneChoice c) { v.push_back(c); };
      if (kind == 0) {
        const int sched[4][2] = {{1, 2}, {1, 4}, {2, 2}, {2, 4}};
        for (auto &sc : sched) {
          TuneChoice c = b;
          c.first = sc[0];
          c.growth = sc[1];
          add(c);
        }
      } else if (kind == 1) {
        for (int t : {0, 256, 512}) {
          TuneChoice c = b;
          c.threads = t;
          add(c);
        }
      } else if (kind == 2) {
        for (int d : {24, 16, 32, 0}) {
          TuneChoice c = b;
          c.dense = d;
          add(c);
        }
      } else if (kind == 3) {
        for (int mw : {128, 64, 256}) {
          TuneChoice c = b;
          c.max_words = mw;
          add(c);
        }
      } else {
        for (int k : {0, 1}) {
          TuneChoice c = b;
          c.kernel = k;
          add(c);
        }
      }
      return v;
    }
    // The next run's configuration: an unmeasured variant of the current
    // dimension, else (moving on to later dimensions) one of theirs, else
    // the best measured. Returns its index in cands.
    int pick() {
      for (size_t i = 0; i < ms.size(); ++i)
        if (ms[i] < 0) return int(i);  // a candidate whose run was not a fair sample: again
      while (dim < dim_kind.size()) {
        const int b = best();
        if (b >= 0) base = cands[size_t(b)];
        for (const TuneChoice &c : variants(dim_kind[dim], base)) {
          if (std::find(cands.begin(), cands.end(), c) != cands.end()) continue;
          cands.push_back(c);
          ms.push_back(-1.0);
          return int(cands.size()) - 1;
        }
        ++dim;
      }
      if (cands.empty()) {
        cands.push_back(base);
        ms.push_back(-1.0);
        return 0;
      }
      return best();
    }
    int tried() const {
      int n = 0;
      for (double x : ms) n += x >= 0 ? 1 : 0;
      return n;
    }
  };
  std::map<std::string, TuneState> tuned;
  // Batched good-machine graphs (most recent last) and the job-set keys seen
  // once (a key is captured when it comes again).
  std::vector<std::unique_ptr<BatchGoodGraph>> batch_graphs;
  std::vector<std::string> batch_keys_seen;
  uint64_t setup_events = 0;  // sweep scratch (re)allocations: a run that made one is not timed for the tuner
  std::string tune_key(const std::vector<JobState *> &jobs, const EngineOptions &eo, bool &any_drop) const;
  TuneState &tune_state(const std::string &key, const EngineOptions &eo, bool any_drop);
  bool plan(JobState &j, bool dev_ok);
  // Word steps planned on the device (persistent pair sweep): a ring of
  // in-flight steps, each with its page-locked tables and figures and four
  // events (step start, sweep start, sweep end, step end). A slot is reused
  // after its step's figures have been read back (ds_resolve).
  struct DsSlot {
    bool busy = false;
    std::vector<JobState *> jobs;
    HostBuf<JobDev> h_jobs;
    HostBuf<StepJob> h_sj;
    HostBuf<int64_t> h_stat;
    HostBuf<uint32_t> h_epoch;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool replayed = false;  // the step's launches replayed captured graphs
  };
  // Captured launches of a device-planned step (plan, sweep, detection), by
  // launch signature: every host-side kernel argument of the step. A
  // signature is captured when it comes a second time (as in repeated runs
  // and batches of one job set) and replayed from then on. Kernel nodes only.
  struct StepGraph {
    std::vector<uint64_t> sig;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    ~StepGraph() {
      for (int k = 0; k < 3; ++k) {
        if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
        if (graph[k]) (void)hipGraphDestroy(graph[k]);
      }
    }
  };
  static constexpr size_t kStepGraphs = 32, kStepSeen = 128;
  std::vector<std::unique_ptr<StepGraph>> step_graphs;  // most recent last
  std::vector<uint64_t> step_seen;                      // signature hashes seen once
  static uint64_t sig_hash(const std::vector<uint64_t> &sig) {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t w : sig) h = (h ^ w) * 1099511628211ull;
    return h;
  }
  StepGraph *find_step_graph(const std::vector<uint64_t> &sig) {
    for (size_t i = 0; i < step_graphs.size(); ++i)
      if (step_graphs[i]->sig == sig) {
        std::unique_ptr<StepGraph> keep = std::move(step_graphs[i]);
        step_graphs.erase(step_graphs.begin() + long(i));
        step_graphs.push_back(std::move(keep));
        return step_graphs.back().get();
      }
    return nullptr;
  }
  template <class Fn>
  void capture_segment(Fn &&fn, hipGraph_t &graph, hipGraphExec_t &exec) {
    LSIM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    fn();
    const hipError_t le = hipGetLastError();
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    LSIM_HIP(le);
    LSIM_HIP(e);
    LSIM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  static constexpr int kDsRing = 2;
  DsSlot ds[kDsRing];
  uint64_t ds_next = 0;
  DevBuf<StepJob> d_sjobs;
  DevBuf<int32_t> d_fflag, d_fpos, d_upos, d_hist;
  DevBuf<int64_t> d_scan_sum, d_ds_stat;
  DevBuf<StepLen> d_len;
  DevBuf<SweepDyn> d_dyn;
  DevBuf<int> d_counters;
  DevBuf<uint32_t> d_epoch;
  bool device_steps_ok(const EngineOptions &eo) const {
    return eo.device_steps && eo.mode != ENGINE_PLAIN && eo.sweep_kernel == 0 && !arena;
  }
  void step_device(const std::vector<JobState *> &act, const EngineOptions &eo);
  void ds_resolve(DsSlot &sl);
  void ds_drain();
  template <class Flag>
  void scan(Flag flag, const int64_t *n_p, int64_t n, int32_t *out);
  std::vector<JobState *> to_finalize;  // jobs plan() found finished
  HostBuf<uint8_t> h_final;
  // Finished jobs whose result copies are in flight (finalize_begin) and not
  // yet unpacked (finalize_end): the unpack runs while the next sweep does.
  std::vector<JobState *> fin_js;
  std::vector<size_t> fin_at;
  struct Scatter {
    int64_t compact_off, host_off, bytes;
  };
  std::vector<Scatter> fin_scatter;  // gathered arrays: place in h_stage -> place in h_final
  HostBuf<uint8_t> h_stage;
  DevBuf<uint8_t> d_final;
  DevBuf<CopySeg> d_segs;
  std::chrono::steady_clock::time_point fin_t0;
  hipEvent_t fin_ev = nullptr;
  void finalize_begin(const std::vector<JobState *> &js);
  void finalize_end();
  void step_batch(const std::vector<JobState *> &act, const EngineOptions &eo);
  void sweep(const EngineOptions &eo, const Lane *lanes, int64_t nlanes, int64_t ntiles, const std::vector<Tile> *tiles,
             int lvl_entries, int64_t nl_max, const CircuitRes *plain_circuit, bool dense, bool nopend,
             bool narrow, lsim_run_stats &st);
  void sweep_pairs(const EngineOptions &eo, int64_t ntiles, int lvl_entries, int64_t nl_max, bool dense, bool nopend,
                   bool narrow, unsigned long long *tstats, lsim_run_stats &st);
  void ensure_slots(const EngineOptions &eo, int64_t want, int64_t nl_max, int per_cu);
  void ensure_pair_scratch(const EngineOptions &eo, int64_t nl_max, int64_t slots_needed, int64_t full_slots);
  void release_slots();
  void release_pair_scratch();
  int64_t scratch_budget(const EngineOptions &eo, int64_t held) const;
};

Engine::Engine(int device) : impl_(new Impl) {
  impl_->device = device;
  int n = device_count();
  if (device < 0 || device >= n) {
    delete impl_;
    impl_ = nullptr;
    throw DeviceError(LSIM_E_DEVICE, "no HIP device " + std::to_string(device) + " (" + std::to_string(n) +
                                         " visible)");
  }
  try {
    LSIM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    LSIM_HIP(hipGetDeviceProperties(&prop, device));
    impl_->cus = prop.multiProcessorCount;
    impl_->lds_cu = int(prop.maxSharedMemoryPerMultiProcessor);
    impl_->lds_block = int(prop.sharedMemPerBlock);
    LSIM_HIP(hipStreamCreateWithFlags(&impl_->stream, hipStreamNonBlocking));
    impl_->holder.s = impl_->stream;
    // Stream-ordered releases stay in the device's pool up to 32 GiB instead
    // of going back at every synchronisation: jobs of successive batches and
    // serving iterations reuse them.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = uint64_t(32) << 30;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    LSIM_HIP(hipEventCreate(&impl_->ev[0]));
    LSIM_HIP(hipEventCreate(&impl_->ev[1]));
    LSIM_HIP(hipEventCreateWithFlags(&impl_->fin_ev, hipEventDisableTiming));
    impl_->d_bytes.alloc(3);
    LSIM_HIP(hipMemsetAsync(impl_->d_bytes.p, 0, 3 * sizeof(unsigned long long), impl_->stream));
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

Engine::Engine(const Engine &host, void *scratch, size_t bytes) : impl_(new Impl) {
  const Impl &h = *host.impl_;
  Impl &m = *impl_;
  m.device = h.device;
  m.cus = h.cus;
  m.lds_cu = h.lds_cu;
  m.lds_block = h.lds_block;
  m.stream = h.stream;  // borrowed: holder stays empty
  m.ev[0] = h.ev[0];
  m.ev[1] = h.ev[1];
  m.fin_ev = h.fin_ev;
  m.borrowed = true;
  m.arena_mem.base = static_cast<uint8_t *>(scratch);
  m.arena_mem.cap = bytes;
  m.arena = &m.arena_mem;
  try {
    LSIM_HIP(hipSetDevice(m.device));
    AllocStream alloc_on(m.stream, m.arena);
    m.d_bytes.alloc(3);
    LSIM_HIP(hipMemsetAsync(m.d_bytes.p, 0, 3 * sizeof(unsigned long long), m.stream));
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

int64_t device_allocations() { return g_allocs.load(std::memory_order_relaxed); }

namespace {
// The no-allocation scratch. In host code `scratch` names a host-side
// stand-in, not device memory: only the per-device address from
// hipGetSymbolAddress is passed to kernels or copies.
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device, null until looked up
}  // namespace

void *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const int n = device_count();
  if (device < 0 || device >= n) throw DeviceError(LSIM_E_DEVICE, "no HIP device " + std::to_string(device));
  if (g_scratch_ptr.size() < size_t(n)) g_scratch_ptr.resize(size_t(n), nullptr);
  if (!g_scratch_ptr[size_t(device)]) {
    int prev = 0;
    LSIM_HIP(hipGetDevice(&prev));
    LSIM_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !ptr)
      throw DeviceError(LSIM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[size_t(device)] = ptr;
  }
  return g_scratch_ptr[size_t(device)];
}

Engine::~Engine() {
  if (impl_) {
    (void)hipSetDevice(impl_->device);
    AllocStream alloc_on(impl_->stream, impl_->arena);
    delete impl_;
  }
}

int Engine::device() const { return impl_->device; }

int64_t Engine::free_bytes() const {
  size_t fr = 0, tot = 0;
  LSIM_HIP(hipSetDevice(impl_->device));
  LSIM_HIP(hipMemGetInfo(&fr, &tot));
  return int64_t(fr);
}

CircuitRes *Engine::add_circuit(const Netlist &nl, const FaultStructure &fs) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  auto cp = std::make_unique<CircuitRes>();
  CircuitRes &c = *cp;
  c.N = nl.N;
  c.P = nl.P;
  c.G = nl.G;
  c.O = nl.O;
  c.E = nl.E();
  c.L = nl.max_level;
  c.has_mux = nl.has_mux;
  const int64_t N = c.N;
  // Device order: live nodes (a PO is reachable) first, dead nodes after, each
  // part sorted by level and then by topological position, so every level is
  // a contiguous range in both parts. (FIFO Kahn order is already level-sorted
  // unless a node reads an unfilled pin as node 0; fanins sit on lower levels,
  // so the level sort stays topological.)
  c.pos2dp.assign(N, 0);
  c.dp2pos.assign(N, 0);
  int64_t nl_count = 0;
  for (int64_t p = 0; p < N; ++p) nl_count += fs.reach_po[p] ? 1 : 0;
  c.NL = nl_count;
  c.live_lvl.assign(c.L + 2, 0);
  c.dead_lvl.assign(c.L + 2, 0);
  {
    std::vector<int64_t> cl(c.L + 2, 0), cd(c.L + 2, 0);
    for (int64_t p = 0; p < N; ++p) (fs.reach_po[p] ? cl : cd)[nl.level[p] + 1]++;
    c.live_lvl[0] = 0;
    c.dead_lvl[0] = int32_t(c.NL);
    for (int l = 0; l <= c.L; ++l) {
      c.live_lvl[l + 1] = int32_t(c.live_lvl[l] + cl[l + 1]);
      c.dead_lvl[l + 1] = int32_t(c.dead_lvl[l] + cd[l + 1]);
    }
    std::vector<int32_t> at_live(c.live_lvl.begin(), c.live_lvl.end() - 1);
    std::vector<int32_t> at_dead(c.dead_lvl.begin(), c.dead_lvl.end() - 1);
    for (int64_t p = 0; p < N; ++p)
      c.pos2dp[p] = fs.reach_po[p] ? at_live[nl.level[p]]++ : at_dead[nl.level[p]]++;
  }
  for (int64_t p = 0; p < N; ++p) c.dp2pos[c.pos2dp[p]] = int32_t(p);
  c.level_dp.assign(N, 0);
  for (int64_t p = 0; p < N; ++p) c.level_dp[c.pos2dp[p]] = nl.level[p];
  {
    int live_levels = 0;
    for (int l = 0; l <= c.L; ++l) live_levels += c.live_lvl[l + 1] > c.live_lvl[l] ? 1 : 0;
    c.wide = c.NL > 256 * int64_t(std::max(live_levels, 1));
  }
  const int64_t nl1 = std::max<int64_t>(c.NL, 1);
  std::vector<uint32_t> info(N);
  std::vector<int4> fin(N);
  c.narrow = true;
  for (int64_t p = 0; p < N; ++p) c.narrow = c.narrow && nl.fin_off[p + 1] - nl.fin_off[p] <= 2;
  std::vector<int32_t> parent(nl1, -1);
  for (int64_t p = 0; p < N; ++p) {
    const int32_t u = c.pos2dp[p];
    const int nin = nl.fin_off[p + 1] - nl.fin_off[p];
    info[u] = uint32_t(nl.type[p]) | (uint32_t(nin) << 4);
    int v[4] = {0, 0, 0, 0};
    if (nl.type[p] == G_PI) v[0] = int(p);  // PI index == position
    for (int k = 0; k < nin; ++k) v[k] = c.pos2dp[nl.fin[nl.fin_off[p] + k]];
    fin[u] = make_int4(v[0], v[1], v[2], v[3]);
    if (fs.reach_po[p] && fs.parent[p] >= 0) parent[u] = c.pos2dp[fs.parent[p]];
  }
  // Live fanouts of live nodes, ascending position, each with the fanout's
  // level in bits 0..23 and, in bits 24..27, the input pins of the fanout
  // that read the node (a node feeding two pins of one gate is listed once).
  if (c.L >= (1 << 24)) throw DeviceError(LSIM_E_INPUT, "circuit too deep (levels must fit in 24 bits)");
  std::vector<int32_t> fo_off(nl1 + 1, 0);
  std::vector<int2> fo;
  std::vector<int32_t> outs;
  for (int64_t u = 0; u < c.NL; ++u) {
    const int32_t p = c.dp2pos[u];
    outs.clear();
    for (int32_t e = fs.fo_off[p]; e < fs.fo_off[p + 1]; ++e)
      if (fs.reach_po[fs.fo[e]]) outs.push_back(c.pos2dp[fs.fo[e]]);
    std::sort(outs.begin(), outs.end());
    for (int32_t v : outs) {
      const int4 f = fin[v];
      const int fv[4] = {f.x, f.y, f.z, f.w};
      const int nin = int((info[v] >> 4) & 7u);
      uint32_t pins = 0;
      for (int k = 0; k < nin; ++k) pins |= fv[k] == int(u) ? (1u << k) : 0u;
      if (!pins) throw DeviceError(LSIM_E_DEVICE, "internal: a live fanout does not read the node");
      fo.push_back(make_int2(v, int32_t(uint32_t(c.level_dp[v]) | (pins << 24))));
    }
    fo_off[u + 1] = int32_t(fo.size());
  }
  // info bits 8..31: level of the node's deepest live fanout (the last one,
  // as positions ascend with level), which bounds how far a sweep must walk.
  for (int64_t u = 0; u < c.NL; ++u)
    if (fo_off[u + 1] > fo_off[u]) info[u] |= uint32_t(c.level_dp[fo[fo_off[u + 1] - 1].x]) << 8;
  if (fo.empty()) fo.push_back(make_int2(0, 0));
  c.head_is_po.assign(nl1, 0);
  for (int64_t p = 0; p < N; ++p)
    if (nl.type[p] == G_PO) c.head_is_po[c.pos2dp[p]] = 1;
  c.head_dp.assign(N, -1);
  for (int64_t p = 0; p < N; ++p)
    if (fs.reach_po[p]) c.head_dp[p] = c.pos2dp[fs.stem_of[p]];

  c.info.alloc(N);
  c.fin.alloc(N);
  std::vector<int4> nrec;
  if (c.narrow) {
    // One 16-byte load per scanned node in the pair sweep instead of four.
    nrec.resize(size_t(c.NL) + 1);
    for (int64_t u = 0; u < c.NL; ++u)
      nrec[size_t(u)] = make_int4(int(info[size_t(u)]), fin[size_t(u)].x, fin[size_t(u)].y, fo_off[size_t(u)]);
    nrec[size_t(c.NL)] = make_int4(0, 0, 0, fo_off[size_t(c.NL)]);
    c.nrec.alloc(nrec.size());
  }
  c.parent.alloc(nl1);
  c.d_live_lvl.alloc(c.live_lvl.size());
  c.d_dead_lvl.alloc(c.dead_lvl.size());
  c.fo_off.alloc(fo_off.size());
  c.fo.alloc(fo.size());
  c.is_po.alloc(nl1);
  m.h2d(c.info.p, info.data(), N * sizeof(uint32_t));
  m.h2d(c.fin.p, fin.data(), N * sizeof(int4));
  if (!nrec.empty()) m.h2d(c.nrec.p, nrec.data(), nrec.size() * sizeof(int4));
  m.h2d(c.parent.p, parent.data(), parent.size() * sizeof(int32_t));
  m.h2d(c.d_live_lvl.p, c.live_lvl.data(), c.live_lvl.size() * sizeof(int32_t));
  m.h2d(c.d_dead_lvl.p, c.dead_lvl.data(), c.dead_lvl.size() * sizeof(int32_t));
  m.h2d(c.fo_off.p, fo_off.data(), fo_off.size() * sizeof(int32_t));
  m.h2d(c.fo.p, fo.data(), fo.size() * sizeof(int2));
  m.h2d(c.is_po.p, c.head_is_po.data(), c.head_is_po.size());
  LSIM_HIP(hipStreamSynchronize(m.stream));
  c.h_info = std::move(info);
  c.h_fin = std::move(fin);
  m.circuits.push_back(std::move(cp));
  return m.circuits.back().get();
}

CircuitRes *Engine::share_circuit(const CircuitRes *src) {
  Impl &m = *impl_;
  auto cp = std::make_unique<CircuitRes>();
  CircuitRes &c = *cp;
  c.N = src->N;
  c.NL = src->NL;
  c.P = src->P;
  c.G = src->G;
  c.O = src->O;
  c.E = src->E;
  c.L = src->L;
  c.h_info = src->h_info;
  c.h_fin = src->h_fin;
  c.wide = src->wide;
  c.narrow = src->narrow;
  c.has_mux = src->has_mux;
  c.pos2dp = src->pos2dp;
  c.dp2pos = src->dp2pos;
  c.level_dp = src->level_dp;
  c.live_lvl = src->live_lvl;
  c.dead_lvl = src->dead_lvl;
  c.head_dp = src->head_dp;
  c.head_is_po = src->head_is_po;
  // The device arrays are read only: views of src's, never freed here.
  auto share = [](auto &dst, const auto &from) { dst.view(from.p, from.n); };
  share(c.info, src->info);
  share(c.fin, src->fin);
  share(c.nrec, src->nrec);
  share(c.parent, src->parent);
  share(c.d_live_lvl, src->d_live_lvl);
  share(c.d_dead_lvl, src->d_dead_lvl);
  share(c.fo_off, src->fo_off);
  share(c.fo, src->fo);
  share(c.is_po, src->is_po);
  m.circuits.push_back(std::move(cp));
  return m.circuits.back().get();
}

void Engine::remove_circuit(CircuitRes *c) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  LSIM_HIP(hipStreamSynchronize(m.stream));
  // Batched graphs and remembered job sets that name the circuit go with it.
  for (size_t i = m.batch_graphs.size(); i-- > 0;) {
    const auto &cs = m.batch_graphs[i]->circuits;
    if (std::find(cs.begin(), cs.end(), static_cast<const void *>(c)) != cs.end())
      m.batch_graphs.erase(m.batch_graphs.begin() + long(i));
  }
  m.batch_keys_seen.clear();
  for (auto it = m.circuits.begin(); it != m.circuits.end(); ++it)
    if (it->get() == c) {
      m.circuits.erase(it);
      return;
    }
}

int64_t Engine::circuit_bytes(const CircuitRes *c) const { return c->bytes(); }

namespace {
// Stable counting sort of j.live by injection unit (site or region head, both
// live positions in [0, NL)).
void order_live(JobState &j) {
  const int64_t NL = j.c->NL;
  const std::vector<int32_t> &key = j.direct ? j.site : j.head;
  std::vector<int32_t> cnt(size_t(NL) + 1, 0);
  for (int32_t f : j.live) cnt[size_t(key[size_t(f)]) + 1]++;
  for (int64_t u = 0; u < NL; ++u) cnt[size_t(u) + 1] += cnt[size_t(u)];
  std::vector<int32_t> sorted(j.live.size());
  for (int32_t f : j.live) sorted[size_t(cnt[size_t(key[size_t(f)])]++)] = f;
  j.live.swap(sorted);
  j.live_by_site = j.direct;
  // Units of a step over every live fault: one per region with a non-PO head.
  j.stems_cap = 0;
  int32_t prev = -1;
  for (int32_t f : j.live) {
    const int32_t h = j.head[size_t(f)];
    if (h != prev && !j.c->head_is_po[size_t(h)]) ++j.stems_cap;
    prev = h;
  }
}
}  // namespace

int64_t Engine::noalloc_min_bytes(const CircuitRes *c, int64_t faults, int64_t R, bool with_last, bool matrix) {
  const int64_t nl = std::max<int64_t>(c->NL, 1);
  const int64_t slot = nl * std::max<int64_t>(int64_t(sizeof(PairRec)) + 4 + 64 * 8,
                                              64 * 8 + 8 + int64_t(sizeof(MaskRec)) + 4);
  return c->bytes() + job_bytes(c, faults, R, !matrix) + (with_last ? faults : 0) + 2 * slot + (int64_t(1) << 20);
}

int64_t Engine::job_bytes(const CircuitRes *c, int64_t faults, int64_t R, bool drop) {
  const int64_t RB = (R + 63) / 64;
  return 8 * c->N * R + 8 * c->P * R + 4 * std::max<int64_t>(c->NL, 1) + faults * (4 + 4 + 1 + 4 + 4 + 1 + 16) +
         (drop ? 0 : 8 * faults * RB);
}

JobState *Engine::add_job(CircuitRes *c, std::vector<int32_t> fault_site, std::vector<uint8_t> fault_sa,
                          int64_t num_patterns, int64_t R, const std::vector<uint64_t> &patterns,
                          const JobOptions &opt) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  auto jp = std::make_unique<JobState>();
  JobState &j = *jp;
  j.c = c;
  j.opt = opt;
  j.direct = opt.direct || (c->has_mux && !opt.mux_safe);
  j.F = int64_t(fault_site.size());
  j.R = R;
  j.num_patterns = num_patterns;
  j.RB = (R + 63) / 64;
  j.site.resize(j.F);
  j.head.resize(j.F);
  j.sa = std::move(fault_sa);
  for (int64_t f = 0; f < j.F; ++f) {
    const int32_t p = fault_site[f];
    j.site[f] = c->pos2dp[p];
    j.head[f] = c->head_dp[p];
    if (j.head[f] >= 0) j.live.push_back(int32_t(f));
  }
  // PI words transposed to [PI][round] for coalesced level-0 reads.
  std::vector<uint64_t> piw(std::max<int64_t>(c->P * R, 1));
  for (int64_t r = 0; r < R; ++r)
    for (int64_t i = 0; i < c->P; ++i) piw[i * R + r] = patterns[r * c->P + i];
  const int64_t F1 = std::max<int64_t>(j.F, 1);
  // Every per-job array in one allocation (one release per job).
  {
    size_t total = 0;
    auto take = [&](size_t bytes) {
      const size_t at = total;
      total += (bytes + 255) & ~size_t(255);
      return at;
    };
    const size_t n_good = size_t(std::max<int64_t>(c->N * R, 1)), n_hs = size_t(std::max<int64_t>(c->NL, 1));
    const size_t n_det = opt.drop ? 0 : size_t(std::max<int64_t>(j.F * j.RB, 1));
    const bool with_last = opt.drop && opt.report_last_round;
    const size_t L1 = std::max<size_t>(j.live.size(), 1);
    const size_t o_good = take(n_good * 8), o_piw = take(piw.size() * 8), o_site = take(size_t(F1) * 4),
                 o_head = take(size_t(F1) * 4), o_sa = take(size_t(F1)), o_first = take(size_t(F1) * 4),
                 o_hs = take(n_hs * 4), o_det = take(n_det * 8), o_last = take(with_last ? size_t(F1) : 0),
                 o_live = take(L1 * 4), o_fl0 = take(L1 * 4), o_fl1 = take(L1 * 4), o_units = take(L1 * 4),
                 o_cnt = take(2 * 4);
    if (!m.arenas.take(total, j.arena)) {
      try {
        j.arena.alloc(total);
      } catch (const DeviceError &e) {
        // Out of memory with released jobs' memory cached: free it and retry.
        if (e.code != LSIM_E_NOMEM || m.arenas.blocks.empty()) throw;
        (void)hipGetLastError();
        m.arenas.clear();
        LSIM_HIP(hipStreamSynchronize(m.stream));
        hipMemPool_t mp = nullptr;
        if (hipDeviceGetDefaultMemPool(&mp, m.device) == hipSuccess && mp) (void)hipMemPoolTrimTo(mp, 0);
        j.arena.alloc(total);
      }
    }
    uint8_t *b = j.arena.p;
    j.good.view(reinterpret_cast<uint64_t *>(b + o_good), n_good);
    j.pi_words.view(reinterpret_cast<uint64_t *>(b + o_piw), piw.size());
    j.d_site.view(reinterpret_cast<int32_t *>(b + o_site), size_t(F1));
    j.d_head.view(reinterpret_cast<int32_t *>(b + o_head), size_t(F1));
    j.d_sa.view(b + o_sa, size_t(F1));
    j.first.view(reinterpret_cast<int32_t *>(b + o_first), size_t(F1));
    j.head_slot.view(reinterpret_cast<int32_t *>(b + o_hs), n_hs);
    if (n_det) j.detbits.view(reinterpret_cast<uint64_t *>(b + o_det), n_det);
    if (with_last) j.last.view(b + o_last, size_t(F1));
    j.d_live.view(reinterpret_cast<int32_t *>(b + o_live), L1);
    j.d_fl0.view(reinterpret_cast<int32_t *>(b + o_fl0), L1);
    j.d_fl1.view(reinterpret_cast<int32_t *>(b + o_fl1), L1);
    j.d_units.view(reinterpret_cast<int32_t *>(b + o_units), L1);
    j.d_cnt.view(reinterpret_cast<int32_t *>(b + o_cnt), 2);
  }
  m.h2d(j.pi_words.p, piw.data(), piw.size() * sizeof(uint64_t));
  m.h2d(j.d_site.p, j.site.data(), j.F * sizeof(int32_t));
  m.h2d(j.d_head.p, j.head.data(), j.F * sizeof(int32_t));
  m.h2d(j.d_sa.p, j.sa.data(), j.F);
  k_fill_i32<<<blocks_for(F1, 256), 256, 0, m.stream>>>(j.first.p, F1, INT32_MAX);
  k_fill_i32<<<blocks_for(int64_t(j.head_slot.n), 256), 256, 0, m.stream>>>(j.head_slot.p, int64_t(j.head_slot.n),
                                                                             -1);
  LSIM_HIP(hipGetLastError());
  if (j.detbits.p) LSIM_HIP(hipMemsetAsync(j.detbits.p, 0, j.detbits.bytes(), m.stream));
  if (j.last.p) LSIM_HIP(hipMemsetAsync(j.last.p, 0, j.last.bytes(), m.stream));
  order_live(j);
  m.h2d(j.d_live.p, j.live.data(), j.live.size() * sizeof(int32_t));
  j.alive = j.live;
  j.Wt = std::max(1, std::min(64, opt.first_words));
  j.Wg = std::max(1, std::min(64, opt.word_growth));
  j.res.stats.live_nodes = c->NL;
  j.res.stats.live_faults = int64_t(j.live.size());
  m.jobs.push_back(std::move(jp));
  return m.jobs.back().get();
}

void Engine::release_job(JobState *j) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  // The job's device memory is one stream-ordered allocation: it goes to the
  // engine's cache, where only work queued later on the same stream reuses
  // it, so no wait.
  AllocStream alloc_on(m.stream, m.arena);
  m.ds_drain();
  for (auto it = m.jobs.begin(); it != m.jobs.end(); ++it)
    if (it->get() == j) {
      m.arenas.put(j->arena);
      m.jobs.erase(it);
      return;
    }
}

bool Engine::reset_job(JobState *j, const JobOptions &opt) {
  Impl &m = *impl_;
  if (opt.drop != j->opt.drop || (opt.drop && opt.report_last_round != j->opt.report_last_round)) return false;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  j->opt = opt;
  j->direct = opt.direct || (j->c->has_mux && !opt.mux_safe);
  j->started = j->done = j->last_done = false;
  j->w0 = 0;
  j->Wt = std::max(1, std::min(64, opt.first_words));
  j->Wg = std::max(1, std::min(64, opt.word_growth));
  j->kind = JobState::MATRIX;
  j->step_faults.clear();
  j->units.clear();
  j->step_w0 = 0;
  j->step_Wn = 0;
  m.ds_drain();
  j->dev = false;
  j->dev_steps = 0;
  j->dev_alive = -1;
  if (j->direct != j->live_by_site) {
    order_live(*j);
    m.h2d(j->d_live.p, j->live.data(), j->live.size() * sizeof(int32_t));
  }
  j->alive = j->live;
  j->res = RunResults();
  j->res.stats.live_nodes = j->c->NL;
  j->res.stats.live_faults = int64_t(j->live.size());
  const int64_t F1 = std::max<int64_t>(j->F, 1);
  k_fill_i32<<<blocks_for(F1, 256), 256, 0, m.stream>>>(j->first.p, F1, INT32_MAX);
  k_fill_i32<<<blocks_for(int64_t(j->head_slot.n), 256), 256, 0, m.stream>>>(j->head_slot.p, int64_t(j->head_slot.n),
                                                                              -1);
  LSIM_HIP(hipGetLastError());
  if (j->detbits.p) LSIM_HIP(hipMemsetAsync(j->detbits.p, 0, j->detbits.bytes(), m.stream));
  if (j->last.p) LSIM_HIP(hipMemsetAsync(j->last.p, 0, j->last.bytes(), m.stream));
  return true;
}

bool Engine::done(const JobState *j) const { return j->done; }
RunResults &Engine::results(JobState *j) { return j->res; }

JobDev Engine::Impl::job_dev(JobState &j) const {
  const CircuitRes &c = *j.c;
  JobDev d;
  std::memset(&d, 0, sizeof(d));
  d.info = c.info.p;
  d.fin = c.fin.p;
  d.nrec = c.nrec.p;
  d.live_lvl = c.d_live_lvl.p;
  d.dead_lvl = c.d_dead_lvl.p;
  d.fo_off = c.fo_off.p;
  d.fo = c.fo.p;
  d.parent = c.parent.p;
  d.is_po = c.is_po.p;
  d.max_level = c.L;
  d.NL = int(c.NL);
  d.good = j.good.p;
  d.pi_words = j.pi_words.p;
  d.RW = j.R;
  d.num_patterns = j.num_patterns;
  d.site = j.d_site.p;
  d.head = j.d_head.p;
  d.sa = j.d_sa.p;
  d.head_slot = j.head_slot.p;
  d.RB = j.RB;
  d.last_word = j.R - 1;
  d.direct = j.direct ? 1 : 0;
  return d;
}

namespace {
constexpr size_t kMaxGraphs = 4;

// Items (node, word) of level l for a job with R rounds on circuit c.
int64_t good_items(const CircuitRes &c, int64_t R, int l) {
  if (R <= 0 || l > c.L) return 0;
  return int64_t(c.live_lvl[l + 1] - c.live_lvl[l] + c.dead_lvl[l + 1] - c.dead_lvl[l]) * R;
}
}  // namespace

// Cuts the levels into windows of at most `levels` levels and splits every
// window into blocks of at most `cap` nodes closed under in-window fanins: the
// window's outputs (nodes with no fanout inside it) are taken in position
// order and each adds its in-window backward cone to the current block, which
// is closed when the cone would overflow it. A window whose largest cone
// exceeds `cap` is cut one level shorter until every cone fits (a one-level
// window's cones are single nodes). A node's first block owns it.
void Engine::Impl::build_window_plan(CircuitRes &c, int levels, int cap) {
  const auto t0 = std::chrono::steady_clock::now();
  auto plan = std::make_unique<WindowPlan>();
  plan->levels = levels;
  plan->cap = cap;
  const int64_t N = c.N;
  auto fanins = [&](int32_t u, int32_t *f) {
    const uint32_t type = c.h_info[size_t(u)] & 15u, nin = (c.h_info[size_t(u)] >> 4) & 7u;
    if (type == G_PI) return 0;
    const int4 v = c.h_fin[size_t(u)];
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
    f[3] = v.w;
    return int(nin);
  };
  // A node is an output of its window when its lowest fanout level lies
  // beyond the window (fanouts sit on higher levels than the node).
  std::vector<int32_t> min_fo_level(size_t(N), INT32_MAX);
  for (int32_t u = 0; u < int32_t(N); ++u) {
    int32_t f[4];
    const int k = fanins(u, f);
    for (int q = 0; q < k; ++q)
      min_fo_level[size_t(f[q])] = std::min(min_fo_level[size_t(f[q])], c.level_dp[size_t(u)]);
  }
  std::vector<WinBlock> blocks;
  std::vector<WinNode> nodes;
  std::vector<uint8_t> owned(size_t(N), 0);
  std::vector<int32_t> in_block(size_t(N), -1);  // index in the current block, by block id stamp
  std::vector<int32_t> stamp(size_t(N), -1), cone_stamp(size_t(N), -1);
  std::vector<int32_t> cur, stack, outs, cone_nodes, cone_off;
  int32_t block_id = 0, cone_id = 0;
  auto flush = [&](int l0, int depth) {
    if (cur.empty()) return;
    std::sort(cur.begin(), cur.end(), [&](int32_t a, int32_t b) {
      const int32_t la = c.level_dp[size_t(a)], lb = c.level_dp[size_t(b)];
      return la != lb ? la < lb : a < b;
    });
    WinBlock B;
    B.node0 = int32_t(nodes.size());
    for (size_t i = 0; i < cur.size(); ++i) in_block[size_t(cur[i])] = int32_t(i);
    int k = 0;
    B.seg[0] = 0;
    for (size_t i = 0; i < cur.size(); ++i) {
      const int32_t u = cur[i];
      while (k < depth - 1 && c.level_dp[size_t(u)] > l0 + k) B.seg[++k] = int32_t(i);
      WinNode n;
      n.u = u;
      const bool own = !owned[size_t(u)];
      owned[size_t(u)] = 1;
      n.info = (c.h_info[size_t(u)] & 0xffu) | (own ? (1u << 8) : 0u);
      int32_t f[4] = {0, 0, 0, 0};
      const int nin = fanins(u, f);
      int e[4] = {0, 0, 0, 0};
      if ((c.h_info[size_t(u)] & 15u) == G_PI) {
        e[0] = c.h_fin[size_t(u)].x;
      } else {
        for (int q = 0; q < nin; ++q)
          e[q] = (stamp[size_t(f[q])] == block_id) ? ~in_block[size_t(f[q])] : f[q];
      }
      n.fin = make_int4(e[0], e[1], e[2], e[3]);
      nodes.push_back(n);
    }
    while (k < kWinLevelsMax) B.seg[++k] = int32_t(cur.size());
    blocks.push_back(B);
    cur.clear();
    ++block_id;
  };
  for (int l0 = 0; l0 <= c.L;) {
    // The window's outputs and their in-window backward cones; a window
    // whose largest cone outgrows a block is cut a level shorter (a one-level
    // window's cones are single nodes).
    int depth = std::min(levels, c.L - l0 + 1);
    for (;;) {
      const int l1 = l0 + depth - 1;
      outs.clear();
      for (int l = l0; l <= l1; ++l) {
        for (int32_t u = c.live_lvl[size_t(l)]; u < c.live_lvl[size_t(l) + 1]; ++u)
          if (min_fo_level[size_t(u)] > l1) outs.push_back(u);
        for (int32_t u = c.dead_lvl[size_t(l)]; u < c.dead_lvl[size_t(l) + 1]; ++u)
          if (min_fo_level[size_t(u)] > l1) outs.push_back(u);
      }
      std::sort(outs.begin(), outs.end());
      cone_nodes.clear();
      cone_off.assign(1, 0);
      bool fits = true;
      for (int32_t o : outs) {
        stack.clear();
        stack.push_back(o);
        cone_stamp[size_t(o)] = cone_id;
        const size_t first = cone_nodes.size();
        while (!stack.empty()) {
          const int32_t u = stack.back();
          stack.pop_back();
          cone_nodes.push_back(u);
          int32_t f[4];
          const int k = fanins(u, f);
          for (int q = 0; q < k; ++q)
            if (c.level_dp[size_t(f[q])] >= l0 && cone_stamp[size_t(f[q])] != cone_id) {
              cone_stamp[size_t(f[q])] = cone_id;
              stack.push_back(f[q]);
            }
        }
        ++cone_id;
        if (cone_nodes.size() - first > size_t(cap)) {
          fits = false;
          break;
        }
        cone_off.push_back(int32_t(cone_nodes.size()));
      }
      if (fits) break;
      if (depth == 1) throw DeviceError(LSIM_E_DEVICE, "window block capacity below one node");
      --depth;
    }
    plan->win_block.push_back(int32_t(blocks.size()));
    plan->win_depth.push_back(depth);
    for (size_t oi = 0; oi + 1 < cone_off.size(); ++oi) {
      const int32_t *cn = cone_nodes.data() + cone_off[oi];
      const int64_t cs = cone_off[oi + 1] - cone_off[oi];
      int64_t fresh = 0;
      for (int64_t q = 0; q < cs; ++q) fresh += stamp[size_t(cn[q])] == block_id ? 0 : 1;
      if (int64_t(cur.size()) + fresh > cap) flush(l0, depth);
      for (int64_t q = 0; q < cs; ++q)
        if (stamp[size_t(cn[q])] != block_id) {
          stamp[size_t(cn[q])] = block_id;
          cur.push_back(cn[q]);
        }
    }
    flush(l0, depth);
    l0 += depth;
  }
  plan->win_block.push_back(int32_t(blocks.size()));
  for (int64_t u = 0; u < N; ++u)
    if (!owned[size_t(u)]) throw DeviceError(LSIM_E_DEVICE, "window partition left a node without a block");
  plan->repeated = int64_t(nodes.size()) - N;
  plan->blocks.alloc(std::max<size_t>(blocks.size(), 1));
  plan->nodes.alloc(std::max<size_t>(nodes.size(), 1));
  h2d(plan->blocks.p, blocks.data(), blocks.size() * sizeof(WinBlock));
  h2d(plan->nodes.p, nodes.data(), nodes.size() * sizeof(WinNode));
  LSIM_HIP(hipStreamSynchronize(stream));
  if (trace_on())
    std::fprintf(stderr, "[lsim] window plan: %zu windows of up to %d levels, %zu blocks of up to %d nodes (%zu KB LDS), "
                 "%lld nodes repeated, built in %.1f ms\n", plan->win_depth.size(), levels, blocks.size(), cap,
                 win_lds_bytes(cap) / 1024, (long long)plan->repeated,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  c.win = std::move(plan);
}

int Engine::Impl::win_cap(const EngineOptions &eo) const {
  // A block's values (cap nodes x 64 words x 8 B) in LDS: half the CU's LDS,
  // so two blocks are resident per CU, within the workgroup limit.
  const int per_node = int(win_lds_bytes(1));
  const int by_lds = std::max(1, std::min(lds_block, lds_cu / 2) / per_node);
  return eo.win_cap > 0 ? std::min(eo.win_cap, std::max(1, lds_block / per_node)) : by_lds;
}

GoodGraph *Engine::Impl::good_graph(CircuitRes &c, int64_t R, int form, const EngineOptions &eo) {
  const int wl = form == 1 ? eo.win_levels : 0, wc = form == 1 ? win_cap(eo) : 0;
  for (size_t i = 0; i < c.graphs.size(); ++i)
    if (c.graphs[i]->R == R && c.graphs[i]->form == form && c.graphs[i]->win_levels == wl &&
        c.graphs[i]->win_cap == wc) {
      std::unique_ptr<GoodGraph> g = std::move(c.graphs[i]);
      c.graphs.erase(c.graphs.begin() + long(i));
      c.graphs.push_back(std::move(g));
      return c.graphs.back().get();
    }
  auto g = std::make_unique<GoodGraph>();
  g->R = R;
  g->form = form;
  g->win_levels = wl;
  g->win_cap = wc;
  g->job.alloc(1);
  std::vector<int64_t> off(size_t(c.L + 1) * 2, 0);
  for (int l = 0; l <= c.L; ++l) off[size_t(l) * 2 + 1] = good_items(c, R, l);
  g->off.alloc(off.size());
  h2d(g->off.p, off.data(), off.size() * sizeof(int64_t));
  const size_t win_lds = win_lds_bytes(wc);
  if (form == 1) {
    if (!c.win || c.win->levels != wl || c.win->cap != wc) build_window_plan(c, wl, wc);
    LSIM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_good_window),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(win_lds)));
  }
  LSIM_HIP(hipStreamSynchronize(stream));
  LSIM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  if (form == 1) {
    const int wg = int((R + 63) / 64);
    const WindowPlan &wp = *c.win;
    for (size_t k = 0; k + 1 < wp.win_block.size(); ++k) {
      const int nb = wp.win_block[k + 1] - wp.win_block[k];
      if (!nb) continue;
      k_good_window<<<unsigned(int64_t(nb) * wg), 512, win_lds, stream>>>(g->job.p, wp.blocks.p, wp.nodes.p,
                                                                         wp.win_block[k], wg, wp.win_depth[k], wc);
    }
  } else {
    for (int l = 0; l <= c.L; ++l) {
      const int64_t total = off[size_t(l) * 2 + 1];
      if (!total) continue;
      k_good_level<<<blocks_for(total, 256), 256, 0, stream>>>(g->job.p, 1, g->off.p + size_t(l) * 2, l);
    }
  }
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &graph);
  if (e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    LSIM_HIP(e);
  }
  g->graph = graph;
  LSIM_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  if (c.graphs.size() >= kMaxGraphs) c.graphs.erase(c.graphs.begin());
  c.graphs.push_back(std::move(g));
  return c.graphs.back().get();
}

void Engine::start(const std::vector<JobState *> &jobs, const EngineOptions &eo) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  std::vector<JobState *> js;
  for (JobState *j : jobs)
    if (!j->started) js.push_back(j);
  if (js.empty()) return;
  if (js.size() == 1 && eo.good_graph && js[0]->R > 0) {
    // One job: replay its circuit's captured good-machine graph.
    JobState &j = *js[0];
    GoodGraph *g = m.good_graph(*j.c, j.R, eo.good_form, eo);
    const JobDev d = m.job_dev(j);
    m.time_begin();
    m.h2d(g->job.p, &d, sizeof(JobDev));
    LSIM_HIP(hipGraphLaunch(g->exec, m.stream));
    const double ms = m.time_end();
    m.finish_start(js, ms);
    j.res.stats.good_replayed = 1;
    return;
  }
  // One launch per level index over all jobs' (node, word) pairs of that level.
  const int nj = int(js.size());
  int Lmax = 0;
  for (JobState *j : js) Lmax = std::max(Lmax, j->R > 0 ? j->c->L : 0);
  std::vector<JobDev> devs(nj);
  for (int i = 0; i < nj; ++i) devs[i] = m.job_dev(*js[i]);
  std::vector<int64_t> off(size_t(Lmax + 1) * (nj + 1), 0);
  for (int l = 0; l <= Lmax; ++l) {
    int64_t *o = &off[size_t(l) * (nj + 1)];
    for (int i = 0; i < nj; ++i) {
      const CircuitRes &c = *js[i]->c;
      int64_t items = 0;
      if (js[i]->R > 0 && l <= c.L)
        items = int64_t(c.live_lvl[l + 1] - c.live_lvl[l] + c.dead_lvl[l + 1] - c.dead_lvl[l]) * js[i]->R;
      o[i + 1] = o[i] + items;
    }
  }
  if (eo.good_graph) {
    // The set's key: every job's circuit (with its shape) and rounds, in order.
    std::string key;
    for (JobState *j : js) {
      char buf[96];
      std::snprintf(buf, sizeof(buf), "%p/%lld/%lld/%d/%lld;", static_cast<const void *>(j->c), (long long)j->c->N,
                    (long long)j->c->NL, j->c->L, (long long)j->R);
      key += buf;
    }
    BatchGoodGraph *g = nullptr;
    for (size_t i = 0; i < m.batch_graphs.size(); ++i)
      if (m.batch_graphs[i]->key == key) {
        std::unique_ptr<BatchGoodGraph> keep = std::move(m.batch_graphs[i]);
        m.batch_graphs.erase(m.batch_graphs.begin() + long(i));
        m.batch_graphs.push_back(std::move(keep));
        g = m.batch_graphs.back().get();
        break;
      }
    if (!g && std::find(m.batch_keys_seen.begin(), m.batch_keys_seen.end(), key) != m.batch_keys_seen.end()) {
      auto ng = std::make_unique<BatchGoodGraph>();
      ng->key = key;
      for (JobState *j : js) ng->circuits.push_back(j->c);
      ng->jobs.alloc(nj);
      ng->off.alloc(off.size());
      m.h2d(ng->off.p, off.data(), off.size() * sizeof(int64_t));
      LSIM_HIP(hipStreamSynchronize(m.stream));
      LSIM_HIP(hipStreamBeginCapture(m.stream, hipStreamCaptureModeThreadLocal));
      for (int l = 0; l <= Lmax; ++l) {
        const int64_t total = off[size_t(l) * (nj + 1) + nj];
        if (!total) continue;
        k_good_level<<<blocks_for(total, 256), 256, 0, m.stream>>>(ng->jobs.p, nj, ng->off.p + size_t(l) * (nj + 1), l);
      }
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(m.stream, &graph);
      if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        LSIM_HIP(e);
      }
      ng->graph = graph;
      LSIM_HIP(hipGraphInstantiate(&ng->exec, ng->graph, nullptr, nullptr, 0));
      if (m.batch_graphs.size() >= kMaxGraphs) m.batch_graphs.erase(m.batch_graphs.begin());
      m.batch_graphs.push_back(std::move(ng));
      g = m.batch_graphs.back().get();
    }
    if (g) {
      m.time_begin();
      m.h2d(g->jobs.p, devs.data(), nj * sizeof(JobDev));
      LSIM_HIP(hipGraphLaunch(g->exec, m.stream));
      m.finish_start(js, m.time_end());
      for (JobState *j : js) j->res.stats.good_replayed = 1;
      return;
    }
    if (std::find(m.batch_keys_seen.begin(), m.batch_keys_seen.end(), key) == m.batch_keys_seen.end()) {
      if (m.batch_keys_seen.size() >= 16) m.batch_keys_seen.erase(m.batch_keys_seen.begin());
      m.batch_keys_seen.push_back(key);
    }
  }
  m.d_jobs.alloc(nj);
  m.d_good_off.alloc(off.size());
  m.h2d(m.d_jobs.p, devs.data(), nj * sizeof(JobDev));
  m.h2d(m.d_good_off.p, off.data(), off.size() * sizeof(int64_t));
  m.time_begin();
  for (int l = 0; l <= Lmax; ++l) {
    const int64_t total = off[size_t(l) * (nj + 1) + nj];
    if (!total) continue;
    k_good_level<<<blocks_for(total, 256), 256, 0, m.stream>>>(m.d_jobs.p, nj, m.d_good_off.p + size_t(l) * (nj + 1),
                                                              l);
    LSIM_HIP(hipGetLastError());
  }
  m.finish_start(js, m.time_end());
}

void Engine::Impl::finish_start(const std::vector<JobState *> &js, double ms) {
  Impl &m = *this;
  for (JobState *j : js) {
    j->res.stats.t_good_ms += ms;
    // One word per node written, one per fanin and per PI pattern read.
    j->res.stats.good_bytes += 8 * (j->c->N + j->c->E + j->c->P) * j->R;
    j->started = true;
    if (j->opt.keep_good_last) {
      const CircuitRes &c = *j->c;
      j->res.good_last.assign(c.N, 0);
      if (j->R > 0) {
        const int64_t bits = (j->num_patterns / (64 * j->R)) ? 64 : (j->num_patterns % 64);
        const uint64_t mask = bits < 64 ? ((uint64_t(1) << bits) - 1) : ~uint64_t(0);
        m.d_gather.alloc(c.N);
        k_gather_round<<<blocks_for(c.N, 256), 256, 0, m.stream>>>(j->good.p, j->R, j->R - 1, int(c.N), mask,
                                                                  m.d_gather.p);
        LSIM_HIP(hipGetLastError());
        std::vector<uint64_t> byd(c.N);
        m.d2h(byd.data(), m.d_gather.p, c.N * sizeof(uint64_t));
        LSIM_HIP(hipStreamSynchronize(m.stream));
        for (int64_t p = 0; p < c.N; ++p) j->res.good_last[p] = byd[c.pos2dp[p]];
      }
    }
  }
}

// Decides the job's next step; false (and the job queued for finalize_begin)
// when none is left.
bool Engine::Impl::plan(JobState &j, bool dev_ok) {
  if (j.done) return false;
  // A run's steps are all planned on the host or all on the device.
  if (j.res.stats.steps == 0) j.dev = dev_ok;
  if (j.R == 0 || j.live.empty()) {
    to_finalize.push_back(&j);
    return false;
  }
  if (!j.opt.drop) {
    if (j.w0 >= j.R) {
      to_finalize.push_back(&j);
      return false;
    }
    j.kind = JobState::MATRIX;
    if (!j.dev) j.step_faults = j.live;
    j.step_w0 = j.w0;
    j.step_Wn = int(std::min<int64_t>(64, j.R - j.w0));
    return true;
  }
  // With steps planned on the device the host learns a step's fault count a
  // step or two later; until it reads 0 the job keeps stepping (an empty step
  // sweeps nothing).
  const bool left = j.dev ? j.dev_alive != 0 : !j.alive.empty();
  if (j.w0 < j.R && left) {
    // Words per step grow geometrically: faults left after the first rounds
    // are the hard ones and are dropped less often.
    j.kind = JobState::DROP;
    if (!j.dev) j.step_faults = j.alive;
    j.step_w0 = j.w0;
    j.step_Wn = int(std::min<int64_t>(j.Wt, j.R - j.w0));
    // A remainder of at most a quarter of this step joins it: a launch for a
    // few last words would cost a whole walk of the deepest tiles.
    const int64_t rest = j.R - j.w0 - j.step_Wn;
    if (rest > 0 && 4 * rest <= j.step_Wn) j.step_Wn = int(j.R - j.w0);
    return true;
  }
  if (j.opt.report_last_round && !j.last_done) {
    j.kind = JobState::LAST;
    if (!j.dev) j.step_faults = j.live;
    j.step_w0 = j.R - 1;
    j.step_Wn = 1;
    return true;
  }
  to_finalize.push_back(&j);
  return false;
}

// Results of finished jobs: every job's device arrays copied into one
// page-locked block (finalize_begin enqueues the copies and records fin_ev),
// then unpacked (finalize_end waits for fin_ev). The iteration that follows
// calls finalize_end once its sweep is launched, so the unpack overlaps it.
void Engine::Impl::finalize_begin(const std::vector<JobState *> &js) {
  finalize_end();
  if (js.empty()) return;
  fin_t0 = std::chrono::steady_clock::now();
  std::vector<size_t> &at = fin_at;
  at.assign(js.size() + 1, 0);
  for (size_t i = 0; i < js.size(); ++i) {
    const JobState &j = *js[i];
    size_t b = size_t(j.F) * sizeof(int32_t);
    if (!j.opt.drop && j.detbits.p) b += size_t(j.F * j.RB) * sizeof(uint64_t);
    if (j.opt.drop && j.last.p && j.last_done) b += size_t(j.F);
    at[i + 1] = at[i] + ((b + 63) & ~size_t(63));
  }
  h_final.alloc(std::max<size_t>(at.back(), 1));
  // Arrays of at least kDirectBytes are copied straight into their place in
  // h_final. The smaller ones are gathered on the device into one compact
  // block (d_final), copied back once into h_stage and moved into place by
  // finalize_end.
  constexpr int64_t kDirectBytes = int64_t(1) << 20;
  std::vector<CopySeg> segs;
  fin_scatter.clear();
  int64_t compact = 0;
  auto put = [&](const uint8_t *src, size_t host_off, int64_t bytes) {
    if (bytes <= 0) return;
    if (bytes >= kDirectBytes) {
      d2h(h_final.p + host_off, src, size_t(bytes));
      return;
    }
    segs.push_back({src, compact, bytes});
    fin_scatter.push_back({compact, int64_t(host_off), bytes});
    compact = (compact + bytes + 7) & ~int64_t(7);
  };
  for (size_t i = 0; i < js.size(); ++i) {
    const JobState &j = *js[i];
    size_t o = at[i];
    const int64_t nf = j.F * int64_t(sizeof(int32_t));
    put(reinterpret_cast<const uint8_t *>(j.first.p), o, nf);
    o += size_t(nf);
    if (!j.opt.drop && j.detbits.p) {
      const int64_t nb = j.F * j.RB * int64_t(sizeof(uint64_t));
      put(reinterpret_cast<const uint8_t *>(j.detbits.p), o, nb);
      o += size_t(nb);
    }
    if (j.opt.drop && j.last.p && j.last_done) put(j.last.p, o, j.F);
  }
  if (!segs.empty()) {
    d_final.alloc(size_t(compact));
    d_segs.alloc(segs.size());
    h_stage.alloc(size_t(compact));
    h2d(d_segs.p, segs.data(), segs.size() * sizeof(CopySeg));
    k_gather_words<<<blocks_for((compact + 3) / 4, 256), 256, 0, stream>>>(d_segs.p, int(segs.size()), compact,
                                                                          d_final.p);
    LSIM_HIP(hipGetLastError());
    d2h(h_stage.p, d_final.p, size_t(compact));
  }
  LSIM_HIP(hipEventRecord(fin_ev, stream));
  fin_js = js;
}

void Engine::Impl::finalize_end() {
  if (fin_js.empty()) return;
  using clk = std::chrono::steady_clock;
  const std::vector<JobState *> js = std::move(fin_js);
  fin_js.clear();
  const std::vector<size_t> &at = fin_at;
  const auto t0 = clk::now();
  LSIM_HIP(hipEventSynchronize(fin_ev));
  for (const Scatter &sc : fin_scatter)
    std::memcpy(h_final.p + sc.host_off, h_stage.p + sc.compact_off, size_t(sc.bytes));
  fin_scatter.clear();
  // The gather block is sized by the small arrays of one finalize; a block
  // left large by an unusual batch is given back.
  if (d_final.bytes() > (size_t(64) << 20)) d_final.release();
  const auto t1 = clk::now();
  // Unpacked in parallel, largest job first. A job of many faults is split
  // into contiguous fault ranges, so that several threads touch and fill its
  // result arrays (sized here without initialisation).
  std::vector<int> order(js.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return js[size_t(x)]->F > js[size_t(y)]->F; });
  struct Piece {
    int job;
    int64_t lo, hi;
  };
  std::vector<Piece> pieces;
  for (int ii : order) {
    JobState &j = *js[size_t(ii)];
    RunResults &res = j.res;
    const int64_t F = j.F;
    res.first_round.resize(size_t(F));
    res.last_round.resize(size_t(F));
    res.RB = j.RB;
    if (!j.opt.drop) res.det_bits.assign(size_t(F * j.RB), 0);
    const int64_t n = std::max<int64_t>(1, std::min<int64_t>(16, F >> 18));
    for (int64_t c = 0; c < n; ++c) pieces.push_back({ii, F * c / n, F * (c + 1) / n});
  }
  std::vector<int> piece_ids(pieces.size());
  std::iota(piece_ids.begin(), piece_ids.end(), 0);
  parallel_for(piece_ids, [&](int pi) {
    const Piece &pc = pieces[size_t(pi)];
    const size_t i = size_t(pc.job);
    JobState &j = *js[i];
    RunResults &res = j.res;
    const int64_t F = j.F, R = j.R, lo = pc.lo, hi = pc.hi;
    const uint8_t *o = h_final.p + at[i];
    const int32_t *first = reinterpret_cast<const int32_t *>(o);
    for (int64_t f = lo; f < hi; ++f) res.first_round[size_t(f)] = first[f] == INT32_MAX ? -1 : first[f];
    o += size_t(F) * sizeof(int32_t);
    if (!j.opt.drop) {
      if (j.detbits.p && hi > lo)
        std::memcpy(res.det_bits.data() + lo * j.RB, o + size_t(lo * j.RB) * sizeof(uint64_t),
                    size_t((hi - lo) * j.RB) * sizeof(uint64_t));
      for (int64_t f = lo; f < hi; ++f)
        res.last_round[size_t(f)] =
            R > 0 ? uint8_t((res.det_bits[size_t(f * j.RB + ((R - 1) >> 6))] >> ((R - 1) & 63)) & 1) : uint8_t(0);
    } else if (j.last.p && j.last_done) {
      if (hi > lo) std::memcpy(res.last_round.data() + lo, o + lo, size_t(hi - lo));
    } else if (hi > lo) {
      std::memset(res.last_round.data() + lo, 0, size_t(hi - lo));
    }
  });
  for (JobState *jp : js) {
    RunResults &res = jp->res;
    res.has_matrix = !jp->opt.drop;
    res.has_last = !jp->opt.drop || jp->opt.report_last_round;
    res.stats.slots = slots;
    jp->done = true;
  }
  if (trace_on()) {
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::fprintf(stderr, "[lsim] finalize jobs=%zu: copies enqueued %.2f ms before, waited %.2f ms, unpack %.2f ms\n",
                 js.size(), ms(fin_t0, t0), ms(t0, t1), ms(t1, clk::now()));
  }
}

int64_t Engine::Impl::scratch_budget(const EngineOptions &eo, int64_t held) const {
  // Without allocation: half of what the arena has left, the rest for the
  // step buffers.
  if (arena) return int64_t(arena->left() / 2);
  if (eo.scratch_mb > 0) return eo.scratch_mb << 20;
  size_t fr = 0, tot = 0;
  LSIM_HIP(hipMemGetInfo(&fr, &tot));
  return (int64_t(fr) + held) / 10 * 7;
}

void Engine::Impl::release_slots() {
  d_rows.release();
  d_masks.release();
  d_mrec.release();
  d_pend.release();
  slots = 0;
  slot_nl = 0;
}

void Engine::Impl::release_pair_scratch() {
  d_prec.release();
  d_ppend.release();
  d_pool.release();
  pair_slots = 0;
  pair_req = 0;
  pair_nl = 0;
}

void Engine::Impl::ensure_pair_scratch(const EngineOptions &eo, int64_t nl_max, int64_t slots_needed,
                                       int64_t full_slots) {
  const int64_t nl = std::max<int64_t>(nl_max, 1);
  if (d_pool.p && pair_nl >= nl && pair_req >= slots_needed) return;
  LSIM_HIP(hipStreamSynchronize(stream));
  const int64_t held = int64_t(d_rows.bytes() + d_masks.bytes() + d_mrec.bytes() + d_pend.bytes() +
                               d_prec.bytes() + d_ppend.bytes() + d_pool.bytes());
  release_slots();
  release_pair_scratch();
  // Cached job memory and what the stream-ordered pool keeps count as free
  // for the scratch budget.
  if (!arena) {
    arenas.clear();
    LSIM_HIP(hipStreamSynchronize(stream));
    hipMemPool_t mp = nullptr;
    if (hipDeviceGetDefaultMemPool(&mp, device) == hipSuccess && mp) (void)hipMemPoolTrimTo(mp, 0);
  }
  const int64_t budget = scratch_budget(eo, held);
  (void)full_slots;
  // Per slot: a record and a pending flag per live node. The pool must hold
  // at least one tile of any size (64 words per live node) and needs no more
  // than that for every slot; a tight budget gets fewer slots.
  const int64_t per_slot = nl * int64_t(sizeof(PairRec) + sizeof(uint32_t));
  const int64_t one_tile = nl * 64 * 8;
  const int64_t fit = (budget - one_tile) / per_slot;
  if (fit < 1)
    throw DeviceError(LSIM_E_NOMEM, "sweep scratch of " + std::to_string(budget >> 20) + " MiB cannot hold a slot of a " +
                                        std::to_string(nl) + "-node circuit");
  const int64_t ns = std::max<int64_t>(1, std::min(slots_needed, fit));
  const int64_t pool_words = std::min((budget - ns * per_slot) / 8, ns * nl * 64);
  ++setup_events;
  const auto t0 = std::chrono::steady_clock::now();
  d_prec.alloc_direct(size_t(ns * nl));
  d_ppend.alloc_direct(size_t(ns * nl));
  d_pool.alloc_direct(size_t(pool_words));
  const auto t1 = std::chrono::steady_clock::now();
  // Epoch 0 is never used by a tile: zeroed records and flags are stale.
  LSIM_HIP(hipMemsetAsync(d_prec.p, 0, d_prec.bytes(), stream));
  LSIM_HIP(hipMemsetAsync(d_ppend.p, 0, d_ppend.bytes(), stream));
  if (trace_on()) {
    LSIM_HIP(hipStreamSynchronize(stream));
    const auto t2 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[lsim] pair scratch: %lld slots x %lld nodes, pool %.1f GiB: alloc %.1f ms, clear %.1f ms\n",
                 (long long)ns, (long long)nl, double(pool_words) * 8 / double(1ull << 30),
                 std::chrono::duration<double, std::milli>(t1 - t0).count(),
                 std::chrono::duration<double, std::milli>(t2 - t1).count());
  }
  pair_slots = ns;
  pair_req = slots_needed;
  pair_nl = nl;
}

void Engine::Impl::ensure_slots(const EngineOptions &eo, int64_t want, int64_t nl_max, int per_cu) {
  if (d_pool.p) {
    LSIM_HIP(hipStreamSynchronize(stream));
    release_pair_scratch();
  }
  // Per slot: faulty rows [NL][64], and the plain form's lane masks or the
  // event form's mask records and pending flags.
  const int64_t nl = std::max<int64_t>(nl_max, 1);
  const int64_t per_slot = nl * (64 * 8 + 8 + int64_t(sizeof(MaskRec)) + 4);
  const int64_t held = int64_t(d_rows.bytes() + d_masks.bytes() + d_mrec.bytes() + d_pend.bytes());
  const int64_t budget = scratch_budget(eo, held);
  const int64_t cap = std::max<int64_t>(1, budget / per_slot);
  // No more slots than workgroups the device keeps resident at once.
  const int64_t s =
      std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(cap, int64_t(cus) * std::max(1, per_cu))));
  if (s > slots || nl > slot_nl || !d_rows.p) {
    const int64_t ns = std::min(cap, std::max(s, slots));
    d_rows.release();
    d_masks.release();
    d_mrec.release();
    d_pend.release();
    ++setup_events;
    d_rows.alloc_direct(size_t(ns) * nl * 64);
    d_masks.alloc_direct(size_t(ns) * nl);
    d_mrec.alloc_direct(size_t(ns) * nl);
    d_pend.alloc_direct(size_t(ns) * nl);
    // Epoch 0 is never used by a tile, so zeroed records read as "no lane differs".
    LSIM_HIP(hipMemsetAsync(d_mrec.p, 0, d_mrec.bytes(), stream));
    LSIM_HIP(hipMemsetAsync(d_pend.p, 0, d_pend.bytes(), stream));
    slots = ns;
    slot_nl = nl;
  }
}

// Lanes and tiles are in d_lanes / d_tiles already, or given on the host
// (`lanes` and `tiles` not null) and copied there first.
void Engine::Impl::sweep(const EngineOptions &eo, const Lane *lanes, int64_t nlanes, int64_t ntiles,
                         const std::vector<Tile> *tiles, int lvl_entries, int64_t nl_max,
                         const CircuitRes *plain_circuit, bool dense, bool nopend, bool narrow,
                         lsim_run_stats &st) {
  if (!ntiles) return;
  if (lanes && tiles) {
    d_lanes.alloc(size_t(nlanes));
    d_tiles.alloc(tiles->size());
    h2d(d_lanes.p, lanes, size_t(nlanes) * sizeof(Lane));
    h2d(d_tiles.p, tiles->data(), tiles->size() * sizeof(Tile));
  } else if (plain_circuit) {
    throw DeviceError(LSIM_E_DEVICE, "the per-level sweep needs its tiles on the host");
  }
  const size_t lds = size_t(lvl_entries) * sizeof(int);
  const bool trace = trace_on();
  if (trace) {
    d_tstats.alloc(kTileStats);
    LSIM_HIP(hipMemsetAsync(d_tstats.p, 0, kTileStats * sizeof(unsigned long long), stream));
  }
  auto print_stats = [&] {
    unsigned long long ts[kTileStats];
    d2h(ts, d_tstats.p, sizeof(ts));
    LSIM_HIP(hipStreamSynchronize(stream));
    const double nt = double(std::max<unsigned long long>(ts[0], 1));
    if (ts[8]) {
      const double cy = double(ts[8]);
      std::fprintf(stderr,
                   "[lsim] pair sweep cycles per tile %.0f: scan and stage %.1f %%, pairs %.1f %%, dense pairs %.1f %%, "
                   "records and fanouts %.1f %%, between passes %.1f %%\n",
                   cy / nt, 100.0 * ts[9] / cy, 100.0 * ts[10] / cy, 100.0 * ts[11] / cy, 100.0 * ts[12] / cy,
                   100.0 * (cy - double(ts[9] + ts[10] + ts[11] + ts[12])) / cy);
    }
    std::fprintf(stderr,
                 "[lsim] tiles %llu: per tile levels walked %.1f, pending levels %.1f, scan passes %.1f, staged %.1f, "
                 "written %.1f (max %llu), active lanes per staged node %.2f\n",
                 ts[0], ts[6] / nt, ts[1] / nt, ts[2] / nt, ts[3] / nt, ts[4] / nt, ts[5],
                 double(ts[7]) / double(std::max<unsigned long long>(ts[3], 1)));
  };
  if (!plain_circuit && eo.sweep_kernel == 0) {
    st.sweep_tiles += ntiles;
    time_begin();
    sweep_pairs(eo, ntiles, lvl_entries, nl_max, dense, nopend, narrow, trace ? d_tstats.p : nullptr, st);
    st.t_sweep_ms += time_end();
    if (trace) print_stats();
    return;
  }
  const int bw = std::max(1, std::min(kChunk / 64, eo.block_waves));
  int per_cu = 3;
  if (!plain_circuit) {
    LSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_events<false>, 64 * bw, lds));
    if (per_cu < 1) throw DeviceError(LSIM_E_DEVICE, "sweep kernel cannot be resident (" + std::to_string(lds) +
                                                         " B of level counters)");
  }
  ensure_slots(eo, ntiles, nl_max, per_cu);
  SweepArgs a;
  a.jobs = d_jobs.p;
  a.tiles = d_tiles.p;
  a.ntiles = int(ntiles);
  a.lvl_entries = lvl_entries;
  a.lanes = d_lanes.p;
  a.rows = d_rows.p;
  a.masks = d_masks.p;
  a.rows_stride = slot_nl * 64;
  a.masks_stride = slot_nl;
  a.out = d_dout.p;
  d_queue.alloc(1);
  a.queue = d_queue.p;
  a.bytes = d_bytes.p;
  a.mrec = d_mrec.p;
  a.pend = d_pend.p;
  a.rec_stride = slot_nl;
  a.pend_stride = slot_nl;
  a.epoch_base = 0;
  a.tstats = trace ? d_tstats.p : nullptr;
  st.sweep_tiles += ntiles;
  time_begin();
  if (!plain_circuit) {
    // Tile epochs are unique over the engine's lifetime; before the 32-bit
    // counter wraps, the records are cleared and numbering restarts.
    if (uint64_t(epoch_base) + uint64_t(ntiles) + 1 >= (uint64_t(1) << 31)) {
      LSIM_HIP(hipMemsetAsync(d_mrec.p, 0, d_mrec.bytes(), stream));
      LSIM_HIP(hipMemsetAsync(d_pend.p, 0, d_pend.bytes(), stream));
      epoch_base = 0;
    }
    a.epoch_base = epoch_base;
    epoch_base += uint32_t(ntiles);
    LSIM_HIP(hipMemsetAsync(d_queue.p, 0, sizeof(int), stream));
    const unsigned grid = unsigned(std::min<int64_t>(std::min<int64_t>(slots, ntiles), int64_t(cus) * per_cu));
    if (a.tstats)
      k_sweep_events<true><<<grid, 64 * bw, lds, stream>>>(a);
    else
      k_sweep_events<false><<<grid, 64 * bw, lds, stream>>>(a);
    LSIM_HIP(hipGetLastError());
    st.sweep_launches++;
  } else {
    const CircuitRes &c = *plain_circuit;
    for (int64_t t0 = 0; t0 < ntiles; t0 += slots) {
      const int nb = int(std::min<int64_t>(slots, ntiles - t0));
      const std::vector<Tile> &tv = *tiles;
      int lmin = tv[size_t(t0)].start_level;
      for (int64_t t = t0; t < t0 + nb; ++t) lmin = std::min(lmin, tv[size_t(t)].start_level);
      for (int l = lmin; l <= c.L; ++l) {
        const int64_t n = c.live_lvl[l + 1] - c.live_lvl[l];
        if (n <= 0) continue;
        const int64_t threads = int64_t(nb) * n * 64;
        k_sweep_level<<<blocks_for(threads, 256), 256, 0, stream>>>(a, l, int(t0), nb);
        LSIM_HIP(hipGetLastError());
        st.sweep_launches++;
      }
      for (int64_t t = t0; t < t0 + nb; ++t) st.sweep_node_steps += c.NL - c.live_lvl[tv[size_t(t)].start_level];
    }
  }
  st.t_sweep_ms += time_end();
  if (trace && !plain_circuit) print_stats();
}

void Engine::Impl::sweep_pairs(const EngineOptions &eo, int64_t ntiles, int lvl_entries, int64_t nl_max, bool dense,
                               bool nopend, bool narrow, unsigned long long *tstats, lsim_run_stats &st) {
  const size_t lds = size_t(lvl_entries);  // one flag byte per level
  int pc256 = 0, pc512 = 0;
  LSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc256, pairs_kernel(256, false, narrow), 256, lds));
  LSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc512, pairs_kernel(512, false, narrow), 512, lds));
  if (pc512 < 1)
    throw DeviceError(LSIM_E_DEVICE, "sweep kernel cannot be resident (" + std::to_string(lds) + " B of level counters)");
  // Sized for the largest resident circuit, so later jobs on it reuse the
  // scratch instead of reallocating it; sized for this step's circuits when
  // the budget cannot hold a slot of the largest resident one.
  int64_t nl_all = nl_max;
  for (const auto &c : circuits) nl_all = std::max(nl_all, c->NL);
  try {
    ensure_pair_scratch(eo, nl_all, int64_t(cus) * std::max(pc256, pc512), int64_t(cus) * pc512);
  } catch (const DeviceError &e) {
    if (e.code != LSIM_E_NOMEM || nl_all == nl_max) throw;
    ensure_pair_scratch(eo, nl_max, int64_t(cus) * std::max(pc256, pc512), int64_t(cus) * pc512);
  }
  // Dense or wide steps use 512-thread workgroups, the others 256-thread
  // ones, twice as many tiles at once. (128- and 192-thread groups, 8 and 5
  // tiles per CU, were slower in every word step of the 2M-gate workload.)
  int threads = eo.pair_threads ? eo.pair_threads : (dense ? 512 : 256);
  if (threads == 256 && pc256 < 1) threads = 512;
  const int per_cu = threads == 256 ? pc256 : pc512;
  const int64_t full = pair_nl * 64;  // pool words one tile can use at most
  const int64_t pool_words = int64_t(d_pool.n);
  // Tile epochs are unique over the engine's lifetime; before the 32-bit
  // counter wraps, the records are cleared and numbering restarts.
  if (uint64_t(epoch_base) + 2 * uint64_t(ntiles) + 2 >= (uint64_t(1) << 31)) {
    LSIM_HIP(hipMemsetAsync(d_prec.p, 0, d_prec.bytes(), stream));
    LSIM_HIP(hipMemsetAsync(d_ppend.p, 0, d_ppend.bytes(), stream));
    epoch_base = 0;
  }
  d_queue.alloc(1);
  d_ovf_count.alloc(1);
  d_ovf_list.alloc(size_t(ntiles));
  PairArgs a;
  a.jobs = d_jobs.p;
  a.tiles = d_tiles.p;
  a.tile_list = nullptr;
  a.ntiles = int(ntiles);
  a.lvl_entries = lvl_entries;
  a.lanes = d_lanes.p;
  a.out = d_dout.p;
  a.queue = d_queue.p;
  a.bytes = d_bytes.p;
  a.rec = d_prec.p;
  a.pend = d_ppend.p;
  a.rec_stride = pair_nl;
  a.pool = d_pool.p;
  a.ovf_count = d_ovf_count.p;
  a.ovf_list = d_ovf_list.p;
  a.tstats = tstats;
  a.dense_lanes = eo.pair_dense;
  a.fo_pre = eo.pair_fo_pre ? 1 : 0;
  a.nopend = nopend ? 1 : 0;
  a.dyn = nullptr;
  a.ntiles_p = nullptr;
  // A share of the device's resident workgroups (another engine on the same
  // device may hold the rest).
  const double share = std::min(1.0, std::max(0.0, eo.slot_share));
  const int64_t resident = std::max<int64_t>(1, int64_t(double(int64_t(cus) * per_cu) * share));
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(ntiles, resident), pair_slots));
  a.pool_words = std::min<int64_t>(std::min<int64_t>(full, pool_words / grid), int64_t(UINT32_MAX));
  if (eo.pool_cap > 0) a.pool_words = std::min<int64_t>(a.pool_words, eo.pool_cap);
  a.epoch_base = epoch_base;
  epoch_base += uint32_t(ntiles);
  LSIM_HIP(hipMemsetAsync(d_queue.p, 0, sizeof(int), stream));
  LSIM_HIP(hipMemsetAsync(d_ovf_count.p, 0, sizeof(int), stream));
  hipLaunchKernelGGL(pairs_kernel(threads, tstats != nullptr, narrow), dim3(unsigned(grid)), dim3(threads), lds,
                     stream, a);
  LSIM_HIP(hipGetLastError());
  st.sweep_launches++;
  finalize_end();  // finished jobs' results unpacked while the sweep runs
  if (tstats)
    std::fprintf(stderr, "[lsim] pair sweep: %d threads, %lld workgroups (%d/%d per CU at 256/512), %lld-word pools\n",
                 threads, (long long)grid, pc256, pc512, (long long)a.pool_words);
  if (a.pool_words >= full) return;  // no tile can outgrow its pool
  int count = 0;
  d2h(&count, d_ovf_count.p, sizeof(int));
  LSIM_HIP(hipStreamSynchronize(stream));
  if (count <= 0) return;
  st.sweep_reruns += count;
  // Tiles that outgrew their pool: again, in slots that hold any tile.
  const int64_t g2 = std::max<int64_t>(
      1, std::min<int64_t>(std::min<int64_t>(count, std::max<int64_t>(1, int64_t(double(int64_t(cus) * pc512) * share))),
                           std::min(pair_slots, pool_words / full)));
  d_ovf_list2.alloc(size_t(count));
  a.tile_list = d_ovf_list.p;
  a.ntiles = count;
  a.ovf_list = d_ovf_list2.p;
  a.pool_words = std::min<int64_t>(std::min<int64_t>(full, pool_words / g2), int64_t(UINT32_MAX));
  a.epoch_base = epoch_base;
  epoch_base += uint32_t(count);
  LSIM_HIP(hipMemsetAsync(d_queue.p, 0, sizeof(int), stream));
  LSIM_HIP(hipMemsetAsync(d_ovf_count.p, 0, sizeof(int), stream));
  hipLaunchKernelGGL(pairs_kernel(512, tstats != nullptr, narrow), dim3(unsigned(g2)), dim3(512), lds, stream, a);
  LSIM_HIP(hipGetLastError());
  st.sweep_launches++;
  if (trace_on())
    std::fprintf(stderr, "[lsim] %d of %lld tiles outgrew %lld-word pools; swept again in %lld slots\n", count,
                 (long long)ntiles, (long long)(pool_words / grid), (long long)g2);
  int again = 0;
  d2h(&again, d_ovf_count.p, sizeof(int));
  LSIM_HIP(hipStreamSynchronize(stream));
  if (again) throw DeviceError(LSIM_E_DEVICE, "a tile outgrew a pool of 64 words per live node");
}

template <class Flag>
void Engine::Impl::scan(Flag flag, const int64_t *n_p, int64_t n, int32_t *out) {
  k_scan_sums<Flag><<<kScanBlocks, kScanThreads, 0, stream>>>(flag, n_p, n, d_scan_sum.p);
  k_scan_offsets<<<1, kScanBlocks, 0, stream>>>(d_scan_sum.p, n_p, n, out);
  k_scan_write<Flag><<<kScanBlocks, kScanThreads, 0, stream>>>(flag, n_p, n, d_scan_sum.p, out);
}

// One word step of every job in `act`, planned on the device. Every buffer is
// sized by the step's capacities (each job's live faults, its units times the
// step's words), which the host knows without reading anything back; the
// kernels read the actual counts from device memory.
void Engine::Impl::step_device(const std::vector<JobState *> &act, const EngineOptions &eo) {
  const int nj = int(act.size());
  DsSlot &sl = ds[ds_next % kDsRing];
  ds_resolve(sl);
  for (hipEvent_t &e : sl.ev)
    if (!e) LSIM_HIP(hipEventCreate(&e));
  sl.h_jobs.alloc(size_t(nj));
  sl.h_sj.alloc(size_t(nj));
  sl.h_stat.alloc(size_t(kDsStat + 2 * nj));
  sl.h_epoch.alloc(1);
  int64_t F = 0, W = 0, T = 0, I = 0, nl_max = 1;
  int L = 0;
  bool dense = false, narrow = true;
  for (int i = 0; i < nj; ++i) {
    JobState &j = *act[i];
    const CircuitRes &c = *j.c;
    narrow = narrow && c.narrow;
    const int64_t L1 = std::max<int64_t>(int64_t(j.live.size()), 1);
    const int64_t U1 = j.direct ? L1 : std::max<int64_t>(1, std::min<int64_t>(L1, j.stems_cap));
    JobDev d = job_dev(j);
    d.w0 = j.step_w0;
    d.Wn = j.step_Wn;
    d.first_round = (j.kind == JobState::LAST) ? nullptr : j.first.p;
    d.det_bits = (j.kind == JobState::MATRIX) ? j.detbits.p : nullptr;
    d.last_flag = (j.kind == JobState::LAST) ? j.last.p : nullptr;
    const int cur = j.dev_steps & 1, prev = cur ^ 1;
    int32_t *dst = cur ? j.d_fl1.p : j.d_fl0.p;
    d.flist = dst;
    d.units = j.d_units.p;
    sl.h_jobs.p[i] = d;
    StepJob S;
    std::memset(&S, 0, sizeof(S));
    if (j.kind == JobState::DROP && j.dev_steps > 0) {
      // The faults of the job's previous step that are still undetected.
      S.src = prev ? j.d_fl1.p : j.d_fl0.p;
      S.src_n = j.d_cnt.p + prev;
      S.filter = 1;
    } else {
      S.src = j.d_live.p;
      S.src_count = int32_t(j.live.size());
      S.filter = j.kind == JobState::DROP ? 1 : 0;
    }
    S.dst = dst;
    S.dst_n = j.d_cnt.p + cur;
    S.units = j.d_units.p;
    S.fbase = F;
    S.wide = c.wide ? 1 : 0;
    sl.h_sj.p[i] = S;
    F += L1;
    I += L1 * j.step_Wn;
    W += U1 * j.step_Wn;
    T += (U1 * j.step_Wn + 63) / 64;
    L = std::max(L, c.L);
    nl_max = std::max(nl_max, c.NL);
    dense = dense || j.kind == JobState::MATRIX || (j.kind == JobState::DROP && j.w0 == 0);
  }
  if (F >= INT32_MAX || W >= INT32_MAX) throw DeviceError(LSIM_E_DEVICE, "too many faults or unit words in one step");
  // Tile epochs are unique over the engine's lifetime (a step's tiles and
  // their reruns take at most 2 T + 2); before the 32-bit counter wraps, the
  // records are cleared and numbering restarts.
  if (uint64_t(epoch_base) + 2 * uint64_t(T) + 4 >= (uint64_t(1) << 31)) {
    LSIM_HIP(hipMemsetAsync(d_prec.p, 0, d_prec.bytes(), stream));
    LSIM_HIP(hipMemsetAsync(d_ppend.p, 0, d_ppend.bytes(), stream));
    epoch_base = 0;
  }
  const uint32_t epoch = epoch_base;
  epoch_base += uint32_t(2 * T + 2);
  // Pair sweep scratch, as for a step planned on the host.
  const int lvl_entries = L + 1;
  const size_t lds = size_t(lvl_entries);
  int pc256 = 0, pc512 = 0;
  LSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc256, pairs_kernel(256, false, narrow), 256, lds));
  LSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc512, pairs_kernel(512, false, narrow), 512, lds));
  if (pc512 < 1)
    throw DeviceError(LSIM_E_DEVICE, "sweep kernel cannot be resident (" + std::to_string(lds) + " B of level counters)");
  int64_t nl_all = nl_max;
  for (const auto &c : circuits) nl_all = std::max(nl_all, c->NL);
  try {
    ensure_pair_scratch(eo, nl_all, int64_t(cus) * std::max(pc256, pc512), int64_t(cus) * pc512);
  } catch (const DeviceError &e) {
    if (e.code != LSIM_E_NOMEM || nl_all == nl_max) throw;
    ensure_pair_scratch(eo, nl_max, int64_t(cus) * std::max(pc256, pc512), int64_t(cus) * pc512);
  }
  // Step buffers at capacity (grow only).
  d_fflag.alloc(size_t(F + 1));
  d_fpos.alloc(size_t(F + 1));
  d_upos.alloc(size_t(F + 1));
  d_need.alloc(size_t(std::max<int64_t>(W, 1)));
  d_dout.alloc(size_t(std::max<int64_t>(W, 1)));
  d_pos.alloc(size_t(W + 1));
  d_lanes.alloc(size_t(std::max<int64_t>(W, 1)));
  d_tiles_raw.alloc(size_t(std::max<int64_t>(T, 1)));
  d_tiles.alloc(size_t(std::max<int64_t>(T, 1)));
  d_tile_key.alloc(size_t(std::max<int64_t>(T, 1)));
  d_ovf_list.alloc(size_t(std::max<int64_t>(T, 1)));
  d_ovf_list2.alloc(size_t(std::max<int64_t>(T, 1)));
  d_hist.alloc(size_t(lvl_entries));
  d_jobs.alloc(size_t(nj));
  d_sjobs.alloc(size_t(nj));
  d_item_off.alloc(size_t(nj + 1));
  d_lane_off.alloc(size_t(nj + 1));
  d_tile_off.alloc(size_t(nj + 1));
  d_scan_sum.alloc(kScanBlocks);
  d_len.alloc(1);
  d_counters.alloc(4);
  d_ds_stat.alloc(size_t(kDsStat + 2 * nj));
  if (!d_dyn.p) {
    d_dyn.alloc(1);
    LSIM_HIP(hipMemsetAsync(d_dyn.p, 0, sizeof(SweepDyn), stream));
  }
  if (!d_epoch.p) d_epoch.alloc(1);
  *sl.h_epoch.p = epoch;
  h2d(d_jobs.p, sl.h_jobs.p, size_t(nj) * sizeof(JobDev));
  h2d(d_sjobs.p, sl.h_sj.p, size_t(nj) * sizeof(StepJob));
  h2d(d_epoch.p, sl.h_epoch.p, sizeof(uint32_t));
  const unsigned gF = std::max(1u, std::min<unsigned>(kGsBlocks, blocks_for(F, 256)));
  // Detection items: up to every step fault times the step's words; the
  // grid covers them up to kItemBlocks workgroups (the device's resident
  // limit several times over: the path walks are chains of dependent loads).
  const unsigned gI = std::max(1u, std::min<unsigned>(kItemBlocks, blocks_for(I, 256)));
  ItemArgs ia;
  std::memset(&ia, 0, sizeof(ia));  // padding too: the launch signature reads every byte
  ia.jobs = d_jobs.p;
  ia.njobs = nj;
  ia.item_off = d_item_off.p;
  ia.need = d_need.p;
  ia.dout = d_dout.p;
  // Which workgroup sizes may run: the one asked for, or both (the device
  // picks one; the other returns at once).
  const bool allow256 = pc256 >= 1;
  const bool run512 = eo.pair_threads != 256 || !allow256;
  const bool run256 = allow256 && eo.pair_threads != 512;
  TilePlanArgs tp;
  std::memset(&tp, 0, sizeof(tp));
  tp.sj = d_sjobs.p;
  tp.jobs = d_jobs.p;
  tp.nj = nj;
  tp.pos = d_pos.p;
  tp.len = d_len.p;
  tp.lane_off = d_lane_off.p;
  tp.tile_off = d_tile_off.p;
  tp.hist = d_hist.p;
  tp.nkeys = lvl_entries;
  tp.dyn = d_dyn.p;
  tp.bytes = d_bytes.p;
  tp.counters = d_counters.p;
  tp.epoch = d_epoch.p;
  tp.dense = dense ? 1 : 0;
  tp.threads = eo.pair_threads;
  tp.allow256 = allow256 ? 1 : 0;
  tp.nopend_mode = eo.pair_nopend;
  tp.nopend_pct = eo.pair_nopend_pct;
  TileSortArgs ts;
  std::memset(&ts, 0, sizeof(ts));
  ts.jobs = d_jobs.p;
  ts.nj = nj;
  ts.lanes = d_lanes.p;
  ts.lane_off = d_lane_off.p;
  ts.tile_off = d_tile_off.p;
  ts.len = d_len.p;
  ts.raw = d_tiles_raw.p;
  ts.keys = d_tile_key.p;
  ts.hist = d_hist.p;
  ts.out = d_tiles.p;
  // The sweep launches: the chosen workgroup size runs the step's tiles;
  // tiles that outgrow their slot's pool are swept again in slots that hold
  // any tile.
  const double share = std::min(1.0, std::max(0.0, eo.slot_share));
  const int64_t full = pair_nl * 64;
  const int64_t pool_all = int64_t(d_pool.n);
  PairArgs a;
  std::memset(&a, 0, sizeof(a));
  a.jobs = d_jobs.p;
  a.tiles = d_tiles.p;
  a.lvl_entries = lvl_entries;
  a.lanes = d_lanes.p;
  a.out = d_dout.p;
  a.bytes = d_bytes.p;
  a.rec = d_prec.p;
  a.pend = d_ppend.p;
  a.rec_stride = pair_nl;
  a.pool = d_pool.p;
  a.dense_lanes = eo.pair_dense;
  a.fo_pre = eo.pair_fo_pre ? 1 : 0;
  a.dyn = d_dyn.p;
  a.queue = d_counters.p + 0;
  a.ovf_count = d_counters.p + 2;
  a.ovf_list = d_ovf_list.p;
  struct SweepLaunch {
    int form;
    int64_t grid;
    PairArgs args;
  };
  std::vector<SweepLaunch> launches;
  bool may_overflow = false;
  for (int form : {512, 256}) {
    if ((form == 512 && !run512) || (form == 256 && !run256)) continue;
    const int per_cu = form == 256 ? pc256 : pc512;
    const int64_t resident = std::max<int64_t>(1, int64_t(double(int64_t(cus) * per_cu) * share));
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(resident, pair_slots));
    PairArgs f = a;
    f.pool_words = std::min<int64_t>(std::min<int64_t>(full, pool_all / grid), int64_t(UINT32_MAX));
    if (eo.pool_cap > 0) f.pool_words = std::min<int64_t>(f.pool_words, eo.pool_cap);
    may_overflow = may_overflow || f.pool_words < full;
    launches.push_back({form, grid, f});
  }
  if (may_overflow) {
    const int64_t resident = std::max<int64_t>(1, int64_t(double(int64_t(cus) * pc512) * share));
    const int64_t g2 = std::max<int64_t>(1, std::min<int64_t>(resident, std::min(pair_slots, pool_all / full)));
    PairArgs r = a;
    r.tile_list = d_ovf_list.p;
    r.ntiles_p = d_counters.p + 2;
    r.queue = d_counters.p + 1;
    r.ovf_count = d_counters.p + 3;
    r.ovf_list = d_ovf_list2.p;
    r.pool_words = std::min<int64_t>(std::min<int64_t>(full, pool_all / g2), int64_t(UINT32_MAX));
    launches.push_back({512, g2, r});
  }
  // Three segments of launches, each one graph when the step recurs: plan
  // (fault lists, units, local terms, lanes, tiles), sweep, detection.
  auto seg_plan = [&] {
    k_ds_keep<<<gF, 256, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_fflag.p);
    scan(FlagI32{d_fflag.p}, nullptr, F, d_fpos.p);
    k_ds_compact<<<gF, 256, 0, stream>>>(d_sjobs.p, nj, F, d_fflag.p, d_fpos.p);
    k_ds_unit_flag<<<gF, 256, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_fpos.p, d_fflag.p);
    scan(FlagI32{d_fflag.p}, nullptr, F, d_upos.p);
    k_ds_units<<<gF, 256, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_fflag.p, d_upos.p);
    k_ds_plan<<<1, kPlanThreads, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_fpos.p, d_upos.p, d_item_off.p, d_len.p,
                                               d_ds_stat.p + kDsStat);
    k_ds_zero<<<kGsBlocks, 256, 0, stream>>>(d_len.p, d_need.p, d_dout.p);
    k_ds_slots<<<gF, 256, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_upos.p, 0);
    k_local_gs<<<gI, 256, 0, stream>>>(ia);
    scan(FlagWord{d_need.p}, &d_len.p->words, 0, d_pos.p);
    k_lane_fill_gs<<<kGsBlocks, 256, 0, stream>>>(d_jobs.p, nj, d_need.p, d_pos.p, d_len.p, d_lanes.p);
    k_ds_tiles<<<1, kPlanThreads, 0, stream>>>(tp);
    k_ds_tile_hist<<<kGsBlocks, 256, 0, stream>>>(ts);
    k_ds_hist_scan<<<1, kPlanThreads, 0, stream>>>(d_hist.p, lvl_entries);
    k_ds_tile_scatter<<<kGsBlocks, 256, 0, stream>>>(ts);
  };
  auto seg_sweep = [&] {
    for (const SweepLaunch &l : launches)
      hipLaunchKernelGGL(pairs_kernel(l.form, false, narrow), dim3(unsigned(l.grid)), dim3(l.form), lds, stream,
                         l.args);
  };
  auto seg_detect = [&] {
    k_detect_gs<<<gI, 256, 0, stream>>>(ia);
    k_ds_slots<<<gF, 256, 0, stream>>>(d_sjobs.p, d_jobs.p, nj, F, d_upos.p, 1);
    k_ds_stats<<<1, 1, 0, stream>>>(d_len.p, d_bytes.p, d_counters.p, d_ds_stat.p);
  };
  // The launch signature: every argument the host gives the step's kernels
  // (the rest is read from device memory, uploaded above).
  std::vector<uint64_t> sig = {uint64_t(nj), uint64_t(F), uint64_t(gF), uint64_t(gI), uint64_t(lvl_entries),
                               uint64_t(narrow)};
  auto add_bytes = [&](const void *p, size_t n) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t o = 0; o < n; o += 8) {
      uint64_t w = 0;
      std::memcpy(&w, b + o, std::min<size_t>(8, n - o));
      sig.push_back(w);
    }
  };
  for (const void *q : {static_cast<const void *>(d_sjobs.p), static_cast<const void *>(d_fflag.p),
                        static_cast<const void *>(d_fpos.p), static_cast<const void *>(d_upos.p),
                        static_cast<const void *>(d_scan_sum.p), static_cast<const void *>(d_ds_stat.p)})
    sig.push_back(uint64_t(reinterpret_cast<uintptr_t>(q)));
  add_bytes(&ia, sizeof(ia));
  add_bytes(&tp, sizeof(tp));
  add_bytes(&ts, sizeof(ts));
  for (const SweepLaunch &l : launches) {
    sig.push_back(uint64_t(l.form));
    sig.push_back(uint64_t(l.grid));
    add_bytes(&l.args, sizeof(PairArgs));
  }
  StepGraph *g = eo.step_graphs ? find_step_graph(sig) : nullptr;
  bool capture = false;
  if (!g && eo.step_graphs) {
    const uint64_t h = sig_hash(sig);
    capture = std::find(step_seen.begin(), step_seen.end(), h) != step_seen.end();
    if (!capture) {
      if (step_seen.size() >= kStepSeen) step_seen.erase(step_seen.begin());
      step_seen.push_back(h);
    }
  }
  if (capture) {
    auto ng = std::make_unique<StepGraph>();
    ng->sig = sig;
    capture_segment(seg_plan, ng->graph[0], ng->exec[0]);
    capture_segment(seg_sweep, ng->graph[1], ng->exec[1]);
    capture_segment(seg_detect, ng->graph[2], ng->exec[2]);
    if (step_graphs.size() >= kStepGraphs) step_graphs.erase(step_graphs.begin());
    step_graphs.push_back(std::move(ng));
    g = step_graphs.back().get();
  }
  LSIM_HIP(hipEventRecord(sl.ev[0], stream));
  if (g) LSIM_HIP(hipGraphLaunch(g->exec[0], stream)); else seg_plan();
  LSIM_HIP(hipGetLastError());
  LSIM_HIP(hipEventRecord(sl.ev[1], stream));
  if (g) LSIM_HIP(hipGraphLaunch(g->exec[1], stream)); else seg_sweep();
  LSIM_HIP(hipGetLastError());
  LSIM_HIP(hipEventRecord(sl.ev[2], stream));
  finalize_end();  // finished jobs' results unpacked while the sweep runs
  if (g) LSIM_HIP(hipGraphLaunch(g->exec[2], stream)); else seg_detect();
  LSIM_HIP(hipGetLastError());
  d2h(sl.h_stat.p, d_ds_stat.p, size_t(kDsStat + 2 * nj) * sizeof(int64_t));
  LSIM_HIP(hipEventRecord(sl.ev[3], stream));
  sl.replayed = g != nullptr;
  sl.busy = true;
  sl.jobs = act;
  ++ds_next;
  for (JobState *jp : act) {
    JobState &j = *jp;
    j.res.stats.steps++;
    j.dev_steps++;
    switch (j.kind) {
      case JobState::MATRIX: j.w0 += j.step_Wn; break;
      case JobState::DROP:
        j.w0 += j.step_Wn;
        j.Wt = std::min(std::max(1, eo.max_step_words), j.Wt * j.Wg);
        break;
      case JobState::LAST: j.last_done = true; break;
    }
  }
}

// Reads back a device step's figures (waiting for the step) and adds them to
// its jobs' statistics; a tile that outgrew even a full pool is an error.
void Engine::Impl::ds_resolve(DsSlot &sl) {
  if (!sl.busy) return;
  sl.busy = false;
  const std::vector<JobState *> js = std::move(sl.jobs);
  sl.jobs.clear();
  LSIM_HIP(hipEventSynchronize(sl.ev[3]));
  float t_pre = 0, t_sweep = 0, t_post = 0;
  LSIM_HIP(hipEventElapsedTime(&t_pre, sl.ev[0], sl.ev[1]));
  LSIM_HIP(hipEventElapsedTime(&t_sweep, sl.ev[1], sl.ev[2]));
  LSIM_HIP(hipEventElapsedTime(&t_post, sl.ev[2], sl.ev[3]));
  const int64_t *st = sl.h_stat.p;
  const int64_t launches = 1 + (st[3] > 0 ? 1 : 0);
  for (size_t i = 0; i < js.size(); ++i) {
    JobState &j = *js[i];
    lsim_run_stats &r = j.res.stats;
    r.t_sweep_ms += t_sweep;
    r.t_detect_ms += t_pre + t_post;
    r.sweep_launches += launches;
    r.sweep_tiles += st[2];
    r.sweep_bytes += st[0];
    r.sweep_reruns += st[3];
    r.steps_replayed += sl.replayed ? 1 : 0;
    const int64_t nf = st[kDsStat + 2 * i], nu = st[kDsStat + 2 * i + 1];
    if (!j.direct) r.stems += nu;
    j.dev_alive = nf;
  }
  if (trace_on())
    std::fprintf(stderr,
                 "[lsim] device step jobs=%zu items=%lld words=%lld lanes=%lld tiles=%lld reruns=%lld | plan+local %.2f "
                 "sweep %.2f detect %.2f ms\n",
                 js.size(), (long long)st[5], (long long)st[6], (long long)st[1], (long long)st[2], (long long)st[3],
                 t_pre, t_sweep, t_post);
  if (st[4]) throw DeviceError(LSIM_E_DEVICE, "a tile outgrew a pool of 64 words per live node");
}

// Every in-flight device step, oldest first.
void Engine::Impl::ds_drain() {
  for (int k = 0; k < kDsRing; ++k) ds_resolve(ds[(ds_next + uint64_t(k)) % kDsRing]);
}

void Engine::Impl::step_batch(const std::vector<JobState *> &act, const EngineOptions &eo) {
  const int nj = int(act.size());
  using clk = std::chrono::steady_clock;
  const bool trace = trace_on();
  const auto h0 = clk::now();
  bool narrow = true;  // every circuit of the step has gates of at most two fanins
  for (const JobState *j : act) narrow = narrow && j->c->narrow;
  std::vector<JobDev> devs(nj);
  std::vector<int64_t> item_off(nj + 1, 0);
  std::vector<int64_t> word0(nj + 1, 0);
  std::vector<int64_t> flist_off(nj + 1, 0);  // step faults of job i at [flist_off[i], flist_off[i+1])
  std::vector<int64_t> fault_off(nj + 1, 0);  // dropping jobs' faults in the first-round gather
  int lvl_entries = 1;
  int64_t nl_max = 1;
  lsim_run_stats st{};
  // Largest jobs first when the per-job host work runs in parallel.
  std::vector<int> order(nj);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return act[x]->step_faults.size() > act[y]->step_faults.size(); });
  // Per job: the step's faults in unit order and its units (stems of the
  // faults' regions, or the fault sites themselves for direct injection).
  parallel_for(order, [&](int i) {
    JobState &j = *act[i];
    const CircuitRes &c = *j.c;
    j.units.clear();
    // The step's faults are in unit order (order_live; dropping keeps it).
    const int64_t nf = int64_t(j.step_faults.size());
    const int C = chunk_count(nf, nj == 1);
    if (j.direct) {
      // One unit per fault, at its site.
      j.units.resize(size_t(nf));
      run_chunks(C, nf, [&](int, int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) j.units[size_t(k)] = j.site[size_t(j.step_faults[size_t(k)])];
      });
    } else {
      // Stems of the step's faults in position order: a fault opens a unit
      // when its region head is not a PO and differs from the previous
      // fault's (counted per chunk, then listed; a lone job splits both over
      // host threads).
      auto head_at = [&](int64_t k) { return j.head[size_t(j.step_faults[size_t(k)])]; };
      std::vector<int64_t> cnt(size_t(C) + 1, 0);
      run_chunks(C, nf, [&](int ch, int64_t a, int64_t b) {
        int64_t m = 0;
        int32_t prev = a > 0 ? head_at(a - 1) : -1;
        for (int64_t k = a; k < b; ++k) {
          const int32_t h = head_at(k);
          m += (h != prev && !c.head_is_po[size_t(h)]) ? 1 : 0;
          prev = h;
        }
        cnt[size_t(ch) + 1] = m;
      });
      for (int ch = 0; ch < C; ++ch) cnt[size_t(ch) + 1] += cnt[size_t(ch)];
      j.units.resize(size_t(cnt[size_t(C)]));
      run_chunks(C, nf, [&](int ch, int64_t a, int64_t b) {
        int64_t o = cnt[size_t(ch)];
        int32_t prev = a > 0 ? head_at(a - 1) : -1;
        for (int64_t k = a; k < b; ++k) {
          const int32_t h = head_at(k);
          if (h != prev && !c.head_is_po[size_t(h)]) j.units[size_t(o++)] = h;
          prev = h;
        }
      });
    }
  });
  const auto h0u = clk::now();
  std::vector<int64_t> lane_unit_off(nj + 1, 0);  // units of job i in the step's unit list
  const bool plain = eo.mode == ENGINE_PLAIN;
  for (int i = 0; i < nj; ++i) {
    JobState &j = *act[i];
    const CircuitRes &c = *j.c;
    const int64_t nf = int64_t(j.step_faults.size());
    flist_off[i + 1] = flist_off[i] + nf;
    fault_off[i + 1] = fault_off[i] + (j.kind == JobState::DROP ? nf : 0);
    JobDev d = job_dev(j);
    d.nf = nf;
    d.w0 = j.step_w0;
    d.Wn = j.step_Wn;
    d.first_round = (j.kind == JobState::LAST) ? nullptr : j.first.p;
    d.det_bits = (j.kind == JobState::MATRIX) ? j.detbits.p : nullptr;
    d.last_flag = (j.kind == JobState::LAST) ? j.last.p : nullptr;
    d.item0 = item_off[i];
    d.word0 = word0[i];
    d.lane_unit0 = lane_unit_off[i];
    lane_unit_off[i + 1] = lane_unit_off[i] + int64_t(j.units.size());
    item_off[i + 1] = item_off[i] + nf * j.step_Wn;
    word0[i + 1] = word0[i] + int64_t(j.units.size()) * j.step_Wn;
    devs[i] = d;
    lvl_entries = std::max(lvl_entries, c.L + 1);
    nl_max = std::max(nl_max, c.NL);
  }
  const int64_t items = item_off[nj], words = std::max<int64_t>(word0[nj], 1);
  // All jobs' step faults in one page-locked block, one copy.
  const int64_t nf_all = flist_off[nj];
  h_flist.alloc(size_t(std::max<int64_t>(nf_all, 1)));
  d_flist.alloc(size_t(std::max<int64_t>(nf_all, 1)));
  const int64_t nu_all = lane_unit_off[nj];
  h_unit_pos.alloc(size_t(std::max<int64_t>(nu_all, 1)));
  const auto h0s = clk::now();
  parallel_for(order, [&](int i) {
    const JobState &j = *act[i];
    std::memcpy(h_flist.p + flist_off[i], j.step_faults.data(), j.step_faults.size() * sizeof(int32_t));
    std::memcpy(h_unit_pos.p + lane_unit_off[i], j.units.data(), j.units.size() * sizeof(int32_t));
  });
  // Only positions travel: the device derives each unit's job, index and
  // injection kind (k_set_slots, k_lane_fill).
  if (nu_all) {
    d_unit_pos.alloc(size_t(nu_all));
    h2d(d_unit_pos.p, h_unit_pos.p, size_t(nu_all) * sizeof(int32_t));
  }
  for (int i = 0; i < nj; ++i) {
    devs[i].flist = d_flist.p + flist_off[i];
    devs[i].units = d_unit_pos.p + lane_unit_off[i];
  }
  h2d(d_flist.p, h_flist.p, size_t(nf_all) * sizeof(int32_t));
  d_jobs.alloc(nj);
  d_item_off.alloc(nj + 1);
  d_fault_off.alloc(nj + 1);
  h2d(d_jobs.p, devs.data(), nj * sizeof(JobDev));
  h2d(d_item_off.p, item_off.data(), (nj + 1) * sizeof(int64_t));
  h2d(d_fault_off.p, fault_off.data(), (nj + 1) * sizeof(int64_t));
  const auto h1 = clk::now();
  d_need.alloc(words);
  d_dout.alloc(words);
  LSIM_HIP(hipMemsetAsync(d_need.p, 0, words * sizeof(uint64_t), stream));
  LSIM_HIP(hipMemsetAsync(d_dout.p, 0, words * sizeof(uint64_t), stream));
  if (nu_all) {
    k_set_slots<<<blocks_for(nu_all, 256), 256, 0, stream>>>(d_jobs.p, nj, d_unit_pos.p, nu_all, 0);
    LSIM_HIP(hipGetLastError());
  }
  ItemArgs ia;
  ia.jobs = d_jobs.p;
  ia.njobs = nj;
  ia.item_off = d_item_off.p;
  ia.need = d_need.p;
  ia.dout = d_dout.p;
  time_begin();
  if (items) {
    k_local<<<blocks_for(items, 256), 256, 0, stream>>>(ia);
    LSIM_HIP(hipGetLastError());
  }
  // Dense steps: a dropping job's first step (all faults, most lanes active
  // at most nodes) and full-matrix steps.
  bool dense = false;
  for (const JobState *j : act)
    dense = dense || j->kind == JobState::MATRIX || (j->kind == JobState::DROP && j->w0 == 0);
  int64_t wide_tiles = 0, ntiles = 0, nlanes = 0;
  clk::time_point h2, h3;
  if (!plain) {
    // Lanes and tiles built on the device: the needed bits never leave it.
    const int64_t nw = word0[nj];
    d_pos.alloc(size_t(words));
    d_lanes.alloc(size_t(words));
    d_lane_off.alloc(size_t(nj + 1));
    h_lane_off.alloc(size_t(nj + 1));
    if (nw >= (int64_t(1) << 31)) throw DeviceError(LSIM_E_DEVICE, "too many unit words in one step");
    if (nw) {
      hipcub::TransformInputIterator<int, NeedNonzero, const uint64_t *> flags(d_need.p, NeedNonzero());
      size_t tmp = 0;
      LSIM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flags, d_pos.p, int(nw), stream));
      d_cub.alloc(tmp);
      LSIM_HIP(hipcub::DeviceScan::ExclusiveSum(d_cub.p, tmp, flags, d_pos.p, int(nw), stream));
      k_lane_fill<<<blocks_for(nw, 256), 256, 0, stream>>>(d_jobs.p, nj, d_need.p, d_pos.p, nw, d_lanes.p);
      LSIM_HIP(hipGetLastError());
    }
    k_lane_bounds<<<blocks_for(nj + 1, 256), 256, 0, stream>>>(d_jobs.p, nj, d_need.p, d_pos.p, nw, d_lane_off.p);
    LSIM_HIP(hipGetLastError());
    d2h(h_lane_off.p, d_lane_off.p, size_t(nj + 1) * sizeof(int64_t));
    st.t_detect_ms += time_end();
    h2 = clk::now();
    nlanes = h_lane_off.p[nj];
    std::vector<int64_t> tile_off(nj + 1, 0);
    for (int i = 0; i < nj; ++i) {
      const int64_t nt = (h_lane_off.p[i + 1] - h_lane_off.p[i] + 63) / 64;
      tile_off[i + 1] = tile_off[i] + nt;
      if (act[i]->c->wide) wide_tiles += nt;
      act[i]->res.stats.stems += act[i]->direct ? 0 : int64_t(act[i]->units.size());
    }
    ntiles = tile_off[nj];
    if (ntiles) {
      if (ntiles >= (int64_t(1) << 31)) throw DeviceError(LSIM_E_DEVICE, "too many sweep tiles in one step");
      d_tile_off.alloc(size_t(nj + 1));
      h2d(d_tile_off.p, tile_off.data(), size_t(nj + 1) * sizeof(int64_t));
      d_tiles_raw.alloc(size_t(ntiles));
      d_tiles.alloc(size_t(ntiles));
      d_tile_key.alloc(size_t(ntiles) * 2);
      d_tile_idx.alloc(size_t(ntiles) * 2);
      int *keys = d_tile_key.p, *keys2 = d_tile_key.p + ntiles, *idx = d_tile_idx.p, *idx2 = d_tile_idx.p + ntiles;
      k_make_tiles<<<blocks_for(ntiles, 256), 256, 0, stream>>>(d_jobs.p, nj, d_lanes.p, d_lane_off.p, d_tile_off.p,
                                                               ntiles, d_tiles_raw.p, keys, idx);
      LSIM_HIP(hipGetLastError());
      // Deep tiles first (stable): the queue then ends with short tiles,
      // balancing the tail.
      size_t tmp = 0;
      LSIM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, keys, keys2, idx, idx2, int(ntiles), 0,
                                                            int(sizeof(int) * 8), stream));
      d_cub.alloc(tmp);
      LSIM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(d_cub.p, tmp, keys, keys2, idx, idx2, int(ntiles), 0,
                                                            int(sizeof(int) * 8), stream));
      k_gather_tiles<<<blocks_for(ntiles, 256), 256, 0, stream>>>(d_tiles_raw.p, idx2, ntiles, d_tiles.p);
      LSIM_HIP(hipGetLastError());
    }
    h3 = clk::now();
    const bool mid = staged_share >= kMidShareLo && staged_share <= kMidShareHi;
    // 512-thread groups when most tiles walk levels wider than 256 nodes, in
    // dense steps and outside a middle staged share; 256 otherwise (narrow
    // levels leave most of a 512-thread scan idle, dense steps included).
    const bool big_groups = 2 * wide_tiles > ntiles && (dense || !mid);
    // Dense steps evaluate every node of the levels a tile walks instead of
    // marking fanouts pending (most nodes are reached anyway): a dropping
    // job's first step, full-matrix steps, and steps after one that staged
    // more than pair_nopend_pct percent of the nodes it walked.
    const bool nopend =
        eo.pair_nopend == 2 || (eo.pair_nopend == 1 && (dense || 100.0 * staged_share > double(eo.pair_nopend_pct)));
    sweep(eo, nullptr, nlanes, ntiles, nullptr, lvl_entries, nl_max, nullptr, big_groups, nopend, narrow, st);
  } else {
    // The needed bits decide which (unit, word) pairs become sweep lanes.
    h_need.alloc(size_t(words));
    const uint64_t *need = h_need.p;
    d2h(h_need.p, d_need.p, size_t(word0[nj]) * sizeof(uint64_t));
    st.t_detect_ms += time_end();
    h2 = clk::now();
    // Lanes: one per (unit, word) with needed bits, per job in unit order
    // (units are level-sorted), counted then filled in parallel over chunks of
    // a job's units.
    constexpr int64_t kUnitChunk = 4096;
    struct LaneChunk {
      int job;
      int64_t k0, k1, off;
    };
    std::vector<LaneChunk> chunks;
    for (int i = 0; i < nj; ++i) {
      const int64_t nu = int64_t(act[i]->units.size());
      for (int64_t k0 = 0; k0 < nu; k0 += kUnitChunk) chunks.push_back({i, k0, std::min(nu, k0 + kUnitChunk), 0});
    }
    std::vector<int> chunk_ids(chunks.size());
    std::iota(chunk_ids.begin(), chunk_ids.end(), 0);
    parallel_for(chunk_ids, [&](int ci) {
      LaneChunk &ch = chunks[size_t(ci)];
      const int64_t Wn = act[ch.job]->step_Wn;
      int64_t n = 0;
      for (int64_t o = word0[ch.job] + ch.k0 * Wn, e = word0[ch.job] + ch.k1 * Wn; o < e; ++o) n += need[o] != 0;
      ch.off = n;
    });
    std::vector<int64_t> lane_off(nj + 1, 0);
    {
      int64_t run = 0;
      size_t c = 0;
      for (int i = 0; i < nj; ++i) {
        lane_off[i] = run;
        for (; c < chunks.size() && chunks[c].job == i; ++c) {
          const int64_t n = chunks[c].off;
          chunks[c].off = run;
          run += n;
        }
      }
      lane_off[nj] = run;
    }
    nlanes = lane_off[nj];
    h_lanes.alloc(size_t(std::max<int64_t>(nlanes, 1)));
    Lane *lanes = h_lanes.p;
    parallel_for(chunk_ids, [&](int ci) {
      const LaneChunk &ch = chunks[size_t(ci)];
      const int i = ch.job;
      const JobState &j = *act[i];
      const int Wn = j.step_Wn;
      Lane *out = lanes + ch.off;
      for (int64_t k = ch.k0; k < ch.k1; ++k) {
        const int32_t u = j.units[size_t(k)];
        const int32_t kind = j.direct ? (j.sa[j.step_faults[size_t(k)]] ? K_SA1 : K_SA0) : K_FLIP;
        const int32_t level = j.c->level_dp[u];
        for (int jw = 0; jw < Wn; ++jw) {
          const int64_t o = word0[i] + k * Wn + jw;
          const uint64_t msk = need[o];
          if (!msk) continue;
          Lane ln;
          ln.mask = msk;
          ln.site = u;
          ln.word = int32_t(j.step_w0 + jw);
          ln.out = int32_t(o);
          ln.kind = kind;
          ln.level = level;
          ln.pad = 0;
          *out++ = ln;
        }
      }
    });
    std::vector<Tile> tiles;
    for (int i = 0; i < nj; ++i) {
      JobState &j = *act[i];
      // The pair sweep finds a node's injecting lanes by binary search: a job's
      // lanes follow its units, which are in ascending position.
      for (int64_t l = lane_off[i] + 1; l < lane_off[i + 1]; ++l)
        if (lanes[l].site < lanes[l - 1].site) throw DeviceError(LSIM_E_DEVICE, "sweep lanes out of site order");
      for (int64_t l0 = lane_off[i]; l0 < lane_off[i + 1]; l0 += 64) {
        Tile t;
        t.lane0 = int32_t(l0);
        t.nlanes = int32_t(std::min<int64_t>(64, lane_off[i + 1] - l0));
        t.start_level = lanes[l0].level;
        t.end_level = lanes[l0 + t.nlanes - 1].level;
        t.job = i;
        tiles.push_back(t);
      }
      j.res.stats.stems += j.direct ? 0 : int64_t(j.units.size());
    }
    // Deep tiles first: the queue then ends with short tiles, balancing the tail.
    std::stable_sort(tiles.begin(), tiles.end(), [&](const Tile &x, const Tile &y) {
      return (devs[x.job].max_level - x.start_level) > (devs[y.job].max_level - y.start_level);
    });
    h3 = clk::now();
    for (const Tile &t : tiles) wide_tiles += act[t.job]->c->wide ? 1 : 0;
    ntiles = int64_t(tiles.size());
    const bool big_groups = dense || 2 * wide_tiles > ntiles;
    sweep(eo, lanes, nlanes, ntiles, &tiles, lvl_entries, nl_max, act[0]->c, big_groups, false, narrow, st);
  }
  const auto h4 = clk::now();
  time_begin();
  if (items) {
    k_detect<<<blocks_for(items, 256), 256, 0, stream>>>(ia);
    LSIM_HIP(hipGetLastError());
  }
  if (nu_all) {
    k_set_slots<<<blocks_for(nu_all, 256), 256, 0, stream>>>(d_jobs.p, nj, d_unit_pos.p, nu_all, 1);
    LSIM_HIP(hipGetLastError());
  }
  // One copy of every dropping job's first rounds for the step's faults.
  const int64_t ng = fault_off[nj];
  if (ng) {
    d_first_out.alloc(size_t(ng));
    h_first_out.alloc(size_t(ng));
    k_gather_first<<<blocks_for(ng, 256), 256, 0, stream>>>(d_jobs.p, nj, d_fault_off.p, d_first_out.p);
    LSIM_HIP(hipGetLastError());
    d2h(h_first_out.p, d_first_out.p, size_t(ng) * sizeof(int32_t));
  }
  h_bytes.alloc(3);
  d2h(h_bytes.p, d_bytes.p, 3 * sizeof(unsigned long long));
  LSIM_HIP(hipMemsetAsync(d_bytes.p, 0, 3 * sizeof(unsigned long long), stream));
  st.t_detect_ms += time_end();
  st.sweep_bytes = int64_t(h_bytes.p[0]);
  // Staged share of the nodes in the levels the pair sweep's tiles walked:
  // the next step drops pending flags while it is high.
  if (h_bytes.p[2]) staged_share = double(h_bytes.p[1]) / double(h_bytes.p[2]);
  const auto h5 = clk::now();
  // Advance every job (in parallel: a dropping step rebuilds each job's list
  // of faults still undetected).
  parallel_for(order, [&](int i) {
    JobState &j = *act[i];
    lsim_run_stats &js = j.res.stats;
    js.steps++;
    js.t_sweep_ms += st.t_sweep_ms;
    js.t_detect_ms += st.t_detect_ms;
    js.sweep_launches += st.sweep_launches;
    js.sweep_tiles += st.sweep_tiles;
    js.sweep_node_steps += st.sweep_node_steps;
    js.sweep_bytes += st.sweep_bytes;
    js.sweep_reruns += st.sweep_reruns;
    switch (j.kind) {
      case JobState::MATRIX: j.w0 += j.step_Wn; break;
      case JobState::DROP: {
        // Faults still undetected, in order (a lone job splits the filter
        // over host threads).
        const int32_t *fr = h_first_out.p + fault_off[i];
        const int64_t nf = int64_t(j.step_faults.size());
        const int C = chunk_count(nf, nj == 1);
        std::vector<int64_t> cnt(size_t(C) + 1, 0);
        run_chunks(C, nf, [&](int ch, int64_t a, int64_t b) {
          int64_t m = 0;
          for (int64_t k = a; k < b; ++k) m += fr[k] == INT32_MAX ? 1 : 0;
          cnt[size_t(ch) + 1] = m;
        });
        for (int ch = 0; ch < C; ++ch) cnt[size_t(ch) + 1] += cnt[size_t(ch)];
        std::vector<int32_t> next(static_cast<size_t>(cnt[size_t(C)]));
        run_chunks(C, nf, [&](int ch, int64_t a, int64_t b) {
          int64_t o = cnt[size_t(ch)];
          for (int64_t k = a; k < b; ++k)
            if (fr[k] == INT32_MAX) next[size_t(o++)] = j.step_faults[size_t(k)];
        });
        j.alive.swap(next);
        j.w0 += j.step_Wn;
        j.Wt = std::min(std::max(1, eo.max_step_words), j.Wt * j.Wg);
        break;
      }
      case JobState::LAST: j.last_done = true; break;
    }
  });
  if (trace) {
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::fprintf(stderr,
                 "[lsim] iteration jobs=%d items=%lld units=%lld lanes=%lld tiles=%lld staged share %.3f | prep %.2f "
                 "(units %.2f, offsets %.2f, lists %.2f) local %.2f lanes %.2f sweep %.2f (dev %.2f) detect %.2f "
                 "advance %.2f ms\n",
                 nj, (long long)items, (long long)word0[nj], (long long)nlanes, (long long)ntiles, staged_share,
                 ms(h0, h1), ms(h0, h0u), ms(h0u, h0s), ms(h0s, h1), ms(h1, h2), ms(h2, h3), ms(h3, h4), st.t_sweep_ms,
                 ms(h4, h5), ms(h5, clk::now()));
  }
}

int64_t Engine::Impl::iterate_jobs(const std::vector<JobState *> &jobs, const EngineOptions &eo) {
  std::vector<JobState *> act;
  to_finalize.clear();
  const bool dev_ok = device_steps_ok(eo);
  for (JobState *j : jobs)
    if (j->started && !j->done && plan(*j, dev_ok)) act.push_back(j);
  // Finishing jobs take the figures of their device steps first.
  if (!to_finalize.empty()) ds_drain();
  finalize_begin(to_finalize);
  to_finalize.clear();
  try {
    std::vector<JobState *> on_dev, on_host;
    for (JobState *j : act) (j->dev ? on_dev : on_host).push_back(j);
    if (!on_host.empty()) {
      if (eo.mode == ENGINE_PLAIN) {
        for (JobState *j : on_host) step_batch({j}, eo);  // the per-level form runs one job at a time
      } else {
        step_batch(on_host, eo);
      }
    }
    if (!on_dev.empty()) step_device(on_dev, eo);
    finalize_end();
  } catch (...) {
    // Jobs left unpacked stay unfinished: the next plan() queues them again.
    // Their result copies may still be in flight into h_final / h_stage:
    // wait for them before the buffers can be reused. In-flight device steps
    // are waited for and their figures dropped.
    if (!fin_js.empty()) (void)hipEventSynchronize(fin_ev);
    fin_js.clear();
    fin_scatter.clear();
    for (DsSlot &sl : ds)
      if (sl.busy) {
        if (sl.ev[3]) (void)hipEventSynchronize(sl.ev[3]);
        sl.busy = false;
        sl.jobs.clear();
      }
    throw;
  }
  int64_t left = 0;
  for (JobState *j : jobs) left += j->done ? 0 : 1;
  return left;
}

int64_t Engine::iterate(const std::vector<JobState *> &jobs, const EngineOptions &eo) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  AllocStream alloc_on(m.stream, m.arena);
  return m.iterate_jobs(jobs, eo);
}

std::string Engine::Impl::tune_key(const std::vector<JobState *> &jobs, const EngineOptions &eo,
                                   bool &any_drop) const {
  std::vector<std::string> parts;
  any_drop = false;
  for (const JobState *j : jobs) {
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%p/%lld/%d/%d", static_cast<const void *>(j->c), (long long)j->R,
                  int(j->opt.drop), int(j->direct));
    parts.push_back(buf);
    any_drop = any_drop || j->opt.drop;
  }
  std::sort(parts.begin(), parts.end());
  std::string key = std::to_string(eo.mode) + ":";
  for (const std::string &x : parts) key += x + ",";
  return key;
}

Engine::Impl::TuneState &Engine::Impl::tune_state(const std::string &key, const EngineOptions &eo, bool any_drop) {
  auto it = tuned.find(key);
  if (it != tuned.end()) return it->second;
  TuneState st;
  st.base.kernel = eo.sweep_kernel;
  st.base.threads = eo.pair_threads;
  st.base.dense = eo.pair_dense;
  st.base.max_words = eo.max_step_words;
  // The word schedule and its cap only matter with dropping; the pair
  // sweep's workgroup size and dense passes and the kernel only for the
  // persistent engine.
  if (any_drop) st.dim_kind.push_back(0);
  if (eo.mode != ENGINE_PLAIN) {
    st.dim_kind.push_back(1);
    st.dim_kind.push_back(2);
  }
  if (any_drop) st.dim_kind.push_back(3);
  if (eo.mode != ENGINE_PLAIN) st.dim_kind.push_back(4);
  for (int k : st.dim_kind) st.planned += int(TuneState::variants(k, st.base).size()) - 1;
  return tuned.emplace(key, std::move(st)).first->second;
}

void Engine::run(const std::vector<JobState *> &jobs, const EngineOptions &eo) {
  using clk = std::chrono::steady_clock;
  Impl &m = *impl_;
  const auto t0 = clk::now();
  start(jobs, eo);
  EngineOptions e2 = eo;
  int first = jobs.empty() ? 1 : jobs[0]->Wt, growth = jobs.empty() ? 2 : jobs[0]->Wg;
  Impl::TuneState *ts = nullptr;
  int cand = -1;
  Impl::TuneChoice tc;
  tc.kernel = eo.sweep_kernel;
  tc.first = first;
  tc.growth = growth;
  tc.threads = eo.pair_threads;
  tc.dense = eo.pair_dense;
  tc.max_words = eo.max_step_words;
  if (eo.autotune && !jobs.empty()) {
    bool any_drop = false;
    ts = &m.tune_state(m.tune_key(jobs, eo, any_drop), eo, any_drop);
    cand = ts->pick();
    tc = ts->cands[size_t(cand)];
    e2.sweep_kernel = tc.kernel;
    e2.pair_threads = tc.threads;
    e2.pair_dense = tc.dense;
    e2.max_step_words = tc.max_words;
    first = tc.first;
    growth = tc.growth;
    for (JobState *j : jobs)
      if (j->opt.drop && j->w0 == 0 && !j->done) {
        j->Wt = tc.first;
        j->Wg = tc.growth;
      }
  }
  const auto t1 = clk::now();
  const uint64_t setups = m.setup_events;
  while (iterate(jobs, e2) > 0) {
  }
  const auto t2 = clk::now();
  // A run that (re)allocated sweep scratch is not a fair sample: the
  // candidate stays untried and runs again.
  if (ts && m.setup_events == setups) {
    const double steps_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    double &slot = ts->ms[size_t(cand)];
    slot = slot < 0 ? steps_ms : std::min(slot, steps_ms);
    if (trace_on())
      std::fprintf(stderr,
                   "[lsim] autotune kernel=%d first=%d growth=%d threads=%d dense=%d max_words=%d steps %.2f ms "
                   "(%d of %d tried)\n",
                   tc.kernel, first, growth, tc.threads, tc.dense, tc.max_words, steps_ms, ts->tried(), ts->planned);
  }
  const double ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  for (JobState *j : jobs) {
    j->res.stats.t_total_ms = ms;
    j->res.stats.tune_trials = ts ? ts->tried() : 0;
    j->res.stats.tune_candidates = ts ? int64_t(ts->planned) : 0;
    j->res.stats.tuned = tc.code();
    j->res.stats.engine = e2.mode == ENGINE_PLAIN ? 0 : (e2.sweep_kernel == 0 ? 1 : 2);
  }
}

}  // namespace lsim
