// GPU stuck-at fault simulation for gfx950 (MI355X).
//
// What is computed is the reference's detection function d(f, r): fault f is
// detected in 64-pattern round r iff some PO's faulty word differs from its
// good word in the round's live bits (fsim/simulation.cpp:118-146,186-228).
// The reference re-simulates the whole circuit per (fault, round); here:
//
//  * Good machine: one pass over all rounds, node-major layout good[node][R]
//    (lane = round word), one launch per level.
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
//    which couples the 64 patterns of a word; circuits with a MUX inject
//    every fault directly (full word), skipping words where the site's good
//    value already equals the stuck value.
//  * Sweeps: a tile holds up to 64 lanes, each lane one (injection site,
//    word, injected bits). A wave evaluates one node for all lanes of a tile;
//    per node the tile keeps a 64-bit mask of the lanes whose value differs
//    from the good machine and a 64-word row of faulty values.
//      - plain form: one launch per level, every live node of the level
//        visited by every tile of the batch;
//      - event form: one persistent launch; a workgroup owns a tile at a
//        time and visits only nodes with a differing fanin, compacted per
//        level into an LDS worklist, ending at the deepest pending level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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
};

// Per-tile record of a node that differs from the good machine in some lane:
// the lane mask, valid only when `epoch` equals the tile's epoch.
struct MaskRec {
  unsigned long long mask;
  uint32_t epoch;
  uint32_t pad;
};

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

__device__ __forceinline__ uint64_t round_mask(int64_t num_patterns, int64_t w) {
  const int64_t bits = (num_patterns / (64 * (w + 1))) ? 64 : (num_patterns % 64);
  return bits < 64 ? ((uint64_t(1) << bits) - 1) : ~uint64_t(0);
}

// ---------------------------------------------------------------- good machine

// Threads cover (node, word) pairs of two device-position ranges (the live
// and the dead nodes of one level); word is the fastest index so a wave
// reads a node's consecutive words in one coalesced access.
__global__ void k_good_level(const uint32_t *__restrict__ info, const int4 *__restrict__ fin,
                             uint64_t *__restrict__ good, const uint64_t *__restrict__ pi_words, int64_t RW, int a0,
                             int n0, int a1, int n1) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n0 + n1) * RW) return;
  const int i = int(t / RW);
  const int64_t w = t - int64_t(i) * RW;
  const int u = i < n0 ? a0 + i : a1 + (i - n0);
  const uint32_t inf = info[u];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = fin[u];
  uint64_t r;
  if (type == G_PI) {
    r = pi_words[int64_t(f.x) * RW + w];
  } else {
    const uint64_t v0 = good[int64_t(f.x) * RW + w];
    const uint64_t v1 = nin > 1 ? good[int64_t(f.y) * RW + w] : 0;
    const uint64_t v2 = nin > 2 ? good[int64_t(f.z) * RW + w] : 0;
    const uint64_t v3 = nin > 3 ? good[int64_t(f.w) * RW + w] : 0;
    r = eval_gate(type, nin, v0, v1, v2, v3);
  }
  good[int64_t(u) * RW + w] = r;
}

// Masked good values of one round, gathered per device position.
__global__ void k_gather_round(const uint64_t *__restrict__ good, int64_t RW, int64_t w, int N, uint64_t mask,
                               uint64_t *__restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < N) out[u] = good[int64_t(u) * RW + w] & mask;
}

// ---------------------------------------------------------------- sweeps

struct SweepArgs {
  const uint32_t *__restrict__ info;
  const int4 *__restrict__ fin;
  const uint64_t *__restrict__ good;
  int64_t RW;
  const int32_t *__restrict__ live_lvl;  // first live device position of each level, L+2 entries
  int max_level;
  const Tile *__restrict__ tiles;
  int ntiles;
  const Lane *__restrict__ lanes;
  uint64_t *rows;    // per slot: [NL][64] faulty words
  uint64_t *masks;   // per slot: [NL] lane masks (plain form)
  int64_t rows_stride, masks_stride;
  uint64_t *out;     // lane results
  int *queue;        // event form: next tile
  unsigned long long *bytes;  // operand bytes touched
  MaskRec *mrec;     // per slot: [NL] (event form)
  uint32_t *pend;    // per slot: [NL], == epoch when the node must be evaluated
  int64_t rec_stride, pend_stride;
  const int32_t *__restrict__ fo_off;  // live fanouts, ascending position
  const int2 *__restrict__ fo;         // {position, level}
  uint32_t epoch_base;
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
__device__ __forceinline__ void sweep_node(const SweepArgs &a, int u, int base, int lane, const LaneSetup &s,
                                           uint64_t *rows, uint64_t *masks, uint64_t &acc, uint64_t &bytes) {
  u = __builtin_amdgcn_readfirstlane(u);
  const uint32_t inf = a.info[u];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = a.fin[u];
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
  const int64_t RW = a.RW;
  const uint64_t gu = a.good[int64_t(u) * RW + s.w];
  uint64_t r = gu;
  if (type != G_PI) {
    const uint64_t v0 = ((m0 >> lane) & 1) ? rows[int64_t(f.x) * 64 + lane] : a.good[int64_t(f.x) * RW + s.w];
    uint64_t v1 = 0, v2 = 0, v3 = 0;
    if (nin > 1) v1 = ((m1 >> lane) & 1) ? rows[int64_t(f.y) * 64 + lane] : a.good[int64_t(f.y) * RW + s.w];
    if (nin > 2) v2 = ((m2 >> lane) & 1) ? rows[int64_t(f.z) * 64 + lane] : a.good[int64_t(f.z) * RW + s.w];
    if (nin > 3) v3 = ((m3 >> lane) & 1) ? rows[int64_t(f.w) * 64 + lane] : a.good[int64_t(f.w) * RW + s.w];
    r = eval_gate(type, nin, v0, v1, v2, v3);
  }
  if (inj) r = inject(s, gu);
  const uint64_t d = r ^ gu;
  const uint64_t nm = __ballot(d != 0);
  if (d) rows[int64_t(u) * 64 + lane] = r;
  if (lane == 0) masks[u] = nm;
  if (type == G_PO) acc |= d;
  bytes += uint64_t(s.active) * 8 * (type == G_PI ? 1 : nin + 1) + 8 * uint64_t(__popcll(nm));
}

// Plain form: one launch per level; a wave handles one (tile, node) pair.
__global__ void k_sweep_level(SweepArgs a, int level, int tile0, int nbatch) {
  const int lane = threadIdx.x & 63;
  const int64_t gw = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int lo = a.live_lvl[level], hi = a.live_lvl[level + 1];
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
  sweep_node(a, u, a.live_lvl[T.start_level], lane, s, rows, masks, acc, bytes);
  if (acc && lane < s.active) atomicOr((unsigned long long *)&a.out[s.out], (unsigned long long)acc);
  if (lane == 0) atomicAdd(a.bytes, (unsigned long long)bytes);
}

// Event form. A level's pending nodes are staged in LDS by the scan that finds
// them: node record, live-fanout range and the lane masks of its fanins (the
// fanins sit on lower levels, so their masks are final). A MaskRec stamped
// with another epoch means "no lane differs".
struct NodeStage {
  uint64_t m[4];    // fanin lane masks
  int4 fin;
  int u;
  uint32_t info;
  int fo0, fo1;     // live fanouts fo[fo0, fo1)
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
template <int NB>
__device__ __forceinline__ void event_nodes(const SweepArgs &a, const NodeStage *st, int cnt, int lane,
                                            const LaneSetup &s, uint32_t epoch, uint64_t *rows, MaskRec *mrec,
                                            uint32_t *pend, uint64_t &acc, uint64_t &bytes, int *s_maxl,
                                            int *s_lvlcnt) {
  int u[NB];
  uint32_t type[NB], nin[NB];
  int4 f[NB];
  uint64_t m[NB][4], gu[NB], r[NB];
  const int64_t RW = a.RW;
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
    gu[k] = a.good[int64_t(u[k]) * RW + s.w];
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    r[k] = gu[k];
    if (k < cnt && type[k] != G_PI) {
      const uint32_t n = nin[k];
      const uint64_t v0 =
          ((m[k][0] >> lane) & 1) ? rows[int64_t(f[k].x) * 64 + lane] : a.good[int64_t(f[k].x) * RW + s.w];
      uint64_t v1 = 0, v2 = 0, v3 = 0;
      if (n > 1)
        v1 = ((m[k][1] >> lane) & 1) ? rows[int64_t(f[k].y) * 64 + lane] : a.good[int64_t(f[k].y) * RW + s.w];
      if (n > 2)
        v2 = ((m[k][2] >> lane) & 1) ? rows[int64_t(f[k].z) * 64 + lane] : a.good[int64_t(f[k].z) * RW + s.w];
      if (n > 3)
        v3 = ((m[k][3] >> lane) & 1) ? rows[int64_t(f[k].w) * 64 + lane] : a.good[int64_t(f[k].w) * RW + s.w];
      r[k] = eval_gate(type[k], n, v0, v1, v2, v3);
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k >= cnt) break;
    if (s.site == u[k]) r[k] = inject(s, gu[k]);
    const uint64_t d = r[k] ^ gu[k];
    const uint64_t nm = __ballot(d != 0);
    bytes += uint64_t(s.active) * 8 * (type[k] == G_PI ? 1 : nin[k] + 1);
    if (nm) {
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
        const int2 v = a.fo[e];
        pend[v.x] = epoch;
        atomicAdd(&s_lvlcnt[v.y], 1);
      }
      if (e1 > e0 && lane == 0) atomicMax(s_maxl, a.fo[e1 - 1].y);
      bytes += 8 * uint64_t(__popcll(nm)) + 16 + 12 * uint64_t(e1 - e0);
    }
    if (type[k] == G_PO) acc |= d;
  }
}

// Event form, persistent: a workgroup pulls tiles from a queue. For each tile
// it walks the levels from the lowest injection, skipping levels with no
// pending node (per-level counts in LDS). At a level with pending nodes, one
// thread per node checks the pending flag; those that find one append the node
// to the LDS stage together with its record and fanin masks; then the waves
// evaluate the staged nodes, two per wave at a time. The walk ends after the
// deepest level any pending node sits on. Pending flags and mask records
// carry the tile's epoch, so nothing in global memory is cleared between
// tiles; each level's LDS count is zeroed once the level is done. Every
// workgroup leaves once the queue passes the last tile.
constexpr int kChunk = 512;
constexpr int kNodesPerWave = 2;
__global__ void __launch_bounds__(kChunk) k_sweep_events(SweepArgs a) {
  __shared__ unsigned long long acc_lds[64];
  __shared__ int s_tile, s_cnt, s_maxl;
  __shared__ NodeStage s_stage[kChunk];
  extern __shared__ int s_lvlcnt[];  // max_level + 1 entries
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6, nwaves = blockDim.x >> 6;
  uint64_t *rows = a.rows + int64_t(blockIdx.x) * a.rows_stride;
  MaskRec *mrec = a.mrec + int64_t(blockIdx.x) * a.rec_stride;
  uint32_t *pend = a.pend + int64_t(blockIdx.x) * a.pend_stride;
  for (int i = tid; i <= a.max_level; i += blockDim.x) s_lvlcnt[i] = 0;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(a.queue, 1);
    if (tid < 64) acc_lds[tid] = 0;
    __syncthreads();
    const int t = s_tile;
    if (t >= a.ntiles) return;
    const Tile T = a.tiles[t];
    const LaneSetup s = lane_setup(a, T, lane);
    const uint32_t epoch = a.epoch_base + uint32_t(t) + 1u;
    if (tid < 64 && s.site >= 0) {
      pend[s.site] = epoch;
      atomicAdd(&s_lvlcnt[s.level], 1);
    }
    if (tid == 0) s_maxl = T.end_level;
    uint64_t acc = 0, bytes = 0, stage_bytes = 0;  // bytes: wave-uniform; stage_bytes: per lane
    __syncthreads();
    for (int l = T.start_level; l <= s_maxl; ++l) {
      if (s_lvlcnt[l] == 0) continue;
      const int lo = a.live_lvl[l], hi = a.live_lvl[l + 1];
      for (int c0 = lo; c0 < hi; c0 += blockDim.x) {
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
            const uint32_t inf = a.info[u];
            const int4 f = a.fin[u];
            const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
            n.u = u;
            n.info = inf;
            n.fin = f;
            n.fo0 = a.fo_off[u];
            n.fo1 = a.fo_off[u + 1];
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
        for (int i = wave * kNodesPerWave; i < n; i += nwaves * kNodesPerWave)
          event_nodes<kNodesPerWave>(a, &s_stage[i], min(kNodesPerWave, n - i), lane, s, epoch, rows, mrec, pend,
                                     acc, bytes, &s_maxl, s_lvlcnt);
        __syncthreads();
      }
      if (tid == 0) s_lvlcnt[l] = 0;
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

// ---------------------------------------------------------------- detection

struct DetectArgs {
  const uint32_t *__restrict__ info;
  const int4 *__restrict__ fin;
  const uint64_t *__restrict__ good;
  int64_t RW;
  const int32_t *__restrict__ parent;     // per live position, -1 at region heads
  const int32_t *__restrict__ head_slot;  // per live position: the stem's unit index in this step, else -1
  const uint8_t *__restrict__ is_po;      // per live position
  uint64_t *need;                         // [unit][Wn]: bits a unit's sweep must cover
  const uint64_t *__restrict__ dout;      // [unit][Wn]: sweep results
  int64_t w0;
  int Wn;
  const int32_t *__restrict__ flist;      // faults of this step (direct: unit k is fault flist[k])
  int nf;
  const int32_t *__restrict__ site;       // per fault, live position
  const int32_t *__restrict__ head;       // per fault, region head (live position)
  const uint8_t *__restrict__ sa;
  int64_t num_patterns;
  int32_t *first_round;                   // atomicMin, or null
  uint64_t *det_bits;                     // [F][RB], or null
  int64_t RB;
  uint8_t *last_flag;                     // d(f, last) bytes, or null
  int64_t last_word;
};

__device__ __forceinline__ void record(const DetectArgs &a, int f, int64_t w, uint64_t det) {
  if (!det) return;
  if (a.first_round) atomicMin(&a.first_round[f], int32_t(w));
  if (a.det_bits)
    atomicOr((unsigned long long *)&a.det_bits[int64_t(f) * a.RB + (w >> 6)], 1ull << (w & 63));
  if (a.last_flag && w == a.last_word) a.last_flag[f] = 1;
}

// Flip sensitivity of gate g to its input node x at word w: the patterns where
// complementing every fanin slot fed by x changes g's output.
__device__ __forceinline__ uint64_t sens(const DetectArgs &a, int g, int x, int64_t w) {
  const uint32_t inf = a.info[g];
  const uint32_t type = inf & 15u, nin = (inf >> 4) & 7u;
  const int4 f = a.fin[g];
  const uint64_t v0 = a.good[int64_t(f.x) * a.RW + w];
  const uint64_t v1 = nin > 1 ? a.good[int64_t(f.y) * a.RW + w] : 0;
  const uint64_t v2 = nin > 2 ? a.good[int64_t(f.z) * a.RW + w] : 0;
  const uint64_t v3 = nin > 3 ? a.good[int64_t(f.w) * a.RW + w] : 0;
  const uint64_t r0 = eval_gate(type, nin, v0, v1, v2, v3);
  const uint64_t r1 = eval_gate(type, nin, f.x == x ? ~v0 : v0, f.y == x ? ~v1 : v1, f.z == x ? ~v2 : v2,
                                f.w == x ? ~v3 : v3);
  return r0 ^ r1;
}

// L = activation & path sensitisation to the region head & live bits.
__device__ __forceinline__ uint64_t local_term(const DetectArgs &a, int f, int64_t w) {
  const int x0 = a.site[f];
  uint64_t D = (a.good[int64_t(x0) * a.RW + w] ^ (a.sa[f] ? ~uint64_t(0) : uint64_t(0))) &
               round_mask(a.num_patterns, w);
  for (int x = x0, g = a.parent[x]; g >= 0 && D; x = g, g = a.parent[x]) D &= sens(a, g, x, w);
  return D;
}

// Stem tracing, phase 1: thread = (fault, word). A fault whose region ends at
// a PO is decided here; otherwise its bits are added to its stem's need.
__global__ void k_local_cpt(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int fi = int(t / a.Wn);
  const int j = int(t - int64_t(fi) * a.Wn);
  const int f = a.flist[fi];
  const int64_t w = a.w0 + j;
  const uint64_t L = local_term(a, f, w);
  if (!L) return;
  const int h = a.head[f];
  if (a.is_po[h]) {
    record(a, f, w, L);
  } else {
    atomicOr((unsigned long long *)&a.need[int64_t(a.head_slot[h]) * a.Wn + j], (unsigned long long)L);
  }
}

// Stem tracing, phase 2: d = L & D_stem.
__global__ void k_detect_cpt(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int fi = int(t / a.Wn);
  const int j = int(t - int64_t(fi) * a.Wn);
  const int f = a.flist[fi];
  const int h = a.head[f];
  if (a.is_po[h]) return;
  const uint64_t D = a.dout[int64_t(a.head_slot[h]) * a.Wn + j];
  if (!D) return;
  const int64_t w = a.w0 + j;
  record(a, f, w, local_term(a, f, w) & D);
}

// Direct injection, phase 1: a word needs a lane iff the stuck value differs
// from the site's good value somewhere in the full word.
__global__ void k_local_direct(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int k = int(t / a.Wn);
  const int j = int(t - int64_t(k) * a.Wn);
  const int f = a.flist[k];
  a.need[t] = a.good[int64_t(a.site[f]) * a.RW + a.w0 + j] ^ (a.sa[f] ? ~uint64_t(0) : uint64_t(0));
}

// Direct injection, phase 2: the sweep result is the PO difference.
__global__ void k_detect_direct(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int k = int(t / a.Wn);
  const int j = int(t - int64_t(k) * a.Wn);
  const int64_t w = a.w0 + j;
  record(a, a.flist[k], w, a.dout[t] & round_mask(a.num_patterns, w));
}

__global__ void k_set_slots(int32_t *head_slot, const int32_t *units, int n, int clear) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head_slot[units[i]] = clear ? -1 : i;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    LSIM_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

inline unsigned blocks_for(int64_t threads, int bs) { return unsigned((threads + bs - 1) / bs); }

}  // namespace

struct Engine::Impl {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  // host view
  int64_t N = 0, NL = 0, P = 0, G = 0, O = 0, R = 0, F = 0, E = 0, num_patterns = 0;
  int L = 0;
  bool has_mux = false;
  std::vector<int32_t> pos2dp, dp2pos, level_dp, live_lvl, dead_lvl;
  std::vector<int32_t> f_site, f_head;  // device positions; head -1 = dead site
  std::vector<uint8_t> f_sa, head_is_po;
  int64_t live_faults = 0;
  // device: circuit
  DevBuf<uint32_t> d_info;
  DevBuf<int4> d_fin;
  DevBuf<int32_t> d_parent, d_live_lvl, d_site, d_head, d_head_slot, d_fo_off;
  DevBuf<uint8_t> d_sa, d_is_po;
  DevBuf<int2> d_fo;
  DevBuf<uint64_t> d_pi_words, d_good;
  // device: per step
  DevBuf<int32_t> d_flist, d_units, d_first;
  DevBuf<uint8_t> d_last;
  DevBuf<Lane> d_lanes;
  DevBuf<Tile> d_tiles;
  DevBuf<uint64_t> d_need, d_dout, d_detbits, d_gather;
  DevBuf<int> d_queue;
  DevBuf<unsigned long long> d_bytes;
  // device: sweep slots
  DevBuf<uint64_t> d_rows, d_masks;
  DevBuf<MaskRec> d_mrec;
  DevBuf<uint32_t> d_pend;
  int64_t slots = 0;
  uint32_t epoch_base = 0;
  std::vector<uint8_t> head_mark;
  hipEvent_t ev[2] = {nullptr, nullptr};

  ~Impl() {
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void time_begin() { LSIM_HIP(hipEventRecord(ev[0], stream)); }
  double time_end() {
    LSIM_HIP(hipEventRecord(ev[1], stream));
    LSIM_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    LSIM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    return ms;
  }

  void step(const RunOptions &opt, const std::vector<int32_t> &faults, int64_t w0, int Wn, bool direct,
            int32_t *first_dev, uint64_t *detbits_dev, int64_t RB, uint8_t *last_dev, RunResults &res);
  void sweep(const RunOptions &opt, const std::vector<Lane> &lanes, std::vector<Tile> &tiles, RunResults &res);
  void ensure_slots(const RunOptions &opt, int64_t want);
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
    LSIM_HIP(hipStreamCreateWithFlags(&impl_->stream, hipStreamNonBlocking));
    LSIM_HIP(hipEventCreate(&impl_->ev[0]));
    LSIM_HIP(hipEventCreate(&impl_->ev[1]));
  } catch (...) {
    delete impl_;
    impl_ = nullptr;
    throw;
  }
}

Engine::~Engine() {
  if (impl_) {
    (void)hipSetDevice(impl_->device);
    delete impl_;
  }
}

int Engine::device() const { return impl_->device; }

void Engine::load(const Netlist &nl, const FaultStructure &fs) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  m.N = nl.N;
  m.P = nl.P;
  m.G = nl.G;
  m.O = nl.O;
  m.E = nl.E();
  m.R = nl.R;
  m.F = int64_t(nl.fault_site.size());
  m.num_patterns = nl.num_patterns;
  m.L = nl.max_level;
  m.has_mux = nl.has_mux;
  const int64_t N = m.N;
  // Device order: live nodes (a PO is reachable) first, dead nodes after, each
  // in topological order so every level is a contiguous range in both parts.
  m.pos2dp.assign(N, 0);
  m.dp2pos.assign(N, 0);
  int64_t nl_count = 0;
  for (int64_t p = 0; p < N; ++p)
    if (fs.reach_po[p]) m.pos2dp[p] = int32_t(nl_count++);
  m.NL = nl_count;
  int64_t nd = nl_count;
  for (int64_t p = 0; p < N; ++p)
    if (!fs.reach_po[p]) m.pos2dp[p] = int32_t(nd++);
  for (int64_t p = 0; p < N; ++p) m.dp2pos[m.pos2dp[p]] = int32_t(p);
  m.level_dp.assign(N, 0);
  for (int64_t p = 0; p < N; ++p) m.level_dp[m.pos2dp[p]] = nl.level[p];
  m.live_lvl.assign(m.L + 2, 0);
  m.dead_lvl.assign(m.L + 2, 0);
  {
    std::vector<int64_t> cl(m.L + 2, 0), cd(m.L + 2, 0);
    for (int64_t p = 0; p < N; ++p) (fs.reach_po[p] ? cl : cd)[nl.level[p] + 1]++;
    m.live_lvl[0] = 0;
    m.dead_lvl[0] = int32_t(m.NL);
    for (int l = 0; l <= m.L; ++l) {
      m.live_lvl[l + 1] = int32_t(m.live_lvl[l] + cl[l + 1]);
      m.dead_lvl[l + 1] = int32_t(m.dead_lvl[l] + cd[l + 1]);
    }
  }
  const int64_t nl1 = std::max<int64_t>(m.NL, 1);
  std::vector<uint32_t> info(N);
  std::vector<int4> fin(N);
  std::vector<int32_t> parent(nl1, -1);
  for (int64_t p = 0; p < N; ++p) {
    const int32_t u = m.pos2dp[p];
    const int nin = nl.fin_off[p + 1] - nl.fin_off[p];
    info[u] = uint32_t(nl.type[p]) | (uint32_t(nin) << 4);
    int v[4] = {0, 0, 0, 0};
    if (nl.type[p] == G_PI) v[0] = int(p);  // PI index == position
    for (int k = 0; k < nin; ++k) v[k] = m.pos2dp[nl.fin[nl.fin_off[p] + k]];
    fin[u] = make_int4(v[0], v[1], v[2], v[3]);
    if (fs.reach_po[p] && fs.parent[p] >= 0) parent[u] = m.pos2dp[fs.parent[p]];
  }
  // Live fanouts of live nodes, ascending position, with the fanout's level.
  std::vector<int32_t> fo_off(nl1 + 1, 0);
  std::vector<int2> fo;
  for (int64_t u = 0; u < m.NL; ++u) {
    const int32_t p = m.dp2pos[u];
    std::vector<int32_t> outs;
    for (int32_t e = fs.fo_off[p]; e < fs.fo_off[p + 1]; ++e)
      if (fs.reach_po[fs.fo[e]]) outs.push_back(m.pos2dp[fs.fo[e]]);
    std::sort(outs.begin(), outs.end());
    for (int32_t v : outs) fo.push_back(make_int2(v, m.level_dp[v]));
    fo_off[u + 1] = int32_t(fo.size());
  }
  if (fo.empty()) fo.push_back(make_int2(0, 0));
  // Faults: site, and the head of the site's fanout-free region.
  m.f_site.resize(m.F);
  m.f_head.resize(m.F);
  m.f_sa.resize(m.F);
  m.head_is_po.assign(nl1, 0);
  for (int64_t p = 0; p < N; ++p)
    if (nl.type[p] == G_PO) m.head_is_po[m.pos2dp[p]] = 1;
  m.live_faults = 0;
  for (int64_t f = 0; f < m.F; ++f) {
    const int32_t p = nl.fault_site[f];
    m.f_site[f] = m.pos2dp[p];
    m.f_sa[f] = nl.fault_sa[f];
    m.f_head[f] = fs.reach_po[p] ? m.pos2dp[fs.stem_of[p]] : -1;
    m.live_faults += fs.reach_po[p] ? 1 : 0;
  }
  // PI words transposed to [PI][round] for coalesced level-0 reads.
  std::vector<uint64_t> piw(std::max<int64_t>(m.P * m.R, 1));
  for (int64_t r = 0; r < m.R; ++r)
    for (int64_t i = 0; i < m.P; ++i) piw[i * m.R + r] = nl.patterns[r * m.P + i];

  const int64_t F1 = std::max<int64_t>(m.F, 1);
  m.d_info.alloc(N);
  m.d_fin.alloc(N);
  m.d_parent.alloc(nl1);
  m.d_live_lvl.alloc(m.live_lvl.size());
  m.d_site.alloc(F1);
  m.d_head.alloc(F1);
  m.d_sa.alloc(F1);
  m.d_is_po.alloc(nl1);
  m.d_head_slot.alloc(nl1);
  m.d_fo_off.alloc(fo_off.size());
  m.d_fo.alloc(fo.size());
  m.d_pi_words.alloc(piw.size());
  m.d_good.alloc(size_t(std::max<int64_t>(N * m.R, 1)));
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    if (bytes) LSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, m.stream));
  };
  h2d(m.d_info.p, info.data(), N * sizeof(uint32_t));
  h2d(m.d_fin.p, fin.data(), N * sizeof(int4));
  h2d(m.d_parent.p, parent.data(), parent.size() * sizeof(int32_t));
  h2d(m.d_live_lvl.p, m.live_lvl.data(), m.live_lvl.size() * sizeof(int32_t));
  h2d(m.d_site.p, m.f_site.data(), m.F * sizeof(int32_t));
  h2d(m.d_head.p, m.f_head.data(), m.F * sizeof(int32_t));
  h2d(m.d_sa.p, m.f_sa.data(), m.F * sizeof(uint8_t));
  h2d(m.d_is_po.p, m.head_is_po.data(), m.head_is_po.size());
  h2d(m.d_fo_off.p, fo_off.data(), fo_off.size() * sizeof(int32_t));
  h2d(m.d_fo.p, fo.data(), fo.size() * sizeof(int2));
  h2d(m.d_pi_words.p, piw.data(), piw.size() * sizeof(uint64_t));
  k_fill_i32<<<blocks_for(nl1, 256), 256, 0, m.stream>>>(m.d_head_slot.p, nl1, -1);
  LSIM_HIP(hipGetLastError());
  LSIM_HIP(hipStreamSynchronize(m.stream));
}

void Engine::Impl::ensure_slots(const RunOptions &opt, int64_t want) {
  // Per slot: faulty rows [NL][64], and the plain form's lane masks or the
  // event form's mask records and pending flags.
  const int64_t per_slot = NL * (64 * 8 + 8 + int64_t(sizeof(MaskRec)) + 4);
  int64_t budget;
  if (opt.scratch_mb > 0) {
    budget = opt.scratch_mb << 20;
  } else {
    size_t fr = 0, tot = 0;
    LSIM_HIP(hipMemGetInfo(&fr, &tot));
    const int64_t held = int64_t(d_rows.n * 8 + d_masks.n * 8 + d_mrec.n * sizeof(MaskRec) + d_pend.n * 4);
    budget = (int64_t(fr) + held) / 10 * 7;
  }
  const int64_t cap = std::max<int64_t>(1, budget / std::max<int64_t>(per_slot, 1));
  // Up to three resident 8-wave workgroups per CU (the event sweep's register
  // use allows 6 waves per SIMD).
  const int64_t s = std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(cap, int64_t(cus) * 3)));
  if (s > slots || !d_rows.p) {
    d_rows.release();
    d_masks.release();
    d_mrec.release();
    d_pend.release();
    const int64_t nl = std::max<int64_t>(NL, 1);
    d_rows.alloc(size_t(s) * nl * 64);
    d_masks.alloc(size_t(s) * nl);
    d_mrec.alloc(size_t(s) * nl);
    d_pend.alloc(size_t(s) * nl);
    // Epoch 0 is never used by a tile, so zeroed records read as "no lane differs".
    LSIM_HIP(hipMemsetAsync(d_mrec.p, 0, size_t(s) * nl * sizeof(MaskRec), stream));
    LSIM_HIP(hipMemsetAsync(d_pend.p, 0, size_t(s) * nl * sizeof(uint32_t), stream));
    slots = s;
  }
}

void Engine::Impl::sweep(const RunOptions &opt, const std::vector<Lane> &lanes, std::vector<Tile> &tiles,
                         RunResults &res) {
  lsim_run_stats &st = res.stats;
  const int64_t ntiles = int64_t(tiles.size());
  if (!ntiles) return;
  d_lanes.alloc(lanes.size());
  d_tiles.alloc(tiles.size());
  LSIM_HIP(hipMemcpyAsync(d_lanes.p, lanes.data(), lanes.size() * sizeof(Lane), hipMemcpyHostToDevice, stream));
  LSIM_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
  ensure_slots(opt, ntiles);
  SweepArgs a;
  a.info = d_info.p;
  a.fin = d_fin.p;
  a.good = d_good.p;
  a.RW = R;
  a.live_lvl = d_live_lvl.p;
  a.max_level = L;
  a.tiles = d_tiles.p;
  a.ntiles = int(ntiles);
  a.lanes = d_lanes.p;
  a.rows = d_rows.p;
  a.masks = d_masks.p;
  a.rows_stride = NL * 64;
  a.masks_stride = NL;
  a.out = d_dout.p;
  d_queue.alloc(1);
  a.queue = d_queue.p;
  a.bytes = d_bytes.p;
  a.mrec = d_mrec.p;
  a.pend = d_pend.p;
  a.rec_stride = NL;
  a.pend_stride = NL;
  a.fo_off = d_fo_off.p;
  a.fo = d_fo.p;
  a.epoch_base = 0;
  st.sweep_tiles += ntiles;
  time_begin();
  if (opt.mode == LSIM_MODE_PARTITIONED) {
    // Tile epochs are unique over the engine's lifetime; before the 32-bit
    // counter wraps, the records are cleared and numbering restarts.
    if (uint64_t(epoch_base) + uint64_t(ntiles) + 1 >= (uint64_t(1) << 31)) {
      LSIM_HIP(hipMemsetAsync(d_mrec.p, 0, d_mrec.n * sizeof(MaskRec), stream));
      LSIM_HIP(hipMemsetAsync(d_pend.p, 0, d_pend.n * sizeof(uint32_t), stream));
      epoch_base = 0;
    }
    a.epoch_base = epoch_base;
    epoch_base += uint32_t(ntiles);
    LSIM_HIP(hipMemsetAsync(d_queue.p, 0, sizeof(int), stream));
    const int bw = std::max(1, std::min(kChunk / 64, opt.block_waves));
    const unsigned grid = unsigned(std::min<int64_t>(slots, ntiles));
    k_sweep_events<<<grid, 64 * bw, size_t(L + 1) * sizeof(int), stream>>>(a);
    LSIM_HIP(hipGetLastError());
    st.sweep_launches++;
  } else {
    for (int64_t t0 = 0; t0 < ntiles; t0 += slots) {
      const int nb = int(std::min<int64_t>(slots, ntiles - t0));
      int lmin = tiles[t0].start_level;
      for (int64_t t = t0; t < t0 + nb; ++t) lmin = std::min(lmin, tiles[t].start_level);
      for (int l = lmin; l <= L; ++l) {
        const int64_t n = live_lvl[l + 1] - live_lvl[l];
        if (n <= 0) continue;
        const int64_t threads = int64_t(nb) * n * 64;
        k_sweep_level<<<blocks_for(threads, 256), 256, 0, stream>>>(a, l, int(t0), nb);
        LSIM_HIP(hipGetLastError());
        st.sweep_launches++;
      }
      for (int64_t t = t0; t < t0 + nb; ++t) st.sweep_node_steps += NL - live_lvl[tiles[t].start_level];
    }
  }
  st.t_sweep_ms += time_end();
}

void Engine::Impl::step(const RunOptions &opt, const std::vector<int32_t> &faults, int64_t w0, int Wn, bool direct,
                        int32_t *first_dev, uint64_t *detbits_dev, int64_t RB, uint8_t *last_dev, RunResults &res) {
  lsim_run_stats &st = res.stats;
  st.steps++;
  const int64_t nf = int64_t(faults.size());
  if (!nf) return;
  // Units: stems of the faults' regions (stem tracing) or the faults
  // themselves in site order (direct injection).
  std::vector<int32_t> flist, units;
  if (direct) {
    std::vector<int32_t> cnt(size_t(NL) + 1, 0);
    for (int32_t f : faults) cnt[f_site[f] + 1]++;
    for (int64_t u = 0; u < NL; ++u) cnt[u + 1] += cnt[u];
    flist.resize(faults.size());
    for (int32_t f : faults) flist[cnt[f_site[f]]++] = f;
    units.resize(flist.size());
    for (size_t k = 0; k < flist.size(); ++k) units[k] = f_site[flist[k]];
  } else {
    flist = faults;
    head_mark.assign(size_t(NL), 0);
    for (int32_t f : flist) {
      const int32_t h = f_head[f];
      if (!head_is_po[h]) head_mark[h] = 1;
    }
    for (int64_t u = 0; u < NL; ++u)
      if (head_mark[u]) units.push_back(int32_t(u));
  }
  const int64_t nu = int64_t(units.size());
  d_flist.alloc(nf);
  LSIM_HIP(hipMemcpyAsync(d_flist.p, flist.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  const int64_t nuw = std::max<int64_t>(nu * Wn, 1);
  d_need.alloc(nuw);
  d_dout.alloc(nuw);
  LSIM_HIP(hipMemsetAsync(d_need.p, 0, nuw * sizeof(uint64_t), stream));
  LSIM_HIP(hipMemsetAsync(d_dout.p, 0, nuw * sizeof(uint64_t), stream));
  if (!direct && nu) {
    d_units.alloc(nu);
    LSIM_HIP(hipMemcpyAsync(d_units.p, units.data(), nu * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    k_set_slots<<<blocks_for(nu, 256), 256, 0, stream>>>(d_head_slot.p, d_units.p, int(nu), 0);
    LSIM_HIP(hipGetLastError());
  }
  DetectArgs d;
  d.info = d_info.p;
  d.fin = d_fin.p;
  d.good = d_good.p;
  d.RW = R;
  d.parent = d_parent.p;
  d.head_slot = d_head_slot.p;
  d.is_po = d_is_po.p;
  d.need = d_need.p;
  d.dout = d_dout.p;
  d.w0 = w0;
  d.Wn = Wn;
  d.flist = d_flist.p;
  d.nf = int(nf);
  d.site = d_site.p;
  d.head = d_head.p;
  d.sa = d_sa.p;
  d.num_patterns = num_patterns;
  d.first_round = first_dev;
  d.det_bits = detbits_dev;
  d.RB = RB;
  d.last_flag = last_dev;
  d.last_word = R - 1;
  time_begin();
  if (direct) k_local_direct<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
  else k_local_cpt<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
  LSIM_HIP(hipGetLastError());
  // The needed bits decide which (unit, word) pairs become sweep lanes.
  std::vector<uint64_t> need(size_t(nu) * Wn);
  if (nu)
    LSIM_HIP(hipMemcpyAsync(need.data(), d_need.p, need.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  st.t_detect_ms += time_end();
  std::vector<Lane> lanes;
  for (int64_t k = 0; k < nu; ++k)
    for (int j = 0; j < Wn; ++j) {
      const uint64_t m = need[size_t(k) * Wn + j];
      if (!m) continue;
      Lane ln;
      ln.mask = m;
      ln.site = units[k];
      ln.word = int32_t(w0 + j);
      ln.out = int32_t(k * Wn + j);
      ln.kind = direct ? (f_sa[flist[k]] ? K_SA1 : K_SA0) : K_FLIP;
      ln.level = level_dp[units[k]];
      ln.pad = 0;
      lanes.push_back(ln);
    }
  std::vector<Tile> tiles;
  for (size_t i = 0; i < lanes.size(); i += 64) {
    Tile t;
    t.lane0 = int32_t(i);
    t.nlanes = int32_t(std::min<size_t>(64, lanes.size() - i));
    t.start_level = level_dp[lanes[i].site];
    t.end_level = level_dp[lanes[i + t.nlanes - 1].site];
    tiles.push_back(t);
  }
  st.stems += direct ? 0 : nu;
  const double before = st.t_sweep_ms;
  sweep(opt, lanes, tiles, res);
  if (std::getenv("LSIM_TRACE"))
    std::fprintf(stderr, "[lsim] step w0=%lld Wn=%d faults=%lld units=%lld lanes=%lld tiles=%lld sweep=%.2fms\n",
                 (long long)w0, Wn, (long long)nf, (long long)nu, (long long)lanes.size(), (long long)tiles.size(),
                 st.t_sweep_ms - before);
  time_begin();
  if (direct) k_detect_direct<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
  else k_detect_cpt<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
  LSIM_HIP(hipGetLastError());
  if (!direct && nu) {
    k_set_slots<<<blocks_for(nu, 256), 256, 0, stream>>>(d_head_slot.p, d_units.p, int(nu), 1);
    LSIM_HIP(hipGetLastError());
  }
  st.t_detect_ms += time_end();
}

void Engine::run(const RunOptions &opt, RunResults &res) {
  Impl &m = *impl_;
  LSIM_HIP(hipSetDevice(m.device));
  res = RunResults();
  lsim_run_stats &st = res.stats;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t N = m.N, R = m.R, F = m.F;
  st.live_nodes = m.NL;
  st.live_faults = m.live_faults;
  m.d_bytes.alloc(1);
  LSIM_HIP(hipMemsetAsync(m.d_bytes.p, 0, sizeof(unsigned long long), m.stream));
  // Good machine over every round.
  if (R > 0) {
    m.time_begin();
    for (int l = 0; l <= m.L; ++l) {
      const int a0 = m.live_lvl[l], n0 = m.live_lvl[l + 1] - a0;
      const int a1 = m.dead_lvl[l], n1 = m.dead_lvl[l + 1] - a1;
      const int64_t threads = int64_t(n0 + n1) * R;
      if (threads == 0) continue;
      k_good_level<<<blocks_for(threads, 256), 256, 0, m.stream>>>(m.d_info.p, m.d_fin.p, m.d_good.p,
                                                                  m.d_pi_words.p, R, a0, n0, a1, n1);
      LSIM_HIP(hipGetLastError());
    }
    st.t_good_ms = m.time_end();
    // One word per node written, one per fanin and per PI pattern read.
    st.good_bytes = 8 * (N + m.E + m.P) * R;
  }
  // Last-round good values, masked (fsim/simulation.cpp:163-178).
  res.good_last.assign(N, 0);
  if (R > 0) {
    const int64_t bits = (m.num_patterns / (64 * R)) ? 64 : (m.num_patterns % 64);
    const uint64_t mask = bits < 64 ? ((uint64_t(1) << bits) - 1) : ~uint64_t(0);
    m.d_gather.alloc(N);
    k_gather_round<<<blocks_for(N, 256), 256, 0, m.stream>>>(m.d_good.p, R, R - 1, int(N), mask, m.d_gather.p);
    LSIM_HIP(hipGetLastError());
    std::vector<uint64_t> byd(N);
    LSIM_HIP(hipMemcpyAsync(byd.data(), m.d_gather.p, N * sizeof(uint64_t), hipMemcpyDeviceToHost, m.stream));
    LSIM_HIP(hipStreamSynchronize(m.stream));
    for (int64_t p = 0; p < N; ++p) res.good_last[p] = byd[m.pos2dp[p]];
  }
  std::vector<int32_t> live;
  live.reserve(m.live_faults);
  for (int64_t f = 0; f < F; ++f)
    if (m.f_head[f] >= 0) live.push_back(int32_t(f));
  const bool direct = opt.direct || m.has_mux;
  const int64_t F1 = std::max<int64_t>(F, 1);
  m.d_first.alloc(F1);
  k_fill_i32<<<blocks_for(F1, 256), 256, 0, m.stream>>>(m.d_first.p, F1, INT32_MAX);
  LSIM_HIP(hipGetLastError());
  res.first_round.assign(F, INT32_MAX);
  const int64_t RB = (R + 63) / 64;
  if (R > 0 && !live.empty()) {
    if (!opt.drop) {
      m.d_detbits.alloc(size_t(std::max<int64_t>(F * RB, 1)));
      LSIM_HIP(hipMemsetAsync(m.d_detbits.p, 0, size_t(std::max<int64_t>(F * RB, 1)) * sizeof(uint64_t), m.stream));
      for (int64_t w0 = 0; w0 < R;) {
        const int Wn = int(std::min<int64_t>(64, R - w0));
        m.step(opt, live, w0, Wn, direct, m.d_first.p, m.d_detbits.p, RB, nullptr, res);
        w0 += Wn;
      }
      res.RB = RB;
      res.det_bits.assign(size_t(F * RB), 0);
      if (F)
        LSIM_HIP(hipMemcpyAsync(res.det_bits.data(), m.d_detbits.p, size_t(F * RB) * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, m.stream));
      LSIM_HIP(hipStreamSynchronize(m.stream));
      res.has_matrix = true;
      res.last_round.assign(F, 0);
      for (int64_t f = 0; f < F; ++f)
        res.last_round[f] = uint8_t((res.det_bits[f * RB + ((R - 1) >> 6)] >> ((R - 1) & 63)) & 1);
      res.has_last = true;
    } else {
      // Words per step grow geometrically: the faults left after the first
      // rounds are the hard ones and are dropped less often.
      std::vector<int32_t> alive = live;
      std::vector<int32_t> first(F);
      int Wt = 1;
      for (int64_t w0 = 0; w0 < R && !alive.empty();) {
        const int Wn = int(std::min<int64_t>(Wt, R - w0));
        m.step(opt, alive, w0, Wn, direct, m.d_first.p, nullptr, 0, nullptr, res);
        LSIM_HIP(hipMemcpyAsync(first.data(), m.d_first.p, F * sizeof(int32_t), hipMemcpyDeviceToHost, m.stream));
        LSIM_HIP(hipStreamSynchronize(m.stream));
        std::vector<int32_t> next;
        next.reserve(alive.size());
        for (int32_t f : alive)
          if (first[f] == INT32_MAX) next.push_back(f);
        alive.swap(next);
        w0 += Wn;
        Wt = std::min(64, Wt * 2);
      }
      if (opt.report_last_round) {
        m.d_last.alloc(F1);
        LSIM_HIP(hipMemsetAsync(m.d_last.p, 0, F1, m.stream));
        m.step(opt, live, R - 1, 1, direct, nullptr, nullptr, 0, m.d_last.p, res);
        res.last_round.assign(F, 0);
        if (F) LSIM_HIP(hipMemcpyAsync(res.last_round.data(), m.d_last.p, F, hipMemcpyDeviceToHost, m.stream));
        res.has_last = true;
      }
    }
    LSIM_HIP(hipMemcpyAsync(res.first_round.data(), m.d_first.p, F * sizeof(int32_t), hipMemcpyDeviceToHost,
                            m.stream));
    LSIM_HIP(hipStreamSynchronize(m.stream));
  } else {
    // No rounds or no observable fault: nothing is ever detected.
    res.has_matrix = !opt.drop;
    res.RB = RB;
    res.det_bits.assign(size_t(F * RB), 0);
    res.last_round.assign(F, 0);
    res.has_last = !opt.drop || opt.report_last_round;
  }
  for (auto &v : res.first_round)
    if (v == INT32_MAX) v = -1;
  {
    unsigned long long b = 0;
    LSIM_HIP(hipMemcpyAsync(&b, m.d_bytes.p, sizeof(b), hipMemcpyDeviceToHost, m.stream));
    LSIM_HIP(hipStreamSynchronize(m.stream));
    st.sweep_bytes = int64_t(b);
  }
  st.slots = m.slots;
  st.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace lsim
