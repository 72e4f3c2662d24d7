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
//  * Stem tracing (circuits without MUX): inside a fanout-free region a
//    flip of node x reaches the region's head h exactly where every gate on
//    the path is sensitised, so d(x/v) = (good_x ^ v) & sens(x->h) & D_h,
//    with D_h the patterns where flipping h changes a PO. D_h of a PO is all
//    ones; D_h of a stem comes from an explicit forward sweep of the flip.
//    Gates are bitwise per pattern, so this is exact. The reference's MUX
//    uses the word-level `!s` (fsim/simulation.cpp:1056), which couples the
//    64 patterns of a word; circuits with a MUX inject every fault directly.
//  * Sweeps: a tile is up to 64 lanes, lane = (injection, word). A wave
//    evaluates one node for all lanes; per node the tile keeps a 64-bit mask
//    of lanes whose value differs from the good machine and a 64-word row of
//    faulty values. A node is skipped when no fanin differs for any lane.
//    Tiles sweep the live nodes level by level from the first injection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

struct Tile {
  int32_t inj_begin;   // first injection of the tile
  int32_t K;           // injections in the tile
  int32_t Wt;          // words per injection; lane = k * Wt + j
  int32_t word0;       // first word
  int32_t start_level; // level of the first (lowest) injection site
  int32_t out_begin;   // out[out_begin + lane]
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
  const int32_t *__restrict__ inj_site;
  const uint8_t *__restrict__ inj_kind;
  uint64_t *rows;    // per slot: [NL][64] faulty words
  uint64_t *masks;   // per slot: [NL] lane masks
  int64_t rows_stride, masks_stride;
  uint64_t *out;     // [injection][word] OR over POs of (faulty ^ good)
  int *queue;        // persistent form: next tile
  unsigned long long *bytes;  // operand bytes touched (see sweep_node)
};

struct LaneSetup {
  int site;
  uint32_t kind;
  int64_t w;
  int active;
};

__device__ __forceinline__ LaneSetup lane_setup(const SweepArgs &a, const Tile &T, int lane) {
  LaneSetup s;
  s.active = T.K * T.Wt;
  const bool on = lane < s.active;
  const int k = on ? lane / T.Wt : 0;
  const int j = on ? lane - k * T.Wt : 0;
  s.site = on ? a.inj_site[T.inj_begin + k] : -1;
  s.kind = on ? a.inj_kind[T.inj_begin + k] : 0u;
  s.w = int64_t(T.word0) + j;
  return s;
}

// Evaluates live node u for the 64 lanes of one tile; `base` is the first
// live position of the tile's start level: positions below it lie outside
// every injection's fanout and hold good values for all lanes.
// `bytes` accumulates (wave-uniformly) the operand bytes the node visit
// touches: node record and fanin masks always; for an evaluated node also one
// good or faulty word per fanin and the node's good word per active lane, and
// the faulty words written.
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
  if (inj) r = s.kind == K_FLIP ? ~gu : (s.kind == K_SA1 ? ~uint64_t(0) : uint64_t(0));
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
  if (acc && lane < s.active) atomicOr((unsigned long long *)&a.out[T.out_begin + lane], (unsigned long long)acc);
  if (lane == 0) atomicAdd(a.bytes, (unsigned long long)bytes);
}

// Persistent form: a workgroup pulls tiles from a queue and sweeps the levels
// of each with a workgroup barrier between levels. Every workgroup leaves the
// loop once the queue index passes the last tile.
__global__ void __launch_bounds__(1024) k_sweep_persistent(SweepArgs a) {
  __shared__ unsigned long long acc_lds[64];
  __shared__ int s_tile;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  uint64_t *rows = a.rows + int64_t(blockIdx.x) * a.rows_stride;
  uint64_t *masks = a.masks + int64_t(blockIdx.x) * a.masks_stride;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(a.queue, 1);
    if (threadIdx.x < 64) acc_lds[threadIdx.x] = 0;
    __syncthreads();
    const int t = s_tile;
    if (t >= a.ntiles) return;
    const Tile T = a.tiles[t];
    const LaneSetup s = lane_setup(a, T, lane);
    const int base = a.live_lvl[T.start_level];
    uint64_t acc = 0, bytes = 0;
    for (int l = T.start_level; l <= a.max_level; ++l) {
      const int lo = a.live_lvl[l], hi = a.live_lvl[l + 1];
      for (int u = lo + wave; u < hi; u += nwaves) sweep_node(a, u, base, lane, s, rows, masks, acc, bytes);
      __syncthreads();
    }
    if (lane == 0 && bytes) atomicAdd(a.bytes, (unsigned long long)bytes);
    if (acc) atomicOr(&acc_lds[lane], (unsigned long long)acc);
    __syncthreads();
    if (threadIdx.x < s.active) a.out[T.out_begin + threadIdx.x] = acc_lds[threadIdx.x];
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
  const int32_t *__restrict__ head_slot;  // per live position: stem's injection index this step, else -1
  const uint64_t *__restrict__ dout;      // [injection][Wn]
  int64_t w0;
  int Wn;
  const int32_t *__restrict__ flist;      // faults of this step (direct: injection order)
  int nf;
  const int32_t *__restrict__ site;       // per fault, live position
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

// Stem tracing: thread = (fault, word).
__global__ void k_cpt_detect(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int fi = int(t / a.Wn);
  const int j = int(t - int64_t(fi) * a.Wn);
  const int f = a.flist[fi];
  const int x0 = a.site[f];
  const int64_t w = a.w0 + j;
  uint64_t D = ~uint64_t(0);
  int x = x0;
  for (int g = a.parent[x]; g >= 0 && D; g = a.parent[x]) {
    D &= sens(a, g, x, w);
    x = g;
  }
  if ((a.info[x] & 15u) != G_PO) {
    const int slot = a.head_slot[x];
    D &= slot >= 0 ? a.dout[int64_t(slot) * a.Wn + j] : 0;
  }
  const uint64_t act = a.good[int64_t(x0) * a.RW + w] ^ (a.sa[f] ? ~uint64_t(0) : uint64_t(0));
  record(a, f, w, act & D & round_mask(a.num_patterns, w));
}

// Direct injection: thread = (injection, word); the sweep output is the PO diff.
__global__ void k_direct_detect(DetectArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(a.nf) * a.Wn) return;
  const int k = int(t / a.Wn);
  const int j = int(t - int64_t(k) * a.Wn);
  const int64_t w = a.w0 + j;
  record(a, a.flist[k], w, a.dout[t] & round_mask(a.num_patterns, w));
}

__global__ void k_set_slots(int32_t *head_slot, const int32_t *inj_site, int n, int clear) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head_slot[inj_site[i]] = clear ? -1 : i;
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
  // device
  DevBuf<uint32_t> d_info;
  DevBuf<int4> d_fin;
  DevBuf<int32_t> d_parent, d_live_lvl, d_site, d_head_slot;
  DevBuf<uint8_t> d_sa;
  DevBuf<uint64_t> d_pi_words, d_good;
  // per step
  DevBuf<int32_t> d_inj_site, d_flist, d_first;
  DevBuf<uint8_t> d_inj_kind, d_last;
  DevBuf<Tile> d_tiles;
  DevBuf<uint64_t> d_out, d_rows, d_masks, d_detbits, d_gather;
  DevBuf<int> d_queue;
  DevBuf<unsigned long long> d_bytes;
  int64_t slots = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};

  ~Impl() {
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (stream) (void)hipStreamDestroy(stream);
  }
  double time_begin() {
    LSIM_HIP(hipEventRecord(ev[0], stream));
    return 0;
  }
  double time_end() {
    LSIM_HIP(hipEventRecord(ev[1], stream));
    LSIM_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    LSIM_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    return ms;
  }

  void step(const RunOptions &opt, const std::vector<int32_t> &faults, int64_t w0, int Wn, bool direct,
            int32_t *first_dev, uint64_t *detbits_dev, int64_t RB, uint8_t *last_dev, RunResults &res);
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
  std::vector<uint32_t> info(N);
  std::vector<int4> fin(N);
  std::vector<int32_t> parent(std::max<int64_t>(m.NL, 1), -1);
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
  // Faults: site, and the head of the site's fanout-free region.
  m.f_site.resize(m.F);
  m.f_head.resize(m.F);
  m.f_sa.resize(m.F);
  m.head_is_po.assign(std::max<int64_t>(m.NL, 1), 0);
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

  m.d_info.alloc(N);
  m.d_fin.alloc(N);
  m.d_parent.alloc(parent.size());
  m.d_live_lvl.alloc(m.live_lvl.size());
  m.d_site.alloc(std::max<int64_t>(m.F, 1));
  m.d_sa.alloc(std::max<int64_t>(m.F, 1));
  m.d_head_slot.alloc(std::max<int64_t>(m.NL, 1));
  m.d_pi_words.alloc(piw.size());
  m.d_good.alloc(size_t(std::max<int64_t>(N * m.R, 1)));
  LSIM_HIP(hipMemcpyAsync(m.d_info.p, info.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice, m.stream));
  LSIM_HIP(hipMemcpyAsync(m.d_fin.p, fin.data(), N * sizeof(int4), hipMemcpyHostToDevice, m.stream));
  LSIM_HIP(hipMemcpyAsync(m.d_parent.p, parent.data(), parent.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                          m.stream));
  LSIM_HIP(hipMemcpyAsync(m.d_live_lvl.p, m.live_lvl.data(), m.live_lvl.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, m.stream));
  if (m.F) {
    LSIM_HIP(hipMemcpyAsync(m.d_site.p, m.f_site.data(), m.F * sizeof(int32_t), hipMemcpyHostToDevice, m.stream));
    LSIM_HIP(hipMemcpyAsync(m.d_sa.p, m.f_sa.data(), m.F * sizeof(uint8_t), hipMemcpyHostToDevice, m.stream));
  }
  LSIM_HIP(hipMemcpyAsync(m.d_pi_words.p, piw.data(), piw.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                          m.stream));
  k_fill_i32<<<blocks_for(std::max<int64_t>(m.NL, 1), 256), 256, 0, m.stream>>>(m.d_head_slot.p,
                                                                              std::max<int64_t>(m.NL, 1), -1);
  LSIM_HIP(hipGetLastError());
  LSIM_HIP(hipStreamSynchronize(m.stream));
}

void Engine::Impl::ensure_slots(const RunOptions &opt, int64_t want) {
  const int64_t per_slot = (NL * 64 + NL) * int64_t(sizeof(uint64_t));
  int64_t budget;
  if (opt.scratch_mb > 0) {
    budget = opt.scratch_mb << 20;
  } else {
    size_t fr = 0, tot = 0;
    LSIM_HIP(hipMemGetInfo(&fr, &tot));
    const int64_t held = int64_t((d_rows.n + d_masks.n) * sizeof(uint64_t));
    budget = std::min<int64_t>(int64_t(fr) / 2 + held, int64_t(64) << 30);
  }
  int64_t cap = std::max<int64_t>(1, budget / std::max<int64_t>(per_slot, 1));
  int64_t s = std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(cap, int64_t(cus) * 2)));
  if (s > slots || !d_rows.p) {
    d_rows.release();
    d_masks.release();
    d_rows.alloc(size_t(s) * NL * 64);
    d_masks.alloc(size_t(s) * NL);
    slots = s;
  }
}

void Engine::Impl::step(const RunOptions &opt, const std::vector<int32_t> &faults, int64_t w0, int Wn, bool direct,
                        int32_t *first_dev, uint64_t *detbits_dev, int64_t RB, uint8_t *last_dev, RunResults &res) {
  // Injections of this step, ordered by device position so that tiles start
  // at increasing levels and neighbouring lanes share most of their sweep.
  std::vector<int32_t> inj_site, flist;
  std::vector<uint8_t> inj_kind;
  flist = faults;
  std::sort(flist.begin(), flist.end(), [&](int32_t a, int32_t b) {
    return f_site[a] != f_site[b] ? f_site[a] < f_site[b] : a < b;
  });
  if (direct) {
    inj_site.resize(flist.size());
    inj_kind.resize(flist.size());
    for (size_t k = 0; k < flist.size(); ++k) {
      inj_site[k] = f_site[flist[k]];
      inj_kind[k] = f_sa[flist[k]] ? K_SA1 : K_SA0;
    }
  } else {
    for (int32_t f : flist) {
      const int32_t h = f_head[f];
      if (!head_is_po[h]) inj_site.push_back(h);
    }
    std::sort(inj_site.begin(), inj_site.end());
    inj_site.erase(std::unique(inj_site.begin(), inj_site.end()), inj_site.end());
    inj_kind.assign(inj_site.size(), K_FLIP);
  }
  const int64_t ninj = int64_t(inj_site.size());
  const int Wt = Wn;
  const int K = std::max(1, 64 / Wt);
  std::vector<Tile> tiles;
  for (int64_t i = 0; i < ninj; i += K) {
    Tile t;
    t.inj_begin = int32_t(i);
    t.K = int32_t(std::min<int64_t>(K, ninj - i));
    t.Wt = Wt;
    t.word0 = int32_t(w0);
    t.start_level = level_dp[inj_site[i]];
    t.out_begin = int32_t(i * Wn);
    tiles.push_back(t);
  }
  const int64_t nf = int64_t(flist.size());
  d_flist.alloc(std::max<int64_t>(nf, 1));
  if (nf)
    LSIM_HIP(hipMemcpyAsync(d_flist.p, flist.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  lsim_run_stats &st = res.stats;
  st.steps++;
  if (ninj > 0) {
    d_inj_site.alloc(ninj);
    d_inj_kind.alloc(ninj);
    d_tiles.alloc(tiles.size());
    d_out.alloc(size_t(ninj) * Wn);
    LSIM_HIP(hipMemcpyAsync(d_inj_site.p, inj_site.data(), ninj * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    LSIM_HIP(hipMemcpyAsync(d_inj_kind.p, inj_kind.data(), ninj, hipMemcpyHostToDevice, stream));
    LSIM_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, stream));
    LSIM_HIP(hipMemsetAsync(d_out.p, 0, size_t(ninj) * Wn * sizeof(uint64_t), stream));
    if (!direct) {
      k_set_slots<<<blocks_for(ninj, 256), 256, 0, stream>>>(d_head_slot.p, d_inj_site.p, int(ninj), 0);
      LSIM_HIP(hipGetLastError());
    }
    const int64_t ntiles = int64_t(tiles.size());
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
    a.inj_site = d_inj_site.p;
    a.inj_kind = d_inj_kind.p;
    a.rows = d_rows.p;
    a.masks = d_masks.p;
    a.rows_stride = NL * 64;
    a.masks_stride = NL;
    a.out = d_out.p;
    d_queue.alloc(1);
    a.queue = d_queue.p;
    a.bytes = d_bytes.p;
    for (const Tile &t : tiles) st.sweep_node_steps += NL - live_lvl[t.start_level];
    st.sweep_tiles += ntiles;
    st.stems += direct ? 0 : ninj;
    time_begin();
    if (opt.mode == LSIM_MODE_PARTITIONED) {
      LSIM_HIP(hipMemsetAsync(d_queue.p, 0, sizeof(int), stream));
      const int bw = std::max(1, std::min(16, opt.block_waves));
      const unsigned grid = unsigned(std::min<int64_t>(slots, ntiles));
      k_sweep_persistent<<<grid, 64 * bw, 0, stream>>>(a);
      LSIM_HIP(hipGetLastError());
      st.sweep_launches++;
    } else {
      for (int64_t t0 = 0; t0 < ntiles; t0 += slots) {
        const int nb = int(std::min<int64_t>(slots, ntiles - t0));
        for (int l = tiles[t0].start_level; l <= L; ++l) {
          const int64_t n = live_lvl[l + 1] - live_lvl[l];
          if (n <= 0) continue;
          const int64_t threads = int64_t(nb) * n * 64;
          k_sweep_level<<<blocks_for(threads, 256), 256, 0, stream>>>(a, l, int(t0), nb);
          LSIM_HIP(hipGetLastError());
          st.sweep_launches++;
        }
      }
    }
    st.t_sweep_ms += time_end();
  }
  if (nf > 0) {
    DetectArgs d;
    d.info = d_info.p;
    d.fin = d_fin.p;
    d.good = d_good.p;
    d.RW = R;
    d.parent = d_parent.p;
    d.head_slot = d_head_slot.p;
    d.dout = ninj > 0 ? d_out.p : nullptr;
    d.w0 = w0;
    d.Wn = Wn;
    d.flist = d_flist.p;
    d.nf = int(nf);
    d.site = d_site.p;
    d.sa = d_sa.p;
    d.num_patterns = num_patterns;
    d.first_round = first_dev;
    d.det_bits = detbits_dev;
    d.RB = RB;
    d.last_flag = last_dev;
    d.last_word = R - 1;
    time_begin();
    if (direct) {
      // Injection k is fault flist[k] (same order).
      k_direct_detect<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
    } else {
      k_cpt_detect<<<blocks_for(nf * Wn, 256), 256, 0, stream>>>(d);
    }
    LSIM_HIP(hipGetLastError());
    if (!direct && ninj > 0) {
      k_set_slots<<<blocks_for(ninj, 256), 256, 0, stream>>>(d_head_slot.p, d_inj_site.p, int(ninj), 1);
      LSIM_HIP(hipGetLastError());
    }
    st.t_detect_ms += time_end();
  }
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
  m.d_first.alloc(std::max<int64_t>(F, 1));
  k_fill_i32<<<blocks_for(std::max<int64_t>(F, 1), 256), 256, 0, m.stream>>>(m.d_first.p, std::max<int64_t>(F, 1),
                                                                            INT32_MAX);
  LSIM_HIP(hipGetLastError());
  res.first_round.assign(F, INT32_MAX);
  if (R > 0 && !live.empty()) {
    if (!opt.drop) {
      const int64_t RB = (R + 63) / 64;
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
      std::vector<int32_t> alive = live;
      std::vector<int32_t> first(F);
      for (int64_t w0 = 0; w0 < R && !alive.empty();) {
        // Fewer live injections -> more words per tile, so the step still
        // offers about two tiles per CU.
        int64_t ninj;
        if (direct) {
          ninj = int64_t(alive.size());
        } else {
          std::vector<int32_t> heads;
          for (int32_t f : alive)
            if (!m.head_is_po[m.f_head[f]]) heads.push_back(m.f_head[f]);
          std::sort(heads.begin(), heads.end());
          ninj = int64_t(std::unique(heads.begin(), heads.end()) - heads.begin());
        }
        int Wt = 1;
        while (Wt < 64 && ninj * Wt < int64_t(64) * 2 * m.cus) Wt *= 2;
        const int Wn = int(std::min<int64_t>(Wt, R - w0));
        m.step(opt, alive, w0, Wn, direct, m.d_first.p, nullptr, 0, nullptr, res);
        LSIM_HIP(hipMemcpyAsync(first.data(), m.d_first.p, F * sizeof(int32_t), hipMemcpyDeviceToHost, m.stream));
        LSIM_HIP(hipStreamSynchronize(m.stream));
        std::vector<int32_t> next;
        for (int32_t f : alive)
          if (first[f] == INT32_MAX) next.push_back(f);
        alive.swap(next);
        w0 += Wn;
      }
      if (opt.report_last_round) {
        m.d_last.alloc(std::max<int64_t>(F, 1));
        LSIM_HIP(hipMemsetAsync(m.d_last.p, 0, std::max<int64_t>(F, 1), m.stream));
        m.step(opt, live, R - 1, 1, direct, nullptr, nullptr, 0, m.d_last.p, res);
        res.last_round.assign(F, 0);
        if (F)
          LSIM_HIP(hipMemcpyAsync(res.last_round.data(), m.d_last.p, F, hipMemcpyDeviceToHost, m.stream));
        res.has_last = true;
      }
    }
    LSIM_HIP(hipMemcpyAsync(res.first_round.data(), m.d_first.p, F * sizeof(int32_t), hipMemcpyDeviceToHost,
                            m.stream));
    LSIM_HIP(hipStreamSynchronize(m.stream));
  } else {
    // No rounds or no observable fault: nothing is ever detected.
    res.has_matrix = !opt.drop;
    res.RB = (R + 63) / 64;
    res.det_bits.assign(size_t(F * res.RB), 0);
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
