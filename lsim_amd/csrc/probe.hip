// Random-access probe: the rate at which the device serves independent 8-byte
// reads at random addresses of a large buffer, and the latency of a dependent
// chain of such reads. The fault sweeps read and write small records at
// random node positions, so this rate, not the streaming bandwidth, is the
// ceiling their memory traffic is measured against.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "probe.hpp"

namespace lsim {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Each thread: `steps` reads, each at an index that depends on the last value.
__global__ void k_probe_chain(const uint64_t *buf, uint64_t n, int steps, uint64_t *out) {
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t v = 0, s = mix64(t + 12345);
  for (int i = 0; i < steps; ++i) {
    s = mix64(s + v);
    v = buf[s % n];
  }
  out[t] = v + s;
}

// Each thread: `rounds` x 8 independent reads at random indices.
__global__ void k_probe_random(const uint64_t *buf, uint64_t n, int rounds, uint64_t *out) {
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t acc = 0, s = mix64(t + 777);
  for (int i = 0; i < rounds; ++i) {
    uint64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      s = mix64(s + uint64_t(k));
      v[k] = buf[s % n];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += v[k];
  }
  out[t] = acc;
}

// Plain streaming copy, 16 bytes per thread per step (grid-stride): the
// known-size reference for the memory counters (n16 16-byte units).
__global__ void k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw ProbeError(std::string(what) + ": " + hipGetErrorString(e));
}

struct Mem {
  void *p = nullptr;
  ~Mem() {
    if (p) (void)hipFree(p);
  }
};

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

ProbeResult probe_random_reads(int device, int64_t bytes) {
  if (bytes < (int64_t(1) << 20)) throw ProbeError("probe buffer must be at least 1 MiB");
  check(hipSetDevice(device), "hipSetDevice");
  constexpr int kChainBlocks = 256, kChainThreads = 64, kChainSteps = 2000;
  constexpr int kRandBlocks = 2048, kRandThreads = 256, kRandRounds = 64;
  Mem buf, out;
  check(hipMalloc(&buf.p, size_t(bytes)), "probe buffer");
  check(hipMalloc(&out.p, sizeof(uint64_t) * size_t(kRandBlocks) * kRandThreads), "probe output");
  check(hipMemset(buf.p, 0, size_t(bytes)), "probe clear");
  const uint64_t n = uint64_t(bytes) / 8;
  const uint64_t *b = static_cast<const uint64_t *>(buf.p);
  uint64_t *o = static_cast<uint64_t *>(out.p);
  Events ev;
  check(hipEventCreate(&ev.a), "event");
  check(hipEventCreate(&ev.b), "event");
  ProbeResult r;
  // Warm-up launches, then the timed ones.
  hipLaunchKernelGGL(k_probe_chain, dim3(kChainBlocks), dim3(kChainThreads), 0, 0, b, n, 100, o);
  check(hipEventRecord(ev.a, 0), "event");
  hipLaunchKernelGGL(k_probe_chain, dim3(kChainBlocks), dim3(kChainThreads), 0, 0, b, n, kChainSteps, o);
  check(hipEventRecord(ev.b, 0), "event");
  check(hipEventSynchronize(ev.b), "chain");
  float ms = 0;
  check(hipEventElapsedTime(&ms, ev.a, ev.b), "event");
  r.latency_ns = double(ms) * 1e6 / kChainSteps;
  hipLaunchKernelGGL(k_probe_random, dim3(kRandBlocks), dim3(kRandThreads), 0, 0, b, n, 4, o);
  check(hipEventRecord(ev.a, 0), "event");
  hipLaunchKernelGGL(k_probe_random, dim3(kRandBlocks), dim3(kRandThreads), 0, 0, b, n, kRandRounds, o);
  check(hipEventRecord(ev.b, 0), "event");
  check(hipEventSynchronize(ev.b), "random reads");
  check(hipEventElapsedTime(&ms, ev.a, ev.b), "event");
  r.reads = double(kRandBlocks) * kRandThreads * kRandRounds * 8;
  r.reads_per_s = r.reads / (double(ms) * 1e-3);
  r.ms = ms;
  return r;
}

}  // namespace lsim

namespace lsim {

double probe_copy(int device, int64_t bytes, int reps) {
  if (bytes < (int64_t(1) << 20) || bytes % 16) throw ProbeError("copy size must be a multiple of 16, at least 1 MiB");
  check(hipSetDevice(device), "hipSetDevice");
  Mem src, dst;
  check(hipMalloc(&src.p, size_t(bytes)), "copy source");
  check(hipMalloc(&dst.p, size_t(bytes)), "copy destination");
  check(hipMemset(src.p, 1, size_t(bytes)), "copy fill");
  Events ev;
  check(hipEventCreate(&ev.a), "event");
  check(hipEventCreate(&ev.b), "event");
  const uint64_t n16 = uint64_t(bytes) / 16;
  check(hipEventRecord(ev.a, 0), "event");
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_probe_copy, dim3(4096), dim3(256), 0, 0, static_cast<const uint4 *>(src.p),
                       static_cast<uint4 *>(dst.p), n16);
  check(hipEventRecord(ev.b, 0), "event");
  check(hipEventSynchronize(ev.b), "copy");
  float ms = 0;
  check(hipEventElapsedTime(&ms, ev.a, ev.b), "event");
  return double(ms) / reps;
}

}  // namespace lsim

// Entry point of the measurement library liblsim_probe.so, which only the
// benchmark loads; the simulator library does not contain the probe.
extern "C" __attribute__((visibility("default"))) int lsim_bench_probe_random_reads(int device, int64_t bytes,
                                                                                 double *reads_per_s,
                                                                                 double *latency_ns) {
  if (!reads_per_s || !latency_ns) return -1;
  try {
    const lsim::ProbeResult r = lsim::probe_random_reads(device, bytes);
    *reads_per_s = r.reads_per_s;
    *latency_ns = r.latency_ns;
    return 0;
  } catch (const std::exception &) {
    return -5;
  }
}

// A plain device copy of `bytes` (read and written once per rep): the
// known-size kernel the memory counters are checked against. Average ms per
// copy in *ms.
extern "C" __attribute__((visibility("default"))) int lsim_bench_copy(int device, int64_t bytes, int reps,
                                                                   double *ms) {
  if (!ms || reps < 1) return -1;
  try {
    *ms = lsim::probe_copy(device, bytes, reps);
    return 0;
  } catch (const std::exception &) {
    return -5;
  }
}
