// fused.hpp -- persistent single-network kernel: the whole timestep-group loop of
// SpikingModel::run (SpikingModel.cpp:197-236) in one cooperative launch.
//
// Owner computes.  The LIF neurons are split into P contiguous partitions, one per
// workgroup (the Poisson inputs likewise).  A workgroup keeps its neurons' slice of the
// input ring in LDS for the whole launch and is the only writer of it, so delivery uses
// LDS atomics and no global atomics.  Per timestep group:
//   1. neuron phase: the group's ring slots are read in parallel over (neuron, sub-step),
//      with the external Poisson counts and input spikes drawn one group earlier; then one
//      lane per neuron runs the G-step LIF recurrence (LIFSpikingNeurons.cu:91-195);
//      spikes are compacted into the workgroup's exchange list (id | g << 24);
//   2. the list is published as tagged 64-bit words in the workgroup's exchange block --
//      by the LIF wavefront itself when one wavefront holds the partition -- and every
//      workgroup polls the blocks it needs (no grid barrier); with plasticity a grid
//      barrier orders the phases instead;
//   3. plasticity of the plastic synapses whose postsynaptic neuron the workgroup owns
//      (they are the only weights its own delivery reads);
//   4. delivery: every spike's row segment that targets this partition
//      (CSR rows sorted by (partition, delay)) is added into the LDS ring.
// The next group's random draws run while the workgroup waits for the others' lists.
// The arithmetic of every phase is identical to the per-phase kernels in kernels.hpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "philox.hpp"

#pragma clang fp contract(off)

namespace spk {

constexpr int kFusedBlock = 1024;      // threads per workgroup
constexpr int kBatchSmallBlock = 512;  // threads per workgroup of the two-per-CU batch kernel
constexpr int kMaxPartitions = 1024;    // workgroups of one network
constexpr uint32_t kIdMask = 0x00FFFFFFu;
constexpr int kBarGroups = 8;          // level-1 barrier counters (blockIdx % 8)
constexpr int kBarStride = 32;         // uints between counters (128 B)
constexpr int kProfPhases = 12;
constexpr int kXcds = 8;               // MI355X: 8 accelerator complex dies, blocks dispatched round-robin
constexpr int kPrefetch = 16;          // sub-steps whose inputs are read ahead

constexpr int kMaxFusedMons = 4;      // spike monitors the persistent kernel records for

// Device store of one spike monitor (SpikingActivityMonitor.cu:22-33): records beyond cap
// are counted but not written, so the host sees the overflow.
struct FusedMon {
  int32_t* ids;
  float* times;
  unsigned long long* count;
  long long cap;
  unsigned int lo, hi;
};

// One plasticity rule of the persistent kernel.
struct FusedRule {
  StdpArgs args;
  const int64_t* part;   // [P+1] offsets of each partition in the rule's synapse order
  int kind;
};

struct FusedArgs {
  // partitions: neurons [nlo[p], nlo[p+1]), inputs [ilo[p], ilo[p+1])
  const int* nlo;
  const int* ilo;
  // spike counts per neuron / input of a rate monitor, or nullptr
  uint32_t* rate_n;
  uint32_t* rate_in;
  // LIF state and constants
  float* V;
  float* last_n;
  uint32_t* hist_n;
  float* gcond;
  int64_t* ring;            // global copy [B][N*L], loaded/stored at launch edges
  const float* v0;
  const float* vreset;
  const float* thr;
  const float* c;
  const float* rp;
  const float* decay_g;
  const float* erev;
  const uint32_t* ext_cdf;
  // inputs
  float* last_in;
  uint32_t* hist_in;
  const uint32_t* in_thr;   // selected stimulus
  // synapses
  const uint32_t* seg;      // [U][P+1] absolute segment offsets
  const uint2* syn;         // per synapse {local target | delay << 20 | label << 28, weight bits}
  const uint32_t* wc;       // 16-bit weights (wfmt 1 bfloat16, 2 binary16) in the low halves, instead of the records'
  int wfmt;
  const float* scale;       // nullptr: uniform scale0
  // compact form (no plasticity, one label, uniform scale): every synapse of a source
  // u < src_split has weight w_lo, every other one w_hi, all have delay src_delay, and the
  // records are the 16-bit local targets alone (null: the full records above)
  const uint16_t* syn16;
  float w_lo, w_hi;
  int src_split, src_delay;
  // compact form with the partition's delivery table in LDS (copied from seg and syn16 when
  // the group loop starts): 16-bit local starts of the lseg_U source units' segments (and
  // the end), at most lrec targets per partition; 0: off
  int lseg_U, lrec;
  // compact form with a count ring: each LDS ring word holds the step's count of
  // deliveries from sources below src_split (low 16 bits) and from the others (high 16
  // bits); the reading neuron rebuilds the exact fixed-point sum from the counts.  The
  // global int64 ring then only carries the inputs pending when the launch starts (read
  // for steps below the launch's first step + BL) and receives the pending ones at its end.
  int cring;
  // plasticity (rules grouped by postsynaptic partition), descriptors in device memory
  const FusedRule* rules;
  int n_rules;
  // exchange blocks [2][P][xstride]: {count, entries (id | g << 24)...}; 32-bit words
  // behind the grid barrier (plastic launches), 64-bit tagged words otherwise (xlist64, the
  // same memory): low half the count or entry, high half the tag of the group that wrote it
  uint32_t* xlist;
  unsigned long long* xlist64;
  unsigned int xtag_base;     // tag of group s of this launch: xtag_base + s + 1
  int xquad;                  // tagged delivery by quads per source (small partitions) or staged
  int early;                  // with xquad and no plasticity: wavefront 0 publishes its list right
                              // after its LIF, the other wavefronts make the next group's draws
  int dahead;                 // groups whose random draws are made in one pass (LDS slots; 1)
  // two-level grid barrier: counters bar[g * kBarStride] (g < 8), bar[8 * kBarStride]
  // (level 2) and bar[9 * kBarStride] (release generation)
  unsigned int* bar;
  int* error;
  // spike monitors: each appends (local id, time) records of units [lo, hi) to its store
  FusedMon mon[kMaxFusedMons];
  int n_mons;
  // barrier generations completed on the counters before this launch
  unsigned int bar_base;
  StatsArgs st;                 // window statistics of the LIF population (st.on = 0: off)
  unsigned long long* stats;    // [0] deliveries [1] external [2] LIF spikes [3] input spikes
  int64_t q_ext;
  int64_t arm_step;
  uint64_t seed;
  long long spin_limit;
  float scale0, two_k, inv_scale;
  float start, dt, tref, ibg;
  int64_t group0;             // index of the first group of this launch within the run
  int N, Nin, L, G, B, W, P, xcap, xstride, nown_max, iown_max, ext_k, ngroups;
  int BL;                       // slots of the LDS ring: max(largest delay, G)
  int plastic;
  unsigned long long* prof;   // optional: wall-clock ticks per phase of workgroup 0
};

// Phase clock of workgroup 0 (no effect on the computation): thread 0 accumulates laps
// in LDS and writes them once when the launch ends.
struct PhaseClock {
  unsigned long long t0;
  unsigned long long* acc;   // LDS, kProfPhases entries
  bool on;
  __device__ __forceinline__ void start(const FusedArgs& a, unsigned long long* lds, int vb) {
    acc = lds;
    on = a.prof != nullptr && vb == 0 && threadIdx.x == 0;
    if (on) {
      for (int i = 0; i < kProfPhases; ++i) acc[i] = 0;
      t0 = wall_clock64();
    }
  }
  __device__ __forceinline__ void lap(const FusedArgs&, int phase) {
    if (!on) return;
    const unsigned long long t1 = wall_clock64();
    acc[phase] += t1 - t0;
    t0 = t1;
  }
  __device__ __forceinline__ void flush(const FusedArgs& a) {
    if (!on) return;
    for (int i = 0; i < kProfPhases; ++i) a.prof[i] += acc[i];
  }
};

// LDS layout of one workgroup: every sub-array starts on a 16-byte boundary (the input
// spike masks take 64-bit LDS atomics, which need 8-byte aligned addresses).
struct FusedLayout {
  size_t ring, hist, ext, imask, inp, xbuf, head, scan, ent, counter, st_cnt, st_nisi, st_prev, st_s1,
      st_s2, st_g, cdf, guide, lseg, lrec, total;
};

constexpr int kHeadWords = 12;         // words of a source block fetched in the first trip
constexpr int kEntCap = 4096;          // spike entries staged in LDS per delivery round
constexpr int kQuad = 4;               // lanes cooperating on one spike's row segment
constexpr int kSpikesPerQuad = 4;      // spikes a quad handles with its loads in flight together

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// `BL` is the number of ring slots kept in LDS (see k_fused: max(largest delay, G)).
// `segU`, `lrec`: source units and 2-byte targets of the LDS-resident delivery table (0: the
// table is read from global memory).
__host__ __device__ inline FusedLayout fused_layout(int BL, int L, int W, int G, int nown, int iown,
                                                    int xcap, int P, int segU = 0, int lrec = 0, int cring = 0,
                                                    int dahead = 1) {
  FusedLayout f;
  size_t o = 0;
  f.ring = o;    o = align16(o + (size_t)BL * nown * L * (cring ? 4 : 8));
  f.hist = o;    o = align16(o + (size_t)W * nown * 4);
  f.ext = o;     o = align16(o + (size_t)dahead * nown * G);   // slots of the drawn groups
  f.imask = o;   o = align16(o + (size_t)dahead * iown * 8);
  // this group's inputs as float; count-ring partitions of more than one wavefront without
  // an LDS delivery table keep them in the ring's own slots instead (in_ring in fused_body)
  f.inp = o;     o = align16(o + (cring && nown > kWave && segU == 0 ? 0 : (size_t)nown * G * L * 4));
  f.xbuf = o;    o = align16(o + (size_t)xcap * 4);
  f.head = o;    o = align16(o + (size_t)P * kHeadWords * 4);
  f.scan = o;    o = align16(o + (size_t)(P + 1) * 4);
  f.ent = o;     o = align16(o + (size_t)kEntCap * 4);
  f.counter = o; o = align16(o + 4 * 4);
  // window statistics of the owned neurons, and spikes per sub-step of the group
  f.st_cnt = o;  o = align16(o + (size_t)nown * 4);
  f.st_nisi = o; o = align16(o + (size_t)nown * 4);
  f.st_prev = o; o = align16(o + (size_t)nown * 4);
  f.st_s1 = o;   o = align16(o + (size_t)nown * 8);
  f.st_s2 = o;   o = align16(o + (size_t)nown * 8);
  f.st_g = o;    o = align16(o + 64 * 4);
  f.cdf = o;     o = align16(o + (size_t)kMaxExtTable * 4);   // external count thresholds
  f.guide = o;   o = align16(o + 256);                          // guide table of the thresholds
  f.lseg = o;    o = align16(o + (segU ? (size_t)(segU + 1) * 2 : 0));   // local starts
  f.lrec = o;    o = align16(o + (size_t)lrec * 2);             // this partition's targets
  f.total = o;
  return f;
}

__host__ __device__ inline size_t fused_lds_bytes(int BL, int L, int W, int G, int nown, int iown,
                                                  int xcap, int P, int segU = 0, int lrec = 0, int cring = 0,
                                                  int dahead = 1) {
  return fused_layout(BL, L, W, G, nown, iown, xcap, P, segU, lrec, cring, dahead).total;
}

__host__ __device__ inline int fused_xstride(int xcap) { return ((1 + xcap + 7) / 8) * 8; }

__device__ __forceinline__ unsigned int atomic_load_agent(const unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exchange data between workgroups is written and read with device-coherent (agent
// scope) operations, which bypass the per-XCD L2's non-coherent copies.  The barrier then
// needs no L2 write-back/invalidate: a wait for completed stores before the arrival
// orders the data before the counter increment.
__device__ __forceinline__ void store_coherent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_coherent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tagged exchange words (launches without plasticity): every 64-bit word carries the tag of
// the group that wrote it next to its payload, and a reader accepts a word only when the
// tag is the group's own.  A word is written and read whole (naturally aligned 64-bit
// accesses at agent scope), so no ordering between the words of a block is needed and the
// writer never waits for its stores: there is no grid barrier.  Safety of the double
// buffer: a workgroup writes the parity-p blocks of group s + 2 only after it has read
// every source's block of group s + 1, which each source publishes only after it has
// finished reading the parity-p blocks of group s.
__device__ __forceinline__ void xstore64(unsigned long long* p, uint32_t v, uint32_t tag) {
  __hip_atomic_store(p, ((unsigned long long)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long xload64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t xtag(unsigned long long w) { return (uint32_t)(w >> 32); }

// Barrier arrival (thread 0, after the workgroup's stores are complete): one increment of
// the level-1 counter of this workgroup's member index vb (0..P-1); the last arrival of a level-1 group increments the
// level-2 counter, and the last of those publishes generation s+1.  With plasticity the
// rules read other partitions' state through ordinary loads, so arrival and departure
// carry agent-scope release/acquire fences.
__device__ __forceinline__ void barrier_arrive(const FusedArgs& a, int s, int vb) {
  const int P = a.P;
  const int ngrp = P < kBarGroups ? P : kBarGroups;
  const int grp = vb % ngrp;
  const unsigned int gsize = (unsigned int)((P - grp + ngrp - 1) / ngrp);
  if (a.plastic) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  const unsigned int gen = a.bar_base + (unsigned int)(s + 1);
  const unsigned int old = atomicAdd(a.bar + grp * kBarStride, 1u);
  if (old + 1u == gen * gsize) {
    const unsigned int old2 = atomicAdd(a.bar + kBarGroups * kBarStride, 1u);
    if (old2 + 1u == gen * (unsigned int)ngrp)
      __hip_atomic_store(a.bar + (kBarGroups + 1) * kBarStride, gen, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Barrier wait (thread 0): bounded spin on the release generation.
__device__ __forceinline__ bool barrier_wait(const FusedArgs& a, int s) {
  const unsigned int* rel = a.bar + (kBarGroups + 1) * kBarStride;
  long long spins = 0;
  const unsigned int gen = a.bar_base + (unsigned int)(s + 1);
  while (atomic_load_agent(rel) < gen) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > a.spin_limit ||
        __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      atomicCAS(a.error, 0, 1);   // keep an earlier error code
      return false;
    }
  }
  if (a.plastic) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

__device__ __forceinline__ void group_time(const FusedArgs& a, int s, float& t, int& tidx) {
  t = a.start + (float)((a.group0 + s) * (int64_t)a.G) * a.dt;
  tidx = (int)roundf(t / a.dt);
}

// Random draws of groups s0 .. s0 + cnt - 1: external Poisson counts per (neuron, sub-step)
// and the spike masks of the Poisson inputs (Bernoulli per sub-step), into LDS slot
// s % dahead of each group.  Independent of network state, so they are made ahead, once
// every dahead groups, while the workgroup waits for the other workgroups' spikes (the host
// sets dahead = 1: one group per pass hides in that wait, eight did not).  One work
// item per (group, unit, 4-step Philox block) uses all four outputs of its call; a count is
// the number of thresholds C[k] <= u, scanned upwards from the guide entry of u (`sguide[j]`
// is the count of u = j * 2^24, a lower bound for every u in [j * 2^24, (j + 1) * 2^24),
// Chen & Asau 1974), so the scan takes about one step.  The input masks are ORed into slots
// that phase A cleared when it read them.
template <int BS>
__device__ __forceinline__ void draw_groups_flat(const FusedArgs& a, int s0, int cnt, uint8_t* sext,
                                                 unsigned long long* simask, int nlo, int nown, int ilo,
                                                 int iown, unsigned long long& st_ext, const uint32_t* scdf,
                                                 const uint8_t* sguide, int t0 = 0) {
  const int G = a.G;
  const int tid = (int)threadIdx.x - t0;   // threads t0 .. BS-1 draw
  const int nbm = (G + 3) / 4 + 1;   // 4-step blocks a group's G steps can touch
  const int per = nown * nbm;
  if (a.ext_k > 0) {
    for (int k = tid; k < cnt * per; k += BS - t0) {
      const int j = k / per;
      const int r0 = k - j * per;
      const int i = r0 / nbm, bb = r0 - i * nbm;
      float t;
      int tidx;
      group_time(a, s0 + j, t, tidx);
      const int64_t b0 = (int64_t)tidx >> 2;
      const int nbk = (int)((((int64_t)tidx + G - 1) >> 2) - b0) + 1;
      if (bb >= nbk) continue;
      const int64_t blk = b0 + bb;
      const u32x4 r = draw_block(a.seed, blk, (uint32_t)(nlo + i), kStreamExternal);
      uint8_t* ext = sext + ((size_t)((s0 + j) % a.dahead) * a.nown_max + i) * G;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int g = (int)(blk * 4 + c - (int64_t)tidx);
        if (g >= 0 && g < G) {
          const uint32_t u = pick(r, c);
          int n = sguide[u >> 24];
          while (n < a.ext_k && u >= scdf[n]) ++n;
          ext[g] = (uint8_t)n;
          st_ext += (unsigned long long)n;
        }
      }
    }
  }
  if (iown > 0) {
    const int peri = iown * nbm;
    for (int k = tid; k < cnt * peri; k += BS - t0) {
      const int j = k / peri;
      const int r0 = k - j * peri;
      const int i = r0 / nbm, bb = r0 - i * nbm;
      float t;
      int tidx;
      group_time(a, s0 + j, t, tidx);
      const int64_t b0 = (int64_t)tidx >> 2;
      const int nbk = (int)((((int64_t)tidx + G - 1) >> 2) - b0) + 1;
      if (bb >= nbk) continue;
      const int64_t blk = b0 + bb;
      const u32x4 r = draw_block(a.seed, blk, (uint32_t)(ilo + i), kStreamInput);
      const uint32_t thr = a.in_thr[ilo + i];
      unsigned long long bits = 0ull;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int64_t step = blk * 4 + c;
        const int g = (int)(step - (int64_t)tidx);
        if (g >= 0 && g < G && pick(r, c) < thr && step != a.arm_step) bits |= 1ull << g;
      }
      if (bits) atomicOr(&simask[(size_t)((s0 + j) % a.dahead) * a.iown_max + i], bits);
    }
  }
}

template <int BS>
__device__ __forceinline__ void draw_groups_seq(const FusedArgs& a, int s0, int cnt, uint8_t* sext,
                                            unsigned long long* simask, int nlo, int nown, int ilo,
                                            int iown, unsigned long long& st_ext, const uint32_t* scdf,
                                            const uint8_t* sguide, int t0 = 0) {
  const int G = a.G;
  const int tid = (int)threadIdx.x - t0;   // threads t0 .. BS-1 draw
  for (int j = 0; j < cnt; ++j) {
    // per group (uniform): its first step and the 4-step blocks its G steps touch
    float t;
    int tidx;
    group_time(a, s0 + j, t, tidx);
    const int64_t b0 = (int64_t)tidx >> 2;
    const int nbk = (int)((((int64_t)tidx + G - 1) >> 2) - b0) + 1;
    const float inv_nbk = 1.0f / (float)nbk;
    const int slot = (s0 + j) % a.dahead;
    if (a.ext_k > 0) {
      uint8_t* ext = sext + (size_t)slot * a.nown_max * G;
      for (int k = tid; k < nown * nbk; k += BS - t0) {
        // k / nbk by a float product: (k + 1/2) / nbk is at least 1/(2 nbk) from an
        // integer and k < 2^16
        const int i = (int)(((float)k + 0.5f) * inv_nbk);
        const int64_t blk = b0 + (k - i * nbk);
        const u32x4 r = draw_block(a.seed, blk, (uint32_t)(nlo + i), kStreamExternal);
        const int gb = (int)(blk * 4 - (int64_t)tidx);
        // the four counts' table scans advance together, so that their LDS reads overlap
        uint32_t u4[4];
        int n4[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          u4[c] = pick(r, c);
          n4[c] = sguide[u4[c] >> 24];
        }
        bool more = true;
        while (more) {
          more = false;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (n4[c] < a.ext_k && u4[c] >= scdf[n4[c]]) {
              ++n4[c];
              more = true;
            }
          }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int g = gb + c;
          if (g >= 0 && g < G) {
            ext[i * G + g] = (uint8_t)n4[c];
            st_ext += (unsigned long long)n4[c];
          }
        }
      }
    }
    if (iown > 0) {
      unsigned long long* im = simask + (size_t)slot * a.iown_max;
      for (int k = tid; k < iown * nbk; k += BS - t0) {
        const int i = (int)(((float)k + 0.5f) * inv_nbk);
        const int64_t blk = b0 + (k - i * nbk);
        const u32x4 r = draw_block(a.seed, blk, (uint32_t)(ilo + i), kStreamInput);
        const uint32_t thr = a.in_thr[ilo + i];
        unsigned long long bits = 0ull;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int64_t step = blk * 4 + c;
          const int g = (int)(step - (int64_t)tidx);
          if (g >= 0 && g < G && pick(r, c) < thr && step != a.arm_step) bits |= 1ull << g;
        }
        if (bits) atomicOr(&im[i], bits);
      }
    }
  }
}

// The draws of several groups for small partitions (at most four items per thread) take one
// pass over (group, unit, block) items, so that the groups' Philox calls run side by side;
// otherwise one group at a time, with each group's constants computed once and no integer
// divisions per item -- also for a single group with external drive (its count scans run
// interleaved there; Poisson inputs alone keep the one-pass form, measured faster).
// Threads t0 .. BS-1 share the items.
template <int BS>
__device__ __forceinline__ void draw_groups(const FusedArgs& a, int s0, int cnt, uint8_t* sext,
                                            unsigned long long* simask, int nlo, int nown, int ilo,
                                            int iown, unsigned long long& st_ext, const uint32_t* scdf,
                                            const uint8_t* sguide, int t0 = 0) {
  const int nbm = (a.G + 3) / 4 + 1;
  if ((cnt > 1 || a.ext_k == 0) && cnt * ((a.ext_k > 0 ? nown : 0) + iown) * nbm <= 4 * (BS - t0))
    draw_groups_flat<BS>(a, s0, cnt, sext, simask, nlo, nown, ilo, iown, st_ext, scdf, sguide, t0);
  else
    draw_groups_seq<BS>(a, s0, cnt, sext, simask, nlo, nown, ilo, iown, st_ext, scdf, sguide, t0);
}

// One quad's share of kSpikesPerQuad spikes: every lane loads its synapses of each spike's
// segment (lane q takes s0+q, s0+q+4, ...) with all loads issued before the LDS atomics.
// With 16-bit weights only the packed word of a record is read, plus the weight's 4-byte word.
__device__ __forceinline__ float syn_weight(const FusedArgs& a, const uint32_t* synw, uint32_t k) {
  return a.wfmt ? dec16((uint16_t)(a.wc[k] & 0xFFFFu), a.wfmt) : __uint_as_float(synw[(size_t)k * 2 + 1]);
}

template <bool PER_SYN_SCALE>
__device__ __forceinline__ void deliver_quad(const FusedArgs& a, const uint32_t* ents, int n, int q,
                                             int p, int slot0, int B, int nownL, int L, int64_t* sring,
                                             unsigned long long& st_del) {
  const uint32_t* synw = (const uint32_t*)a.syn;   // word 2k: packed, word 2k+1: FP32 weight
  uint32_t s0[kSpikesPerQuad], s1[kSpikesPerQuad];
  int gs[kSpikesPerQuad];
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
    s0[r] = 0u;
    s1[r] = 0u;
    gs[r] = 0;
    if (r < n) {
      const uint32_t e = ents[r];
      const size_t sb = (size_t)(e & kIdMask) * (a.P + 1) + p;
      s0[r] = a.seg[sb];
      s1[r] = a.seg[sb + 1];
      gs[r] = slot0 + (int)(e >> 24);
    }
  }
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r)
    if (q == 0) st_del += (unsigned long long)(s1[r] - s0[r]);
  // first two synapses per lane of every segment, loads in flight together
  uint32_t pkh[kSpikesPerQuad][2];
  float wvh[kSpikesPerQuad][2];
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t k = s0[r] + (uint32_t)(q + h * kQuad);
      pkh[r][h] = 0u;
      wvh[r][h] = 0.0f;
      if (k < s1[r]) {
        pkh[r][h] = synw[(size_t)k * 2];
        wvh[r][h] = syn_weight(a, synw, k) * (PER_SYN_SCALE ? a.scale[k] : a.scale0);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t k = s0[r] + (uint32_t)(q + h * kQuad);
      if (k < s1[r]) {
        const uint32_t pk = pkh[r][h];
        const long long qv = __float2ll_rn(wvh[r][h] * a.two_k);
        int slot = gs[r] + (int)((pk >> 20) & 0xFFu);
        while (slot >= B) slot -= B;
        atomicAdd((unsigned long long*)&sring[(size_t)slot * nownL + (size_t)(pk & 0xFFFFFu) * L + (pk >> 28)],
                  (unsigned long long)qv);
      }
    }
    // longer segments: the rest, four synapses per lane per step with the loads in flight
    // together
    for (uint32_t k0 = s0[r] + (uint32_t)(q + 2 * kQuad); k0 < s1[r]; k0 += 4 * kQuad) {
      uint32_t pk4[4];
      float wv4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t k = k0 + (uint32_t)(u * kQuad);
        pk4[u] = 0u;
        wv4[u] = 0.0f;
        if (k < s1[r]) {
          pk4[u] = synw[(size_t)k * 2];
          wv4[u] = syn_weight(a, synw, k) * (PER_SYN_SCALE ? a.scale[k] : a.scale0);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t k = k0 + (uint32_t)(u * kQuad);
        if (k < s1[r]) {
          const long long qv = __float2ll_rn(wv4[u] * a.two_k);
          int slot = gs[r] + (int)((pk4[u] >> 20) & 0xFFu);
          while (slot >= B) slot -= B;
          atomicAdd((unsigned long long*)&sring[(size_t)slot * nownL + (size_t)(pk4[u] & 0xFFFFFu) * L + (pk4[u] >> 28)],
                    (unsigned long long)qv);
        }
      }
    }
  }
}

// Compact form of deliver_quad: 2-byte records, the fixed-point weight and the slot are
// those of the spike's source (q_lo below src_split, q_hi from it on), computed once per
// spike with the same expression as the full form.  The targets are read eight at a time:
// a spike's segment [s0, s1) of syn16 is covered by the 16-byte aligned chunks it touches
// (the array is padded to whole chunks), lane q of the quad takes chunks q, q + 4, ... and
// adds the chunk's targets that lie inside the segment; the first chunk of every spike is
// loaded before any is used.  With the count ring the increment is 1 (sources below
// src_split) or 1 << 16.
template <bool CR>
__device__ __forceinline__ void deliver_quad_compact(const FusedArgs& a, const uint32_t* ents, int n, int q,
                                                     int p, int slot0, int B, int nown, int64_t* sring,
                                                     unsigned long long& st_del, long long q_lo, long long q_hi) {
  uint32_t s0[kSpikesPerQuad], s1[kSpikesPerQuad];
  long long qv[kSpikesPerQuad];
  int slot[kSpikesPerQuad];
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
    s0[r] = 0u;
    s1[r] = 0u;
    qv[r] = 0;
    slot[r] = 0;
    if (r < n) {
      const uint32_t e = ents[r];
      const uint32_t id = e & kIdMask;
      const size_t sb = (size_t)id * (a.P + 1) + p;
      s0[r] = a.seg[sb];
      s1[r] = a.seg[sb + 1];
      qv[r] = CR ? ((int)id < a.src_split ? 1 : 0x10000) : ((int)id < a.src_split ? q_lo : q_hi);
      int sl = slot0 + (int)(e >> 24) + a.src_delay;
      while (sl >= B) sl -= B;
      slot[r] = sl;
    }
  }
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r)
    if (q == 0) st_del += (unsigned long long)(s1[r] - s0[r]);
  const uint4* chunks = (const uint4*)a.syn16;
  uint4 v[kSpikesPerQuad];
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
    const uint32_t c = (s0[r] >> 3) + (uint32_t)q;
    v[r] = c < ((s1[r] + 7u) >> 3) && s0[r] < s1[r] ? chunks[c] : make_uint4(0u, 0u, 0u, 0u);
  }
#pragma unroll
  for (int r = 0; r < kSpikesPerQuad; ++r) {
    if (s0[r] >= s1[r]) continue;
    int64_t* row = sring + (size_t)slot[r] * nown;
    uint32_t* crow = (uint32_t*)sring + (size_t)slot[r] * nown;
    const uint32_t cend = (s1[r] + 7u) >> 3;
    uint32_t c = (s0[r] >> 3) + (uint32_t)q;
    uint4 w = v[r];
    while (c < cend) {
      const uint32_t cn = c + kQuad;
      const uint4 wn = cn < cend ? chunks[cn] : make_uint4(0u, 0u, 0u, 0u);
      const uint32_t words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        const uint32_t k = c * 8u + (uint32_t)h;
        if (k >= s0[r] && k < s1[r]) {
          const uint32_t t = (words[h >> 1] >> ((h & 1) * 16)) & 0xFFFFu;
          if (CR) atomicAdd(&crow[t], (uint32_t)qv[r]);
          else atomicAdd((unsigned long long*)&row[t], (unsigned long long)qv[r]);
        }
      }
      c = cn;
      w = wn;
    }
  }
}

// Inclusive prefix sum over the 64 lanes of a wavefront by DPP: shifts by 1, 2, 4, 8 inside
// each row of 16 lanes (lanes past the row start read 0), then row 0's total added to row 1
// and row 2's to row 3 (row_bcast:15), then the total of rows 0-1 added to rows 2 and 3
// (row_bcast:31).  Lanes outside a step's row mask add 0.
__device__ __forceinline__ int wave_inclusive_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
  return v;
}

// Compact delivery from the LDS-resident table: entry -> local start and end (the next
// unit's start) -> targets, all in LDS; the quad's lanes take every kQuad-th target.
template <bool CR>
__device__ __forceinline__ void deliver_quad_lds(const FusedArgs& a, const uint32_t* ents, int n, int q,
                                                 int slot0, int B, int nown, int64_t* sring,
                                                 unsigned long long& st_del, long long q_lo, long long q_hi,
                                                 const uint16_t* slseg, const uint16_t* slrec) {
  for (int r = 0; r < n; ++r) {
    const uint32_t e = ents[r];
    const uint32_t id = e & kIdMask;
    const uint32_t s0 = slseg[id], s1 = slseg[id + 1];
    int sl = slot0 + (int)(e >> 24) + a.src_delay;
    while (sl >= B) sl -= B;
    if (q == 0) st_del += (unsigned long long)(s1 - s0);
    if (CR) {
      const uint32_t inc = (int)id < a.src_split ? 1u : 0x10000u;
      uint32_t* row = (uint32_t*)sring + (size_t)sl * nown;
      for (uint32_t k = s0 + (uint32_t)q; k < s1; k += kQuad) atomicAdd(&row[slrec[k]], inc);
    } else {
      const unsigned long long qv = (unsigned long long)((int)id < a.src_split ? q_lo : q_hi);
      unsigned long long* row = (unsigned long long*)(sring + (size_t)sl * nown);
      for (uint32_t k = s0 + (uint32_t)q; k < s1; k += kQuad) atomicAdd(&row[slrec[k]], qv);
    }
  }
}

// One lane's delivery of one entry from the LDS-resident table: the entry's segment
// (local start and end) and its targets, read four at a time with the loads in flight.
template <bool CR>
__device__ __forceinline__ void deliver_lane_lds(const FusedArgs& a, uint32_t e, int slot0, int B, int nown,
                                                 int64_t* sring, unsigned long long& st_del, long long q_lo,
                                                 long long q_hi, const uint16_t* slseg, const uint16_t* slrec) {
  const uint32_t id = e & kIdMask;
  const uint32_t s0 = slseg[id], s1 = slseg[id + 1];
  int sl = slot0 + (int)(e >> 24) + a.src_delay;
  while (sl >= B) sl -= B;
  st_del += (unsigned long long)(s1 - s0);
  const bool lo = (int)id < a.src_split;
  uint32_t* crow = (uint32_t*)sring + (size_t)sl * nown;
  unsigned long long* row = (unsigned long long*)(sring + (size_t)sl * nown);
  const uint32_t inc = lo ? 1u : 0x10000u;
  const unsigned long long qv = (unsigned long long)(lo ? q_lo : q_hi);
  uint32_t k = s0;
  for (; k + 4 <= s1; k += 4) {
    const uint16_t t0 = slrec[k], t1 = slrec[k + 1], t2 = slrec[k + 2], t3 = slrec[k + 3];
    if (CR) {
      atomicAdd(&crow[t0], inc); atomicAdd(&crow[t1], inc); atomicAdd(&crow[t2], inc); atomicAdd(&crow[t3], inc);
    } else {
      atomicAdd(&row[t0], qv); atomicAdd(&row[t1], qv); atomicAdd(&row[t2], qv); atomicAdd(&row[t3], qv);
    }
  }
  for (; k < s1; ++k) {
    if (CR) atomicAdd(&crow[slrec[k]], inc);
    else atomicAdd(&row[slrec[k]], qv);
  }
}

template <bool CR>
__device__ __forceinline__ void deliver_lane_lds2(const FusedArgs& a, bool has0, uint32_t e0, bool has1, uint32_t e1,
                                                  int slot0, int B, int nown, int64_t* sring,
                                                  unsigned long long& st_del, long long q_lo, long long q_hi,
                                                  const uint16_t* slseg, const uint16_t* slrec) {
  if (has0) deliver_lane_lds<CR>(a, e0, slot0, B, nown, sring, st_del, q_lo, q_hi, slseg, slrec);
  if (has1) deliver_lane_lds<CR>(a, e1, slot0, B, nown, sring, st_del, q_lo, q_hi, slseg, slrec);
}

constexpr int kQuadEnt = 8;   // entries of one source a quad stages in LDS at a time

// Delivery forms a kernel is compiled for: every form chosen at run time, or only the
// compact form with the LDS-resident table, or only the compact form read from global
// memory (a kernel holding only the form it runs keeps its registers for it).
enum { kDeliverAny = 0, kDeliverLdsTable = 1, kDeliverCompact = 2, kDeliverCompactAny = 3 };

// One quad's delivery of n entries staged in LDS, kSpikesPerQuad at a time, by the form of
// the delivery table the launch uses.
template <int DM>
__device__ __forceinline__ void deliver_staged(const FusedArgs& a, const uint32_t* ents, int n, int q, int p,
                                               int slot0, int BL, int nownL, int L, int64_t* sring,
                                               unsigned long long& st_del, long long q_lo, long long q_hi,
                                               const uint16_t* slseg, const uint16_t* slrec) {
  for (int j0 = 0; j0 < n; j0 += kSpikesPerQuad) {
    const int m = min(kSpikesPerQuad, n - j0);
    if (DM != kDeliverCompact && a.lseg_U > 0) {
      if (a.cring) deliver_quad_lds<true>(a, ents + j0, m, q, slot0, BL, nownL, sring, st_del, q_lo, q_hi, slseg, slrec);
      else deliver_quad_lds<false>(a, ents + j0, m, q, slot0, BL, nownL, sring, st_del, q_lo, q_hi, slseg, slrec);
    } else if (DM != kDeliverLdsTable && a.syn16 != nullptr) {
      if (a.cring) deliver_quad_compact<true>(a, ents + j0, m, q, p, slot0, BL, nownL, sring, st_del, q_lo, q_hi);
      else deliver_quad_compact<false>(a, ents + j0, m, q, p, slot0, BL, nownL, sring, st_del, q_lo, q_hi);
    } else if (DM == kDeliverAny) {
      if (a.scale == nullptr) deliver_quad<false>(a, ents + j0, m, q, p, slot0, BL, nownL, L, sring, st_del);
      else deliver_quad<true>(a, ents + j0, m, q, p, slot0, BL, nownL, L, sring, st_del);
    }
  }
}

// Delivery of group s without a grid barrier (tagged exchange words).  Quad k of the
// workgroup takes sources k, k + BS/4, ...: lane q polls words q and q + 4 of the source's
// block -- the count and the first seven entries -- until the count and every entry below
// it carry the group's tag, the entries are staged in the quad's LDS slice and delivered;
// longer lists follow eight entries at a time with their loads in flight together.
// Returns false (in every lane of the wave) when a wait exceeded the spin limit or another
// workgroup reported an error; the error flag is then set.
template <int BS, int DM>
__device__ __forceinline__ bool tagged_gather_deliver(const FusedArgs& a, int p, int par, uint32_t tag, int slot0,
                                                      int BL, int nownL, int L, int64_t* sring,
                                                      unsigned long long& st_del, long long q_lo, long long q_hi,
                                                      const uint16_t* slseg, const uint16_t* slrec, uint32_t* sent) {
  const int tid = threadIdx.x;
  const int lane = (int)__lane_id();
  const int quad = tid / kQuad, q = tid % kQuad;
  constexpr int nquads = BS / kQuad;
  const int qbase = lane & ~(kQuad - 1);
  uint32_t* qent = sent + (size_t)quad * kQuadEnt;
  const unsigned long long* xb = a.xlist64 + (size_t)par * a.P * a.xstride;
  bool ok = true;
  for (int src0 = 0; src0 < a.P; src0 += nquads) {
    // uniform point of the loop: stop every lane once one has failed
    if (__any(!ok)) {
      ok = false;
      break;
    }
    const int src = src0 + quad;
    const bool active = src < a.P;
    const unsigned long long* blk = xb + (size_t)(active ? src : 0) * a.xstride;
    unsigned long long w0 = active ? xload64(blk + q) : 0ull;
    unsigned long long w1 = active ? xload64(blk + q + kQuad) : 0ull;
    long long spins = 0;
    int c = 0;
    // wave-uniform loop: every lane shuffles the count of its quad's source
    while (true) {
      const bool t0 = xtag(w0) == tag, t1 = xtag(w1) == tag;
      const int cw = __shfl((int)(uint32_t)w0, qbase);
      const bool cok = __shfl(t0 ? 1 : 0, qbase) != 0;
      c = min(cw, a.xcap);
      const bool need0 = active && !t0 && (q == 0 || !cok || q <= c);
      const bool need1 = active && !t1 && (!cok || q + kQuad <= c);
      if (!__any(need0 || need1)) break;
      if (++spins > a.spin_limit) {
        atomicCAS(a.error, 0, 1);
        ok = false;
        break;
      }
      if ((spins & 63) == 0 && __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      if (need0) w0 = xload64(blk + q);
      if (need1) w1 = xload64(blk + q + kQuad);
    }
    if (!ok) break;
    if (!active) continue;
    if (DM != kDeliverCompact && a.lseg_U > 0) {
      // LDS table: every lane delivers the entries it holds -- word q is entry q - 1, word
      // q + 4 entry q + 3 -- and longer lists eight at a time, two entries per lane
      const bool h0 = q >= 1 && q - 1 < c, h1 = q + 3 < c;
      if (a.cring)
        deliver_lane_lds2<true>(a, h0, (uint32_t)w0, h1, (uint32_t)w1, slot0, BL, nownL, sring, st_del, q_lo, q_hi,
                                slseg, slrec);
      else
        deliver_lane_lds2<false>(a, h0, (uint32_t)w0, h1, (uint32_t)w1, slot0, BL, nownL, sring, st_del, q_lo,
                                 q_hi, slseg, slrec);
      for (int j0 = kQuadEnt - 1; j0 < c; j0 += kQuadEnt) {
        const int m = min(kQuadEnt, c - j0);
        const unsigned long long* eb = blk + 1 + j0;
        unsigned long long e0 = q < m ? xload64(eb + q) : 0ull;
        unsigned long long e1 = q + kQuad < m ? xload64(eb + q + kQuad) : 0ull;
        long long sp = 0;
        while ((q < m && xtag(e0) != tag) || (q + kQuad < m && xtag(e1) != tag)) {
          if (++sp > a.spin_limit) {
            atomicCAS(a.error, 0, 1);
            ok = false;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          if (q < m && xtag(e0) != tag) e0 = xload64(eb + q);
          if (q + kQuad < m && xtag(e1) != tag) e1 = xload64(eb + q + kQuad);
        }
        if (!ok) break;
        if (a.cring)
          deliver_lane_lds2<true>(a, q < m, (uint32_t)e0, q + kQuad < m, (uint32_t)e1, slot0, BL, nownL, sring,
                                  st_del, q_lo, q_hi, slseg, slrec);
        else
          deliver_lane_lds2<false>(a, q < m, (uint32_t)e0, q + kQuad < m, (uint32_t)e1, slot0, BL, nownL, sring,
                                   st_del, q_lo, q_hi, slseg, slrec);
      }
      continue;
    }
    // word q holds entry q - 1, word q + 4 entry q + 3
    if (q >= 1 && q - 1 < c) qent[q - 1] = (uint32_t)w0;
    if (q + 3 < c) qent[q + 3] = (uint32_t)w1;
    __builtin_amdgcn_wave_barrier();
    deliver_staged<DM>(a, qent, min(c, kQuadEnt - 1), q, p, slot0, BL, nownL, L, sring, st_del, q_lo, q_hi, slseg,
                       slrec);
    for (int j0 = kQuadEnt - 1; j0 < c; j0 += kQuadEnt) {
      const int m = min(kQuadEnt, c - j0);
      const unsigned long long* eb = blk + 1 + j0;
      unsigned long long e0 = q < m ? xload64(eb + q) : 0ull;
      unsigned long long e1 = q + kQuad < m ? xload64(eb + q + kQuad) : 0ull;
      long long sp = 0;
      while ((q < m && xtag(e0) != tag) || (q + kQuad < m && xtag(e1) != tag)) {
        if (++sp > a.spin_limit) {
          atomicCAS(a.error, 0, 1);
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (q < m && xtag(e0) != tag) e0 = xload64(eb + q);
        if (q + kQuad < m && xtag(e1) != tag) e1 = xload64(eb + q + kQuad);
      }
      if (!ok) break;
      __builtin_amdgcn_wave_barrier();
      if (q < m) qent[q] = (uint32_t)e0;
      if (q + kQuad < m) qent[q + kQuad] = (uint32_t)e1;
      __builtin_amdgcn_wave_barrier();
      deliver_staged<DM>(a, qent, m, q, p, slot0, BL, nownL, L, sring, st_del, q_lo, q_hi, slseg, slrec);
    }
    __builtin_amdgcn_wave_barrier();
  }
  return !__any(!ok);
}

// Delivery of group s without a grid barrier when partitions are large (long lists, few
// sources): (1) thread t polls source t's count word until it carries the group's tag;
// (2) wavefront 0 scans the counts; (3) the entries of the concatenated lists are staged
// into LDS kEntCap at a time, spread over the whole workgroup (entry index -> source by
// binary search over the scan), every thread's loads in flight together and then checked
// against the tag; (4) quads deliver kSpikesPerQuad staged entries at a time.  `scnt` is
// LDS for P counts, `sscan` for P + 1 offsets.  Returns false when a wait failed (the error
// flag is then set); uniform across the workgroup.
template <int BS, int DM>
__device__ __forceinline__ bool tagged_gather_stage(const FusedArgs& a, int p, int par, uint32_t tag, int slot0,
                                                    int BL, int nownL, int L, int64_t* sring,
                                                    unsigned long long& st_del, long long q_lo, long long q_hi,
                                                    const uint16_t* slseg, const uint16_t* slrec, uint32_t* sent,
                                                    int* scnt, int* sscan, int* s_ok) {
  constexpr int kPer = kEntCap / BS;   // staged entries per thread and round
  const int tid = threadIdx.x;
  const int lane = (int)__lane_id();
  const unsigned long long* xb = a.xlist64 + (size_t)par * a.P * a.xstride;
  for (int src = tid; src < a.P; src += BS) {
    const unsigned long long* blk = xb + (size_t)src * a.xstride;
    unsigned long long w = xload64(blk);
    long long spins = 0;
    bool ok = true;
    while (xtag(w) != tag) {
      if (++spins > a.spin_limit) {
        atomicCAS(a.error, 0, 1);
        ok = false;
        break;
      }
      if ((spins & 63) == 0 && __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      w = xload64(blk);
    }
    if (!ok) *s_ok = 0;
    scnt[src] = ok ? min((int)(uint32_t)w, a.xcap) : 0;
  }
  __syncthreads();
  if (!*s_ok) return false;
  if (tid < kWave) {
    int base = 0;
    for (int k = 0; k < a.P; k += kWave) {
      const int src = k + lane;
      const int c = src < a.P ? scnt[src] : 0;
      const int incl = wave_inclusive_sum(c);
      if (src < a.P) sscan[src] = base + incl - c;
      base += __builtin_amdgcn_readlane(incl, kWave - 1);
    }
    if (tid == 0) sscan[a.P] = base;
  }
  __syncthreads();
  const int total = sscan[a.P];
  const int quad = tid / kQuad, q = tid % kQuad;
  constexpr int nquads = BS / kQuad;
  for (int r0 = 0; r0 < total; r0 += kEntCap) {
    const int nst = min(kEntCap, total - r0);
    uint32_t wo[kPer];            // word offsets from xb
    unsigned long long w[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + u * BS;
      wo[u] = 0u;
      w[u] = 0ull;
      if (e < nst) {
        const int gidx = r0 + e;
        int lo = 0, hi = a.P;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (sscan[mid] <= gidx) lo = mid;
          else hi = mid;
        }
        wo[u] = (uint32_t)lo * (uint32_t)a.xstride + 1u + (uint32_t)(gidx - sscan[lo]);
        w[u] = xload64(xb + wo[u]);
      }
    }
    bool ok = true;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      if (tid + u * BS >= nst) continue;
      long long spins = 0;
      while (ok && xtag(w[u]) != tag) {
        if (++spins > a.spin_limit) {
          atomicCAS(a.error, 0, 1);
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        w[u] = xload64(xb + wo[u]);
      }
      sent[tid + u * BS] = (uint32_t)w[u];
    }
    if (!ok) *s_ok = 0;
    __syncthreads();
    if (!*s_ok) return false;
    for (int j0 = quad * kSpikesPerQuad; j0 < nst; j0 += nquads * kSpikesPerQuad)
      deliver_staged<DM>(a, sent + j0, min(kSpikesPerQuad, nst - j0), q, p, slot0, BL, nownL, L, sring, st_del,
                         q_lo, q_hi, slseg, slrec);
    __syncthreads();
  }
  return true;
}

// Plasticity of the synapses onto partition p (one group).
template <int BS>
__device__ __forceinline__ void fused_plasticity(const FusedRule* rules, int n_rules, int p, int tidx,
                                              float t) {
  for (int r = 0; r < n_rules; ++r) {
    StdpArgs ra = rules[r].args;
    ra.tidx = tidx;
    ra.t = t;
    const int kind = rules[r].kind;
    const int64_t j0 = rules[r].part[p], j1 = rules[r].part[p + 1];
    for (int64_t j = j0 + threadIdx.x; j < j1; j += BS) {
      if (kind == 0) stdp_wd_one(ra, j);
      else if (kind == 1) stdp_vogels_one(ra, j);
      else if (kind == 2) stdp_evans_one(ra, j);
      else stdp_custom_one(ra, j);
    }
  }
}

// G sub-steps of one neuron's LIF recurrence (LIFSpikingNeurons.cu:115-190) with voltage
// input `in[g]`; returns the spike mask.  The same expressions, in the same order, as the
// generic loop in fused_body.
template <int GG>
__device__ __forceinline__ uint32_t lif_block(const float* in, float t, float dt, float tref, float half_dt,
                                              float cv0, float omc, float cib, float thr, float vr, float& V,
                                              float& last) {
  float dvs[GG];
#pragma unroll
  for (int x = 0; x < GG; ++x) dvs[x] = in[x];
  // keep (float)x * dt computed here: hoisted out of the group loop, the GG products would
  // be spilled to scratch and reloaded one per sub-step (an empty asm, no instruction)
  float dtl = dt;
  asm volatile("" : "+v"(dtl));
  uint32_t mk = 0u;
#pragma unroll
  for (int x = 0; x < GG; ++x) {
    const float tg = t + (float)x * dtl;
    const bool ready = ((tg - last) - tref) > half_dt;
    const float vn = ((cv0 + omc * V) + cib) + dvs[x];
    const bool fire = ready && (vn >= thr);
    V = ready ? (fire ? vr : vn) : V;
    last = fire ? tg : last;
    mk |= (fire ? 1u : 0u) << x;
  }
  return mk;
}

// Partition of a workgroup.  Workgroups are dispatched to the XCDs round-robin, so block
// b runs on XCD b % 8; mapping the blocks of one XCD to consecutive partitions keeps that
// XCD's synapse segments contiguous inside every CSR row (its L2 then fetches only its own
// part of a spiking neuron's row instead of every line of it).
__device__ __forceinline__ int partition_of_block(int b, int P) {
  if (P % kXcds != 0) return b;
  return (b % kXcds) * (P / kXcds) + b / kXcds;
}

// The group loop of one network for partition p, run by the workgroup that is member vb of
// the network's P workgroups.
template <int KIND, bool L1, bool PLASTIC, int BS, int DM>
__device__ __forceinline__ void fused_body(const FusedArgs& a, const int p, const int vb, unsigned char* smem) {
  const int tid = threadIdx.x;
  const int lane = (int)__lane_id();
  const int nlo = a.nlo[p], nhi = a.nlo[p + 1];
  const int ilo = a.ilo[p], ihi = a.ilo[p + 1];
  const int nown = nhi - nlo, iown = ihi - ilo;
  const int L = L1 ? 1 : a.L, G = a.G, B = a.B, BL = a.BL, W = a.W;
  const FusedLayout f = fused_layout(BL, L, W, G, a.nown_max, a.iown_max, a.xcap, a.P, a.lseg_U, a.lrec, a.cring,
                                     a.dahead);
  int64_t* sring = (int64_t*)(smem + f.ring);
  uint32_t* scring = (uint32_t*)(smem + f.ring);   // count ring (a.cring)
  uint32_t* shist = (uint32_t*)(smem + f.hist);
  uint8_t* sext = (uint8_t*)(smem + f.ext);
  unsigned long long* simask = (unsigned long long*)(smem + f.imask);
  float* sinp = (float*)(smem + f.inp);
  uint32_t* sxbuf = (uint32_t*)(smem + f.xbuf);
  uint32_t* shead = (uint32_t*)(smem + f.head);
  int* sscan = (int*)(smem + f.scan);
  uint32_t* sent = (uint32_t*)(smem + f.ent);
  int* scounter = (int*)(smem + f.counter);
  unsigned int* sst_cnt = (unsigned int*)(smem + f.st_cnt);
  unsigned int* sst_nisi = (unsigned int*)(smem + f.st_nisi);
  float* sst_prev = (float*)(smem + f.st_prev);
  double* sst_s1 = (double*)(smem + f.st_s1);
  double* sst_s2 = (double*)(smem + f.st_s2);
  unsigned int* sst_g = (unsigned int*)(smem + f.st_g);
  uint32_t* scdf = (uint32_t*)(smem + f.cdf);
  uint8_t* sguide = (uint8_t*)(smem + f.guide);
  uint16_t* slseg = (uint16_t*)(smem + f.lseg);
  uint16_t* slrec = (uint16_t*)(smem + f.lrec);
  __shared__ int s_ok;
  __shared__ int s_wsum[BS / kWave];
  __shared__ unsigned long long s_prof[kProfPhases];

  const int NL = a.N * L;
  const int nownL = nown * L;
  // load the owned ring slice.  The LDS ring keeps BL = max(largest delay, G) slots: when a
  // group starts at step t every pending input is for a step in [t, t + largest delay), so
  // step s lives in LDS slot s % BL, and in slot s % B of the global ring (B = the
  // reference's D + 2G + 1 slots, SpikingSynapses.cu:48), which holds zeros elsewhere.
  // The count ring starts empty: the pending inputs stay in the global ring, read where
  // their steps are (steps below carry_end).
  int carry_end;
  {
    float tl;
    int t0;
    group_time(a, 0, tl, t0);
    carry_end = t0 + BL;
    if (a.cring) {
      for (int k = tid; k < BL * nown; k += BS) scring[k] = 0u;
    } else {
      const int r0 = t0 % BL;
      for (int j = 0; j < BL; ++j) {
        const int s = t0 + (j - r0 + BL) % BL;
        const int64_t* src = a.ring + (size_t)(s % B) * NL + (size_t)nlo * L;
        for (int k = tid; k < nownL; k += BS) sring[(size_t)j * nownL + k] = src[k];
      }
    }
  }
  for (int k = tid; k < W * nown; k += BS) {
    const int wd = k / nown, i = k - wd * nown;
    shist[k] = a.hist_n[(size_t)wd * a.N + nlo + i];
  }
  if (a.st.on) {
    for (int i = tid; i < nown; i += BS) {
      sst_cnt[i] = a.st.count[nlo + i];
      sst_nisi[i] = a.st.nisi[nlo + i];
      sst_prev[i] = a.st.prev[nlo + i];
      sst_s1[i] = a.st.s1[nlo + i];
      sst_s2[i] = a.st.s2[nlo + i];
    }
    if (tid < 64) sst_g[tid] = 0u;
  }
  // LDS-resident delivery table: this partition's segment of every source row, concatenated
  // in source order (so start[u + 1] ends segment u).  Thread t takes a run of consecutive
  // sources: lengths, a workgroup scan of the run sums, then the segments' targets copied
  // to their local starts.  The host allows it only when a partition has < 2^16 targets.
  if (a.lseg_U > 0) {
    const int U = a.lseg_U;
    const int per = (U + BS - 1) / BS;
    const int u0 = min(U, tid * per), u1 = min(U, u0 + per);
    int run = 0;
    for (int u = u0; u < u1; ++u) {
      const size_t sb = (size_t)u * (a.P + 1) + p;
      const uint32_t len = a.seg[sb + 1] - a.seg[sb];
      slseg[u] = (uint16_t)len;
      run += (int)len;
    }
    int incl = run;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (lane == kWave - 1) s_wsum[tid / kWave] = incl;
    __syncthreads();
    int off = incl - run;
    for (int k = 0; k < tid / kWave; ++k) off += s_wsum[k];
    for (int u = u0; u < u1; ++u) {
      const uint32_t len = slseg[u];
      const uint32_t g0 = a.seg[(size_t)u * (a.P + 1) + p];
      for (uint32_t k = 0; k < len; ++k) slrec[off + (int)k] = a.syn16[g0 + k];
      slseg[u] = (uint16_t)off;
      off += (int)len;
    }
    if (u0 < u1 && u1 == U) slseg[U] = (uint16_t)off;
  }
  // per-lane neuron state (lane i owns neuron nlo+i; nown <= BS)
  const bool own = tid < nown;
  const int gi = nlo + (own ? tid : 0);
  float V = 0.f, last = 0.f, cv0 = 0.f, omc = 0.f, cib = 0.f, thr = 0.f, vr = 0.f, rp = 0.f;
  float gl[kMaxLabels];
#pragma unroll
  for (int l = 0; l < kMaxLabels; ++l) gl[l] = 0.f;
  if (own) {
    V = a.V[gi];
    last = a.last_n[gi];
    const float cc = a.c[gi];
    cv0 = cc * a.v0[gi];
    omc = 1.0f - cc;
    cib = cc * a.ibg;
    thr = a.thr[gi];
    vr = a.vreset[gi];
    rp = a.rp[gi];
    if (KIND != 0)
      for (int l = 0; l < L; ++l) gl[l] = a.gcond[(size_t)gi * L + l];
  }
  unsigned long long st_del = 0, st_ext = 0, st_spk = 0, st_in = 0;
  const float half_dt = 0.5f * a.dt;
  // fixed-point weights of the compact form: the full form's expression for one record
  const long long q_lo = __float2ll_rn((a.w_lo * a.scale0) * a.two_k);
  const long long q_hi = __float2ll_rn((a.w_hi * a.scale0) * a.two_k);
  const double qd_lo = (double)q_lo, qd_hi = (double)q_hi, qd_ext = (double)a.q_ext;
  const float inv_G = 1.0f / (float)G;
  const int bits = W * 32;
  const int nown_waves = ((nown + kWave - 1) / kWave) * kWave;
  // count-ring partitions of more than one wavefront without an LDS delivery table: B1's
  // inputs stay in the ring's slots (no separate LDS array, so larger partitions fit; with
  // one LIF wavefront the extra LDS traffic would lengthen the critical path, so small
  // partitions keep the array; kernels of the LDS-table form never take this path)
  const bool in_ring = DM != kDeliverLdsTable && a.cring && a.nown_max > kWave && a.lseg_U == 0;
  // one wavefront holds every unit of the partition (nown_waves <= kWave)
  const bool early = !PLASTIC && a.xquad && a.early && BS > kWave;
  // spike counters of the exchange list, one per group parity (the other parity's is
  // cleared while the group runs)
  if (tid < 2) scounter[tid] = 0;
  if (tid == 0) s_ok = 1;
  for (int j = tid; j < a.ext_k; j += BS) scdf[j] = a.ext_cdf[j];
  if (!PLASTIC) {
    // this launch's exchange blocks start empty (tag 0 is never a group's tag), before any
    // of its publications
    for (int k = tid; k < 2 * a.xstride; k += BS) {
      const int par2 = k / a.xstride, wd = k - par2 * a.xstride;
      xstore64(a.xlist64 + ((size_t)par2 * a.P + p) * a.xstride + wd, 0u, 0u);
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  if (tid < 256) {
    const uint32_t u0 = (uint32_t)tid << 24;
    int c = 0;
    while (c < a.ext_k && u0 >= scdf[c]) ++c;
    sguide[tid] = (uint8_t)c;
  }
  __syncthreads();
  for (int k = tid; k < a.dahead * a.iown_max; k += BS) simask[k] = 0ull;
  __syncthreads();
  draw_groups<BS>(a, 0, min(a.dahead, a.ngroups), sext, simask, nlo, nown, ilo, iown, st_ext, scdf, sguide);
  __syncthreads();
  PhaseClock clk;
  clk.start(a, s_prof, vb);
  // start and end of this group loop (wall clock of workgroup 0 of the network)
  if (a.prof != nullptr && vb == 0 && tid == 0) a.prof[kProfPhases - 2] = wall_clock64();

  for (int s = 0; s < a.ngroups; ++s) {
    float t;
    int tidx;
    group_time(a, s, t, tidx);
    const int par = s & 1;
    int* scnt = scounter + par;
    const int dslot = s % a.dahead;
    const uint8_t* ext = sext + (size_t)dslot * a.nown_max * G;
    unsigned long long* imask = simask + (size_t)dslot * a.iown_max;
    const int b0 = tidx % bits;
    const int slot0 = tidx % BL;
    // -- A. Poisson inputs: append this group's spikes, update histories ------------
    for (int i = tid; i < iown; i += BS) {
      const unsigned long long m = imask[i];
      imask[i] = 0ull;   // cleared for the draws of a later group
      const int c = __popcll(m);
      int pos = c ? atomicAdd(scnt, c) : 0;
      unsigned long long mm = m;
      while (mm) {
        const int g = __ffsll((unsigned long long)mm) - 1;
        mm &= mm - 1ull;
        if (pos < a.xcap) sxbuf[pos] = ((uint32_t)g << 24) | (uint32_t)(a.N + ilo + i);
        ++pos;
      }
      for (int g = 0; g < G; ++g) {
        int b = b0 + g;
        if (b >= bits) b -= bits;
        uint32_t* wp = a.hist_in + (size_t)(b >> 5) * a.Nin + ilo + i;
        uint32_t word = *wp & ~(1u << (b & 31));
        if ((m >> g) & 1ull) word |= (1u << (b & 31));
        *wp = word;
      }
      if (m) a.last_in[ilo + i] = t + (float)(63 - __clzll((long long)m)) * a.dt;
      st_in += (unsigned long long)c;
      if (a.rate_in != nullptr && c) a.rate_in[ilo + i] += (uint32_t)c;
    }
    // -- B1. this group's synaptic input per (neuron, sub-step[, label]) as float ------
    // The ring slots of the group's steps are complete (every delay is >= G): read and
    // clear them in parallel.  Voltage synapses keep the label sum, conductance synapses
    // the per-label values.
    for (int k = tid; k < nown * G; k += BS) {
      // i = k / G without an integer division: (k + 1/2) / G lies at least 1/(2G) from an
      // integer and k < 2^16, far inside float precision
      const int i = (int)(((float)k + 0.5f) * inv_G), g = k - i * G;
      int slot = slot0 + g;
      if (slot >= BL) slot -= BL;
      int64_t* rs = sring + (size_t)slot * nownL + (size_t)i * L;
      if (a.cring) {
        // one label: the counts of the two source types times their fixed-point weights,
        // the external events and, early in a launch, the input pending at its start --
        // integers below 2^53, so their double sum is the exact fixed-point sum
        uint32_t* cs = scring + (size_t)slot * nown + i;
        const uint32_t cw = *cs;
        if (!in_ring) *cs = 0u;
        double qs = (double)(cw & 0xFFFFu) * qd_lo + (double)(cw >> 16) * qd_hi;
        if (a.ext_k > 0) qs += (double)ext[i * G + g] * qd_ext;
        const int step = tidx + g;
        if (step < carry_end) qs += (double)a.ring[(size_t)(step % B) * NL + (size_t)(nlo + i)];
        const float v = (float)qs * a.inv_scale;
        // in the ring: the input stays in its slot as a float, and the neuron's LIF lane
        // takes it and clears the slot
        if (in_ring) *cs = __float_as_uint(v);
        else sinp[k] = v;
        continue;
      }
      const int64_t ext_q = (a.ext_k > 0) ? (int64_t)ext[i * G + g] * a.q_ext : 0;
      if (KIND == 0) {
        float total = 0.0f;
        for (int l = 0; l < L; ++l) {
          int64_t q = rs[l];
          if (l == 0) q += ext_q;
          rs[l] = 0;
          total += fixed_to_float(q) * a.inv_scale;
        }
        sinp[k] = total;
      } else {
        for (int l = 0; l < L; ++l) {
          int64_t q = rs[l];
          if (l == 0) q += ext_q;
          rs[l] = 0;
          sinp[(size_t)k * L + l] = fixed_to_float(q) * a.inv_scale;
        }
      }
    }
    __syncthreads();
    clk.lap(a, 0);
    // -- B2. LIF recurrence, one lane per neuron (LIFSpikingNeurons.cu:115-190) -------
    if (early && tid >= kWave) {
      // the next group's random draws on the other wavefronts while wavefront 0 integrates
      if ((s + 1) % a.dahead == 0 && s + 1 < a.ngroups)
        draw_groups<BS>(a, s + 1, min(a.dahead, a.ngroups - s - 1), sext, simask, nlo, nown, ilo, iown, st_ext,
                        scdf, sguide, kWave);
    } else if (tid < nown_waves || (early && tid < kWave)) {
      // the LIF wavefront of the early form first at the SIMD's issue arbiter while the
      // draws run beside it
      if (early) __builtin_amdgcn_s_setprio(3);
      unsigned long long smask = 0ull;
      if (own) {
        // with in_ring, this group's inputs are floats in the count ring's slots (B1): take
        // them and clear the slots for later deliveries
        auto take = [&](int g) -> float {
          int sl = slot0 + g;
          if (sl >= BL) sl -= BL;
          uint32_t* w = scring + (size_t)sl * nown + tid;
          const float v = __uint_as_float(*w);
          *w = 0u;
          return v;
        };
        if (KIND == 0 && G <= kPrefetch) {
          // G known at compile time: straight-line code, reading the LDS array at immediate
          // offsets or the values taken from the ring into registers
          uint32_t mk = 0u;
#define SPK_LIF_CASE(GG, IN) \
            case GG: mk = lif_block<GG>(IN, t, a.dt, a.tref, half_dt, cv0, omc, cib, thr, vr, V, last); break;
#define SPK_LIF_SWITCH(IN) \
          switch (G) { \
            SPK_LIF_CASE(1, IN) SPK_LIF_CASE(2, IN) SPK_LIF_CASE(3, IN) SPK_LIF_CASE(4, IN) SPK_LIF_CASE(5, IN) \
            SPK_LIF_CASE(6, IN) SPK_LIF_CASE(7, IN) SPK_LIF_CASE(8, IN) SPK_LIF_CASE(9, IN) SPK_LIF_CASE(10, IN) \
            SPK_LIF_CASE(11, IN) SPK_LIF_CASE(12, IN) SPK_LIF_CASE(13, IN) SPK_LIF_CASE(14, IN) \
            SPK_LIF_CASE(15, IN) SPK_LIF_CASE(16, IN) \
            default: break; \
          }
          if (in_ring) {
            float cin[kPrefetch];
#pragma unroll
            for (int x = 0; x < kPrefetch; ++x) cin[x] = (x < G) ? take(x) : 0.0f;
            SPK_LIF_SWITCH(cin)
          } else {
            const float* in = sinp + tid * G;
            SPK_LIF_SWITCH(in)
          }
#undef SPK_LIF_SWITCH
#undef SPK_LIF_CASE
          smask = mk;
        }
        for (int g0 = 0; g0 < G && !(KIND == 0 && G <= kPrefetch); g0 += kPrefetch) {
          const int gn = min(kPrefetch, G - g0);
          if (KIND == 0) {
            float dvs[kPrefetch];
#pragma unroll
            for (int x = 0; x < kPrefetch; ++x)
              dvs[x] = (x < gn) ? (in_ring ? take(g0 + x) : sinp[tid * G + g0 + x]) : 0.0f;
            const float tref = a.tref, dt = a.dt, tg0 = t;
            unsigned long long mk = 0ull;
#pragma unroll
            for (int x = 0; x < kPrefetch; ++x) {
              if (x < gn) {
                const int g = g0 + x;
                const float tg = tg0 + (float)g * dt;
                const bool ready = ((tg - last) - tref) > half_dt;
                const float vn = ((cv0 + omc * V) + cib) + dvs[x];
                const bool fire = ready && (vn >= thr);
                V = ready ? (fire ? vr : vn) : V;
                last = fire ? tg : last;
                mk |= (fire ? 1ull : 0ull) << g;
              }
            }
            smask |= mk;
          } else {
            for (int x = 0; x < gn; ++x) {
              const int g = g0 + x;
              float total = 0.0f;
#pragma unroll
              for (int l = 0; l < kMaxLabels; ++l) {
                if (l < L) {
                  const float in = sinp[((size_t)tid * G + g) * L + l];
                  gl[l] = a.decay_g[l] * gl[l] + in;
                  total += (KIND == 1) ? gl[l] * (a.erev[l] - V) : gl[l];
                }
              }
              const float dV = total * rp;
              const float tg = t + (float)g * a.dt;
              if (((tg - last) - a.tref) > half_dt) {
                V = ((cv0 + omc * V) + cib) + dV;
                if (V >= thr) {
                  last = tg;
                  V = vr;
                  smask |= 1ull << g;
                }
              }
            }
          }
        }
        clk.lap(a, 9);
      }
      if (own) {
        // spike history: clear this group's G bits [b0, b0 + G) of the bit ring and set
        // those of the spikes, one 32-bit word (or the part up to the ring's end) at a time
        {
          int pos = b0, done = 0;
          while (done < G) {
            const int off = pos & 31;
            const int n = min(min(32 - off, G - done), bits - pos);
            const uint32_t ones = n == 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
            const uint32_t sp = (uint32_t)(smask >> done) & ones;
            uint32_t* wp = shist + (pos >> 5) * nown + tid;
            *wp = (*wp & ~(ones << off)) | (sp << off);
            done += n;
            pos += n;
            if (pos >= bits) pos -= bits;
          }
        }
        if (a.rate_n != nullptr && smask) a.rate_n[gi] += (uint32_t)__popcll(smask);
        if (a.st.on && smask) {
          // window statistics of this neuron's spikes, in time order
          unsigned int c0 = sst_cnt[tid], c1 = sst_nisi[tid];
          float pv = sst_prev[tid];
          double d1 = sst_s1[tid], d2 = sst_s2[tid];
          unsigned long long mm = smask;
          while (mm) {
            const int g = __ffsll((unsigned long long)mm) - 1;
            mm &= mm - 1ull;
            stats_spike(a.st, (int64_t)tidx + g, t + (float)g * a.dt, c0, c1, pv, d1, d2);
            atomicAdd(&sst_g[g], 1u);
          }
          sst_cnt[tid] = c0; sst_nisi[tid] = c1; sst_prev[tid] = pv;
          sst_s1[tid] = d1; sst_s2[tid] = d2;
        }
      }
      if (early) {
        // wavefront 0 holds the partition: its spikes' list positions by a wavefront prefix
        // sum after the inputs' entries of phase A, and the list goes to the exchange block
        // at once (phase C publishes nothing); after the history update, whose loop holds a
        // vector-memory counter wait that would otherwise wait for these stores to complete
        const int c = __popcll(smask);
        const int incl = wave_inclusive_sum(c);
        const int base = *scnt;
        int pos = base + incl - c;
        const int nxr = base + __builtin_amdgcn_readlane(incl, kWave - 1);
        unsigned long long* xl = a.xlist64 + ((size_t)par * a.P + p) * a.xstride;
        const uint32_t tag = a.xtag_base + (uint32_t)s + 1u;
        st_spk += (unsigned long long)c;
        unsigned long long mm = smask;
        while (mm) {
          const int g = __ffsll((unsigned long long)mm) - 1;
          mm &= mm - 1ull;
          if (pos < a.xcap) {
            const uint32_t e = ((uint32_t)g << 24) | (uint32_t)(nlo + tid);
            sxbuf[pos] = e;
            xstore64(xl + 1 + pos, e, tag);
          }
          ++pos;
        }
        for (int k = lane; k < min(base, a.xcap); k += kWave) xstore64(xl + 1 + k, sxbuf[k], tag);
        if (lane == 0) {
          xstore64(xl, (uint32_t)min(nxr, a.xcap), tag);
          *scnt = nxr;
        }
        __builtin_amdgcn_s_setprio(0);
      }
      if (!early && smask) {
        const int c = __popcll(smask);
        int pos = atomicAdd(scnt, c);
        st_spk += (unsigned long long)c;
        unsigned long long mm = smask;
        while (mm) {
          const int g = __ffsll((unsigned long long)mm) - 1;
          mm &= mm - 1ull;
          if (pos < a.xcap) sxbuf[pos] = ((uint32_t)g << 24) | (uint32_t)(nlo + tid);
          ++pos;
        }
      }
    }
    __syncthreads();
    if (a.st.on && tid < G) {
      // population spikes of this partition per sub-step into the window's bins
      const unsigned int c = sst_g[tid];
      const int64_t step = (int64_t)tidx + tid;
      if (c && step >= a.st.t0 && step < a.st.t0 + a.st.nsteps)
        atomicAdd(a.st.hist + (step - a.st.t0) / a.st.bin, c);
      sst_g[tid] = 0u;
    }
    clk.lap(a, 1);
    // -- C. publish: {count, entries} block, and state other workgroups may read ------
    const int nxr = scnt[0];
    const int nx = min(nxr, a.xcap);
    if (tid == 0) scounter[par ^ 1] = 0;
    // more spikes than the exchange list holds: stop the launch with error 2
    if (tid == 0 && nxr > a.xcap) atomicExch(a.error, 2);
    const uint32_t tag = a.xtag_base + (uint32_t)s + 1u;
    if (!PLASTIC) {
      if (!early) {
        unsigned long long* xl = a.xlist64 + ((size_t)par * a.P + p) * a.xstride;
        for (int k = tid; k < nx; k += BS) xstore64(xl + 1 + k, sxbuf[k], tag);
        if (tid == 0) xstore64(xl, (uint32_t)nx, tag);
      }
    } else {
      uint32_t* xl = a.xlist + ((size_t)par * a.P + p) * a.xstride;
      for (int k = tid; k < nx; k += BS) store_coherent(xl + 1 + k, sxbuf[k]);
      if (tid == 0) store_coherent(xl, (uint32_t)nx);
    }
    if (PLASTIC) {
      if (own) a.last_n[gi] = last;
      for (int k = tid; k < W * nown; k += BS) {
        const int wd = k / nown, i = k - wd * nown;
        a.hist_n[(size_t)wd * a.N + nlo + i] = shist[k];
      }
    }
    // spike monitors (SpikingActivityMonitor.cu:69-98): one store slot per wavefront
    // ballot; time = group time + g * dt as the reference's collection kernel computes it
    // (the guard keeps the monitors' argument loads out of launches without monitors)
    if (a.n_mons > 0)
#pragma unroll
    for (int mi = 0; mi < kMaxFusedMons; ++mi) {
      if (mi >= a.n_mons) continue;
      const FusedMon m = a.mon[mi];
      for (int k0 = 0; k0 < nx; k0 += BS) {
        const int k = k0 + tid;
        const uint32_t e = k < nx ? sxbuf[k] : 0u;
        const uint32_t id = e & kIdMask;
        const bool in = k < nx && id >= m.lo && id < m.hi;
        const unsigned long long bal = __ballot(in);
        unsigned long long base = 0;
        if (lane == 0 && bal) base = atomicAdd(m.count, (unsigned long long)__popcll(bal));
        base = __shfl(base, 0);
        if (in) {
          const unsigned long long pos = base + (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
          if (pos < (unsigned long long)m.cap) {
            m.ids[pos] = (int32_t)(id - m.lo);
            m.times[pos] = t + (float)(e >> 24) * a.dt;
          }
        }
      }
    }
    if (!PLASTIC) {
      clk.lap(a, 2);
      // -- D. the next group's random draws while the other workgroups finish ----------
      if (!early && (s + 1) % a.dahead == 0 && s + 1 < a.ngroups)
        draw_groups<BS>(a, s + 1, min(a.dahead, a.ngroups - s - 1), sext, simask, nlo, nown, ilo, iown, st_ext,
                        scdf, sguide);
      clk.lap(a, 3);
      // -- F. every source's list of this group, polled by its tags, delivered into the
      //    LDS ring (no grid barrier)
      if (a.xquad) {
        if (!tagged_gather_deliver<BS, DM>(a, p, par, tag, slot0, BL, nownL, L, sring, st_del, q_lo, q_hi, slseg,
                                           slrec, sent))
          s_ok = 0;
        __syncthreads();
      } else {
        tagged_gather_stage<BS, DM>(a, p, par, tag, slot0, BL, nownL, L, sring, st_del, q_lo, q_hi, slseg, slrec,
                                    sent, (int*)shead, sscan, &s_ok);
      }
      if (!s_ok) return;
      clk.lap(a, 8);
      continue;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    clk.lap(a, 2);
    // -- D. barrier: arrive, draw the next group while others finish, wait ----------
    if (tid == 0) barrier_arrive(a, s, vb);
    if ((s + 1) % a.dahead == 0 && s + 1 < a.ngroups)
      draw_groups<BS>(a, s + 1, min(a.dahead, a.ngroups - s - 1), sext, simask, nlo, nown, ilo, iown, st_ext,
                      scdf, sguide);
    clk.lap(a, 3);
    if (tid == 0) s_ok = barrier_wait(a, s) ? 1 : 0;
    __syncthreads();
    if (!s_ok) return;
    clk.lap(a, 4);
    // -- E. plasticity of the synapses onto this partition ---------------------------
    if (PLASTIC) {
      fused_plasticity<BS>(a.rules, a.n_rules, p, tidx, t);
      __syncthreads();
    }
    clk.lap(a, 5);
    // -- F. delivery into the LDS ring ------------------------------------------------
    {
      const uint32_t* xbase = a.xlist + (size_t)par * a.P * a.xstride;
      // (1) heads of all source blocks: count and first entries, one coalesced trip
      for (int k = tid; k < a.P * kHeadWords; k += BS) {
        const int src = k / kHeadWords, wd = k - src * kHeadWords;
        shead[k] = (wd <= a.xcap) ? load_coherent(xbase + (size_t)src * a.xstride + wd) : 0u;
      }
      __syncthreads();
      clk.lap(a, 6);
      // (2) wavefront 0: exclusive scan of the counts, kWave consecutive sources at a time
      const int cps = (a.P + kWave - 1) / kWave;
      if (tid < kWave) {
        int base = 0;
        for (int k = 0; k < cps; ++k) {
          const int src = k * kWave + lane;
          const int c = src < a.P ? min((int)shead[src * kHeadWords], a.xcap) : 0;
          const int incl = wave_inclusive_sum(c);
          if (src < a.P) sscan[src] = base + incl - c;
          base += __builtin_amdgcn_readlane(incl, kWave - 1);
        }
        if (tid == 0) sscan[a.P] = base;
      }
      __syncthreads();
      const int total = sscan[a.P];
      int lps = 1;   // lanes per source: the largest power of two (up to 64) with lps * P <= BS
      while (lps < kWave && lps * 2 * a.P <= BS) lps *= 2;
      // lane groups when the sources are many (short lists, mostly in the heads)
      const bool by_source = lps <= 16;
      for (int r0 = 0; r0 < total; r0 += kEntCap) {
        // (3) stage entries [r0, r0 + kEntCap) of the concatenated lists; the first
        //     kHeadWords-1 entries of a source come from its head in LDS, the rest from its
        //     block.  Either a group of lps lanes per source, or one entry per thread with
        //     its source found by binary search (the last with sscan[src] <= index; empty
        //     sources share their start with the next one), which spreads long lists over
        //     the whole workgroup.
        if (by_source) {
          for (int src = tid / lps; src < a.P; src += BS / lps) {
            const int b0 = sscan[src], b1 = sscan[src + 1];
            const int j0 = max(0, r0 - b0), j1 = min(min(b1 - b0, a.xcap), r0 + kEntCap - b0);
            for (int j = j0 + (tid & (lps - 1)); j < j1; j += lps)
              sent[b0 + j - r0] = (j + 1 < kHeadWords) ? shead[src * kHeadWords + 1 + j]
                                                       : load_coherent(xbase + (size_t)src * a.xstride + 1 + j);
          }
        } else {
          const int nst = min(kEntCap, total - r0);
          for (int e = tid; e < nst; e += BS) {
            const int gidx = r0 + e;
            int lo = 0, hi = a.P;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (sscan[mid] <= gidx) lo = mid;
              else hi = mid;
            }
            const int j = gidx - sscan[lo];
            sent[e] = (j + 1 < kHeadWords) ? shead[lo * kHeadWords + 1 + j]
                    : (j < a.xcap ? load_coherent(xbase + (size_t)lo * a.xstride + 1 + j) : 0u);
          }
        }
        __syncthreads();
        clk.lap(a, 7);
        // (4) quads of lanes over the spikes, kSpikesPerQuad at a time
        const int nr = min(kEntCap, total - r0);
        const int quad = tid / kQuad, q = tid % kQuad;
        const int nquads = BS / kQuad;
        for (int j0 = quad * kSpikesPerQuad; j0 < nr; j0 += nquads * kSpikesPerQuad) {
          const int n = min(kSpikesPerQuad, nr - j0);
          if (a.lseg_U > 0 || a.syn16 != nullptr)
            deliver_staged<DM>(a, sent + j0, n, q, p, slot0, BL, nownL, L, sring, st_del, q_lo, q_hi, slseg, slrec);
          else if (a.scale == nullptr)
            deliver_quad<false>(a, sent + j0, n, q, p, slot0, BL, nownL, L, sring, st_del);
          else
            deliver_quad<true>(a, sent + j0, n, q, p, slot0, BL, nownL, L, sring, st_del);
        }
        __syncthreads();
        clk.lap(a, 8);
      }
    }
  }
  clk.flush(a);
  if (a.prof != nullptr && vb == 0 && tid == 0) a.prof[kProfPhases - 1] = wall_clock64();
  // write back state
  if (own) {
    a.V[gi] = V;
    a.last_n[gi] = last;
    if (KIND != 0)
      for (int l = 0; l < L; ++l) a.gcond[(size_t)gi * L + l] = gl[l];
  }
  // write the whole global ring slice back in its B-slot layout (zeros past the LDS window)
  {
    float tl;
    int t1;
    group_time(a, a.ngroups, tl, t1);
    const int r1 = t1 % B;
    for (int k = 0; k < B; ++k) {
      const int s = t1 + (k - r1 + B) % B;
      int64_t* dst = a.ring + (size_t)k * NL + (size_t)nlo * L;
      const bool live = s - t1 < BL;
      if (a.cring) {
        // counts back to fixed point, plus the pending input of the launch's start still
        // unread (a launch shorter than BL steps)
        const uint32_t* src = scring + (size_t)(s % BL) * nown;
        for (int j = tid; j < nown; j += BS) {
          int64_t v = 0;
          if (live) {
            const uint32_t cw = src[j];
            v = (int64_t)(cw & 0xFFFFu) * q_lo + (int64_t)(cw >> 16) * q_hi;
            if (s < carry_end) v += dst[j];
          }
          dst[j] = v;
        }
      } else {
        const int64_t* src = sring + (size_t)(s % BL) * nownL;
        for (int j = tid; j < nownL; j += BS) dst[j] = live ? src[j] : 0;
      }
    }
  }
  for (int k = tid; k < W * nown; k += BS) {
    const int wd = k / nown, i = k - wd * nown;
    a.hist_n[(size_t)wd * a.N + nlo + i] = shist[k];
  }
  if (a.st.on) {
    for (int i = tid; i < nown; i += BS) {
      a.st.count[nlo + i] = sst_cnt[i];
      a.st.nisi[nlo + i] = sst_nisi[i];
      a.st.prev[nlo + i] = sst_prev[i];
      a.st.s1[nlo + i] = sst_s1[i];
      a.st.s2[nlo + i] = sst_s2[i];
    }
  }
  // statistics: one atomic per wavefront
  unsigned long long v4[4] = {st_del, st_ext, st_spk, st_in};
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    unsigned long long v = v4[qq];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(a.stats + qq, v);
  }
}

template <int KIND, bool L1, bool PLASTIC, int DM = kDeliverAny>
__global__ __launch_bounds__(kFusedBlock) void k_fused(FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  fused_body<KIND, L1, PLASTIC, kFusedBlock, DM>(a, partition_of_block(blockIdx.x, gridDim.x), blockIdx.x, smem);
}

// Many independent networks in one cooperative launch.  Workgroup b runs the tasks
// tasks[task_off[b] .. task_off[b+1]) in order, each the whole group loop of one partition
// of one network (x = network, y = partition).  Every workgroup's list is in increasing
// network order, so a workgroup waiting at a network's barrier waits only for workgroups
// that are finishing earlier networks: every wait ends.  BS threads per workgroup: with
// kBatchSmallBlock, two workgroups share a compute unit, so one network's partition runs
// while the other's waits at its exchange (partitions of at most BS neurons).
template <int KIND, bool L1, bool PLASTIC, int BS, int DM = kDeliverAny>
__global__ __launch_bounds__(BS, 4) void k_fused_batch(const FusedArgs* __restrict__ nets,
                                                    const int2* __restrict__ tasks,
                                                    const int* __restrict__ task_off) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int j0 = __builtin_amdgcn_readfirstlane(task_off[blockIdx.x]);
  const int j1 = __builtin_amdgcn_readfirstlane(task_off[blockIdx.x + 1]);
  for (int j = j0; j < j1; ++j) {
    const int k = __builtin_amdgcn_readfirstlane(tasks[j].x);
    const int p = __builtin_amdgcn_readfirstlane(tasks[j].y);
    fused_body<KIND, L1, PLASTIC, BS, DM>(nets[k], p, p, smem);
    __syncthreads();
  }
}

}  // namespace spk
