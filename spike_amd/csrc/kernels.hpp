// kernels.hpp -- per-phase gfx950 kernels of the Spike backend.
//
// One lane per neuron for the neuron phases; for delivery, each spiking unit's row is
// shared by several 64-lane wavefronts.  State is structure-of-arrays FP32; the input ring
// is int64 fixed point (exact, order-independent sums); spike lists are compacted with a
// wavefront ballot + popcount prefix and one atomic per wavefront.  Networks with G = 1 run
// a group as two launches (k_units, k_deliver) built from the same per-unit functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "philox.hpp"

#pragma clang fp contract(off)

namespace spk {

constexpr int kWave = 64;
// Synapse records are {packed target/delay/label, weight} pairs: weight k is w[k * kWStride]
// of a float view offset by one word.
constexpr int kWStride = 2;
constexpr int kMaxLabels = 8;
constexpr int kMaxExtTable = 64;

// An input-ring sum (|q| <= 2^52 by the choice of fraction bits) as float: through double,
// which holds q exactly, so the one rounding is the same as that of (float)q and the
// conversion is a few instructions instead of a 64-bit integer-to-float sequence.
__host__ __device__ __forceinline__ float fixed_to_float(int64_t q) { return (float)(double)q; }

// ---------------------------------------------------------------------------------
// 16-bit weight and trace storage (format 1: bfloat16, format 2: IEEE binary16), both
// rounded to nearest, ties to even, by integer operations (no dependence on the device's
// conversion modes); decoding is exact.

__host__ __device__ __forceinline__ uint32_t f32_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
__host__ __device__ __forceinline__ float bits_f32(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const uint32_t u = f32_bits(f);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN stays NaN
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__host__ __device__ __forceinline__ uint16_t f32_to_f16(float f) {
  const uint32_t x = f32_bits(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);   // rounds to 65520 or more: inf
  if (ax <= 0x33000000u) return (uint16_t)sign;                // at most 2^-25: zero
  const int e = (int)(ax >> 23) - 127;
  const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
  uint32_t q, rem, half;
  if (e < -14) {                                               // binary16 subnormal
    const int sh = -(e + 1);                                   // 14..24
    q = m >> sh;
    rem = m & ((1u << sh) - 1u);
    half = 1u << (sh - 1);
  } else {
    q = ((uint32_t)(e + 15) << 10) | ((m >> 13) & 0x3FFu);
    rem = m & 0x1FFFu;
    half = 0x1000u;
  }
  if (rem > half || (rem == half && (q & 1u))) ++q;            // a carry may raise the exponent
  return (uint16_t)(sign | q);
}

__host__ __device__ __forceinline__ float f16_to_f32(uint16_t h) {
  const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  const uint32_t e = ((uint32_t)h >> 10) & 0x1Fu, m = (uint32_t)h & 0x3FFu;
  if (e == 0u) return bits_f32(f32_bits((float)m * 5.9604644775390625e-8f) | sign);   // m 2^-24
  if (e == 31u) return bits_f32(sign | 0x7F800000u | (m << 13));
  return bits_f32(sign | ((e + 112u) << 23) | (m << 13));
}

__host__ __device__ __forceinline__ uint16_t enc16(float f, int fmt) {
  return fmt == 1 ? f32_to_bf16(f) : f32_to_f16(f);
}
__host__ __device__ __forceinline__ float dec16(uint16_t h, int fmt) {
  return fmt == 1 ? bits_f32((uint32_t)h << 16) : f16_to_f32(h);
}

// Synaptic weights: FP32 in the synapse records (w[k * kWStride]), or, with fmt != 0, one
// 32-bit word per synapse: the 16-bit weight in the low half and its 16-bit Kahan
// compensation for plastic updates in the high half (one memory line per random update).
struct Weights {
  float* w;
  uint32_t* wc;
  int fmt;
};

__device__ __forceinline__ float wload(const Weights& W, int64_t k) {
  return W.fmt ? dec16((uint16_t)(W.wc[k] & 0xFFFFu), W.fmt) : W.w[(size_t)k * kWStride];
}

// The weight and (16-bit formats) its compensation word, in one load.
__device__ __forceinline__ float wload_c(const Weights& W, int64_t k, uint16_t& c16) {
  if (!W.fmt) {
    c16 = 0;
    return W.w[(size_t)k * kWStride];
  }
  const uint32_t word = W.wc[k];
  c16 = (uint16_t)(word >> 16);
  return dec16((uint16_t)(word & 0xFFFFu), W.fmt);
}

// Store a plastic weight that a rule moved from w0 (the stored value) to nw (already
// clamped by the rule to [lo, hi]).  FP32: nw as it is.  16-bit: one step of Kahan
// summation of the increment, carrying what the 16-bit word cannot hold in c16; a value
// that the compensation would push outside [lo, hi] is clamped and the compensation
// cleared.  A zero increment is no step at all (the weights a spike-driven rule skips
// must be those the per-synapse form leaves).
// Returns the weight as stored afterwards.  c16 is the synapse's compensation word, when
// the caller has loaded it already (wupdate loads it otherwise).
__device__ __forceinline__ float wupdate_c(const Weights& W, int64_t k, float w0, uint16_t c16, float nw,
                                           float lo, float hi) {
  if (!W.fmt) {
    W.w[(size_t)k * kWStride] = nw;
    return nw;
  }
  if (nw == w0) return w0;   // a zero increment leaves the word and its compensation as they are
  const float c = dec16(c16, W.fmt);
  const float d = nw - w0;
  const float y = d - c;
  const float t = w0 + y;
  uint16_t s1 = enc16(t, W.fmt);
  const float v1 = dec16(s1, W.fmt);
  uint16_t c1;
  if (v1 < lo) {
    s1 = enc16(lo, W.fmt);
    c1 = 0;
  } else if (v1 > hi) {
    s1 = enc16(hi, W.fmt);
    c1 = 0;
  } else {
    c1 = enc16((v1 - w0) - y, W.fmt);
  }
  W.wc[k] = (uint32_t)s1 | ((uint32_t)c1 << 16);
  return dec16(s1, W.fmt);
}

__device__ __forceinline__ float wupdate(const Weights& W, int64_t k, float w0, float nw, float lo, float hi) {
  if (!W.fmt) return wupdate_c(W, k, w0, 0, nw, lo, hi);
  if (nw == w0) return w0;
  return wupdate_c(W, k, w0, (uint16_t)(W.wc[k] >> 16), nw, lo, hi);
}

// Plasticity traces per plastic synapse: FP32, or 16-bit words of the weight format.
struct Traces {
  float* a;
  float* b;
  uint16_t* a16;
  uint16_t* b16;
  int fmt;
};
__device__ __forceinline__ float tload(const Traces& T, int which, int64_t j) {
  if (!T.fmt) return which ? T.b[j] : T.a[j];
  return dec16(which ? T.b16[j] : T.a16[j], T.fmt);
}
__device__ __forceinline__ void tstore(const Traces& T, int which, int64_t j, float v) {
  if (!T.fmt) {
    (which ? T.b : T.a)[j] = v;
    return;
  }
  (which ? T.b16 : T.a16)[j] = enc16(v, T.fmt);
}

// ---------------------------------------------------------------------------------
// spike history: bit (step mod 32W) of neuron i lives in word hist[(b/32)*npop + i]

__device__ __forceinline__ int hist_bitpos(int64_t step, int W) {
  const int64_t bits = (int64_t)W * 32;
  int64_t b = step % bits;
  if (b < 0) b += bits;
  return (int)b;
}

__device__ __forceinline__ bool hist_get(const uint32_t* __restrict__ hist, int npop, int W,
                                         int i, int64_t step) {
  const int b = hist_bitpos(step, W);
  return (hist[(size_t)(b >> 5) * npop + i] >> (b & 31)) & 1u;
}

// Append `id` to list (when `pred`) with one atomic per wavefront.  Every lane of the
// wavefront must call this (the ballot spans the whole wavefront).
__device__ __forceinline__ void wave_append(bool pred, uint32_t id, uint32_t* __restrict__ list,
                                            int* __restrict__ count, int cap,
                                            unsigned long long* __restrict__ stat) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) {
    base = atomicAdd(count, __popcll(m));
    atomicAdd(stat, (unsigned long long)__popcll(m));
  }
  base = __shfl(base, leader);
  if (pred) {
    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
    if (pos < cap) list[pos] = id;
  }
}

// ---------------------------------------------------------------------------------
// group bookkeeping: clear the G per-sub-step counts of a list slot, record its time

__global__ void k_begin_group(int* __restrict__ counts, int G, float* __restrict__ slot_time,
                              int* __restrict__ slot_tidx, int slot, float t, int tidx) {
  const int i = threadIdx.x;
  for (int g = i; g < G; g += blockDim.x) counts[g] = 0;
  if (i == 0) {
    slot_time[slot] = t;
    slot_tidx[slot] = tidx;
  }
}

// ---------------------------------------------------------------------------------
// LIF neurons (LIFSpikingNeurons.cu:91-195; injection InlineDeviceFunctions.hpp:22-111)

// ---------------------------------------------------------------------------------
// Spike statistics of the LIF population over a window of steps [t0, t0 + nsteps):
// per neuron the spike count and the sums of its inter-spike intervals (FP64), and the
// population spike count per bin of `bin` steps.  Accumulated where the spikes are made,
// reduced in FP64 by k_stats_reduce.
struct StatsArgs {
  unsigned int* count;   // [N] spikes in the window
  unsigned int* nisi;    // [N] intervals between consecutive in-window spikes
  float* prev;           // [N] time of the neuron's last in-window spike
  double* s1;            // [N] sum of intervals (s)
  double* s2;            // [N] sum of squared intervals
  unsigned int* hist;    // [nbins] population spikes per bin
  int64_t t0, nsteps;
  int bin, on;
};

// One spike of neuron i at step `step`, time tg: update its running statistics.
__device__ __forceinline__ void stats_spike(const StatsArgs& st, int64_t step, float tg, unsigned int& cnt,
                                            unsigned int& nisi, float& prev, double& s1, double& s2) {
  if (step < st.t0 || step >= st.t0 + st.nsteps) return;
  if (cnt > 0u) {
    const double isi = (double)tg - (double)prev;
    s1 += isi;
    s2 += isi * isi;
    ++nisi;
  }
  ++cnt;
  prev = tg;
}

struct LifArgs {
  float* V;
  float* last;
  uint32_t* hist;
  int64_t* ring;
  float* gcond;               // [N*L] conductance synapses only
  const float* v0;
  const float* vreset;
  const float* thr;
  const float* c;
  const float* rp;
  const float* decay_g;       // [L]
  const float* erev;          // [L]
  const uint32_t* ext_cdf;    // [ext_k]
  uint32_t* lists;            // [G][cap]
  int* counts;                // [G]
  unsigned long long* ext_events;
  unsigned long long* spike_stat;
  int N, L, G, B, W, cap, ext_k, tidx;
  float t, dt, tref, ibg, inv_scale;
  int64_t q_ext;              // fixed-point value of one external event
  uint64_t seed;
  StatsArgs st;
  uint32_t* rate;             // [N] spike counts of a rate monitor, or nullptr
  int uni;                    // 1: every neuron has the parameters u_* (the arrays are not read)
  float u_v0, u_vreset, u_thr, u_c, u_rp;
};

// One lane of the LIF update for neuron i (lanes with i >= N take part in the wavefront's
// list appends only).  Returns the neuron's last spike time after the group.
template <int KIND>
__device__ __forceinline__ float lif_unit(const LifArgs& a, int i) {
  const bool active = i < a.N;
  const int ii = active ? i : 0;
  float V = a.V[ii];
  float last = a.last[ii];
  const float cc = a.uni ? a.u_c : a.c[ii];
  const float cv0 = cc * (a.uni ? a.u_v0 : a.v0[ii]);
  const float omc = 1.0f - cc;
  const float cib = cc * a.ibg;
  const float thr = a.uni ? a.u_thr : a.thr[ii];
  const float vr = a.uni ? a.u_vreset : a.vreset[ii];
  const float rp = a.uni ? a.u_rp : a.rp[ii];
  const size_t NL = (size_t)a.N * a.L;
  float gl[kMaxLabels];
  if (KIND != 0) {
#pragma unroll
    for (int l = 0; l < kMaxLabels; ++l) gl[l] = (l < a.L) ? a.gcond[(size_t)ii * a.L + l] : 0.0f;
  }
  int cur_w = -1;
  uint32_t word = 0;
  unsigned long long ext_sum = 0;
  const float half_dt = 0.5f * a.dt;
  int64_t ext_blk = -1;
  u32x4 ext_r{0u, 0u, 0u, 0u};
  unsigned int st_cnt = 0u, st_nisi = 0u;
  uint32_t nspk = 0u;
  float st_prev = 0.0f;
  double st_s1 = 0.0, st_s2 = 0.0;
  if (a.st.on && active) {
    st_cnt = a.st.count[i]; st_nisi = a.st.nisi[i]; st_prev = a.st.prev[i];
    st_s1 = a.st.s1[i]; st_s2 = a.st.s2[i];
  }
  for (int g = 0; g < a.G; ++g) {
    const int64_t step = (int64_t)a.tidx + g;
    const int b = hist_bitpos(step, a.W);
    const int w = b >> 5;
    if (w != cur_w) {
      if (cur_w >= 0 && active) a.hist[(size_t)cur_w * a.N + i] = word;
      word = active ? a.hist[(size_t)w * a.N + i] : 0u;
      cur_w = w;
    }
    word &= ~(1u << (b & 31));
    // external per-neuron Poisson drive, joined to label 0 of this step's input
    int64_t ext = 0;
    if (a.ext_k > 0) {
      if ((step >> 2) != ext_blk) {
        ext_blk = step >> 2;
        ext_r = draw_block(a.seed, ext_blk, (uint32_t)ii, kStreamExternal);
      }
      const uint32_t u = pick(ext_r, (int)(step & 3));
      int k = 0;
      for (int j = 0; j < a.ext_k; ++j) k += (u >= a.ext_cdf[j]) ? 1 : 0;
      ext = (int64_t)k * a.q_ext;
      ext_sum += (unsigned long long)k;
    }
    const size_t slot = (size_t)((a.tidx + g) % a.B);
    int64_t* rs = a.ring + slot * NL + (size_t)ii * a.L;
    float dV;
    if (KIND == 0) {
      float total = 0.0f;
      for (int l = 0; l < a.L; ++l) {
        int64_t q = active ? rs[l] : 0;
        if (l == 0) q += ext;
        if (active) rs[l] = 0;
        total += fixed_to_float(q) * a.inv_scale;
      }
      dV = total;
    } else {
      float total = 0.0f;
#pragma unroll
      for (int l = 0; l < kMaxLabels; ++l) {
        if (l < a.L) {
          int64_t q = active ? rs[l] : 0;
          if (l == 0) q += ext;
          if (active) rs[l] = 0;
          const float in = fixed_to_float(q) * a.inv_scale;
          gl[l] = a.decay_g[l] * gl[l] + in;
          // conductance: g (E - V) (InlineDeviceFunctions.hpp:22-51); current: the
          // current trace itself (:52-84)
          total += (KIND == 1) ? gl[l] * (a.erev[l] - V) : gl[l];
        }
      }
      dV = total * rp;
    }
    const float tg = a.t + (float)g * a.dt;
    bool spk = false;
    if (((tg - last) - a.tref) > half_dt) {
      V = ((cv0 + omc * V) + cib) + dV;
      if (V >= thr) {
        spk = true;
        last = tg;
        V = vr;
      }
    }
    spk = spk && active;
    if (spk) word |= (1u << (b & 31));
    if (spk && a.st.on) {
      stats_spike(a.st, step, tg, st_cnt, st_nisi, st_prev, st_s1, st_s2);
      if (step >= a.st.t0 && step < a.st.t0 + a.st.nsteps) atomicAdd(a.st.hist + (step - a.st.t0) / a.st.bin, 1u);
    }
    wave_append(spk, (uint32_t)i, a.lists + (size_t)g * a.cap, a.counts + g, a.cap, a.spike_stat);
    nspk += spk ? 1u : 0u;
  }
  if (active) {
    if (cur_w >= 0) a.hist[(size_t)cur_w * a.N + i] = word;
    if (a.rate != nullptr && nspk) a.rate[i] += nspk;
    a.V[i] = V;
    a.last[i] = last;
    if (KIND != 0) {
      for (int l = 0; l < a.L; ++l) a.gcond[(size_t)i * a.L + l] = gl[l];
    }
    if (a.st.on) {
      a.st.count[i] = st_cnt; a.st.nisi[i] = st_nisi; a.st.prev[i] = st_prev;
      a.st.s1[i] = st_s1; a.st.s2[i] = st_s2;
    }
  }
  if (a.ext_k > 0) {
    // one atomic per wavefront for the event counter
    unsigned long long s = active ? ext_sum : 0ull;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == 0 && s) atomicAdd(a.ext_events, s);
  }
  return last;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_lif(LifArgs a) {
  lif_unit<KIND>(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------------
// Poisson inputs (PoissonInputSpikingNeurons.cu:73-130), Bernoulli form

struct PoissonArgs {
  float* last;
  uint32_t* hist;
  const uint32_t* thr;   // [Nin] of the selected stimulus
  uint32_t* lists;
  int* counts;
  unsigned long long* spike_stat;
  int N, Nin, G, W, cap, tidx;
  int64_t arm_step;
  float t, dt;
  uint64_t seed;
  uint32_t* rate;          // [Nin] spike counts of a rate monitor, or nullptr
};

// One lane of the Poisson inputs for input i; returns its last spike time after the group.
__device__ __forceinline__ float poisson_unit(const PoissonArgs& a, int i) {
  const bool active = i < a.Nin;
  const int ii = active ? i : 0;
  const uint32_t T = a.thr[ii];
  float last = a.last[ii];
  int64_t in_blk = -1;
  u32x4 in_r{0u, 0u, 0u, 0u};
  int cur_w = -1;
  uint32_t word = 0;
  uint32_t nspk = 0u;
  for (int g = 0; g < a.G; ++g) {
    const int64_t step = (int64_t)a.tidx + g;
    const int b = hist_bitpos(step, a.W);
    const int w = b >> 5;
    if (w != cur_w) {
      if (cur_w >= 0 && active) a.hist[(size_t)cur_w * a.Nin + i] = word;
      word = active ? a.hist[(size_t)w * a.Nin + i] : 0u;
      cur_w = w;
    }
    word &= ~(1u << (b & 31));
    if ((step >> 2) != in_blk) {
      in_blk = step >> 2;
      in_r = draw_block(a.seed, in_blk, (uint32_t)ii, kStreamInput);
    }
    const uint32_t u = pick(in_r, (int)(step & 3));
    const bool spk = active && (u < T) && (step != a.arm_step);
    if (spk) {
      word |= (1u << (b & 31));
      last = a.t + (float)g * a.dt;
      ++nspk;
    }
    wave_append(spk, (uint32_t)(a.N + i), a.lists + (size_t)g * a.cap, a.counts + g, a.cap,
                a.spike_stat);
  }
  if (active) {
    if (cur_w >= 0) a.hist[(size_t)cur_w * a.Nin + i] = word;
    a.last[i] = last;
    if (a.rate != nullptr && nspk) a.rate[i] += nspk;
  }
  return last;
}

__global__ __launch_bounds__(256) void k_poisson(PoissonArgs a) {
  poisson_unit(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------------
// Generator inputs (GeneratorInputSpikingNeurons.cu:71-128): scheduled spikes of the
// selected stimulus.  The host passes the stimulus entries [lo, hi) (sorted by time) that
// can fall in this group; entry e fires input ids[e] at sub-step g when
// |(t - onset + g*dt) - times[e]| < 0.5*dt, the reference's test (its float difference
// compared in double with 0.5*dt).  One lane per input: it clears its G history bits,
// then walks the entries in time order, so its last spike time is its latest one; an
// input spikes at most once per step (a second entry for the same step finds the history
// bit set), like every other unit.

struct GeneratorArgs {
  float* last;
  uint32_t* hist;
  const int32_t* ids;
  const float* times;
  uint32_t* lists;
  int* counts;
  unsigned long long* spike_stat;
  int64_t lo, hi;
  int N, Nin, G, W, cap, tidx;
  float t, onset, dt;
  uint32_t* rate;          // [Nin] spike counts of a rate monitor, or nullptr
};

__global__ __launch_bounds__(256) void k_generator(GeneratorArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < a.Nin;
  if (active) {
    for (int g = 0; g < a.G; ++g) {
      const int b = hist_bitpos((int64_t)a.tidx + g, a.W);
      a.hist[(size_t)(b >> 5) * a.Nin + i] &= ~(1u << (b & 31));
    }
  }
  const float base = a.t - a.onset;
  float last = active ? a.last[i] : 0.0f;
  bool fired = false;
  uint32_t nspk = 0u;
  for (int64_t e = a.lo; e < a.hi; ++e) {
    const int id = a.ids[e];
    const float te = a.times[e];
    for (int g = 0; g < a.G; ++g) {
      const float d = (base + (float)g * a.dt) - te;
      bool spk = active && id == i && (double)fabsf(d) < 0.5 * (double)a.dt;
      if (spk) {
        const int b = hist_bitpos((int64_t)a.tidx + g, a.W);
        uint32_t* wp = a.hist + (size_t)(b >> 5) * a.Nin + i;
        spk = !((*wp >> (b & 31)) & 1u);
        *wp |= (1u << (b & 31));
      }
      if (spk) {
        last = a.t + (float)g * a.dt;
        fired = true;
        ++nspk;
      }
      wave_append(spk, (uint32_t)(a.N + i), a.lists + (size_t)g * a.cap, a.counts + g, a.cap,
                  a.spike_stat);
    }
  }
  if (active && fired) a.last[i] = last;
  if (active && a.rate != nullptr && nspk) a.rate[i] += nspk;
}

// ---------------------------------------------------------------------------------
// Delivery (SpikingSynapses.cu:168-202): one wavefront per spike, CSR rows sorted by
// delay, int64 fixed-point atomics into the ring.

struct PropArgs {
  const uint32_t* lists;
  const int* counts;
  const int64_t* rowptr;
  const int32_t* tgt;
  Weights wts;
  const float* scale;       // nullptr: uniform scale0
  const uint16_t* delay;
  const uint8_t* label;     // nullptr: single label
  int64_t* ring;
  unsigned long long* events;
  float scale0, two_k;
  int N, L, G, B, cap, tidx;
};

// A row is shared by kRowWaves wavefronts (wavefront c of a row takes its synapses
// c*64 + lane, stepping by kRowWaves*64), so that a few long rows still keep enough
// loads in flight: each lane's iterations are chains of dependent loads.
constexpr int kRowWaves = 8;

__global__ __launch_bounds__(256) void k_propagate(PropArgs a) {
  const int lane = (int)__lane_id();
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const int nwaves = (int)((gridDim.x * blockDim.x) / kWave);
  const size_t NL = (size_t)a.N * a.L;
  const int base_slot = a.tidx % a.B;
  unsigned long long ev = 0;
  for (int g = 0; g < a.G; ++g) {
    int n = a.counts[g];
    if (n > a.cap) n = a.cap;
    const uint32_t* list = a.lists + (size_t)g * a.cap;
    for (int w = wave; w < n * kRowWaves; w += nwaves) {
      const int j = w / kRowWaves, c = w - j * kRowWaves;
      const uint32_t u = list[j];
      const int64_t s0 = a.rowptr[u], s1 = a.rowptr[u + 1];
      if (lane == 0 && c == 0) ev += (unsigned long long)(s1 - s0);
      for (int64_t s = s0 + c * kWave + lane; s < s1; s += kRowWaves * kWave) {
        const int post = a.tgt[s];
        const int d = a.delay[s];
        const float sc = a.scale ? a.scale[s] : a.scale0;
        const float wv = wload(a.wts, s) * sc;
        const long long q = __float2ll_rn(wv * a.two_k);
        const int slot = (base_slot + d + g) % a.B;
        const int l = a.label ? (int)a.label[s] : 0;
        atomicAdd((unsigned long long*)(a.ring + (size_t)slot * NL + (size_t)post * a.L + l),
                  (unsigned long long)q);
      }
    }
  }
  if (lane == 0 && ev) atomicAdd(a.events, ev);
}

// ---------------------------------------------------------------------------------
// Plasticity rules, one lane per plastic synapse (rule's own order)

struct StdpArgs {
  const int64_t* pos;      // CSR position of each plastic synapse
  const int32_t* pre;      // unified presynaptic id (inputs at N + i)
  const int32_t* tgt;
  const uint16_t* delay;
  Weights wts;
  Traces tr;
  const float* last_n;
  const float* last_in;
  const uint32_t* hist_n;
  const uint32_t* hist_in;
  const float* tab;        // rule constants / tables
  double dterm;
  int64_t P;
  int N, Nin, G, W, tidx, nearest;
  float t, dt;
};

// WeightDependentSTDPPlasticity.cu:67-163. tab = {am, ap, lambda, alpha, wmax,
// Em[0..G], Ep[0..G], iEm[0..G], iEp[0..G]}

// Step of the presynaptic neuron's last spike within the group starting at t; a last
// spike G or more steps ahead is taken as before the group (WeightDependentSTDPPlasticity.cu:98-101).
__device__ __forceinline__ int wd_pre_step(float last, float t, float dt, int G) {
  int pg = (int)roundf((last - t) / dt);
  if (pg >= G) pg *= -1;
  return pg;
}
// Step of the postsynaptic neuron's last spike within the group.  While time advances it
// is below G; the bound keeps the table lookups inside [0, G] for any other input.
__device__ __forceinline__ int wd_post_step(float last, float t, float dt, int G) {
  return min((int)roundf((last - t) / dt), G);
}

// A trace over one group: decay by E[G], then the group's spike at step k (if k >= 0)
// adds A*E[G-k], or with nearest-spike pairing replaces the trace by it.
__device__ __forceinline__ float wd_trace(float x, int k, float A, const float* E, int G, int nearest) {
  x *= E[G];
  if (nearest) {
    if (k >= 0) x = A * E[G - k];
  } else {
    x += (k >= 0) ? A * E[G - k] : 0.0f;
  }
  return x;
}

// New weight of a synapse from the stored weight, the group's updated traces and the
// spike steps of its two neurons (WeightDependentSTDPPlasticity.cu:125-158), clamped at 0.
__device__ __forceinline__ float wd_weight(float w, float xpre, float xpost, int pg, int qg,
                                           const float* tab, int G) {
  const float am = tab[0], ap = tab[1], lam = tab[2], alp = tab[3], wmax = tab[4];
  const float* Em = tab + 5;
  const float* Ep = Em + (G + 1);
  const float* iEm = Ep + (G + 1);
  const float* iEp = iEm + (G + 1);
  float upd = 0.0f;
  if (pg >= 0) {
    float tmp = xpost;
    tmp += (qg > pg) ? -am * Em[G - qg] : 0.0f;
    tmp *= iEm[G - pg];
    upd -= ((lam * alp) * w) * tmp;
  }
  if (qg >= 0) {
    float tmp = xpre;
    tmp += (pg > qg) ? -ap * Ep[G - pg] : 0.0f;
    tmp *= iEp[G - qg];
    upd += (lam * (wmax - w)) * tmp;
  }
  float nw = w + upd;
  if (nw < 0.0f) nw = 0.0f;
  return nw;
}

// Every plastic synapse of the rule, with its own pair of traces.
__device__ __forceinline__ void stdp_wd_one(const StdpArgs& a, int64_t j) {
  const int64_t s = a.pos[j];
  const int u = a.pre[j];
  const int post = a.tgt[s];
  const int G = a.G;
  const float* Em = a.tab + 5;
  const float* Ep = Em + (G + 1);
  const float lpre = (u < a.N) ? a.last_n[u] : a.last_in[u - a.N];
  const int pg = wd_pre_step(lpre, a.t, a.dt, G);
  const int qg = wd_post_step(a.last_n[post], a.t, a.dt, G);
  const float w = wload(a.wts, s);
  const float xpost = wd_trace(tload(a.tr, 1, j), qg, a.tab[0], Em, G, a.nearest);
  const float xpre = wd_trace(tload(a.tr, 0, j), pg, a.tab[1], Ep, G, a.nearest);
  wupdate(a.wts, s, w, wd_weight(w, xpre, xpost, pg, qg, a.tab, G), 0.0f, __builtin_inff());
  tstore(a.tr, 0, j, xpre);
  tstore(a.tr, 1, j, xpost);
}

__global__ __launch_bounds__(256) void k_stdp_wd(StdpArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.P) stdp_wd_one(a, j);
}

// ---------------------------------------------------------------------------------
// The weight-dependent rule driven by spikes.  A synapse's two traces depend only on its
// presynaptic and its postsynaptic neuron (every synapse of a neuron sees the same
// decays and the same spike steps), so they are kept once per neuron; and a weight moves
// only in a group where its presynaptic neuron has pg >= 0 or its postsynaptic neuron has
// qg >= 0 -- otherwise the update adds exactly 0 to a weight that an earlier group has
// already clamped at 0.  Per group: (1) one lane per neuron advances both traces and
// lists the neurons with pg >= 0 (presynaptic side) and qg >= 0 (postsynaptic side);
// (2) one wavefront per listed presynaptic neuron updates all its rule synapses, and
// one per listed postsynaptic neuron the rest of its incoming rule synapses (those whose
// presynaptic neuron is not listed), so each synapse is updated by at most one lane.

struct WdEventArgs {
  Traces trn;              // stored traces per neuron: a over unified presynaptic ids, b over neurons
  float* xa;               // this group's traces before rounding to storage: [N + Nin], [N]
  float* xb;
  Weights wts;
  const float* last_n;
  const float* last_in;
  const float* tab;
  const int64_t* pos;      // rule index -> CSR position
  const int32_t* pre;      // rule index -> unified presynaptic id
  const int32_t* tgt;      // CSR position -> postsynaptic neuron
  const int64_t* pre_ptr;  // [N + Nin + 1]: rule indices of presynaptic neuron u are
  const int32_t* pre_idx;  //   pre_idx[pre_ptr[u] .. pre_ptr[u+1]) (nullptr: the range itself)
  const int64_t* post_ptr; // [N + 1]: CSR positions and presynaptic ids of the rule's
  const uint32_t* post_pos; //   synapses onto neuron v at [post_ptr[v], post_ptr[v+1])
  const int32_t* post_pre;
  uint32_t* act;           // listed neurons: presynaptic ids from 0, postsynaptic from N + Nin
  int* cnt;                // [2][2]: list lengths, double-buffered by `par`
  int par;
  int N, Nin, G, nearest;
  float t, dt;
};

__device__ __forceinline__ void wave_push(bool pred, uint32_t id, uint32_t* __restrict__ list,
                                          int* __restrict__ count, int cap) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader);
  const int p = base + __popcll(m & ((1ull << lane) - 1ull));
  if (pred && p < cap) list[p] = id;
}

// Per-neuron traces of unit u as a presynaptic source (l: its last spike time); lists it
// (when `push`) if it spiked in the group and has rule synapses.  Wavefront-uniform call.
__device__ __forceinline__ void wd_pre_unit(const WdEventArgs& a, int u, bool active, float l, bool push) {
  const int G = a.G;
  const float* Ep = a.tab + 5 + (G + 1);
  bool listed = false;
  if (active) {
    const int pg = wd_pre_step(l, a.t, a.dt, G);
    const float x = wd_trace(tload(a.trn, 0, u), pg, a.tab[1], Ep, G, a.nearest);
    a.xa[u] = x;
    tstore(a.trn, 0, u, x);
    listed = pg >= 0 && a.pre_ptr[u + 1] > a.pre_ptr[u];
  }
  if (push) wave_push(listed, (uint32_t)u, a.act, a.cnt + 2 * a.par, a.N + a.Nin);
}

// The same for neuron v as a postsynaptic target (l: its last spike time).
__device__ __forceinline__ void wd_post_unit(const WdEventArgs& a, int v, bool active, float l) {
  const int G = a.G;
  const float* Em = a.tab + 5;
  bool listed = false;
  if (active) {
    const int qg = wd_post_step(l, a.t, a.dt, G);
    const float x = wd_trace(tload(a.trn, 1, v), qg, a.tab[0], Em, G, a.nearest);
    a.xb[v] = x;
    tstore(a.trn, 1, v, x);
    listed = qg >= 0 && a.post_ptr[v + 1] > a.post_ptr[v];
  }
  wave_push(listed, (uint32_t)v, a.act + a.N + a.Nin, a.cnt + 2 * a.par + 1, a.N);
}

__global__ __launch_bounds__(256) void k_wd_neurons(WdEventArgs a) {
  const int U = a.N + a.Nin;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool pre = i < U;
  wd_pre_unit(a, i, pre, pre ? (i < a.N ? a.last_n[i] : a.last_in[i - a.N]) : 0.0f, true);
  const int v = i - U;
  const bool post = v >= 0 && v < a.N;
  wd_post_unit(a, v, post, post ? a.last_n[v] : 0.0f);
}

__global__ __launch_bounds__(256) void k_wd_events(WdEventArgs a) {
  const int lane = (int)__lane_id();
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const int nwaves = (int)((gridDim.x * blockDim.x) / kWave);
  const int U = a.N + a.Nin;
  const int G = a.G;
  const int npre = min(a.cnt[2 * a.par], U), npost = min(a.cnt[2 * a.par + 1], a.N);
  // the other buffer is next group's: clear it (nothing reads it in this group)
  if (blockIdx.x == 0 && threadIdx.x < 2) a.cnt[2 * (a.par ^ 1) + threadIdx.x] = 0;
  for (int w = wave; w < (npre + npost) * kRowWaves; w += nwaves) {
    const int k = w / kRowWaves, c = w - k * kRowWaves;
    if (k < npre) {
      const int u = (int)a.act[k];
      const float l = u < a.N ? a.last_n[u] : a.last_in[u - a.N];
      const int pg = wd_pre_step(l, a.t, a.dt, G);
      const float xpre = a.xa[u];
      const int64_t e1 = a.pre_ptr[u + 1];
      for (int64_t e = a.pre_ptr[u] + c * kWave + lane; e < e1; e += kRowWaves * kWave) {
        const int64_t j = a.pre_idx ? (int64_t)a.pre_idx[e] : e;
        const int64_t s = a.pos[j];
        const int post = a.tgt[s];
        const int qg = wd_post_step(a.last_n[post], a.t, a.dt, G);
        const float w = wload(a.wts, s);
        wupdate(a.wts, s, w, wd_weight(w, xpre, a.xb[post], pg, qg, a.tab, G), 0.0f, __builtin_inff());
      }
    } else {
      const int v = (int)a.act[U + (k - npre)];
      const int qg = wd_post_step(a.last_n[v], a.t, a.dt, G);
      const float xpost = a.xb[v];
      const int64_t e1 = a.post_ptr[v + 1];
      for (int64_t e = a.post_ptr[v] + c * kWave + lane; e < e1; e += kRowWaves * kWave) {
        const int u = a.post_pre[e];
        const float l = u < a.N ? a.last_n[u] : a.last_in[u - a.N];
        const int pg = wd_pre_step(l, a.t, a.dt, G);
        if (pg >= 0) continue;              // updated from the presynaptic side
        const int64_t s = (int64_t)a.post_pos[e];
        const float w = wload(a.wts, s);
        wupdate(a.wts, s, w, wd_weight(w, a.xa[u], xpost, pg, qg, a.tab, G), 0.0f, __builtin_inff());
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// One group (G = 1) of the per-phase path in two launches, for networks with Poisson (or
// no) inputs and at most one rule, a weight-dependent one in its spike-driven form:
//   k_units   LIF neurons and Poisson inputs (one lane each), the rule's per-neuron traces
//             and its list of spiking targets; clears the spike counts of the next group's
//             list slot (so no separate group-start launch is needed);
//   k_deliver propagation of this group's spikes and both sides of the rule:
//             - source side: each spiking unit's row is walked once; a synapse the rule
//               lists (bit of `mask`) is updated first and delivered with its new weight
//               (the reference applies plasticity before propagation, SpikingModel.cpp:228-233);
//             - target side: the rule synapses onto each listed spiking neuron whose source
//               did not spike in the group (those the source side does not reach).
//             The two sides touch disjoint synapses, and propagation reads only the rows of
//             this group's spikers, so all of it runs in one launch.
// Results equal those of the separate kernels (k_begin_group, k_lif, k_poisson,
// k_wd_neurons, k_wd_events, k_propagate) bit for bit.

struct UnitsArgs {
  LifArgs lif;
  PoissonArgs in;
  WdEventArgs wd;
  int* next_counts;     // [G] counts of the next group's list slot
  float* slot_time;     // this group's slot: its start time and step (monitors read them)
  int* slot_tidx;
  int slot, nb_neurons; // blocks [0, nb_neurons) run neurons, the rest inputs
};

template <int KIND, bool WD>
__global__ __launch_bounds__(256) void k_units(UnitsArgs a) {
  if (blockIdx.x == 0) {
    if ((int)threadIdx.x < a.lif.G) a.next_counts[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      a.slot_time[a.slot] = a.lif.t;
      a.slot_tidx[a.slot] = a.lif.tidx;
    }
  }
  if ((int)blockIdx.x < a.nb_neurons) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const float last = lif_unit<KIND>(a.lif, i);
    if (WD) {
      const bool on = i < a.lif.N;
      wd_pre_unit(a.wd, i, on, last, false);
      wd_post_unit(a.wd, i, on, last);
    }
  } else {
    const int i = (int)((blockIdx.x - a.nb_neurons) * blockDim.x + threadIdx.x);
    const float last = poisson_unit(a.in, i);
    if (WD) wd_pre_unit(a.wd, a.in.N + i, i < a.in.Nin, last, false);
  }
}

struct DeliverArgs {
  PropArgs p;
  WdEventArgs wd;
  const uint32_t* mask;   // bit s: CSR position s belongs to the rule
  int row_waves;          // wavefronts per source row (ceil(longest row / 64))
  int post_waves;         // wavefronts per target list
};

template <bool WD>
__global__ __launch_bounds__(256) void k_deliver(DeliverArgs a) {
  const int lane = (int)__lane_id();
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const int nwaves = (int)((gridDim.x * blockDim.x) / kWave);
  const PropArgs& p = a.p;
  const WdEventArgs& r = a.wd;
  const size_t NL = (size_t)p.N * p.L;
  const int slot0 = p.tidx % p.B;
  const int nspk = min(p.counts[0], p.cap);
  int npost = 0;
  if (WD) {
    npost = min(r.cnt[2 * r.par + 1], r.N);
    // the other buffer is next group's: clear it (nothing reads it in this group)
    if (blockIdx.x == 0 && threadIdx.x < 2) r.cnt[2 * (r.par ^ 1) + threadIdx.x] = 0;
  }
  const int tasks_src = nspk * a.row_waves;
  const int tasks = tasks_src + npost * a.post_waves;
  unsigned long long ev = 0;
  for (int w = wave; w < tasks; w += nwaves) {
    if (w < tasks_src) {
      const int j = w / a.row_waves, c = w - j * a.row_waves;
      const uint32_t u = p.lists[j];
      const int64_t s0 = p.rowptr[u], s1 = p.rowptr[u + 1];
      if (lane == 0 && c == 0) ev += (unsigned long long)(s1 - s0);
      float xpre = 0.0f;
      int pg = 0;
      if (WD) {
        xpre = r.xa[u];
        pg = wd_pre_step((int)u < r.N ? r.last_n[u] : r.last_in[u - r.N], r.t, r.dt, r.G);
      }
      for (int64_t s = s0 + c * kWave + lane; s < s1; s += (int64_t)a.row_waves * kWave) {
        const int post = p.tgt[s];
        const int d = p.delay[s];
        const float sc = p.scale ? p.scale[s] : p.scale0;
        uint16_t c16;
        float wt = wload_c(p.wts, s, c16);
        if (WD && ((a.mask[s >> 5] >> (s & 31)) & 1u)) {
          const float lpost = r.last_n[post], xpost = r.xb[post];
          const int qg = wd_post_step(lpost, r.t, r.dt, r.G);
          wt = wupdate_c(p.wts, s, wt, c16, wd_weight(wt, xpre, xpost, pg, qg, r.tab, r.G), 0.0f,
                         __builtin_inff());
        }
        const long long q = __float2ll_rn(wt * sc * p.two_k);
        const int slot = (slot0 + d) % p.B;
        const int l = p.label ? (int)p.label[s] : 0;
        atomicAdd((unsigned long long*)(p.ring + (size_t)slot * NL + (size_t)post * p.L + l),
                  (unsigned long long)q);
      }
    } else if (WD) {
      const int k = (w - tasks_src) / a.post_waves, c = (w - tasks_src) - k * a.post_waves;
      const int v = (int)r.act[r.N + r.Nin + k];
      const int qg = wd_post_step(r.last_n[v], r.t, r.dt, r.G);
      const float xpost = r.xb[v];
      const int64_t e1 = r.post_ptr[v + 1];
      for (int64_t e = r.post_ptr[v] + c * kWave + lane; e < e1; e += (int64_t)a.post_waves * kWave) {
        // every load of the update issued before the test (two levels of dependent loads)
        const int u = r.post_pre[e];
        const int64_t s = (int64_t)r.post_pos[e];
        const float l = u < r.N ? r.last_n[u] : r.last_in[u - r.N];
        const float xu = r.xa[u];
        uint16_t c16;
        const float w0 = wload_c(r.wts, s, c16);
        const int pgu = wd_pre_step(l, r.t, r.dt, r.G);
        if (pgu >= 0) continue;              // updated from the source side
        wupdate_c(r.wts, s, w0, c16, wd_weight(w0, xu, xpost, pgu, qg, r.tab, r.G), 0.0f, __builtin_inff());
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off);
  if (lane == 0 && ev) atomicAdd(p.events, ev);
}

// Marks the rule's synapses in a bitmask over CSR positions (k_deliver's source side).
__global__ __launch_bounds__(256) void k_mark_rule(const int64_t* __restrict__ pos, int64_t P,
                                                   uint32_t* __restrict__ mask) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < P; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = pos[j];
    atomicOr(mask + (s >> 5), 1u << (s & 31));
  }
}

// Traces between the two forms: per synapse (the per-synapse kernels) and per neuron.
// Every synapse of a neuron holds that neuron's trace, so either copy is exact.
struct TraceCopyArgs {
  Traces syn;
  Traces neu;
  const int64_t* pos;
  const int32_t* pre;
  const int32_t* tgt;
  int64_t P;
  int to_neurons;
};

__global__ __launch_bounds__(256) void k_trace_copy(TraceCopyArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.P) return;
  const int u = a.pre[j];
  const int v = a.tgt[a.pos[j]];
  if (a.to_neurons) {   // all lanes that write one neuron's entry write the same value
    tstore(a.neu, 0, u, tload(a.syn, 0, j));
    tstore(a.neu, 1, v, tload(a.syn, 1, j));
  } else {
    tstore(a.syn, 0, j, tload(a.neu, 0, u));
    tstore(a.syn, 1, j, tload(a.neu, 1, v));
  }
}

__device__ __forceinline__ bool pre_bit(const StdpArgs& a, int u, int64_t step) {
  return (u < a.N) ? hist_get(a.hist_n, a.N, a.W, u, step)
                   : hist_get(a.hist_in, a.Nin, a.W, u - a.N, step);
}

// InhibitorySTDPPlasticity.cu:72-155. tab = {decay, lr}; dterm = lr*(2*rate*tau) (double)
__device__ __forceinline__ void stdp_vogels_one(const StdpArgs& a, int64_t j) {
  const int64_t s = a.pos[j];
  const int u = a.pre[j];
  const int post = a.tgt[s];
  const int d = a.delay[s];
  const float decay = a.tab[0], lr = a.tab[1];
  float xpre = tload(a.tr, 0, j), xpost = tload(a.tr, 1, j);
  const float w0 = wload(a.wts, s);
  float w = w0;
  for (int g = 0; g < a.G; ++g) {
    xpre *= decay;
    xpost *= decay;
    const int64_t step = (int64_t)a.tidx + g;
    const bool pr = pre_bit(a, u, step - d);
    const bool po = hist_get(a.hist_n, a.N, a.W, post, step);
    if (pr) xpre += 1.0f;
    if (po) xpost += 1.0f;
    float upd = 0.0f;
    if (pr) {
      upd += lr * xpost;
      upd = (float)((double)upd + (-a.dterm));
    }
    if (po) upd += lr * xpre;
    w += upd;
    if (w < 0.0f) w = 0.0f;
  }
  wupdate(a.wts, s, w0, w, 0.0f, __builtin_inff());
  tstore(a.tr, 0, j, xpre);
  tstore(a.tr, 1, j, xpost);
}

__global__ __launch_bounds__(256) void k_stdp_vogels(StdpArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.P) stdp_vogels_one(a, j);
}

// EvansSTDPPlasticity.cu:59-143. tab = {dC, dD, aC, aD}
__device__ __forceinline__ void stdp_evans_one(const StdpArgs& a, int64_t j) {
  const int64_t s = a.pos[j];
  const int u = a.pre[j];
  const int post = a.tgt[s];
  const int d = a.delay[s];
  const float dC = a.tab[0], dD = a.tab[1], aC = a.tab[2], aD = a.tab[3];
  float C = tload(a.tr, 0, j), D = tload(a.tr, 1, j);
  const float w0 = wload(a.wts, s);
  float w = w0;
  for (int g = 0; g < a.G; ++g) {
    C = dC * C;
    D = dD * D;
    const int64_t step = (int64_t)a.tidx + g;
    const bool pr = pre_bit(a, u, step - d);
    const bool po = hist_get(a.hist_n, a.N, a.W, post, step);
    if (pr) C += aC * (1.0f - C);
    if (po) D += aD * (1.0f - D);
    float upd = 0.0f;
    const float wo = w;
    if (pr) upd -= wo * D;
    if (po) upd += (1.0f - wo) * C;
    w = wo + upd;
    if (w < 0.0f) w = 0.0f;
    if (w > 1.0f) w = 1.0f;
  }
  wupdate(a.wts, s, w0, w, 0.0f, 1.0f);
  tstore(a.tr, 0, j, C);
  tstore(a.tr, 1, j, D);
}

__global__ __launch_bounds__(256) void k_stdp_evans(StdpArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.P) stdp_evans_one(a, j);
}

// CustomSTDPPlasticity.cu:67-155. tab = {post decay expf(-dt/tau_minus), pre decay
// expf(-dt/tau_plus), a_plus, a_minus, a_star, w_max, learning_rate}; nearest: a spike
// sets its trace to the amplitude instead of adding it.  Weights clamped at 0 only.
__device__ __forceinline__ void stdp_custom_one(const StdpArgs& a, int64_t j) {
  const int64_t s = a.pos[j];
  const int u = a.pre[j];
  const int post = a.tgt[s];
  const int d = a.delay[s];
  const float dpost = a.tab[0], dpre = a.tab[1], ap = a.tab[2], am = a.tab[3], ast = a.tab[4];
  const float wmax = a.tab[5], lr = a.tab[6];
  float xpre = tload(a.tr, 0, j), xpost = tload(a.tr, 1, j);
  const float w0 = wload(a.wts, s);
  float w = w0;
  for (int g = 0; g < a.G; ++g) {
    xpost *= dpost;
    xpre *= dpre;
    const int64_t step = (int64_t)a.tidx + g;
    const bool pr = pre_bit(a, u, step - d);
    const bool po = hist_get(a.hist_n, a.N, a.W, post, step);
    if (pr) xpre = a.nearest ? ap : xpre + ap;
    if (po) xpost = a.nearest ? am : xpost + am;
    float upd = 0.0f;
    const float wo = w;
    if (pr) upd -= ((lr * (wo / wmax)) * xpost + lr * ast);
    if (po) upd += lr * xpre;
    w = wo + upd;
    if (w < 0.0f) w = 0.0f;
  }
  wupdate(a.wts, s, w0, w, 0.0f, __builtin_inff());
  tstore(a.tr, 0, j, xpre);
  tstore(a.tr, 1, j, xpost);
}

__global__ __launch_bounds__(256) void k_stdp_custom(StdpArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.P) stdp_custom_one(a, j);
}

// WeightNormSTDPPlasticity.cu:68-172.  One lane per postsynaptic neuron of the rule: the
// change of its afferent squared norm, u = sum of d*d + 2*w0*d over its synapses with
// d = w - w0 != 0 (the reference's float atomicAdd, here summed in the caller's order of
// the synapses), then w /= sqrtf(sum_sq + u) / sqrtf(sum_sq) when that is not 1.
struct WeightNormArgs {
  Weights wts;
  const int64_t* pos;       // rule index -> CSR position
  const float* w0;          // rule index -> initial weight
  const float* sum_sq;      // [N]
  const uint8_t* in_set;    // [N]
  const int64_t* ptr;       // [N + 1] rule indices of neuron v: idx[ptr[v] .. ptr[v+1]),
  const int32_t* idx;       //   in the caller's order
  int N;
};

__global__ __launch_bounds__(256) void k_weightnorm(WeightNormArgs a) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.N || !a.in_set[v]) return;
  const int64_t e0 = a.ptr[v], e1 = a.ptr[v + 1];
  float u = 0.0f;
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t j = a.idx[e];
    const float w0 = a.w0[j];
    const float d = wload(a.wts, a.pos[j]) - w0;
    if (d != 0.0f) u += d * d + (2.0f * w0) * d;
  }
  const float sq = a.sum_sq[v];
  const float div = sqrtf(sq + u) / sqrtf(sq);
  if (div == 1.0f) return;
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t s = a.pos[a.idx[e]];
    const float w = wload(a.wts, s);
    wupdate(a.wts, s, w, w / div, -__builtin_inff(), __builtin_inff());
  }
}

// ---------------------------------------------------------------------------------
// Monitor collection (SpikingActivityMonitor.cu:69-98): copy the spikes of one
// population from list slots [slot0, slot0+nslots) into a bounded store.

struct CollectArgs {
  const uint32_t* lists;      // [slots][G][cap]
  const int* counts;          // [slots][G]
  const float* slot_time;
  int32_t* ids;
  float* times;
  unsigned long long* count;
  int64_t store_cap;
  int slot0, nslots, G, Gs, cap, lo, hi;   // G active lists of Gs per slot; population [lo, hi)
  float dt;
};

__global__ __launch_bounds__(256) void k_collect(CollectArgs a) {
  // blockIdx.y = slot offset * G + g; blockIdx.x strides over the list entries
  const int sg = blockIdx.y;
  const int s = a.slot0 + sg / a.G;
  const int g = sg % a.G;
  int n = a.counts[(size_t)s * a.Gs + g];
  if (n > a.cap) n = a.cap;
  const uint32_t* list = a.lists + ((size_t)s * a.Gs + g) * a.cap;
  const float tg = a.slot_time[s] + (float)g * a.dt;
  const int nthreads = gridDim.x * blockDim.x;
  const int n_round = ((n + kWave - 1) / kWave) * kWave;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n_round; j += nthreads) {
    const bool in_range = j < n;
    const uint32_t u = in_range ? list[j] : 0u;
    const bool pred = in_range && (int)u >= a.lo && (int)u < a.hi;
    const uint64_t m = __ballot(pred);
    if (m == 0) continue;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(a.count, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    if (pred) {
      const int64_t pos = (int64_t)base + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < a.store_cap) {
        a.ids[pos] = (int32_t)u - a.lo;
        a.times[pos] = tg;
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// FP64 reduction of the window statistics (one workgroup; fixed summation order, so the
// result is reproducible):
//   rate      = spikes / (N * nsteps * dt)
//   isi_cv    = mean over neurons with >= 2 intervals of sqrt(var)/mean of their intervals
//               (population variance s2/n - (s1/n)^2)
//   synchrony = variance / mean of the population spike count per bin (Fano factor)
struct StatsOut {
  double rate_hz, isi_cv, synchrony, mean_bin_count;
  long long spikes, cv_neurons, bins, steps;
};

constexpr int kStatsBlock = 1024;

__device__ double block_sum_f64(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kStatsBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kStatsBlock) void k_stats_reduce(StatsArgs st, int N, int64_t steps, float dt,
                                                              StatsOut* out) {
  __shared__ double sh[kStatsBlock];
  double spikes = 0.0, cv = 0.0, ncv = 0.0;
  for (int i = threadIdx.x; i < N; i += kStatsBlock) {
    spikes += (double)st.count[i];
    const unsigned int n = st.nisi[i];
    if (n >= 2u) {
      const double m = st.s1[i] / (double)n;
      double var = st.s2[i] / (double)n - m * m;
      if (var < 0.0) var = 0.0;
      cv += sqrt(var) / m;
      ncv += 1.0;
    }
  }
  const int64_t nb = steps > 0 ? (steps + st.bin - 1) / st.bin : 0;
  double hs = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kStatsBlock) hs += (double)st.hist[b];
  const double S = block_sum_f64(spikes, sh);
  const double CV = block_sum_f64(cv, sh);
  const double NCV = block_sum_f64(ncv, sh);
  const double HS = block_sum_f64(hs, sh);
  const double mean = nb > 0 ? HS / (double)nb : 0.0;
  double hv = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kStatsBlock) {
    const double d = (double)st.hist[b] - mean;
    hv += d * d;
  }
  const double HV = block_sum_f64(hv, sh);
  if (threadIdx.x == 0) {
    StatsOut o;
    o.spikes = (long long)S;
    o.cv_neurons = (long long)NCV;
    o.bins = (long long)nb;
    o.steps = (long long)steps;
    o.rate_hz = (N > 0 && steps > 0) ? S / ((double)N * (double)steps * (double)dt) : 0.0;
    o.isi_cv = NCV > 0.0 ? CV / NCV : 0.0;
    o.mean_bin_count = mean;
    o.synchrony = (nb > 0 && mean > 0.0) ? (HV / (double)nb) / mean : 0.0;
    *out = o;
  }
}

}  // namespace spk
