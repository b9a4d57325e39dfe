// gen.hpp -- connectivity generated on the device: the Brunel fixed in-degree rule (below)
// and the fixed out-degree rule of large networks (further down).
//
// Fixed in-degree (Brunel 2000; the rule of Examples/UtilityFunctions.hpp:51-82, sources
// drawn with replacement): every neuron i receives ce synapses from the excitatory
// population [0, ne) and ci from the inhibitory population [ne, ne+ni).  Slot j of neuron
// i draws u = draw j of (i, stream 3) (output j mod 4 of Philox(seed, ctr = (j/4, i, 3))) and
// takes
//   pre = floor(u * ne / 2^32)            for j <  ce
//   pre = ne + floor(u * ni / 2^32)       for j >= ce
// The synapses are counting-sorted on the device by key (presynaptic unit, partition of
// the postsynaptic neuron): the exclusive scan of the key counts is exactly the
// persistent kernel's segment table seg[u][P+1], and the scatter writes the synapse
// records in place.  Entries of one key land in atomic order; deliveries are exact
// fixed-point sums, so that order does not change any result.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"   // enc16
#include "philox.hpp"

namespace spk {

constexpr uint32_t kStreamConnect = 3;

struct BrunelGenArgs {
  uint64_t seed;
  int ne, ni, ce, ci, P;
  const int* owner;         // [N] partition of each neuron
  const int* nlo;           // [P+1]
  unsigned int* count;      // [N*P] per key
  const unsigned int* off;  // [N*P] exclusive scan of count
  unsigned int* cursor;     // [N*P]
  uint2* syn;               // records {local target | delay << 20, weight bits}
  uint16_t* syn16;          // the same records' local targets (compact delivery)
  uint32_t delay_bits;      // delay << 20
  float we, wi;
};

__device__ __forceinline__ int brunel_pre(const BrunelGenArgs& a, int i, int j) {
  const uint32_t u = draw_u32(a.seed, (int64_t)j, (uint32_t)i, kStreamConnect);
  if (j < a.ce) return (int)(((uint64_t)u * (uint64_t)a.ne) >> 32);
  return a.ne + (int)(((uint64_t)u * (uint64_t)a.ni) >> 32);
}

__global__ __launch_bounds__(256) void k_gen_count(BrunelGenArgs a, int64_t total) {
  const int C = a.ce + a.ci;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / C), j = (int)(k - (int64_t)i * C);
    const int pre = brunel_pre(a, i, j);
    atomicAdd(a.count + (size_t)pre * a.P + a.owner[i], 1u);
  }
}

__global__ __launch_bounds__(256) void k_gen_scatter(BrunelGenArgs a, int64_t total) {
  const int C = a.ce + a.ci;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / C), j = (int)(k - (int64_t)i * C);
    const int pre = brunel_pre(a, i, j);
    const int q = a.owner[i];
    const size_t key = (size_t)pre * a.P + q;
    const unsigned int pos = a.off[key] + atomicAdd(a.cursor + key, 1u);
    const float w = j < a.ce ? a.we : a.wi;
    a.syn[pos] = make_uint2((uint32_t)(i - a.nlo[q]) | a.delay_bits, __float_as_uint(w));
    a.syn16[pos] = (uint16_t)(i - a.nlo[q]);
  }
}

// ---------------------------------------------------------------------------------------
// Fixed out-degree connectivity generated on the device (the large-network builder).
//
// Units u in [0, N + Nin) are the neurons, then the Poisson inputs.  Unit u has K_u
// outgoing synapses (K for a neuron, Kin for an input), drawn with replacement.  Synapse j
// of unit u takes one Philox block r = philox(ctr = {j, 0, u, 5}, key = seed):
//   post   = floor(r.x * N / 2^32)
//   delay  = 1 + floor(r.y * dmax / 2^32)                  (steps)
//   weight = (r.z >> 8) * 2^-24 for excitatory sources (u < ne), U[0, 1) in float32;
//            w_inh for inhibitory neurons, w_in for inputs
//   scale  = scale_e / scale_i / scale_in by the same source classes
// and sits at CSR position rowptr[u] + j, rowptr[u] = u K (u <= N), N K + (u - N) Kin
// after.  The caller order of the synapses is that CSR order, so the row of a source
// needs no sort.  With `plastic`, every excitatory-to-excitatory synapse (u < ne,
// post < ne) belongs to one weight-dependent STDP rule, listed in CSR order.
//
// Two passes over the same draws: k_od_count gathers the delay range, the fixed-point
// bound of each neuron's per-step input and the rule's synapse counts per source and per
// target; after the host scans those counts, k_od_fill writes the synapse arrays and the
// rule's index tables (by source in CSR order; by target in atomic order -- the kernels
// that read that table update each synapse from one lane, so its order does not matter).
constexpr uint32_t kStreamOutDegree = 5;

struct OutDegreeArgs {
  uint64_t seed;
  int N, Nin, ne, K, Kin, dmax;
  float scale_e, scale_i, scale_in, w_inh, w_in;
  int plastic;          // E->E synapses carry the weight-dependent rule
  float cap_plastic;    // |w_max| of that rule (a bound of its weights)
  int wfmt;             // 16-bit weight words to write (0: none)
  // k_od_count
  double* bound;            // [N] sum over afferents of max|w| |scale|
  unsigned int* post_cnt;   // [N] rule synapses onto each neuron
  unsigned int* row_cnt;    // [U] rule synapses of each source
  int* drange;              // [2] min, max delay
  // k_od_fill
  const int64_t* pre_ptr;   // [U+1] exclusive scan of row_cnt
  const int64_t* post_ptr;  // [N+1] exclusive scan of post_cnt
  unsigned int* cursor;     // [N] zeroed
  int32_t* tgt;
  uint2* syn;               // {0, weight bits}
  uint16_t* delay;
  float* scale;
  uint32_t* wc;             // 16-bit weight words (compensation zero)
  int64_t* rpos;            // rule: CSR position, in CSR order
  int32_t* rpre;            //       its source unit
  uint32_t* epos;           // rule by target: CSR position
  int32_t* epre;            //                 source unit
};

struct OdSynapse {
  int post, delay;
  float w, scale;
  bool plastic;
};

__host__ __device__ __forceinline__ int64_t od_row(const OutDegreeArgs& a, int u) {
  return u <= a.N ? (int64_t)u * a.K : (int64_t)a.N * a.K + (int64_t)(u - a.N) * a.Kin;
}

__host__ __device__ __forceinline__ OdSynapse od_synapse(const OutDegreeArgs& a, int u, int j) {
  const u32x4 r = draw_block(a.seed, (int64_t)j, (uint32_t)u, kStreamOutDegree);
  OdSynapse s;
  s.post = (int)(((uint64_t)r.x * (uint64_t)a.N) >> 32);
  s.delay = 1 + (int)(((uint64_t)r.y * (uint64_t)a.dmax) >> 32);
  if (u < a.ne) {
    s.w = (float)(r.z >> 8) * 0x1p-24f;
    s.scale = a.scale_e;
  } else if (u < a.N) {
    s.w = a.w_inh;
    s.scale = a.scale_i;
  } else {
    s.w = a.w_in;
    s.scale = a.scale_in;
  }
  s.plastic = a.plastic && u < a.ne && s.post < a.ne;
  return s;
}

// one workgroup of 256 per source row (grid-stride over rows)
__global__ __launch_bounds__(256) void k_od_count(OutDegreeArgs a) {
  __shared__ unsigned int swave[4];
  const int U = a.N + a.Nin;
  int lo = 1 << 30, hi = 0;
  for (int u = blockIdx.x; u < U; u += gridDim.x) {
    const int len = u < a.N ? a.K : a.Kin;
    unsigned int pc = 0;
    for (int j = threadIdx.x; j < len; j += blockDim.x) {
      const OdSynapse s = od_synapse(a, u, j);
      lo = min(lo, s.delay);
      hi = max(hi, s.delay);
      const float cap = s.plastic ? fmaxf(fabsf(s.w), a.cap_plastic) : fabsf(s.w);
      unsafeAtomicAdd(a.bound + s.post, (double)cap * fabs((double)s.scale));
      if (s.plastic) {
        ++pc;
        atomicAdd(a.post_cnt + s.post, 1u);
      }
    }
    for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o);
    if ((threadIdx.x & 63) == 0) swave[threadIdx.x >> 6] = pc;
    __syncthreads();
    if (threadIdx.x == 0) a.row_cnt[u] = swave[0] + swave[1] + swave[2] + swave[3];
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((threadIdx.x & 63) == 0 && hi > 0) {
    atomicMin(a.drange, lo);
    atomicMax(a.drange + 1, hi);
  }
}

__global__ __launch_bounds__(256) void k_od_fill(OutDegreeArgs a) {
  __shared__ unsigned int swave[4];
  const int U = a.N + a.Nin;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int u = blockIdx.x; u < U; u += gridDim.x) {
    const int len = u < a.N ? a.K : a.Kin;
    const int64_t k0 = od_row(a, u);
    int64_t rbase = a.plastic ? a.pre_ptr[u] : 0;
    for (int j0 = 0; j0 < len; j0 += 256) {
      const int j = j0 + (int)threadIdx.x;
      OdSynapse s{};
      if (j < len) {
        s = od_synapse(a, u, j);
        const int64_t k = k0 + j;
        a.tgt[k] = s.post;
        a.syn[k] = make_uint2(0u, __float_as_uint(s.w));
        a.delay[k] = (uint16_t)s.delay;
        a.scale[k] = s.scale;
        if (a.wfmt) a.wc[k] = enc16(s.w, a.wfmt);
      }
      // rule index: the row's rule synapses before this one, in j order
      const uint64_t m = __ballot(s.plastic);
      const unsigned int before_w = (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) swave[wv] = (unsigned int)__popcll(m);
      __syncthreads();
      unsigned int before = before_w, tot = 0;
      for (int w = 0; w < 4; ++w) {
        if (w < wv) before += swave[w];
        tot += swave[w];
      }
      __syncthreads();
      if (s.plastic) {
        const int64_t k = k0 + j;
        const int64_t r = rbase + before;
        a.rpos[r] = k;
        a.rpre[r] = u;
        const int64_t e = a.post_ptr[s.post] + atomicAdd(a.cursor + s.post, 1u);
        a.epos[e] = (uint32_t)k;
        a.epre[e] = u;
      }
      rbase += tot;
    }
  }
}

}  // namespace spk
