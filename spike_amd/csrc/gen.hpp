// gen.hpp -- Brunel-type connectivity generated on the device.
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

}  // namespace spk
