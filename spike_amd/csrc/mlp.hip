// mlp.hip -- the sweep surrogate: a 5 -> 256 -> 256 -> 256 -> 256 -> 3 SiLU MLP trained and
// evaluated with hand-written gfx950 kernels.  Every matrix product runs on the BF16 MFMA
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the optimiser keeps FP32 master weights.
//
// One step of training = four launches on the surrogate's stream:
//   k_mlp_forward   rows in tiles of 128 per workgroup; the tile's activations stay in LDS
//                   from layer to layer, the weights come from L2; saves the activations
//                   and SiLU slopes the backward pass needs.
//   k_mlp_backward  the same tiles backwards: output gradient -> dZ4 -> ... -> dZ1, each a
//                   product with the transposed weights (kept as a BF16 copy).
//   k_mlp_wgrad     dW_l = dZ_l^T A_(l-1) and db_l, reduced over row chunks (split K);
//                   partial sums go to a [chunk][parameter] array.
//   k_mlp_adam      sums the chunks in a fixed order (deterministic), Adam on the masters,
//                   writes the BF16 copies (W, W^T, padded first and last layer).
// Every ensemble member is grid dimension y (z for the weight gradient), so one launch
// trains all members.
//
// MFMA 16x16x32 lane layout (64 lanes, l = lane): A operand 8 bf16 per lane with row
// l % 16 and reduction index 8 (l / 16) + j; B operand 8 bf16 with column l % 16 and
// reduction index 8 (l / 16) + j; result 4 floats with row 4 (l / 16) + r, column l % 16.
// A and B use the same reduction-index assignment, so a product does not depend on how
// the hardware orders the 32 reduction indices within one instruction.
//
// Tensors that a later kernel reads as MFMA operands are stored in "fragment order" FL:
// for a matrix X[m][c] (m = batch row, c = column; M rows in all)
//   FL(m, c) = ((c / 16) (M / 32) + m / 32) * 512 + ((c % 16) + 16 ((m % 32) / 8)) * 8 + m % 8
// so that one 16-byte load per lane fetches an operand fragment with the reduction running
// over rows -- what the weight gradient needs for both dZ and A.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spike_hip.h"
#include "../../include/spike_mlp.h"
#include "philox.hpp"

namespace spk {
void set_last_error(const std::string& msg);
}

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = SPK_MLP_HIDDEN;
constexpr int DIN = SPK_MLP_INPUTS;
constexpr int DOUT = SPK_MLP_OUTPUTS;
constexpr int K1 = 32;     // first-layer reduction: features hi (0..4), lo (5..9), zeros
constexpr int NO = 16;     // output layer padded to one 16-column block
constexpr int MT = SPK_MLP_ROW_TILE;  // rows per workgroup
constexpr int NT = 256;    // threads per workgroup (4 waves; wave w owns columns 64w..64w+63)
constexpr int LS = H + 8;  // LDS row stride in bf16 (16-byte pad: conflict-free operand reads)
constexpr int LDS_BYTES = MT * LS * 2 + 64;
constexpr uint32_t kStreamBatch = 4;

// FP32 master layout (per member)
constexpr int oW1 = 0, ob1 = oW1 + H * DIN;
constexpr int oW2 = ob1 + H, ob2 = oW2 + H * H;
constexpr int oW3 = ob2 + H, ob3 = oW3 + H * H;
constexpr int oW4 = ob3 + H, ob4 = oW4 + H * H;
constexpr int oW5 = ob4 + H, ob5 = oW5 + DOUT * H;
constexpr int P = ob5 + DOUT;
static_assert(P == SPK_MLP_PARAMS, "parameter count");

// BF16 working copies (per member)
constexpr int bW1 = 0;               // [256][32]: W1 in columns 0..4 and again in 5..9
constexpr int bW = bW1 + H * K1;     // layers 2..4, [256][256] each, W[out][in]
constexpr int bWT = bW + 3 * H * H;  // layers 2..4 transposed, WT[in][out]
constexpr int bW5 = bWT + 3 * H * H; // [16][256]: rows 0..2 = W5
constexpr int bW5T = bW5 + NO * H;   // [256][32]: W5T[k][o] = W5[o][k], o < 3
constexpr int BF = bW5T + H * 32;

constexpr int oW[5] = {oW1, oW2, oW3, oW4, oW5};
constexpr int ob[5] = {ob1, ob2, ob3, ob4, ob5};

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Four floats to four bf16 (round to nearest even; v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ u16x4 pack4(float a, float b, float c, float d) {
  const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
  const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{c, d}, bf16x2));
  return u16x4{(unsigned short)lo, (unsigned short)(lo >> 16), (unsigned short)hi, (unsigned short)(hi >> 16)};
}

// logistic function with the hardware exp2 and reciprocal (about 1 ulp each)
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ size_t fl_index(int M, int m, int c) {
  return ((size_t)(c >> 4) * (size_t)(M >> 5) + (size_t)(m >> 5)) * 512 +
         (size_t)(((c & 15) + 16 * ((m & 31) >> 3)) * 8 + (m & 7));
}

// Columns [0, C) of the LDS tile (rows 0..MT-1 = batch rows m0..) into fragment order.
__device__ __forceinline__ void store_fl(const uint16_t* lds, int C, uint16_t* dst, int M, int m0,
                                         int w, int lane) {
  const int nmb = MT / 32;
  const int nblk = (C / 16) * nmb;
  for (int b = w; b < nblk; b += 4) {
    const int cb = b / nmb, mb = b - cb * nmb;
    const int c = cb * 16 + (lane & 15), r = mb * 32 + (lane >> 4) * 8;
    u16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = lds[(r + j) * LS + c];
    *(u16x8*)(dst + fl_index(M, m0 + r, c)) = v;
  }
}

// Transposed tile product: acc[nb][mb] += W[64 w + 16 nb.., k] . lds[16 mb.., k] over KS steps
// of 32 (K = 32 KS), i.e. Z^T = W A^T.  Result element r of acc[nb][mb] belongs to batch
// row m = 16 mb + lane % 16 and output n = 64 w + 16 nb + 4 (lane / 16) + r: every lane
// holds four consecutive outputs of one row, which it writes back as one 8-byte LDS store.
template <int KS>
__device__ __forceinline__ void mm_tile(const uint16_t* lds, const uint16_t* __restrict__ W, int w, int lane,
                                        f32x4 (&acc)[4][8]) {
  constexpr int K = 32 * KS;
  const int r16 = lane & 15, kg = (lane >> 4) * 8;
  const uint16_t* wp = W + (size_t)(64 * w + r16) * K + kg;
  // weight fragments of step ks + 1 are in flight while step ks multiplies
  u16x8 wf[4], wn[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) wf[nb] = *(const u16x8*)(wp + (size_t)16 * nb * K);
#pragma unroll 2
  for (int ks = 0; ks < KS; ++ks) {
    if (ks + 1 < KS) {
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) wn[nb] = *(const u16x8*)(wp + (size_t)16 * nb * K + (ks + 1) * 32);
    }
#pragma unroll
    for (int mb = 0; mb < 8; ++mb) {
      const u16x8 af = *(const u16x8*)(lds + (16 * mb + r16) * LS + ks * 32 + kg);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) acc[nb][mb] = mfma(wf[nb], af, acc[nb][mb]);
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) wf[nb] = wn[nb];
  }
}

__device__ __forceinline__ size_t slope_index(int tile, int w, int nb, int mb, int lane) {
  return ((((size_t)tile * 4 + w) * 4 + nb) * 8 + mb) * 256 + (size_t)lane * 4;
}

__device__ __forceinline__ uint16_t* lds_quad(uint16_t* lds, int w, int nb, int mb, int lane) {
  return lds + (16 * mb + (lane & 15)) * LS + 64 * w + 16 * nb + 4 * (lane >> 4);
}

struct FwdArgs {
  const uint16_t* wb;    // [E][BF]
  const float* master;   // [E][P]
  const float* x;        // [n][5]
  int64_t n;             // rows of x
  int M;                 // rows per member in this launch
  int mode;              // 0: row r; 1: Philox-drawn rows
  uint64_t seed;
  const int* step;       // optimiser step (sampling counter)
  int* idx;              // [E][M] rows used (written in mode 1)
  float* y;              // [E][M][3]
  uint16_t* acts;        // [E][5][H*M] FL(A_l), l = 0..4; null in inference
  uint16_t* slopes;      // [E][4][H*M] SiLU slopes in result-fragment order
};

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_mlp_forward(FwdArgs a) {
  extern __shared__ __align__(16) uint16_t lds[];
  const int tile = blockIdx.x, e = blockIdx.y;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int m0 = tile * MT;
  const int M = a.M;
  const uint16_t* wb = a.wb + (size_t)e * BF;
  const float* mst = a.master + (size_t)e * P;
  const bool train = a.acts != nullptr;

  // features -> hi/lo bf16 pairs in LDS columns 0..9, zeros up to 31
  if (tid < MT) {
    const int r = m0 + tid;
    int64_t i;
    if (a.mode == 1) {
      const uint32_t u = spk::draw_u32(a.seed, (int64_t)(*a.step) * M + r, (uint32_t)e, kStreamBatch);
      i = (int64_t)(((uint64_t)u * (uint64_t)a.n) >> 32);
      if (r < M) a.idx[(size_t)e * M + r] = (int)i;
    } else {
      i = r < M ? r : M - 1;
    }
    uint16_t* row = lds + tid * LS;
#pragma unroll
    for (int j = 0; j < DIN; ++j) {
      const float f = a.x[i * DIN + j];
      const uint16_t hi = to_bf16(f);
      row[j] = hi;
      row[DIN + j] = to_bf16(f - from_bf16(hi));
    }
#pragma unroll
    for (int j = 2 * DIN; j < K1; ++j) row[j] = 0;
  }
  __syncthreads();
  if (train) store_fl(lds, K1, a.acts + ((size_t)e * 5 + 0) * H * M, M, m0, w, lane);

  const int ng = 4 * (lane >> 4);
  for (int l = 1; l <= 4; ++l) {
    f32x4 acc[4][8];
    const float* bias = mst + ob[l - 1] + 64 * w + ng;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const float* bq = bias + 16 * nb;  // member stride P is odd: four scalar loads
      const f32x4 bv = {bq[0], bq[1], bq[2], bq[3]};
#pragma unroll
      for (int mb = 0; mb < 8; ++mb) acc[nb][mb] = bv;
    }
    if (l == 1)
      mm_tile<1>(lds, wb + bW1, w, lane, acc);
    else
      mm_tile<8>(lds, wb + bW + (size_t)(l - 2) * H * H, w, lane, acc);
    __syncthreads();  // every wave has read the layer input
    uint16_t* sl = train ? a.slopes + ((size_t)e * 4 + (l - 1)) * H * M : nullptr;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int mb = 0; mb < 8; ++mb) {
        const f32x4 z = acc[nb][mb];
        const f32x4 sg = {sigmoid(z[0]), sigmoid(z[1]), sigmoid(z[2]), sigmoid(z[3])};
        *(u16x4*)lds_quad(lds, w, nb, mb, lane) = pack4(z[0] * sg[0], z[1] * sg[1], z[2] * sg[2], z[3] * sg[3]);
        if (train)
          *(u16x4*)(sl + slope_index(tile, w, nb, mb, lane)) =
              pack4(sg[0] * (1.0f + z[0] * (1.0f - sg[0])), sg[1] * (1.0f + z[1] * (1.0f - sg[1])),
                    sg[2] * (1.0f + z[2] * (1.0f - sg[2])), sg[3] * (1.0f + z[3] * (1.0f - sg[3])));
      }
    __syncthreads();
    if (train) store_fl(lds, H, a.acts + ((size_t)e * 5 + l) * H * M, M, m0, w, lane);
  }

  // output layer, transposed: Y^T = W5 A4^T; wave w takes row blocks 2w, 2w+1; lanes 0..15
  // hold outputs 0..3 (3 used) of row 16 mb + lane
  {
    const int r16 = lane & 15, kg = (lane >> 4) * 8;
    f32x4 o0 = {0.0f, 0.0f, 0.0f, 0.0f};
    if (lane < 16) o0 = f32x4{mst[ob5], mst[ob5 + 1], mst[ob5 + 2], 0.0f};
    f32x4 o1 = o0;
    const uint16_t* wp = wb + bW5 + (size_t)r16 * H + kg;
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const u16x8 wf = *(const u16x8*)(wp + ks * 32);
      const u16x8 a0 = *(const u16x8*)(lds + (32 * w + r16) * LS + ks * 32 + kg);
      const u16x8 a1 = *(const u16x8*)(lds + (32 * w + 16 + r16) * LS + ks * 32 + kg);
      o0 = mfma(wf, a0, o0);
      o1 = mfma(wf, a1, o1);
    }
    if (lane < 16) {
      const int ra = m0 + 32 * w + lane, rb2 = ra + 16;
      if (ra < M) {
        float* yp = a.y + ((size_t)e * M + ra) * DOUT;
        yp[0] = o0[0]; yp[1] = o0[1]; yp[2] = o0[2];
      }
      if (rb2 < M) {
        float* yp = a.y + ((size_t)e * M + rb2) * DOUT;
        yp[0] = o1[0]; yp[1] = o1[1]; yp[2] = o1[2];
      }
    }
  }
}

struct BwdArgs {
  const uint16_t* wb;
  const float* y;         // [E][M][3]
  const float* t;         // targets [n][3]
  const int* idx;         // [E][M] rows of t (null: row r)
  float w0, w1, w2;       // loss weights
  float scale;            // 2 / M
  const uint16_t* slopes; // [E][4][H*M]
  uint16_t* grads;        // [E][5][H*M]: FL(dZ_1..dZ_4) in slots 0..3, FL(dY) (16 columns) in 4
  float* loss_part;       // [E][tiles]
  int* step;              // incremented once when tick != 0
  int tick;
  int M;
};

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_mlp_backward(BwdArgs a) {
  extern __shared__ __align__(16) uint16_t lds[];
  float* red = (float*)(lds + MT * LS);
  const int tile = blockIdx.x, e = blockIdx.y, tiles = gridDim.x;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int m0 = tile * MT, M = a.M;
  const uint16_t* wb = a.wb + (size_t)e * BF;

  if (a.tick && tile == 0 && e == 0 && tid == 0) *a.step += 1;

  // output gradient dy = bf16(2 w_o (y - t) / M) in LDS columns 0..2, zeros up to 31
  float lsum = 0.0f;
  if (tid < MT) {
    const int r = m0 + tid;
    const int64_t i = a.idx ? a.idx[(size_t)e * M + r] : r;
    const float wo[3] = {a.w0, a.w1, a.w2};
    uint16_t* row = lds + tid * LS;
#pragma unroll
    for (int o = 0; o < DOUT; ++o) {
      const float res = a.y[((size_t)e * M + r) * DOUT + o] - a.t[i * DOUT + o];
      lsum += wo[o] * res * res;
      row[o] = to_bf16(a.scale * wo[o] * res);
    }
#pragma unroll
    for (int j = DOUT; j < K1; ++j) row[j] = 0;
  }
  // per-tile loss, summed in a fixed order
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) lsum += __shfl_xor(lsum, s);
  if (lane == 0) red[w] = lsum;
  __syncthreads();
  if (tid == 0) a.loss_part[(size_t)e * tiles + tile] = (red[0] + red[1]) + (red[2] + red[3]);
  store_fl(lds, NO, a.grads + ((size_t)e * 5 + 4) * H * M, M, m0, w, lane);

  for (int l = 4; l >= 1; --l) {
    // dA_l^T = W_(l+1)^T dZ_(l+1)^T  (for l = 4: W5^T dY^T)
    f32x4 acc[4][8];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int mb = 0; mb < 8; ++mb) acc[nb][mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (l == 4)
      mm_tile<1>(lds, wb + bW5T, w, lane, acc);
    else
      mm_tile<8>(lds, wb + bWT + (size_t)(l - 1) * H * H, w, lane, acc);
    __syncthreads();
    const uint16_t* sl = a.slopes + ((size_t)e * 4 + (l - 1)) * H * M;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int mb = 0; mb < 8; ++mb) {
        const u16x4 sv = *(const u16x4*)(sl + slope_index(tile, w, nb, mb, lane));
        const f32x4 d = acc[nb][mb];
        *(u16x4*)lds_quad(lds, w, nb, mb, lane) =
            pack4(d[0] * from_bf16(sv[0]), d[1] * from_bf16(sv[1]), d[2] * from_bf16(sv[2]), d[3] * from_bf16(sv[3]));
      }
    __syncthreads();
    store_fl(lds, H, a.grads + ((size_t)e * 5 + (l - 1)) * H * M, M, m0, w, lane);
  }
}

struct WgradArgs {
  const uint16_t* acts;   // [E][5][H*M]
  const uint16_t* grads;  // [E][5][H*M]
  float* part;            // [E][S][P]
  int M, chunk;           // rows, rows per split
};

// 56 wave jobs per member and split: dW1 (4 row tiles of 64 x 32 columns), dW2..dW4
// (16 tiles of 64 x 64 each), dW5 (4 tiles of 16 rows x 64 columns); the first column
// tile of each layer also sums the bias gradient.
__global__ __launch_bounds__(NT) void k_mlp_wgrad(WgradArgs a) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int job = blockIdx.x * 4 + w, s = blockIdx.y, e = blockIdx.z;
  const int S = gridDim.y, M = a.M;
  int layer, n0, k0, nrb, ncb;
  if (job < 4) {
    layer = 1; n0 = 64 * job; k0 = 0; nrb = 4; ncb = 2;
  } else if (job < 52) {
    const int q = job - 4;
    layer = 2 + q / 16; n0 = 64 * ((q % 16) / 4); k0 = 64 * (q % 4); nrb = 4; ncb = 4;
  } else {
    layer = 5; n0 = 0; k0 = 64 * (job - 52); nrb = 1; ncb = 4;
  }
  const uint16_t* dz = a.grads + ((size_t)e * 5 + (layer - 1)) * H * M;  // layer 5: dY
  const uint16_t* ac = a.acts + ((size_t)e * 5 + (layer - 1)) * H * M;
  const int nmb = M >> 5;
  const int mb0 = (s * a.chunk) >> 5, mb1 = mb0 + (a.chunk >> 5);
  f32x4 acc[4][4];
  float rs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const bool bias = k0 == 0;
  for (int mb = mb0; mb < mb1; ++mb) {
    u16x8 bv[4], av[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
      if (cb < ncb) bv[cb] = *(const u16x8*)(ac + ((size_t)((k0 >> 4) + cb) * nmb + mb) * 512 + lane * 8);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
      if (rb < nrb) av[rb] = *(const u16x8*)(dz + ((size_t)((n0 >> 4) + rb) * nmb + mb) * 512 + lane * 8);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      if (rb >= nrb) continue;
      if (bias) {
#pragma unroll
        for (int j = 0; j < 8; ++j) rs[rb] += from_bf16(av[rb][j]);
      }
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
        if (cb < ncb) acc[rb][cb] = mfma(av[rb], bv[cb], acc[rb][cb]);
    }
  }
  float* out = a.part + ((size_t)e * S + s) * P;
  const int r16 = lane & 15, rg = (lane >> 4) * 4;
  const int ndim = layer == 5 ? DOUT : H;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) {
    if (rb >= nrb) continue;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      if (cb >= ncb) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + 16 * rb + rg + r, k = k0 + 16 * cb + r16;
        float v = acc[rb][cb][r];
        if (layer == 1) {
          // W1 column k is used by the hi (k) and lo (k + 5) feature halves
          const float v5 = __shfl(v, (lane & 48) | ((r16 + DIN) & 15));
          if (cb == 0 && r16 < DIN && n < ndim) out[oW1 + n * DIN + r16] = v + v5;
        } else if (n < ndim) {
          out[oW[layer - 1] + n * H + k] = v;
        }
      }
    }
    if (bias) {
      float v = rs[rb];
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      const int n = n0 + 16 * rb + lane;
      if (lane < 16 && n < ndim) out[ob[layer - 1] + n] = v;
    }
  }
}

struct AdamArgs {
  float* master; float* mom; float* vel;
  uint16_t* wb;
  const float* part; int S;   // partial gradients [E][S][P] (or)
  const float* gin;           // caller gradients [E][P]
  float* gout;                // gradients [E][P] (optional)
  const int* step;
  float lr0, lr_min; int total;
  int apply;                  // 0: gradients only; 1: Adam step; 2: write the BF16 copies only
};

__device__ __forceinline__ void write_copies(uint16_t* wb, int p, float v) {
  const uint16_t h = to_bf16(v);
  if (p < ob1) {
    const int n = p / DIN, k = p - n * DIN;
    wb[bW1 + n * K1 + k] = h;
    wb[bW1 + n * K1 + DIN + k] = h;
  } else if (p >= oW2 && p < oW5) {
    int L = -1, q = 0;
    if (p < ob2) {
      L = 0; q = p - oW2;
    } else if (p >= oW3 && p < ob3) {
      L = 1; q = p - oW3;
    } else if (p >= oW4 && p < ob4) {
      L = 2; q = p - oW4;
    }
    if (L >= 0) {  // biases of layers 2..4 have no working copy
      const int n = q >> 8, k = q & 255;
      wb[bW + L * H * H + n * H + k] = h;
      wb[bWT + L * H * H + k * H + n] = h;
    }
  } else if (p >= oW5 && p < ob5) {
    const int o = (p - oW5) >> 8, k = (p - oW5) & 255;
    wb[bW5 + o * H + k] = h;
    wb[bW5T + k * 32 + o] = h;
  }
}

__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x, e = blockIdx.y;
  if (p >= P) return;
  const size_t ep = (size_t)e * P + p;
  float* mst = a.master + ep;
  uint16_t* wb = a.wb + (size_t)e * BF;
  if (a.apply == 2) {
    write_copies(wb, p, *mst);
    return;
  }
  float g;
  if (a.gin) {
    g = a.gin[ep];
  } else {
    g = 0.0f;
    const float* pp = a.part + (size_t)e * a.S * P + p;
    for (int s = 0; s < a.S; ++s) g += pp[(size_t)s * P];
  }
  if (a.gout) a.gout[ep] = g;
  if (!a.apply) return;
  const int t = *a.step;
  const float u = (float)min(max(t - 1, 0), a.total) / (float)max(a.total, 1);
  const float lr = a.lr_min + 0.5f * (a.lr0 - a.lr_min) * (1.0f + cosf(3.14159265358979f * u));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * a.mom[ep] + (1.0f - b1) * g;
  const float v = b2 * a.vel[ep] + (1.0f - b2) * g * g;
  a.mom[ep] = m;
  a.vel[ep] = v;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  const float val = *mst - lr * (m / c1) / (sqrtf(v / c2) + eps);
  *mst = val;
  write_copies(wb, p, val);
}

__global__ void k_mlp_tick(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// host side

struct spk_mlp {
  int E = 0, max_batch = 0;
  float *master = nullptr, *mom = nullptr, *vel = nullptr;
  uint16_t* wb = nullptr;
  float *x = nullptr, *t = nullptr;
  int64_t n = 0, cap = 0;
  int* idx = nullptr;
  float* y = nullptr;
  uint16_t *acts = nullptr, *slopes = nullptr, *grads = nullptr;
  float *part = nullptr, *loss_part = nullptr, *gbuf = nullptr, *gin = nullptr;
  int* step = nullptr;
  float *px = nullptr, *py = nullptr;
  int64_t pcap = 0;
  float* scratch = nullptr;  // this device's copy of the file-scope staging array
  float wout[3] = {1.0f, 1.0f, 1.0f};
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};
  hipGraphExec_t graph = nullptr;
  int graph_batch = 0, graph_mode = -1;
  uint64_t graph_seed = 0;
  float graph_lr0 = 0, graph_lrmin = 0;
  int graph_total = 0;
  spk_mlp_timing timing = {};
};

// Staging memory of spk_mlp_predict_noalloc: features then outputs, in a file-scope
// device array of kNoAllocBytes.  Every GPU the library runs on holds its own copy until
// the process ends.  The array's name is a host-side stand-in in host code: its device
// address (one per GPU) comes from hipGetSymbolAddress at spk_mlp_create.
constexpr size_t kNoAllocBytes = size_t(4) << 20;
constexpr size_t kNoAllocFloats = kNoAllocBytes / sizeof(float);
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

int mfail(int code, const std::string& msg);

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev`, looked up once per device.
int scratch_address(int dev, float** out) {
  if (dev < 0 || dev >= kMaxDevices) return mfail(SPK_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mfail(SPK_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_addr[dev] = (float*)p;
  }
  *out = g_scratch_addr[dev];
  return SPK_OK;
}

}  // namespace

namespace {

int mfail(int code, const std::string& msg) {
  spk::set_last_error(msg);
  return code;
}

#define MLP_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return mfail(SPK_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int malloc_d(T** p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return mfail(SPK_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return SPK_OK;
}

template <typename T>
void free_d(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int lds_attributes() {
  static int done = 0;
  if (done) return SPK_OK;
  MLP_HIP(hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES));
  MLP_HIP(hipFuncSetAttribute((const void*)k_mlp_backward, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES));
  done = 1;
  return SPK_OK;
}

// Row chunks of the weight gradient: chunks of about 1024 rows (8 tiles), at most
// kMaxSplits, dividing the batch evenly.  Every chunk writes a full parameter-sized partial
// sum, so fewer, longer chunks cost less traffic than many short ones.
constexpr int kMaxSplits = 16;
int splits(int batch) {
  const int q = batch / MT;
  for (int s = std::min(std::max(q / 8, 1), kMaxSplits); s > 1; --s)
    if (q % s == 0) return s;
  return 1;
}

// one training step (forward, backward, weight gradient, Adam) on the surrogate's stream
int enqueue_step(spk_mlp* m, int batch, int mode, uint64_t seed, int apply, float lr0, float lr_min, int total,
                 float* gout, float* loss_part, bool timed) {
  const int tiles = batch / MT;
  FwdArgs f{};
  f.wb = m->wb; f.master = m->master; f.x = m->x; f.n = m->n; f.M = batch; f.mode = mode; f.seed = seed;
  f.step = m->step; f.idx = m->idx; f.y = m->y; f.acts = m->acts; f.slopes = m->slopes;
  if (timed) MLP_HIP(hipEventRecord(m->ev[0], m->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(tiles, m->E), dim3(NT), LDS_BYTES, m->stream, f);
  if (timed) MLP_HIP(hipEventRecord(m->ev[1], m->stream));
  BwdArgs b{};
  b.wb = m->wb; b.y = m->y; b.t = m->t; b.idx = mode == 1 ? m->idx : nullptr;
  b.w0 = m->wout[0]; b.w1 = m->wout[1]; b.w2 = m->wout[2];
  b.scale = 2.0f / (float)batch;
  b.slopes = m->slopes; b.grads = m->grads; b.loss_part = loss_part; b.step = m->step; b.tick = apply == 1;
  b.M = batch;
  hipLaunchKernelGGL(k_mlp_backward, dim3(tiles, m->E), dim3(NT), LDS_BYTES, m->stream, b);
  if (timed) MLP_HIP(hipEventRecord(m->ev[2], m->stream));
  const int S = splits(batch);
  const int chunk = batch / S;
  WgradArgs g{};
  g.acts = m->acts; g.grads = m->grads; g.part = m->part; g.M = batch; g.chunk = chunk;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(14, S, m->E), dim3(NT), 0, m->stream, g);
  if (timed) MLP_HIP(hipEventRecord(m->ev[3], m->stream));
  AdamArgs ad{};
  ad.master = m->master; ad.mom = m->mom; ad.vel = m->vel; ad.wb = m->wb; ad.part = m->part; ad.S = S;
  ad.gin = nullptr; ad.gout = gout; ad.step = m->step; ad.lr0 = lr0; ad.lr_min = lr_min; ad.total = total;
  ad.apply = apply;
  hipLaunchKernelGGL(k_mlp_adam, dim3((P + 255) / 256, m->E), dim3(256), 0, m->stream, ad);
  if (timed) MLP_HIP(hipEventRecord(m->ev[4], m->stream));
  MLP_HIP(hipGetLastError());
  return SPK_OK;
}

// The captured training step holds the data pointers, the row count and the loss weights
// as kernel arguments: any change to them drops it (it is captured again when needed).
void drop_graph(spk_mlp* m) {
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  m->graph = nullptr;
}

int check_batch(spk_mlp* m, int batch) {
  if (batch <= 0 || batch % MT || batch > m->max_batch)
    return mfail(SPK_ERR_ARG, "batch must be a positive multiple of " + std::to_string(MT) + " up to max_batch " +
                                  std::to_string(m->max_batch));
  return SPK_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace

extern "C" {

int spk_mlp_param_count(void) { return P; }

int spk_mlp_create(int models, int max_batch, spk_mlp** out) {
  if (!out) return mfail(SPK_ERR_ARG, "null output handle");
  *out = nullptr;
  if (models < 1 || models > 64) return mfail(SPK_ERR_ARG, "models must be 1..64");
  if (max_batch < 1 || max_batch > (1 << 20)) return mfail(SPK_ERR_ARG, "max_batch must be 1..2^20");
  int dev = 0, ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return mfail(SPK_ERR_NODEVICE, "no HIP device");
  MLP_HIP(hipGetDevice(&dev));
  int rc = lds_attributes();
  if (rc) return rc;
  float* scr = nullptr;
  if ((rc = scratch_address(dev, &scr))) return rc;
  spk_mlp* m = new spk_mlp();
  m->scratch = scr;
  m->E = models;
  m->max_batch = (max_batch + MT - 1) / MT * MT;
  const size_t E = models, Mx = m->max_batch;
  const size_t S = kMaxSplits;
  auto bail = [&](int r) { spk_mlp_destroy(m); return r; };
#define MLP_ALLOC(p, cnt) if ((rc = malloc_d(&(p), (cnt)))) return bail(rc)
  MLP_ALLOC(m->master, E * P);
  MLP_ALLOC(m->mom, E * P);
  MLP_ALLOC(m->vel, E * P);
  MLP_ALLOC(m->wb, E * BF);
  MLP_ALLOC(m->idx, E * Mx);
  MLP_ALLOC(m->y, E * Mx * DOUT);
  MLP_ALLOC(m->acts, E * 5 * H * Mx);
  MLP_ALLOC(m->slopes, E * 4 * H * Mx);
  MLP_ALLOC(m->grads, E * 5 * H * Mx);
  MLP_ALLOC(m->part, E * S * P);
  MLP_ALLOC(m->loss_part, E * (Mx / MT));
  MLP_ALLOC(m->gbuf, E * P);
  MLP_ALLOC(m->step, 1);
#undef MLP_ALLOC
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(mfail(SPK_ERR_HIP, "hipStreamCreate failed"));
  for (auto& ev : m->ev)
    if (hipEventCreate(&ev) != hipSuccess) return bail(mfail(SPK_ERR_HIP, "hipEventCreate failed"));
  if (hipMemset(m->master, 0, E * P * sizeof(float)) != hipSuccess ||
      hipMemset(m->wb, 0, E * BF * sizeof(uint16_t)) != hipSuccess)
    return bail(mfail(SPK_ERR_HIP, "hipMemset failed"));
  if ((rc = spk_mlp_reset_optimizer(m))) return bail(rc);
  *out = m;
  return SPK_OK;
}

int spk_mlp_destroy(spk_mlp* m) {
  if (!m) return SPK_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  free_d(m->master); free_d(m->mom); free_d(m->vel); free_d(m->wb);
  free_d(m->x); free_d(m->t); free_d(m->idx); free_d(m->y);
  free_d(m->acts); free_d(m->slopes); free_d(m->grads);
  free_d(m->part); free_d(m->loss_part); free_d(m->gbuf); free_d(m->gin); free_d(m->step);
  free_d(m->px); free_d(m->py);
  for (auto& ev : m->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return SPK_OK;
}

int spk_mlp_reset_optimizer(spk_mlp* m) {
  if (!m) return mfail(SPK_ERR_ARG, "null handle");
  MLP_HIP(hipMemsetAsync(m->mom, 0, (size_t)m->E * P * sizeof(float), m->stream));
  MLP_HIP(hipMemsetAsync(m->vel, 0, (size_t)m->E * P * sizeof(float), m->stream));
  MLP_HIP(hipMemsetAsync(m->step, 0, sizeof(int), m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return SPK_OK;
}

int spk_mlp_set_params(spk_mlp* m, const float* params) {
  if (!m || !params) return mfail(SPK_ERR_ARG, "null argument");
  for (size_t i = 0; i < (size_t)m->E * P; ++i)
    if (!std::isfinite(params[i])) return mfail(SPK_ERR_ARG, "non-finite parameter at " + std::to_string(i));
  MLP_HIP(hipMemcpyAsync(m->master, params, (size_t)m->E * P * sizeof(float), hipMemcpyHostToDevice, m->stream));
  // padding entries of the working copies stay zero from creation; every used entry is rewritten
  AdamArgs ad{};
  ad.master = m->master; ad.wb = m->wb; ad.apply = 2;
  hipLaunchKernelGGL(k_mlp_adam, dim3((P + 255) / 256, m->E), dim3(256), 0, m->stream, ad);
  MLP_HIP(hipGetLastError());
  return spk_mlp_reset_optimizer(m);
}

int spk_mlp_get_params(spk_mlp* m, float* params) {
  if (!m || !params) return mfail(SPK_ERR_ARG, "null argument");
  MLP_HIP(hipMemcpyAsync(params, m->master, (size_t)m->E * P * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  return SPK_OK;
}

int spk_mlp_set_data(spk_mlp* m, const float* features, const float* targets, int64_t n) {
  if (!m || !features || !targets || n < 1 || n > ((int64_t)1 << 31) - 1)
    return mfail(SPK_ERR_ARG, "bad training set");
  MLP_HIP(hipStreamSynchronize(m->stream));
  drop_graph(m);
  if (n > m->cap) {
    MLP_HIP(hipStreamSynchronize(m->stream));
    free_d(m->x);
    free_d(m->t);
    int rc;
    if ((rc = malloc_d(&m->x, (size_t)n * DIN)) || (rc = malloc_d(&m->t, (size_t)n * DOUT))) return rc;
    m->cap = n;
  }
  MLP_HIP(hipMemcpyAsync(m->x, features, (size_t)n * DIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipMemcpyAsync(m->t, targets, (size_t)n * DOUT * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  m->n = n;
  return SPK_OK;
}

int spk_mlp_set_loss_weights(spk_mlp* m, const float* w3) {
  if (!m || !w3) return mfail(SPK_ERR_ARG, "null argument");
  for (int i = 0; i < 3; ++i) {
    if (!(w3[i] >= 0.0f) || !std::isfinite(w3[i])) return mfail(SPK_ERR_ARG, "loss weights must be finite and >= 0");
  }
  MLP_HIP(hipStreamSynchronize(m->stream));
  drop_graph(m);
  for (int i = 0; i < 3; ++i) m->wout[i] = w3[i];
  return SPK_OK;
}

int spk_mlp_gradients(spk_mlp* m, int rows, float* grads, float* loss) {
  if (!m) return mfail(SPK_ERR_ARG, "null handle");
  int rc = check_batch(m, rows);
  if (rc) return rc;
  if (rows > m->n) return mfail(SPK_ERR_ARG, "rows exceed the training set");
  if ((rc = enqueue_step(m, rows, 0, 0, 0, 0.0f, 0.0f, 1, m->gbuf, m->loss_part, true))) return rc;
  MLP_HIP(hipStreamSynchronize(m->stream));
  if (grads) MLP_HIP(hipMemcpy(grads, m->gbuf, (size_t)m->E * P * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) {
    const int tiles = rows / MT;
    std::vector<float> lp((size_t)m->E * tiles);
    MLP_HIP(hipMemcpy(lp.data(), m->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int e = 0; e < m->E; ++e) {
      double s = 0.0;
      for (int k = 0; k < tiles; ++k) s += lp[(size_t)e * tiles + k];
      loss[e] = (float)(s / rows);
    }
  }
  m->timing = spk_mlp_timing{elapsed(m->ev[0], m->ev[4]), elapsed(m->ev[0], m->ev[1]), elapsed(m->ev[1], m->ev[2]),
                             elapsed(m->ev[2], m->ev[3]), elapsed(m->ev[3], m->ev[4]), 0, rows};
  return SPK_OK;
}

int spk_mlp_apply_gradients(spk_mlp* m, const float* grads, float lr) {
  if (!m || !grads) return mfail(SPK_ERR_ARG, "null argument");
  if (!(lr >= 0.0f) || !std::isfinite(lr)) return mfail(SPK_ERR_ARG, "bad learning rate");
  int rc;
  if (!m->gin && (rc = malloc_d(&m->gin, (size_t)m->E * P))) return rc;
  MLP_HIP(hipMemcpyAsync(m->gin, grads, (size_t)m->E * P * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, m->stream, m->step);
  AdamArgs ad{};
  ad.master = m->master; ad.mom = m->mom; ad.vel = m->vel; ad.wb = m->wb; ad.gin = m->gin; ad.step = m->step;
  ad.lr0 = lr; ad.lr_min = lr; ad.total = 1; ad.apply = 1;
  hipLaunchKernelGGL(k_mlp_adam, dim3((P + 255) / 256, m->E), dim3(256), 0, m->stream, ad);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(m->stream));
  return SPK_OK;
}

int spk_mlp_train(spk_mlp* m, int steps, int batch, int mode, uint64_t seed, float lr0, float lr_min,
                  int total_steps, int use_graph, float* loss_per_step) {
  if (!m) return mfail(SPK_ERR_ARG, "null handle");
  int rc = check_batch(m, batch);
  if (rc) return rc;
  if (steps < 0 || total_steps < 1) return mfail(SPK_ERR_ARG, "steps must be >= 0, total_steps >= 1");
  if (mode != SPK_MLP_BATCH_FIRST_ROWS && mode != SPK_MLP_BATCH_SAMPLED) return mfail(SPK_ERR_ARG, "bad mode");
  if (m->n < 1 || (mode == SPK_MLP_BATCH_FIRST_ROWS && batch > m->n))
    return mfail(SPK_ERR_STATE, "training set missing or smaller than the batch");
  if (!(lr0 >= 0.0f) || !(lr_min >= 0.0f) || !std::isfinite(lr0) || !std::isfinite(lr_min))
    return mfail(SPK_ERR_ARG, "bad learning rate");
  const int tiles = batch / MT;
  float* lhist = nullptr;
  if (loss_per_step && steps > 0 && (rc = malloc_d(&lhist, (size_t)steps * m->E * tiles))) return rc;
  MLP_HIP(hipEventRecord(m->ev[5], m->stream));
  if (use_graph && !loss_per_step) {
    const bool same = m->graph && m->graph_batch == batch && m->graph_mode == mode && m->graph_seed == seed &&
                      m->graph_lr0 == lr0 && m->graph_lrmin == lr_min && m->graph_total == total_steps;
    if (!same) {
      if (m->graph) (void)hipGraphExecDestroy(m->graph);
      m->graph = nullptr;
      hipGraph_t g;
      MLP_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_step(m, batch, mode, seed, 1, lr0, lr_min, total_steps, nullptr, m->loss_part, false);
      hipError_t ce = hipStreamEndCapture(m->stream, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return mfail(SPK_ERR_HIP, "graph capture failed");
      hipError_t ie = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return mfail(SPK_ERR_HIP, "graph instantiation failed");
      m->graph_batch = batch; m->graph_mode = mode; m->graph_seed = seed;
      m->graph_lr0 = lr0; m->graph_lrmin = lr_min; m->graph_total = total_steps;
    }
    for (int s = 0; s < steps; ++s) MLP_HIP(hipGraphLaunch(m->graph, m->stream));
  } else {
    for (int s = 0; s < steps; ++s) {
      float* lp = lhist ? lhist + (size_t)s * m->E * tiles : m->loss_part;
      if ((rc = enqueue_step(m, batch, mode, seed, 1, lr0, lr_min, total_steps, nullptr, lp, false))) {
        free_d(lhist);
        return rc;
      }
    }
  }
  MLP_HIP(hipEventRecord(m->ev[4], m->stream));
  hipError_t se = hipStreamSynchronize(m->stream);
  if (se != hipSuccess) {
    free_d(lhist);
    return mfail(SPK_ERR_HIP, std::string("training failed: ") + hipGetErrorString(se));
  }
  if (lhist) {
    std::vector<float> lp((size_t)steps * m->E * tiles);
    hipError_t ce = hipMemcpy(lp.data(), lhist, lp.size() * sizeof(float), hipMemcpyDeviceToHost);
    free_d(lhist);
    if (ce != hipSuccess) return mfail(SPK_ERR_HIP, "loss copy failed");
    for (int s = 0; s < steps; ++s)
      for (int e = 0; e < m->E; ++e) {
        double acc = 0.0;
        for (int k = 0; k < tiles; ++k) acc += lp[((size_t)s * m->E + e) * tiles + k];
        loss_per_step[(size_t)s * m->E + e] = (float)(acc / batch);
      }
  }
  m->timing = spk_mlp_timing{elapsed(m->ev[5], m->ev[4]), 0, 0, 0, 0, steps, batch};
  return SPK_OK;
}

int spk_mlp_predict(spk_mlp* m, const float* features, int64_t n, float* out) {
  if (!m || !features || !out || n < 1 || n > ((int64_t)1 << 28)) return mfail(SPK_ERR_ARG, "bad arguments");
  if (n > m->pcap) {
    MLP_HIP(hipStreamSynchronize(m->stream));
    free_d(m->px);
    free_d(m->py);
    int rc;
    if ((rc = malloc_d(&m->px, (size_t)n * DIN)) || (rc = malloc_d(&m->py, (size_t)m->E * n * DOUT))) return rc;
    m->pcap = n;
  }
  MLP_HIP(hipMemcpyAsync(m->px, features, (size_t)n * DIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
  FwdArgs f{};
  f.wb = m->wb; f.master = m->master; f.x = m->px; f.n = n; f.M = (int)n; f.mode = 0;
  f.y = m->py;
  const int tiles = (int)((n + MT - 1) / MT);
  MLP_HIP(hipEventRecord(m->ev[0], m->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(tiles, m->E), dim3(NT), LDS_BYTES, m->stream, f);
  MLP_HIP(hipEventRecord(m->ev[1], m->stream));
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(out, m->py, (size_t)m->E * n * DOUT * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  const float ms = elapsed(m->ev[0], m->ev[1]);
  m->timing = spk_mlp_timing{ms, ms, 0, 0, 0, 0, n};
  return SPK_OK;
}

int spk_mlp_predict_noalloc(spk_mlp* m, const float* features, int64_t n, float* out) {
  if (!m || !features || !out || n < 1) return mfail(SPK_ERR_ARG, "bad arguments");
  // features at the start of the array, outputs from the next 256-byte boundary
  const int64_t xfl = (n * DIN + 63) / 64 * 64;
  const int64_t need = xfl + (int64_t)m->E * n * DOUT;
  if (n > (int64_t)kNoAllocFloats || need > (int64_t)kNoAllocFloats)
    return mfail(SPK_ERR_ARG, "spk_mlp_predict_noalloc: " + std::to_string(n) + " rows x " +
                                  std::to_string(m->E) + " members need " + std::to_string(need * 4) +
                                  " bytes of staging; the static limit is " + std::to_string(kNoAllocBytes));
  float* px = m->scratch;
  float* py = m->scratch + xfl;
  MLP_HIP(hipMemcpyAsync(px, features, (size_t)n * DIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
  FwdArgs f{};
  f.wb = m->wb; f.master = m->master; f.x = px; f.n = n; f.M = (int)n; f.mode = 0;
  f.y = py;
  const int tiles = (int)((n + MT - 1) / MT);
  MLP_HIP(hipEventRecord(m->ev[0], m->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(tiles, m->E), dim3(NT), LDS_BYTES, m->stream, f);
  MLP_HIP(hipEventRecord(m->ev[1], m->stream));
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(out, py, (size_t)m->E * n * DOUT * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(hipStreamSynchronize(m->stream));
  const float ms = elapsed(m->ev[0], m->ev[1]);
  m->timing = spk_mlp_timing{ms, ms, 0, 0, 0, 0, n};
  return SPK_OK;
}

int spk_mlp_last_timing(const spk_mlp* m, spk_mlp_timing* t) {
  if (!m || !t) return mfail(SPK_ERR_ARG, "null argument");
  *t = m->timing;
  return SPK_OK;
}

}  // extern "C"
