// spike_oracle.cpp -- CPU restatement of Spike's per-timestep update equations.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the reported CPU
// baseline; it is never linked into, loaded by, or called from the product library.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and only
// as the checker.
//
// It restates, in plain multithreaded C++, the update equations of the reference's CUDA
// backend, templated on the real type (float = the FP32 oracle, double = the FP64
// oracle):
//   LIF sub-step ........ Spike/Backend/CUDA/Neurons/LIFSpikingNeurons.cu:91-195
//   injection ........... Spike/Backend/CUDA/InlineDeviceFunctions.hpp:22-51 (conductance),
//                         :86-111 (voltage)
//   spike append/propag.  Spike/Backend/CUDA/Synapses/SpikingSynapses.cu:151-202
//   Poisson inputs ...... Spike/Backend/CUDA/Neurons/PoissonInputSpikingNeurons.cu:73-130
//   spike history ....... Spike/Backend/CUDA/Neurons/SpikingNeurons.cu:23-94
//   WD STDP ............. Spike/Backend/CUDA/Plasticity/WeightDependentSTDPPlasticity.cu:67-163
//   Vogels iSTDP ........ Spike/Backend/CUDA/Plasticity/InhibitorySTDPPlasticity.cu:72-155
//   Evans STDP .......... Spike/Backend/CUDA/Plasticity/EvansSTDPPlasticity.cu:59-143
//   Custom STDP ......... Spike/Backend/CUDA/Plasticity/CustomSTDPPlasticity.cu:67-155
//   nearest-spike traces  Spike/Backend/CUDA/Plasticity/CustomSTDPPlasticity.cu:117-128
//   raster .............. Spike/Backend/CUDA/ActivityMonitor/SpikingActivityMonitor.cu:69-98
//   group loop .......... Spike/Models/SpikingModel.cpp:197-236
//
// Decisions shared with the product (they define the semantics both must meet):
//   * FP32 oracle: the input ring holds exact int64 fixed-point sums with `frac_bits`
//     fraction bits (the reference's float atomicAdd order is undefined).  A delivery adds
//     llrint(ldexp(w*scale, frac_bits)); the neuron reads (float)(double)sum * 2^-frac_bits
//     (the same as (float)sum while |sum| <= 2^53, which the choice of frac_bits ensures;
//     the device converts the same way because it is cheaper there).
//     FP64 oracle: the ring holds plain double sums.
//   * Poisson inputs: the reference's next-spike-time sampler (exponential waiting time,
//     one spike per step at most, first processed step only arms the neuron) is the
//     per-step Bernoulli process p = 1 - exp(-rate*dt) started one step after arming.
//     The draw for step n of input i is output n mod 4 of
//     philox4x32_10(ctr={n/4 lo, n/4 hi, i, 1}, key=seed), compared < floor(p * 2^32)
//     (p in FP64 on the host).
//   * External per-neuron drive: count = #{k : u >= C[k]}, C[k] = floor(F(k) * 2^32), F the
//     Poisson(rate*dt) CDF, u = output step mod 4 of philox(ctr={step/4, neuron, 2}, key=seed); the
//     count times the event weight joins label 0 of that step's ring input.
//   * All float expressions are evaluated as written, without FMA contraction
//     (built with -ffp-contract=off).
//   * last_spike_time of a Poisson input is its last emitted spike (the reference keeps
//     its next scheduled time there; see the design notes on weight-dependent STDP).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// Round: (L,R) pairs multiplied by M0/M1, hi/lo swapped and xored with the key, key
// bumped by the Weyl constants W0/W1 between rounds.
struct Philox {
  static void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
  }
  static void block(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
      uint32_t hi0, lo0, hi1, lo1;
      mulhilo(M0, c0, hi0, lo0);
      mulhilo(M1, c2, hi1, lo1);
      uint32_t n0 = hi1 ^ c1 ^ k0;
      uint32_t n1 = lo1;
      uint32_t n2 = hi0 ^ c3 ^ k1;
      uint32_t n3 = lo0;
      c0 = n0; c1 = n1; c2 = n2; c3 = n3;
      k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  }
  // The four outputs of one call are four consecutive draws: draw n of (id, stream) is
  // output n mod 4 of the call with counter (n / 4 as 64 bits, id, stream), key = seed.
  static void draws4(uint64_t seed, int64_t blk, uint32_t id, uint32_t stream, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)(uint64_t)blk, (uint32_t)((uint64_t)blk >> 32), id, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    block(ctr, key, out);
  }
  static uint32_t draw(uint64_t seed, int64_t n, uint32_t id, uint32_t stream) {
    uint32_t out[4];
    draws4(seed, n >> 2, id, stream, out);
    return out[n & 3];
  }
};

// Draws of one (id, stream) taken in increasing order: one Philox call per four draws.
struct DrawCursor {
  uint64_t seed;
  uint32_t id, stream;
  int64_t blk = -1;
  uint32_t out[4] = {0, 0, 0, 0};
  DrawCursor(uint64_t s, uint32_t i, uint32_t st) : seed(s), id(i), stream(st) {}
  uint32_t operator()(int64_t n) {
    if ((n >> 2) != blk) {
      blk = n >> 2;
      Philox::draws4(seed, blk, id, stream, out);
    }
    return out[n & 3];
  }
};

enum { STREAM_INPUT = 1, STREAM_EXTERNAL = 2 };

std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

// Bernoulli threshold of a Poisson input: floor((1-exp(-rate*dt)) * 2^32), saturated.
uint32_t bernoulli_threshold(float rate, float dt) {
  double p = -std::expm1(-(double)rate * (double)dt);
  if (!(p > 0.0)) return 0u;
  double t = std::floor(p * 4294967296.0);
  if (t >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

// Inverse-CDF thresholds of a Poisson(lambda) count: C[k] = floor(F(k) * 2^32), k < K,
// stopping once the tail mass is below 2^-34.
std::vector<uint32_t> poisson_count_table(double lambda) {
  std::vector<uint32_t> c;
  if (!(lambda > 0.0)) return c;
  double pmf = std::exp(-lambda), cdf = pmf;
  for (int k = 0; k < 255; ++k) {
    double t = std::floor(cdf * 4294967296.0);
    c.push_back(t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t);
    if (1.0 - cdf < 5.820766091346741e-11) break;  // 2^-34
    pmf = pmf * lambda / (double)(k + 1);
    cdf += pmf;
  }
  return c;
}

// Fraction bits of the fixed-point ring: the largest per-(post,label) sum of |w*scale|
// over all incoming synapses (plastic synapses counted at their rule's cap if larger),
// plus the largest external count times its weight, must stay below 2^52.
int choose_frac_bits(double max_entry) {
  if (!(max_entry > 0.0)) return 32;
  int e = (int)std::ceil(std::log2(max_entry));
  int k = 52 - e;
  if (k > 120) k = 120;
  if (k < 0) k = 0;
  return k;
}

// ---- 16-bit storage of weights and traces (format 1 bfloat16, 2 IEEE binary16), round to
// nearest with ties to even.  bfloat16 by the usual bit rounding of the upper half of the
// float; binary16 from the value itself (frexp/ldexp and nearbyint in the default
// to-nearest mode), a different route from the device's integer conversion.
static uint16_t o_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
static float o_from_bf16(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static uint16_t o_to_half(float f) {
  const uint16_t sign = std::signbit(f) ? 0x8000 : 0;
  const double a = std::fabs((double)f);
  if (std::isnan(f)) return sign | 0x7E00;
  if (a >= 65520.0) return sign | 0x7C00;                      // rounds past 65504: inf
  if (a < std::ldexp(1.0, -14)) {                               // subnormal: multiples of 2^-24
    const double q = std::nearbyint(std::ldexp(a, 24));         // 1024 is the smallest normal
    return (uint16_t)(sign | (uint16_t)q);
  }
  int ex;
  std::frexp(a, &ex);                                           // a = m 2^ex, m in [0.5, 1)
  int e = ex - 1;                                               // a in [2^e, 2^(e+1))
  double q = std::nearbyint(std::ldexp(a, 10 - e));             // 1024 .. 2048
  if (q >= 2048.0) { q = 1024.0; ++e; }
  if (e > 15) return sign | 0x7C00;
  return (uint16_t)(sign | (uint16_t)((e + 15) << 10) | (uint16_t)((int)q - 1024));
}
static float o_from_half(uint16_t h) {
  const int e = (h >> 10) & 0x1F, m = h & 0x3FF;
  double v;
  if (e == 0) v = std::ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((double)(1024 + m), e - 25);
  return (float)((h & 0x8000) ? -v : v);
}
static float o_round16(float x, int fmt) {
  return fmt == 1 ? o_from_bf16(o_to_bf16(x)) : o_from_half(o_to_half(x));
}

struct Rule {
  int kind;  // 0 WD, 1 Vogels, 2 Evans, 3 Custom
  std::vector<int64_t> syn;  // caller's synapse indices
  float wd[8]; int nearest;
  float vg[4];
  float ev[5];
  float cu[7];  // a_minus, a_plus, a_star, tau_minus, tau_plus, w_max, learning_rate
  std::vector<float> wn_sumsq; std::vector<uint8_t> wn_inset;  // weight normalization
  std::vector<double> wn_w0;                                    // its initial weights
  std::vector<double> ta, tb;  // traces per plastic synapse
};

struct Monitor { int pop; std::vector<int32_t> ids; std::vector<float> times; };

template <typename R>
struct Net {
  static constexpr bool kFixed = std::is_same<R, float>::value;
  using Acc = typename std::conditional<kFixed, int64_t, double>::type;

  int N = 0, Nin = 0, kind = 0, L = 1;
  float dt = 1e-4f;
  int nthreads = 1;
  // LIF
  std::vector<float> v0, vreset, thr, c, Rp;
  float tref = 0.002f, Ibg = 0.0f;
  bool lif_set = false;
  // Poisson
  std::vector<float> rates; int n_stim = 0, stim = 0;
  // generator inputs: per stimulus the entries sorted by time (stable), offsets [n_stim+1]
  bool gen = false;
  float gen_onset = 0.0f;
  std::vector<int64_t> gen_off;
  std::vector<int32_t> gen_ids;
  std::vector<float> gen_times;
  uint64_t seed = 42;
  int64_t arm_step = -1;  // step at which inputs arm (no spike); -1: next processed step
  bool pending_arm = true;
  float ext_rate = 0.0f, ext_w = 0.0f;
  std::vector<uint32_t> ext_cdf;
  // labels
  std::vector<float> decay_g, Erev;
  // synapses (caller's order)
  std::vector<int32_t> spre, spost, sdelay, slabel;
  std::vector<R> sw;
  std::vector<float> sscale;
  // 16-bit weight storage of the FP32 oracle: sw holds the stored 16-bit values, sc the
  // Kahan compensation of each synapse (also a 16-bit value)
  int wfmt = 0;
  std::vector<float> sc;
  // per-thread sub-CSR by target range: rows[t][u] -> list of synapse ids
  std::vector<int> trange;  // nthreads+1 boundaries over posts
  std::vector<std::vector<int64_t>> rp;   // [t][u] row pointer into rs[t]
  std::vector<std::vector<int64_t>> rs;   // [t][...] synapse ids
  // derived
  int G = 1, Gprep = 1, B = 0, W = 1, frac = 32, dmin = 0, dmax = 0;
  bool prepared = false;
  std::vector<uint32_t> in_thr;  // [n_stim][Nin]
  // state
  std::vector<R> V;
  std::vector<float> last_n, last_in;
  std::vector<uint32_t> hist_n, hist_in;  // [w][pop]
  std::vector<Acc> ring;                   // [B][N*L]
  std::vector<R> gcond;                    // [N*L]
  std::vector<std::vector<int32_t>> spikes;  // [g] unified ids, ascending
  std::vector<Rule> rules;
  std::vector<Monitor> mons;
  // counters
  int64_t cnt_n = 0, cnt_in = 0, cnt_events = 0, cnt_ext = 0;

  int nu() const { return N + Nin; }

  template <typename F>
  void parallel(int n, F f) {
    int T = std::max(1, std::min(nthreads, n));
    if (T == 1) { f(0, 0, n); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
      int a = (int)((int64_t)n * t / T), b = (int)((int64_t)n * (t + 1) / T);
      th.emplace_back([=, &f] { f(t, a, b); });
    }
    for (auto& x : th) x.join();
  }

  template <typename F>
  void parallel64(int64_t n, F f) {
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n / 4096));
    if (T == 1) { f((int64_t)0, n); return; }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back([=, &f] { f(n * t / T, n * (t + 1) / T); });
    for (auto& x : th) x.join();
  }

  // conversion of one delivery to the ring's accumulator
  static Acc to_acc(R v, int k) {
    if constexpr (kFixed) return (Acc)std::llrint(std::ldexp((double)v, k));
    else return (Acc)v;
  }
  static R from_acc(Acc a, int k) {
    if constexpr (kFixed) return (float)(double)a * std::ldexp(1.0f, -k);
    else return (R)a;
  }

  // Store a plastic weight moved by a rule from w0 (stored) to nw (clamped to [lo, hi]):
  // FP32 as it is; with 16-bit storage one Kahan step of a nonzero increment, clamped.
  void wstore(int64_t s, R w0, R nw, float lo, float hi) {
    if constexpr (kFixed) {
      if (wfmt) {
        if (nw == w0) return;   // a zero increment leaves word and compensation as they are
        const float c = sc[s];
        const float d = nw - w0;
        const float y = d - c;
        const float t = w0 + y;
        const float v1 = o_round16(t, wfmt);
        if (v1 < lo) { sw[s] = o_round16(lo, wfmt); sc[s] = 0.0f; }
        else if (v1 > hi) { sw[s] = o_round16(hi, wfmt); sc[s] = 0.0f; }
        else { sw[s] = v1; sc[s] = o_round16((v1 - w0) - y, wfmt); }
        return;
      }
    }
    (void)lo; (void)hi; (void)w0;
    sw[s] = nw;
  }
  // a trace as stored: with either 16-bit weight format the traces are binary16 (a
  // bfloat16 trace would decay in uneven steps: its spacing is near the 0.5 % decay per
  // step of a 20 ms trace at dt = 0.1 ms)
  double tstore(R x) const {
    if constexpr (kFixed) {
      if (wfmt) return (double)o_round16(x, 2);
    }
    return (double)x;
  }

  int prepare(int g) {
    if (!lif_set && N > 0) return fail("LIF parameters not set");
    if (g < 1) return fail("timestep grouping must be >= 1");
    G = Gprep = g;
    int64_t S = (int64_t)spre.size();
    dmin = S ? 1 << 30 : 0; dmax = 0;
    for (int64_t s = 0; s < S; ++s) { dmin = std::min(dmin, sdelay[s]); dmax = std::max(dmax, sdelay[s]); }
    B = dmax + 2 * G + 1;                    // SpikingSynapses.cu:48
    W = (dmax + 2 * G) / 32 + 1;
    // fixed-point scale
    std::vector<double> cap(S, 0.0);
    for (int64_t s = 0; s < S; ++s) cap[s] = std::fabs((double)sw[s]);
    for (auto& r : rules) for (int64_t s : r.syn) {
      double c0 = r.kind == 0 ? std::fabs((double)r.wd[6])
                : (r.kind == 2 ? 1.0 : (r.kind == 3 ? std::fabs((double)r.cu[5]) : 0.0));
      if (r.kind == 4) c0 = std::sqrt(std::fabs((double)r.wn_sumsq[spost[s]]));
      cap[s] = std::max(cap[s], c0);
    }
    std::vector<double> ent((size_t)N * L, 0.0);
    for (int64_t s = 0; s < S; ++s)
      ent[(size_t)spost[s] * L + slabel[s]] += cap[s] * std::fabs((double)sscale[s]);
    // 16-bit storage (FP32 oracle only; the bound above uses the caller's weights)
    if constexpr (kFixed) {
      if (wfmt) {
        for (int64_t s = 0; s < S; ++s) sw[s] = o_round16(sw[s], wfmt);
        sc.assign(S, 0.0f);
      }
    }
    double mx = 0.0;
    for (double e : ent) mx = std::max(mx, e);
    ext_cdf = poisson_count_table((double)ext_rate * (double)dt);
    if (!ext_cdf.empty()) mx += (double)ext_cdf.size() * std::fabs((double)ext_w);
    frac = choose_frac_bits(mx);
    // thresholds
    in_thr.assign((size_t)std::max(n_stim, 1) * Nin, 0u);
    for (size_t i = 0; i < rates.size(); ++i) in_thr[i] = bernoulli_threshold(rates[i], dt);
    // per-thread target-partitioned CSR
    int T = std::max(1, nthreads);
    trange.assign(T + 1, 0);
    for (int t = 0; t <= T; ++t) trange[t] = (int)((int64_t)N * t / T);
    rp.assign(T, {}); rs.assign(T, {});
    std::vector<int> owner(N, 0);
    for (int t = 0; t < T; ++t) for (int i = trange[t]; i < trange[t + 1]; ++i) owner[i] = t;
    for (int t = 0; t < T; ++t) rp[t].assign(nu() + 1, 0);
    auto uid = [&](int32_t pre) { return pre >= 0 ? pre : N + (-pre - 1); };
    for (int64_t s = 0; s < S; ++s) rp[owner[spost[s]]][uid(spre[s]) + 1]++;
    for (int t = 0; t < T; ++t) {
      for (int u = 0; u < nu(); ++u) rp[t][u + 1] += rp[t][u];
      rs[t].assign(rp[t][nu()], 0);
    }
    std::vector<std::vector<int64_t>> fill(T);
    for (int t = 0; t < T; ++t) fill[t].assign(rp[t].begin(), rp[t].end() - 1);
    for (int64_t s = 0; s < S; ++s) { int t = owner[spost[s]]; rs[t][fill[t][uid(spre[s])]++] = s; }
    // weight normalization: the initial weights are the weights as stored now
    for (auto& r : rules)
      if (r.kind == 4) {
        r.wn_w0.resize(r.syn.size());
        for (size_t j = 0; j < r.syn.size(); ++j) r.wn_w0[j] = (double)sw[r.syn[j]];
      }
    prepared = true;
    return reset_state();
  }

  int reset_inputs() {
    last_in.assign(Nin, -1000.0f);
    hist_in.assign((size_t)W * Nin, 0u);
    pending_arm = true; arm_step = -1;
    return 0;
  }

  int reset_state() {
    if (!prepared) return fail("not prepared");
    V.resize(N);
    for (int i = 0; i < N; ++i) V[i] = (R)v0[i];
    last_n.assign(N, -1000.0f);
    hist_n.assign((size_t)W * N, 0u);
    ring.assign((size_t)B * N * L, (Acc)0);
    gcond.assign((size_t)N * L, (R)0);
    reset_inputs();
    for (auto& r : rules) { r.ta.assign(r.syn.size(), 0.0); r.tb.assign(r.syn.size(), 0.0); }
    return 0;
  }

  static void set_bit(std::vector<uint32_t>& h, int npop, int W, int i, int64_t step, bool v) {
    int64_t bits = (int64_t)W * 32;
    int64_t b = ((step % bits) + bits) % bits;
    uint32_t& word = h[(size_t)(b / 32) * npop + i];
    if (v) word |= (1u << (b % 32)); else word &= ~(1u << (b % 32));
  }
  static bool get_bit(const std::vector<uint32_t>& h, int npop, int W, int i, int64_t step) {
    int64_t bits = (int64_t)W * 32;
    int64_t b = ((step % bits) + bits) % bits;
    return (h[(size_t)(b / 32) * npop + i] >> (b % 32)) & 1u;
  }

  // LIFSpikingNeurons.cu:91-195 with the injection of InlineDeviceFunctions.hpp
  void update_neurons(float t) {
    int tidx = (int)std::roundf(t / dt);
    spikes.assign(G, {});
    int T = std::max(1, nthreads);
    std::vector<std::vector<std::vector<int32_t>>> part(T, std::vector<std::vector<int32_t>>(G));
    std::vector<int64_t> ext_counts(T, 0);
    const float inv = std::ldexp(1.0f, -frac);
    const size_t NL = (size_t)N * L;
    parallel(N, [&](int th, int a, int b) {
      for (int i = a; i < b; ++i) {
        R Vi = V[i];
        DrawCursor ext_draw(seed, (uint32_t)i, STREAM_EXTERNAL);
        for (int g = 0; g < G; ++g) {
          int64_t step = (int64_t)tidx + g;
          set_bit(hist_n, N, W, i, step, false);
          size_t slot = (size_t)((tidx + g) % B);
          Acc ext = (Acc)0;
          if (!ext_cdf.empty()) {
            uint32_t u = ext_draw(step);
            int k = 0;
            for (uint32_t cth : ext_cdf) k += (u >= cth) ? 1 : 0;
            ext_counts[th] += k;
            if constexpr (kFixed) ext = (Acc)k * to_acc((R)ext_w, frac);
            else ext = (Acc)k * (Acc)ext_w;
          }
          R dV = (R)0;
          if (kind == 0) {
            // InlineDeviceFunctions.hpp:86-111
            R total = (R)0;
            for (int l = 0; l < L; ++l) {
              Acc a = ring[slot * NL + (size_t)i * L + l];
              if (l == 0) a += ext;
              ring[slot * NL + (size_t)i * L + l] = (Acc)0;
              if constexpr (kFixed) total += (float)(double)a * inv;
              else total += (R)a;
            }
            dV = total;
          } else {
            // conductance: InlineDeviceFunctions.hpp:22-51; current (kind 2): :52-84, the
            // trace itself instead of g (E - V), the slot it reads cleared
            R total = (R)0;
            for (int l = 0; l < L; ++l) {
              Acc a = ring[slot * NL + (size_t)i * L + l];
              if (l == 0) a += ext;
              ring[slot * NL + (size_t)i * L + l] = (Acc)0;
              R in;
              if constexpr (kFixed) in = (float)(double)a * inv; else in = (R)a;
              R gl = (R)decay_g[l] * gcond[(size_t)i * L + l] + in;
              if (kind == 1) total += gl * ((R)Erev[l] - Vi);
              else total += gl;
              gcond[(size_t)i * L + l] = gl;
            }
            dV = total * (R)Rp[i];
          }
          float tg = t + (float)g * dt;
          if (((tg - last_n[i]) - tref) > 0.5f * dt) {                 // :168
            R cc = (R)c[i];
            Vi = ((cc * (R)v0[i] + ((R)1 - cc) * Vi) + cc * (R)Ibg) + dV;  // :169
            if (Vi >= (R)thr[i]) {                                     // :173
              set_bit(hist_n, N, W, i, step, true);
              last_n[i] = tg;
              Vi = (R)vreset[i];
              part[th][g].push_back(i);
            }
          }
        }
        V[i] = Vi;
      }
    });
    for (int g = 0; g < G; ++g)
      for (int th = 0; th < T; ++th)
        for (int32_t x : part[th][g]) spikes[g].push_back(x);
    for (int64_t e : ext_counts) cnt_ext += e;
    for (int g = 0; g < G; ++g) cnt_n += (int64_t)spikes[g].size();
  }

  // GeneratorInputSpikingNeurons.cu:71-128: entry e fires input ids[e] at sub-step g when
  // |(t - onset + g*dt) - times[e]| < 0.5*dt (float difference, compared in double); an
  // input spikes at most once per step, entries taken in time order
  void update_generator(float t) {
    if (spikes.size() != (size_t)G) spikes.assign(G, {});
    const int tidx = (int)std::roundf(t / dt);
    for (int i = 0; i < Nin; ++i)
      for (int g = 0; g < G; ++g) set_bit(hist_in, Nin, W, i, (int64_t)tidx + g, false);
    const float base = t - gen_onset;
    for (int64_t e = gen_off[stim]; e < gen_off[stim + 1]; ++e) {
      const int id = gen_ids[e];
      for (int g = 0; g < G; ++g) {
        const float d = (base + (float)g * dt) - gen_times[e];
        if (!((double)std::fabs(d) < 0.5 * (double)dt)) continue;
        const int64_t step = (int64_t)tidx + g;
        if (get_bit(hist_in, Nin, W, id, step)) continue;
        set_bit(hist_in, Nin, W, id, step, true);
        last_in[id] = t + (float)g * dt;
        spikes[g].push_back(N + id);
        cnt_in++;
      }
    }
  }

  // PoissonInputSpikingNeurons.cu:73-130 in its Bernoulli form
  void update_inputs(float t) {
    if (gen) { update_generator(t); return; }
    if (spikes.size() != (size_t)G) spikes.assign(G, {});
    int tidx = (int)std::roundf(t / dt);
    if (pending_arm) { arm_step = tidx; pending_arm = false; }
    const uint32_t* thr_s = in_thr.data() + (size_t)stim * Nin;
    for (int i = 0; i < Nin; ++i) {
      DrawCursor in_draw(seed, (uint32_t)i, STREAM_INPUT);
      for (int g = 0; g < G; ++g) {
        int64_t step = (int64_t)tidx + g;
        float tg = t + (float)g * dt;
        set_bit(hist_in, Nin, W, i, step, false);
        uint32_t u = in_draw(step);
        if (u < thr_s[i] && step != arm_step) {
          set_bit(hist_in, Nin, W, i, step, true);
          last_in[i] = tg;
          spikes[g].push_back(N + i);
          cnt_in++;
        }
      }
    }
  }

  // SpikingSynapses.cu:168-202, restated per target range (owner computes)
  void update_synapses(float t) {
    int tidx = (int)std::roundf(t / dt);
    const size_t NL = (size_t)N * L;
    int T = (int)rp.size();
    std::vector<int64_t> ev(T, 0);
    parallel(T, [&](int, int a, int b) {
      for (int th = a; th < b; ++th) {
        int64_t e = 0;
        for (int g = 0; g < (int)spikes.size(); ++g)
          for (int32_t u : spikes[g]) {
            for (int64_t k = rp[th][u]; k < rp[th][u + 1]; ++k) {
              int64_t s = rs[th][k];
              int slot = (tidx % B + sdelay[s] + g) % B;  // :196
              R wv;
              if constexpr (kFixed) wv = (float)sw[s] * sscale[s];
              else wv = sw[s] * (R)sscale[s];
              ring[(size_t)slot * NL + (size_t)spost[s] * L + slabel[s]] += to_acc(wv, frac);
              ++e;
            }
          }
        ev[th] = e;
      }
    });
    for (int64_t e : ev) cnt_events += e;
  }

  bool pre_bit(int32_t pre, int64_t step) const {
    if (pre >= 0) return get_bit(hist_n, N, W, pre, step);
    return get_bit(hist_in, Nin, W, -pre - 1, step);
  }
  float pre_last(int32_t pre) const { return pre >= 0 ? last_n[pre] : last_in[-pre - 1]; }

  void update_plasticity(int rid, float t) {
    Rule& r = rules[rid];
    int tidx = (int)std::roundf(t / dt);
    int64_t P = (int64_t)r.syn.size();
    if (r.kind == 0) {
      // WeightDependentSTDPPlasticity.cu:45-163; decays precomputed in float as the
      // host-side expressions -(k*dt)/tau.
      const float am = r.wd[0], ap = r.wd[1], tm = r.wd[2], tp = r.wd[3];
      const float lam = r.wd[4], alp = r.wd[5], wmax = r.wd[6];
      std::vector<float> Em(G + 1), Ep(G + 1), iEm(G + 1), iEp(G + 1);
      for (int k = 0; k <= G; ++k) {
        Em[k] = std::exp(-((float)k * dt) / tm);
        Ep[k] = std::exp(-((float)k * dt) / tp);
        iEm[k] = 1.0f / Em[k];
        iEp[k] = 1.0f / Ep[k];
      }
      // every rule synapse is updated from its own state only: split over threads
      parallel64(P, [&](int64_t jlo, int64_t jhi) {
      for (int64_t j = jlo; j < jhi; ++j) {
        int64_t s = r.syn[j];
        R xpre = (R)r.ta[j], xpost = (R)r.tb[j];
        int pg = (int)std::roundf((pre_last(spre[s]) - t) / dt);
        // qg is below G while time advances; bounded at G like the device's table index
        int qg = std::min((int)std::roundf((last_n[spost[s]] - t) / dt), G);
        if (pg >= G) pg *= -1;
        R w = sw[s];
        xpost *= (R)Em[G];
        xpre *= (R)Ep[G];
        if (r.nearest) {
          if (qg >= 0) xpost = (R)am * (R)Em[G - qg];
          if (pg >= 0) xpre = (R)ap * (R)Ep[G - pg];
        } else {
          xpost += (qg >= 0) ? (R)am * (R)Em[G - qg] : (R)0;
          xpre += (pg >= 0) ? (R)ap * (R)Ep[G - pg] : (R)0;
        }
        R upd = (R)0;
        if (pg >= 0) {
          R tmp = xpost;
          tmp += (qg > pg) ? -(R)am * (R)Em[G - qg] : (R)0;
          tmp *= (R)iEm[G - pg];
          upd -= (((R)lam * (R)alp) * w) * tmp;
        }
        if (qg >= 0) {
          R tmp = xpre;
          tmp += (pg > qg) ? -(R)ap * (R)Ep[G - pg] : (R)0;
          tmp *= (R)iEp[G - qg];
          upd += ((R)lam * ((R)wmax - w)) * tmp;
        }
        R nw = w + upd;
        if (nw < (R)0) nw = (R)0;
        wstore(s, w, nw, 0.0f, INFINITY);
        r.ta[j] = tstore(xpre); r.tb[j] = tstore(xpost);
      }
      });
    } else if (r.kind == 1) {
      // InhibitorySTDPPlasticity.cu:72-155 (the target term is evaluated in double)
      const float decay = std::exp(-dt / r.vg[0]);
      const float lr = r.vg[1];
      const double cterm = (double)lr * (2.0 * (double)r.vg[2] * (double)r.vg[0]);
      for (int64_t j = 0; j < P; ++j) {
        int64_t s = r.syn[j];
        R xpre = (R)r.ta[j], xpost = (R)r.tb[j];
        const R w0 = sw[s];
        R w = w0;
        for (int g = 0; g < G; ++g) {
          xpre *= (R)decay; xpost *= (R)decay;
          int64_t step = (int64_t)tidx + g;
          bool pre = pre_bit(spre[s], step - sdelay[s]);
          bool post = get_bit(hist_n, N, W, spost[s], step);
          if (pre) xpre += (R)1;
          if (post) xpost += (R)1;
          R upd = (R)0;
          if (pre) {
            upd += (R)lr * xpost;
            if constexpr (kFixed) upd = (float)((double)upd + (-cterm));
            else upd = upd + (-cterm);
          }
          if (post) upd += (R)lr * xpre;
          w += upd;
          if (w < (R)0) w = (R)0;
        }
        wstore(s, w0, w, 0.0f, INFINITY);
        r.ta[j] = tstore(xpre); r.tb[j] = tstore(xpost);
      }
    } else if (r.kind == 4) {
      // WeightNormSTDPPlasticity.cu:68-172: per neuron of the set, u = sum of
      // d*d + 2*w0*d (d = w - w0 != 0) over its synapses in the caller's order, then
      // w /= sqrtf(sum_sq + u) / sqrtf(sum_sq) unless that is 1
      std::vector<R> u(N, (R)0);
      for (int64_t j = 0; j < P; ++j) {
        const int64_t s = r.syn[j];
        const R w0 = (R)r.wn_w0[j];
        const R d = sw[s] - w0;
        if (d != (R)0) u[spost[s]] += d * d + ((R)2 * w0) * d;
      }
      std::vector<R> dv(N, (R)1);
      for (int v = 0; v < N; ++v)
        if (r.wn_inset[v]) dv[v] = std::sqrt((R)r.wn_sumsq[v] + u[v]) / std::sqrt((R)r.wn_sumsq[v]);
      for (int64_t j = 0; j < P; ++j) {
        const int64_t s = r.syn[j];
        const int v = spost[s];
        if (!r.wn_inset[v] || dv[v] == (R)1) continue;
        const R w = sw[s];
        wstore(s, w, w / dv[v], -INFINITY, INFINITY);
      }
    } else if (r.kind == 3) {
      // CustomSTDPPlasticity.cu:67-155: per sub-step, traces decayed (post by tau_minus,
      // pre by tau_plus) then set or raised by the delayed pre / the post spike; a pre spike
      // depresses by lr*(w/w_max)*x_post + lr*a_star, a post spike potentiates by lr*x_pre;
      // weights clamped at 0 only.  Decays computed in float on the host.
      const float am = r.cu[0], ap = r.cu[1], ast = r.cu[2], wmax = r.cu[5], lr = r.cu[6];
      const float post_decay = std::exp(-dt / r.cu[3]);
      const float pre_decay = std::exp(-dt / r.cu[4]);
      for (int64_t j = 0; j < P; ++j) {
        int64_t s = r.syn[j];
        R xpre = (R)r.ta[j], xpost = (R)r.tb[j];
        const R w0 = sw[s];
        R w = w0;
        for (int g = 0; g < G; ++g) {
          xpost *= (R)post_decay;
          xpre *= (R)pre_decay;
          int64_t step = (int64_t)tidx + g;
          bool pre = pre_bit(spre[s], step - sdelay[s]);
          bool post = get_bit(hist_n, N, W, spost[s], step);
          if (pre) { xpre += (R)ap; if (r.nearest) xpre = (R)ap; }
          if (post) { xpost += (R)am; if (r.nearest) xpost = (R)am; }
          R upd = (R)0;
          const R wo = w;
          if (pre) upd -= (((R)lr * (wo / (R)wmax)) * xpost + (R)lr * (R)ast);
          if (post) upd += (R)lr * xpre;
          w = wo + upd;
          if (w < (R)0) w = (R)0;
        }
        wstore(s, w0, w, 0.0f, INFINITY);
        r.ta[j] = tstore(xpre); r.tb[j] = tstore(xpost);
      }
    } else {
      // EvansSTDPPlasticity.cu:59-143
      const float dC = 1.0f - (dt / r.ev[3]);
      const float dD = 1.0f - (dt / r.ev[0]);
      const float aC = dt * r.ev[2];
      const float aD = dt * r.ev[1];
      for (int64_t j = 0; j < P; ++j) {
        int64_t s = r.syn[j];
        R C = (R)r.ta[j], D = (R)r.tb[j];
        const R w0 = sw[s];
        R w = w0;
        for (int g = 0; g < G; ++g) {
          C = (R)dC * C;
          D = (R)dD * D;
          int64_t step = (int64_t)tidx + g;
          bool pre = pre_bit(spre[s], step - sdelay[s]);
          bool post = get_bit(hist_n, N, W, spost[s], step);
          if (pre) C += (R)aC * ((R)1 - C);
          if (post) D += (R)aD * ((R)1 - D);
          R upd = (R)0;
          R wo = w;
          if (pre) upd -= wo * D;
          if (post) upd += ((R)1 - wo) * C;
          w = wo + upd;
          if (w < (R)0) w = (R)0;
          if (w > (R)1) w = (R)1;
        }
        wstore(s, w0, w, 0.0f, 1.0f);
        r.ta[j] = tstore(C); r.tb[j] = tstore(D);
      }
    }
  }

  // SpikingActivityMonitor.cu:69-98: (id, t + g*dt) of every spike of the group
  void collect(int mid, float t) {
    Monitor& m = mons[mid];
    for (int g = 0; g < (int)spikes.size(); ++g) {
      float tg = t + (float)g * dt;
      std::vector<int32_t> ids;
      for (int32_t u : spikes[g]) {
        if (m.pop == 0 && u < N) ids.push_back(u);
        if (m.pop == 1 && u >= N) ids.push_back(u - N);
      }
      std::sort(ids.begin(), ids.end());
      for (int32_t x : ids) { m.ids.push_back(x); m.times.push_back(tg); }
    }
  }

  // SpikingModel.cpp:197-236
  void run(float start, int64_t ngroups, int flags) {
    for (int64_t s = 0; s < ngroups; ++s) {
      float t = start + (float)(s * G) * dt;
      update_neurons(t);
      update_inputs(t);
      if (flags & 1)
        for (int r = 0; r < (int)rules.size(); ++r) update_plasticity(r, t);
      update_synapses(t);
      for (int m = 0; m < (int)mons.size(); ++m) collect(m, t);
    }
  }
};

struct Handle {
  int prec;
  Net<float>* f = nullptr;
  Net<double>* d = nullptr;
};

template <typename Fn>
int dispatch(Handle* h, Fn fn) {
  if (!h) return fail("null handle");
  if (h->prec == 32) return fn(*h->f);
  return fn(*h->d);
}

}  // namespace

// ---------------------------------------------------------------------------------
// C interface (mirrors include/spike_hip.h with an `or_` prefix; readouts in double)
extern "C" {

const char* or_last_error() { return g_err.c_str(); }

void* or_create(int n_neurons, int n_inputs, int kind, int n_labels, float dt, int precision) {
  if (n_neurons < 0 || n_inputs < 0 || n_labels < 1 || !(dt > 0.0f)) { fail("bad sizes"); return nullptr; }
  Handle* h = new Handle();
  h->prec = precision == 64 ? 64 : 32;
  auto init = [&](auto* n) {
    n->N = n_neurons; n->Nin = n_inputs; n->kind = kind; n->L = n_labels; n->dt = dt;
    n->decay_g.assign(n_labels, 1.0f); n->Erev.assign(n_labels, 0.0f);
  };
  if (h->prec == 32) { h->f = new Net<float>(); init(h->f); }
  else { h->d = new Net<double>(); init(h->d); }
  return h;
}

int or_destroy(void* p) {
  Handle* h = (Handle*)p;
  if (!h) return 0;
  delete h->f; delete h->d; delete h;
  return 0;
}

int or_set_threads(void* p, int n) {
  return dispatch((Handle*)p, [&](auto& net) { net.nthreads = std::max(1, n); return 0; });
}

int or_set_lif(void* p, const float* v0, const float* vr, const float* th, const float* c,
               const float* rp, float tref, float ibg) {
  return dispatch((Handle*)p, [&](auto& n) {
    n.v0.assign(v0, v0 + n.N); n.vreset.assign(vr, vr + n.N); n.thr.assign(th, th + n.N);
    n.c.assign(c, c + n.N); n.Rp.assign(rp, rp + n.N); n.tref = tref; n.Ibg = ibg;
    n.lif_set = true;
    return 0;
  });
}

int or_set_poisson_rates(void* p, const float* rates, int n_stim) {
  return dispatch((Handle*)p, [&](auto& n) {
    n.rates.assign(rates, rates + (size_t)n_stim * n.Nin); n.n_stim = n_stim; n.gen = false;
    return 0;
  });
}

int or_set_generator_stimuli(void* p, const int32_t* counts, int n_stim, const int32_t* ids,
                             const float* times) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (n_stim < 1) return fail("bad generator stimuli");
    n.gen_off.assign(1, 0); n.gen_ids.clear(); n.gen_times.clear();
    int64_t o = 0;
    for (int k = 0; k < n_stim; ++k) {
      std::vector<int64_t> idx(counts[k]);
      for (int j = 0; j < counts[k]; ++j) idx[j] = o + j;
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return times[x] < times[y]; });
      for (int64_t e : idx) {
        if (ids[e] < 0 || ids[e] >= n.Nin) return fail("stimulus neuron id out of range");
        n.gen_ids.push_back(ids[e]);
        n.gen_times.push_back(times[e]);
      }
      o += counts[k];
      n.gen_off.push_back(o);
    }
    n.gen = true; n.n_stim = n_stim; n.stim = 0;
    n.rates.assign((size_t)n_stim * n.Nin, 0.0f);
    return 0;
  });
}

int or_set_generator_onset(void* p, float onset) {
  return dispatch((Handle*)p, [&](auto& n) { n.gen_onset = onset; return 0; });
}

int or_select_stimulus(void* p, int s) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (s < 0 || s >= std::max(1, n.n_stim)) return fail("bad stimulus");
    n.stim = s;
    return n.prepared ? n.reset_inputs() : 0;
  });
}

int or_set_seed(void* p, uint64_t seed) {
  return dispatch((Handle*)p, [&](auto& n) { n.seed = seed; return 0; });
}

int or_set_weight_format(void* p, int fmt) {
  if (fmt < 0 || fmt > 2) return fail("weight format must be 0, 1 or 2");
  return dispatch((Handle*)p, [&](auto& n) { n.wfmt = fmt; return 0; });
}

uint16_t or_round16_bits(float x, int fmt) { return fmt == 1 ? o_to_bf16(x) : o_to_half(x); }
float or_decode16(uint16_t h, int fmt) { return fmt == 1 ? o_from_bf16(h) : o_from_half(h); }

int or_set_external_drive(void* p, float rate, float w) {
  return dispatch((Handle*)p, [&](auto& n) { n.ext_rate = rate; n.ext_w = w; return 0; });
}

int or_set_labels(void* p, const float* decay, const float* rev) {
  return dispatch((Handle*)p, [&](auto& n) {
    n.decay_g.assign(decay, decay + n.L); n.Erev.assign(rev, rev + n.L); return 0;
  });
}

int or_set_synapses(void* p, int64_t S, const int32_t* pre, const int32_t* post,
                    const float* w, const float* scale, const int32_t* delay,
                    const int32_t* label) {
  return dispatch((Handle*)p, [&](auto& n) {
    using R = typename std::remove_reference<decltype(n)>::type::Acc;
    (void)sizeof(R);
    n.spre.assign(pre, pre + S); n.spost.assign(post, post + S);
    n.sdelay.assign(delay, delay + S);
    n.slabel.assign(S, 0);
    if (label) n.slabel.assign(label, label + S);
    n.sw.resize(S);
    for (int64_t s = 0; s < S; ++s) n.sw[s] = w[s];
    n.sscale.assign(S, 1.0f);
    if (scale) n.sscale.assign(scale, scale + S);
    for (int64_t s = 0; s < S; ++s) {
      if (post[s] < 0 || post[s] >= n.N) return fail("post out of range");
      if (pre[s] >= n.N || (pre[s] < 0 && -pre[s] - 1 >= n.Nin)) return fail("pre out of range");
      if (delay[s] < 1) return fail("delay < 1");
      if (n.slabel[s] < 0 || n.slabel[s] >= n.L) return fail("label out of range");
    }
    return 0;
  });
}

static int add_rule(void* p, int kind, const int64_t* syn, int64_t n, const float* par, int npar,
                    int nearest) {
  return dispatch((Handle*)p, [&](auto& net) {
    Rule r;
    r.kind = kind;
    r.syn.assign(syn, syn + n);
    for (int64_t s : r.syn) if (s < 0 || s >= (int64_t)net.spre.size()) return fail("bad synapse index");
    if (kind == 0) std::memcpy(r.wd, par, sizeof(float) * npar);
    if (kind == 1) std::memcpy(r.vg, par, sizeof(float) * npar);
    if (kind == 2) std::memcpy(r.ev, par, sizeof(float) * npar);
    if (kind == 3) std::memcpy(r.cu, par, sizeof(float) * npar);
    r.nearest = nearest;
    net.rules.push_back(r);
    return 0;
  });
}

// par: a_minus, a_plus, tau_minus, tau_plus, lambda, alpha, w_max
int or_add_wd_stdp(void* p, const int64_t* syn, int64_t n, const float* par, int nearest) {
  return add_rule(p, 0, syn, n, par, 7, nearest);
}
// par: tau_istdp, learningrate, targetrate, momentumrate
int or_add_vogels_stdp(void* p, const int64_t* syn, int64_t n, const float* par) {
  return add_rule(p, 1, syn, n, par, 4, 0);
}
// par: tau_D, alpha_D, alpha_C, tau_C, rho
int or_add_evans_stdp(void* p, const int64_t* syn, int64_t n, const float* par) {
  return add_rule(p, 2, syn, n, par, 5, 0);
}
// par: a_minus, a_plus, a_star, tau_minus, tau_plus, w_max, learning_rate
int or_add_custom_stdp(void* p, const int64_t* syn, int64_t n, const float* par, int nearest) {
  return add_rule(p, 3, syn, n, par, 7, nearest);
}

int or_add_weightnorm(void* p, const int64_t* syn, int64_t n, const float* sum_sq, const uint8_t* in_set) {
  const float none[1] = {0.0f};
  int rc = add_rule(p, 4, syn, n, none, 0, 0);
  if (rc) return rc;
  return dispatch((Handle*)p, [&](auto& net) {
    Rule& r = net.rules.back();
    r.wn_sumsq.assign(sum_sq, sum_sq + net.N);
    r.wn_inset.assign(in_set, in_set + net.N);
    return 0;
  });
}

int or_weightnorm_set_initial(void* p, int rid, const float* w0) {
  return dispatch((Handle*)p, [&](auto& net) {
    if (rid < 0 || rid >= (int)net.rules.size() || net.rules[rid].kind != 4) return fail("not a weight normalization rule");
    Rule& r = net.rules[rid];
    for (size_t j = 0; j < r.syn.size(); ++j) r.wn_w0[j] = (double)w0[j];
    return 0;
  });
}

int or_add_monitor(void* p, int pop) {
  return dispatch((Handle*)p, [&](auto& n) {
    Monitor m; m.pop = pop; n.mons.push_back(m); return (int)n.mons.size() - 1;
  });
}

int or_prepare(void* p, int G) {
  return dispatch((Handle*)p, [&](auto& n) { return n.prepare(G); });
}

int or_set_timestep_grouping(void* p, int G) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (G < 1 || G > n.Gprep) return fail("bad grouping");
    n.G = G; return 0;
  });
}

// info: G, B, W, frac, dmin, dmax
int or_get_info(void* p, int* info) {
  return dispatch((Handle*)p, [&](auto& n) {
    info[0] = n.G; info[1] = n.B; info[2] = n.W; info[3] = n.frac; info[4] = n.dmin;
    info[5] = n.dmax; info[6] = (int)n.ext_cdf.size();
    return 0;
  });
}

int or_reset_state(void* p) { return dispatch((Handle*)p, [&](auto& n) { return n.reset_state(); }); }
int or_reset_inputs(void* p) { return dispatch((Handle*)p, [&](auto& n) { return n.reset_inputs(); }); }

int or_update_neurons(void* p, float t) { return dispatch((Handle*)p, [&](auto& n) { n.update_neurons(t); return 0; }); }
int or_update_inputs(void* p, float t) { return dispatch((Handle*)p, [&](auto& n) { n.update_inputs(t); return 0; }); }
int or_update_plasticity(void* p, int r, float t) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (r < 0 || r >= (int)n.rules.size()) return fail("bad rule");
    n.update_plasticity(r, t); return 0;
  });
}
int or_update_synapses(void* p, float t) { return dispatch((Handle*)p, [&](auto& n) { n.update_synapses(t); return 0; }); }
int or_collect_spikes(void* p, int m, float t) { return dispatch((Handle*)p, [&](auto& n) { n.collect(m, t); return 0; }); }

int or_run(void* p, float start, int64_t ngroups, int flags) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (!n.prepared) return fail("not prepared");
    n.run(start, ngroups, flags); return 0;
  });
}

// counters: spikes_neurons, spikes_inputs, synaptic_events, external_events
int or_get_counters(void* p, int64_t* c) {
  return dispatch((Handle*)p, [&](auto& n) {
    c[0] = n.cnt_n; c[1] = n.cnt_in; c[2] = n.cnt_events; c[3] = n.cnt_ext; return 0;
  });
}

int64_t or_monitor_count(void* p, int m) {
  Handle* h = (Handle*)p;
  if (h->prec == 32) return (int64_t)h->f->mons[m].ids.size();
  return (int64_t)h->d->mons[m].ids.size();
}

int or_monitor_take(void* p, int m, int32_t* ids, float* times) {
  return dispatch((Handle*)p, [&](auto& n) {
    Monitor& mo = n.mons[m];
    std::copy(mo.ids.begin(), mo.ids.end(), ids);
    std::copy(mo.times.begin(), mo.times.end(), times);
    mo.ids.clear(); mo.times.clear();
    return 0;
  });
}

int or_get_weights(void* p, double* w) {
  return dispatch((Handle*)p, [&](auto& n) {
    for (size_t s = 0; s < n.sw.size(); ++s) w[s] = (double)n.sw[s];
    return 0;
  });
}

int or_set_weights(void* p, const double* w) {
  return dispatch((Handle*)p, [&](auto& n) {
    using RT = typename std::remove_reference<decltype(n.sw[0])>::type;
    for (size_t s = 0; s < n.sw.size(); ++s) n.sw[s] = (RT)w[s];
    if constexpr (std::is_same<RT, float>::value) {
      if (n.wfmt && n.prepared) {
        for (size_t s = 0; s < n.sw.size(); ++s) n.sw[s] = o_round16((float)w[s], n.wfmt);
        n.sc.assign(n.sw.size(), 0.0f);
      }
    }
    return 0;
  });
}

int or_get_membrane(void* p, double* v) {
  return dispatch((Handle*)p, [&](auto& n) {
    for (int i = 0; i < n.N; ++i) v[i] = (double)n.V[i];
    return 0;
  });
}

int or_get_last_spike(void* p, int pop, float* out) {
  return dispatch((Handle*)p, [&](auto& n) {
    const auto& src = pop == 0 ? n.last_n : n.last_in;
    std::copy(src.begin(), src.end(), out);
    return 0;
  });
}

int or_get_conductances(void* p, double* g) {
  return dispatch((Handle*)p, [&](auto& n) {
    for (size_t i = 0; i < n.gcond.size(); ++i) g[i] = (double)n.gcond[i];
    return 0;
  });
}

// fixed-point ring (FP32 oracle) or double ring (FP64 oracle)
int or_get_ring_fixed(void* p, int64_t* out) {
  Handle* h = (Handle*)p;
  if (!h || h->prec != 32) return fail("fixed ring only in the FP32 oracle");
  std::copy(h->f->ring.begin(), h->f->ring.end(), out);
  return 0;
}

int or_get_ring_real(void* p, double* out) {
  return dispatch((Handle*)p, [&](auto& n) {
    for (size_t i = 0; i < n.ring.size(); ++i) out[i] = (double)n.ring[i];
    return 0;
  });
}

int or_get_history(void* p, int pop, uint32_t* out) {
  return dispatch((Handle*)p, [&](auto& n) {
    const auto& src = pop == 0 ? n.hist_n : n.hist_in;
    std::copy(src.begin(), src.end(), out);
    return 0;
  });
}

int or_get_traces(void* p, int r, double* a, double* b) {
  return dispatch((Handle*)p, [&](auto& n) {
    if (r < 0 || r >= (int)n.rules.size()) return fail("bad rule");
    if (n.rules[r].kind == 4) {   // weight normalization: its initial weights, and zeros
      std::copy(n.rules[r].wn_w0.begin(), n.rules[r].wn_w0.end(), a);
      std::fill(b, b + n.rules[r].syn.size(), 0.0);
      return 0;
    }
    std::copy(n.rules[r].ta.begin(), n.rules[r].ta.end(), a);
    std::copy(n.rules[r].tb.begin(), n.rules[r].tb.end(), b);
    return 0;
  });
}

// Philox block for known-answer tests
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  Philox::block(ctr, key, out);
}

uint32_t or_bernoulli_threshold(float rate, float dt) { return bernoulli_threshold(rate, dt); }

int or_poisson_count_table(double lambda, uint32_t* out, int cap) {
  std::vector<uint32_t> t = poisson_count_table(lambda);
  for (int i = 0; i < (int)t.size() && i < cap; ++i) out[i] = t[i];
  return (int)t.size();
}

int or_choose_frac_bits(double m) { return choose_frac_bits(m); }

// Fixed in-degree Brunel connectivity drawn from Philox (the rule the library generates on
// the device): neuron i, slot j < ce + ci draws u = output j mod 4 of
// philox(ctr = {j/4, 0, i, 3}, key = seed);
// the source is floor(u*ne/2^32) for j < ce, else ne + floor(u*ni/2^32).  pre[i*(ce+ci)+j].
int or_brunel_connectivity(uint64_t seed, int ne, int ni, int ce, int ci, int n, int32_t* pre) {
  const int C = ce + ci;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < C; ++j) {
      const uint32_t u = Philox::draw(seed, (int64_t)j, (uint32_t)i, 3u);
      pre[(size_t)i * C + j] = j < ce ? (int32_t)(((uint64_t)u * (uint64_t)ne) >> 32)
                                      : (int32_t)(ne + (((uint64_t)u * (uint64_t)ni) >> 32));
    }
  return 0;
}

// Fixed out-degree connectivity drawn from Philox (the rule the library generates on the
// device for large networks): unit u (neurons [0, n), then inputs) has k (k_in) synapses;
// synapse j of unit u takes r = philox(ctr = {j, 0, u, 5}, key = seed) and gets
//   post = floor(r0*n/2^32), delay = 1 + floor(r1*dmax/2^32),
//   weight = (r2 >> 8) * 2^-24 for u < ne, else w_inh (neurons) / w_in (inputs),
//   scale = s_e / s_i / s_in by the same classes,
// at index row(u) + j, row(u) = u*k for u <= n, then k_in per input; pre is u for a
// neuron and -(u-n)-1 for input u-n.  Rows are split over `threads` threads.
int or_outdegree_connectivity(uint64_t seed, int n, int n_in, int ne, int k, int k_in, int dmax,
                              float s_e, float s_i, float s_in, float w_inh, float w_in, int threads,
                              int32_t* pre, int32_t* post, float* weight, float* scale, int32_t* delay) {
  const int U = n + n_in;
  auto row = [&](int u) { return u <= n ? (int64_t)u * k : (int64_t)n * k + (int64_t)(u - n) * k_in; };
  auto rows = [&](int u0, int u1) {
    for (int u = u0; u < u1; ++u) {
      const int len = u < n ? k : k_in;
      const int64_t base = row(u);
      for (int j = 0; j < len; ++j) {
        uint32_t r[4];
        Philox::draws4(seed, (int64_t)j, (uint32_t)u, 5u, r);
        const int64_t s = base + j;
        pre[s] = u < n ? u : -(u - n) - 1;
        post[s] = (int32_t)(((uint64_t)r[0] * (uint64_t)n) >> 32);
        delay[s] = 1 + (int32_t)(((uint64_t)r[1] * (uint64_t)dmax) >> 32);
        if (u < ne) {
          weight[s] = (float)(r[2] >> 8) * 0x1p-24f;
          scale[s] = s_e;
        } else if (u < n) {
          weight[s] = w_inh;
          scale[s] = s_i;
        } else {
          weight[s] = w_in;
          scale[s] = s_in;
        }
      }
    }
  };
  const int T = std::max(1, std::min(threads, std::max(1, U)));
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back(rows, (int)((int64_t)U * t / T), (int)((int64_t)U * (t + 1) / T));
  for (auto& x : th) x.join();
  return 0;
}

}  // extern "C"
