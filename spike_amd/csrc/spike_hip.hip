// spike_hip.hip -- C ABI of the MI355X Spike backend (declared in include/spike_hip.h).
//
// A network object holds everything one SpikingModel needs on one GPU: SoA neuron state,
// the int64 fixed-point input ring, the delay-sorted CSR of outgoing synapses, per-group
// spike lists, plasticity traces and monitor stores.  Host code here only validates,
// lays out and uploads data, and sequences kernels on the network's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spike_hip.h"
#include "kernels.hpp"
#include "fused.hpp"
#include "gen.hpp"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define SPK_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SPK_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(SPK_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return SPK_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Bernoulli threshold of a Poisson input with rate r: floor((1 - exp(-r*dt)) * 2^32).
uint32_t bernoulli_threshold(float rate, float dt) {
  double p = -std::expm1(-(double)rate * (double)dt);
  if (!(p > 0.0)) return 0u;
  double t = std::floor(p * 4294967296.0);
  if (t >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

// Inverse-CDF thresholds of a Poisson(lambda) count, C[k] = floor(F(k) 2^32), until the
// tail is below 2^-34.
std::vector<uint32_t> poisson_count_table(double lambda) {
  std::vector<uint32_t> c;
  if (!(lambda > 0.0)) return c;
  double pmf = std::exp(-lambda), cdf = pmf;
  for (int k = 0; k < 255; ++k) {
    double t = std::floor(cdf * 4294967296.0);
    c.push_back(t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t);
    if (1.0 - cdf < 5.820766091346741e-11) break;
    pmf = pmf * lambda / (double)(k + 1);
    cdf += pmf;
  }
  return c;
}

// Spikes one LIF neuron can emit in a group of G steps: consecutive spikes are at least
// max(1, round(tref/dt)) steps apart (the refractory test ((t - last) - tref) > dt/2 needs
// a gap above tref/dt + 1/2 steps; one step less covers float rounding).
int max_spikes_per_group(float tref, float dt, int G) {
  const double r = std::floor((double)tref / (double)dt + 0.5);
  const int gap = (int)std::max(1.0, std::min(r, (double)G));
  return std::min(G, (G + gap - 1) / gap);
}

// Fraction bits of the fixed-point ring: the largest possible per-step input of one
// (neuron, label) entry stays below 2^52, leaving 11 bits of headroom in an int64.
int choose_frac_bits(double max_entry) {
  if (!(max_entry > 0.0)) return 32;
  int e = (int)std::ceil(std::log2(max_entry));
  int k = 52 - e;
  return std::max(0, std::min(120, k));
}

// Host arrays of synapse size: resize leaves elements uninitialised, so that the first
// touch of their pages happens in the threaded passes that fill them.
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind { using other = NoInit<U>; };
  template <typename U>
  void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) { ::new (static_cast<void*>(p)) U(std::forward<A>(a)...); }
};
template <typename T>
using hvec = std::vector<T, NoInit<T>>;

struct RuleHost {
  int kind = 0;  // 0 WD, 1 Vogels, 2 Evans, 3 Custom
  std::vector<int64_t> syn;   // caller synapse indices (empty for a rule generated on the device)
  int64_t cnt = 0;            // synapses of the rule
  spk_wd_stdp_params wd{};
  spk_vogels_stdp_params vg{};
  spk_evans_stdp_params ev{};
  spk_custom_stdp_params cu{};
  int nearest = 0;   // nearest-spike traces (WD, Custom)
  // weight normalization (kind 4): the front-end's squared norms and neuron set, the
  // initial weights (in the per-synapse trace array a), rule indices per neuron
  std::vector<float> wn_sumsq;
  std::vector<uint8_t> wn_inset;
  float* d_wn_sumsq = nullptr;
  uint8_t* d_wn_inset = nullptr;
  int64_t* d_wn_ptr = nullptr;
  int32_t* d_wn_idx = nullptr;
  // device
  int64_t* d_pos = nullptr;
  int32_t* d_pre = nullptr;
  float* d_ta = nullptr;
  float* d_tb = nullptr;
  uint16_t* d_ta16 = nullptr;   // traces in the network's 16-bit weight format
  uint16_t* d_tb16 = nullptr;
  float* d_tab = nullptr;
  int tab_G = -1;
  double dterm = 0.0;
  hvec<int64_t> order;            // device order j -> index into syn (grouped by post partition)
  std::vector<int64_t> part;      // [P+1] offsets of each postsynaptic partition
  int64_t* d_part = nullptr;
  // spike-driven form of the weight-dependent rule (per-phase kernels): rule indices by
  // presynaptic and by postsynaptic neuron, traces per neuron, the listed neurons
  bool ev_ok = false;
  int64_t* d_ev_pre_ptr = nullptr;
  int32_t* d_ev_pre_idx = nullptr;   // nullptr: rule order already groups by presynaptic neuron
  int64_t* d_ev_post_ptr = nullptr;
  uint32_t* d_ev_post_pos = nullptr;  // CSR position of each synapse, grouped by postsynaptic neuron
  int32_t* d_ev_post_pre = nullptr;   // its presynaptic unit
  int64_t max_post = 0;               // longest list of one postsynaptic neuron
  float* d_xn_a = nullptr;           // stored per-neuron traces (FP32 or binary16)
  float* d_xn_b = nullptr;
  uint16_t* d_xn_a16 = nullptr;
  uint16_t* d_xn_b16 = nullptr;
  float* d_xcur_a = nullptr;         // this group's traces before rounding
  float* d_xcur_b = nullptr;
  uint32_t* d_act = nullptr;
  int* d_act_cnt = nullptr;
  int par = 0;
  // which trace form is current: per synapse (per-synapse kernels, the persistent kernel,
  // read-out) and/or per neuron (spike-driven kernels)
  bool syn_tr = true, neu_tr = true;
  // the next group runs the per-synapse form: it also clamps weights that were set
  // below 0, which the spike-driven form (touching only synapses with spikes) would not
  bool dense_next = true;
};

struct MonitorHost {
  int pop = 0;
  int64_t capacity = 0;
  int32_t* d_ids = nullptr;
  float* d_times = nullptr;
  unsigned long long* d_count = nullptr;
  std::vector<int32_t> h_ids;  // drained by spk_run
  std::vector<float> h_times;
};

}  // namespace

struct spk_network {
  int N = 0, Nin = 0, kind = 0, L = 1;
  float dt = 1e-4f;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  // host parameters
  std::vector<float> v0, vreset, thr, c, rp;
  float tref = 0.002f, ibg = 0.0f;
  bool lif_set = false;
  bool lif_uniform = false;   // every neuron has the same parameters
  std::vector<float> rates;
  int n_stim = 0, stim = 0;
  // generator inputs (scheduled spikes) instead of Poisson inputs: per stimulus the
  // entries sorted by time, concatenated; offsets [n_stim + 1]
  bool gen = false;
  float gen_onset = 0.0f;
  std::vector<int64_t> gen_off;
  std::vector<int32_t> gen_ids;
  std::vector<float> gen_times;
  int32_t* d_gen_ids = nullptr;
  float* d_gen_times = nullptr;
  uint64_t seed = 42;
  float ext_rate = 0.0f, ext_w = 0.0f;
  std::vector<float> decay_g, erev;
  hvec<int32_t> h_pre, h_post, h_delay, h_label;
  hvec<float> h_w, h_scale;
  bool syn_set = false;
  // fixed out-degree synapses generated on the device by spk_prepare (no host arrays;
  // caller synapse index = CSR position, rule order = CSR order)
  bool od = false;
  spk_outdegree_params odp{};
  std::vector<RuleHost> rules;
  std::vector<MonitorHost> mons;
  // derived sizes
  int G = 1, Gprep = 1, B = 1, W = 1, frac = 32, dmin = 0, dmax = 0;
  int BL = 1;   // ring slots the persistent kernel keeps in LDS: max(dmax, G)
  int64_t S = 0;
  bool prepared = false;
  bool uniform_scale = true;
  float scale0 = 1.0f;
  std::vector<uint32_t> ext_cdf;
  int64_t q_ext = 0;
  int64_t arm_step = -1;
  bool pending_arm = true;
  hvec<int64_t> perm;  // caller synapse index -> CSR position
  // device state
  float *d_V = nullptr, *d_v0 = nullptr, *d_vreset = nullptr, *d_thr = nullptr, *d_c = nullptr,
        *d_rp = nullptr;
  float *d_last_n = nullptr, *d_last_in = nullptr;
  uint32_t *d_hist_n = nullptr, *d_hist_in = nullptr;
  int64_t* d_ring = nullptr;
  float* d_g = nullptr;
  float *d_decay = nullptr, *d_erev = nullptr;
  uint32_t* d_in_thr = nullptr;
  uint32_t* d_ext_cdf = nullptr;
  int64_t* d_rowptr = nullptr;
  int32_t* d_tgt = nullptr;
  uint2* d_syn = nullptr;   // per synapse {packed local target/delay/label, weight bits}
  // compact form of the persistent kernel's delivery (FusedArgs::syn16): 16-bit local
  // targets; sources below csplit weigh cw_lo, the others cw_hi; one delay cdelay
  uint16_t* d_syn16 = nullptr;
  bool compact = false;
  // per-phase two-launch groups (k_units, k_deliver): bitmask of the CSR positions of a
  // spike-driven weight-dependent rule, and the longest source row
  uint32_t* d_wd_mask = nullptr;
  int64_t max_row = 0;
  // 16-bit weights (spk_set_weight_format): words and Kahan compensations in device order
  int wfmt = SPK_WEIGHTS_FP32;
  uint32_t* d_wc = nullptr;   // per synapse: 16-bit weight | compensation << 16
  float cw_lo = 0.0f, cw_hi = 0.0f;
  int csplit = 0, cdelay = 0;
  float* d_scale = nullptr;
  uint16_t* d_delay = nullptr;
  uint8_t* d_label = nullptr;
  // spike lists: [slots][G][cap], counts [slots][G]
  int slots = 1, cap = 1;
  uint32_t* d_lists = nullptr;
  int* d_counts = nullptr;
  float* d_slot_time = nullptr;
  int* d_slot_tidx = nullptr;
  unsigned long long* d_stats = nullptr;  // [0] events, [1] external events
  // fused persistent path
  int P = 1, nown_max = 0, iown_max = 0, xcap = 0;
  int req_partitions = 0;
  bool fused_ok = false;
  std::string fused_why;
  size_t fused_lds = 0;        // dynamic LDS of the persistent kernel
  size_t fused_lds_base = 0;   // the same without the LDS-resident delivery table
  size_t fused_lds_i64 = 0;    // the same with the int64 ring (after caller weights end the compact form)
  bool cring_ok = false;       // compact form: per-type in-degrees below 2^16 (count ring)
  int lseg_U = 0, lrec = 0;    // that table: source units, targets per partition (0: off)
  int cu_count = 0;
  std::vector<int> h_nlo, h_ilo;
  int *d_nlo = nullptr, *d_ilo = nullptr;
  uint32_t* d_seg = nullptr;
  uint32_t* d_xlist = nullptr;
  unsigned int* d_bar = nullptr;
  int* d_err = nullptr;
  unsigned int bar_gen = 0;      // barrier generations completed since the counters were cleared
  unsigned int xtag = 0;         // tag base of the next launch's tagged exchange words
  bool bar_clear = true;         // counters (and the error flag) must be cleared first
  bool rules_uploaded = false;   // device descriptors of the rules are current
  bool generated = false;        // connectivity generated on the device (persistent kernel only)
  size_t lds_attr[2] = {0, 0};  // dynamic LDS limit set for the non-plastic/plastic kernel
  const void* lds_fn[2] = {nullptr, nullptr};  // ... and the kernel it was set for
  // window statistics (spk_stats_begin)
  unsigned int *d_st_cnt = nullptr, *d_st_nisi = nullptr, *d_st_hist = nullptr;
  float* d_st_prev = nullptr;
  double *d_st_s1 = nullptr, *d_st_s2 = nullptr;
  spk::StatsOut* d_st_out = nullptr;
  bool st_on = false, st_pending = false;
  int st_bin = 10;
  int64_t st_t0 = 0, st_max = 0, st_steps = 0, st_nbins = 0;
  unsigned long long* d_prof = nullptr;
  spk::FusedRule* d_rules = nullptr;
  bool profiling = false;
  double prof_us[spk::kProfPhases] = {0};
  // per-phase bookkeeping
  int cur_tidx = -1;
  // start time of the latest group simulated since spk_reset_state: time only advances
  // (SpikingModel::run continues from the current time), so last-spike times never lie
  // ahead of a group, which the plasticity tables rely on
  float t_hi = -INFINITY;
  // rate monitors: one spike counter per neuron / input (shared by the monitors of a
  // population; each monitor keeps the counts at its last reset as a baseline)
  uint32_t* d_rate[2] = {nullptr, nullptr};
  std::vector<int> rate_pop;
  std::vector<std::vector<uint32_t>> rate_base;
  int cur_slot = 0;
  // timing
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  std::vector<hipEvent_t> ev_kernel;
  spk_timing timing{};

  int nu() const { return N + Nin; }
};

namespace {

// Weight k of the synapse records is at float index 2k+1.
float* weights_view(spk_network* n) { return reinterpret_cast<float*>(n->d_syn) + 1; }

spk::Weights weights_of(spk_network* n) {
  spk::Weights w{};
  w.w = weights_view(n);
  w.wc = n->d_wc;
  w.fmt = n->wfmt;
  return w;
}

spk::Traces traces_of(spk_network* n, RuleHost& r) {
  spk::Traces t{};
  // traces are binary16 in both 16-bit modes: a trace decays by about 0.5 % per 0.1 ms
  // step, near the spacing of bfloat16 values, which would make bfloat16 traces decay in
  // uneven (biased) steps; binary16's 11-bit significand keeps that error to 0.05 %
  t.a = r.d_ta; t.b = r.d_tb; t.a16 = r.d_ta16; t.b16 = r.d_tb16;
  t.fmt = n->wfmt ? SPK_WEIGHTS_FP16 : SPK_WEIGHTS_FP32;
  return t;
}

int check_net(const spk_network* n) {
  if (!n) return fail(SPK_ERR_ARG, "null network");
  return SPK_OK;
}

int check_prepared(const spk_network* n) {
  if (!n) return fail(SPK_ERR_ARG, "null network");
  if (!n->prepared) return fail(SPK_ERR_STATE, "network not prepared (call spk_prepare)");
  return SPK_OK;
}

// Entry points that need the host-built CSR (per-phase kernels, weight and trace access).
int check_listed(const spk_network* n) {
  if (int rc = check_prepared(n)) return rc;
  if (n->generated)
    return fail(SPK_ERR_STATE, "network generated on the device: persistent kernel only");
  return SPK_OK;
}

void free_device(spk_network* n) {
  dfree(n->d_V); dfree(n->d_v0); dfree(n->d_vreset); dfree(n->d_thr); dfree(n->d_c);
  dfree(n->d_rp); dfree(n->d_last_n); dfree(n->d_last_in); dfree(n->d_hist_n);
  dfree(n->d_hist_in); dfree(n->d_ring); dfree(n->d_g); dfree(n->d_decay); dfree(n->d_erev);
  dfree(n->d_in_thr); dfree(n->d_ext_cdf); dfree(n->d_gen_ids); dfree(n->d_gen_times);
  dfree(n->d_rate[0]); dfree(n->d_rate[1]); dfree(n->d_rowptr); dfree(n->d_tgt); dfree(n->d_syn); dfree(n->d_syn16);
  dfree(n->d_wc);
  dfree(n->d_scale); dfree(n->d_delay); dfree(n->d_label); dfree(n->d_lists);
  dfree(n->d_counts); dfree(n->d_slot_time); dfree(n->d_slot_tidx); dfree(n->d_stats);
  for (auto& r : n->rules) {
    dfree(r.d_pos); dfree(r.d_pre); dfree(r.d_ta); dfree(r.d_tb); dfree(r.d_tab); dfree(r.d_part);
    dfree(r.d_ta16); dfree(r.d_tb16);
    dfree(r.d_ev_pre_ptr); dfree(r.d_ev_pre_idx); dfree(r.d_ev_post_ptr); dfree(r.d_ev_post_pos);
    dfree(r.d_ev_post_pre);
    dfree(r.d_xn_a); dfree(r.d_xn_b); dfree(r.d_xn_a16); dfree(r.d_xn_b16);
    dfree(r.d_xcur_a); dfree(r.d_xcur_b); dfree(r.d_act); dfree(r.d_act_cnt);
    dfree(r.d_wn_sumsq); dfree(r.d_wn_inset); dfree(r.d_wn_ptr); dfree(r.d_wn_idx);
  }
  dfree(n->d_nlo); dfree(n->d_ilo); dfree(n->d_seg); dfree(n->d_xlist); dfree(n->d_wd_mask);
  dfree(n->d_bar); dfree(n->d_err); dfree(n->d_prof); dfree(n->d_rules);
  dfree(n->d_st_cnt); dfree(n->d_st_nisi); dfree(n->d_st_hist); dfree(n->d_st_prev);
  dfree(n->d_st_s1); dfree(n->d_st_s2); dfree(n->d_st_out);
  for (auto& m : n->mons) {
    dfree(m.d_ids); dfree(m.d_times); dfree(m.d_count);
  }
}

template <typename T>
int upload(T* d, const T* h, size_t n, hipStream_t s) {
  if (n == 0) return SPK_OK;
  SPK_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return SPK_OK;
}

int grid_for(int64_t n, int block) { return (int)std::max<int64_t>(1, (n + block - 1) / block); }

// Host threads for the O(S) passes of preparation: f(t, lo, hi) on contiguous chunks
// [lo, hi) of [0, n), chunk t to thread t; each pass writes disjoint outputs, so results
// do not depend on the thread count.
template <typename F>
void host_chunks(int64_t n, int threads, F f, int64_t serial_below = (int64_t)1 << 20) {
  if (threads <= 1 || n < serial_below) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] { f(t, n * t / threads, n * (t + 1) / threads); });
  for (auto& th : pool) th.join();
}

// Stable counting sort of the items [0, n) by key(i) in [0, K): ptr (K + 1 entries) gets
// the bucket starts and place(i, k) is called once per item with its sorted position k.
// Items of one chunk keep their order and the chunks follow each other, so the result is
// that of the sequential sort whatever the thread count.
template <typename Key, typename Place>
void counting_sort(int64_t n, int64_t K, int threads, Key key, Place place, std::vector<int64_t>& ptr) {
  const int T = n >= ((int64_t)1 << 20) ? std::max(1, threads) : 1;
  std::vector<std::vector<int64_t>> cnt(T);
  host_chunks(n, T, [&](int t, int64_t a, int64_t b) {
    cnt[t].assign((size_t)K, 0);
    for (int64_t i = a; i < b; ++i) cnt[t][key(i)]++;
  });
  ptr.assign((size_t)K + 1, 0);
  for (int64_t k = 0; k < K; ++k) {
    int64_t o = ptr[k];
    for (int t = 0; t < T; ++t) {
      const int64_t c = cnt[t][k];
      cnt[t][k] = o;
      o += c;
    }
    ptr[k + 1] = o;
  }
  host_chunks(n, T, [&](int t, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) place(i, cnt[t][key(i)]++);
  });
}

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h == 0 ? 1u : h, 16u));
}

// Per-rule device constants that depend on G.
int refresh_rule_tables(spk_network* n, RuleHost& r) {
  if (r.tab_G == n->G && r.d_tab) return SPK_OK;
  std::vector<float> tab;
  const float dt = n->dt;
  const int G = n->G;
  if (r.kind == 0) {
    // host-side float expressions of WeightDependentSTDPPlasticity.cu:125-145
    tab = {r.wd.a_minus, r.wd.a_plus, r.wd.lambda, r.wd.alpha, r.wd.w_max};
    std::vector<float> Em(G + 1), Ep(G + 1);
    for (int k = 0; k <= G; ++k) {
      Em[k] = expf(-((float)k * dt) / r.wd.tau_minus);
      Ep[k] = expf(-((float)k * dt) / r.wd.tau_plus);
    }
    for (int k = 0; k <= G; ++k) tab.push_back(Em[k]);
    for (int k = 0; k <= G; ++k) tab.push_back(Ep[k]);
    for (int k = 0; k <= G; ++k) tab.push_back(1.0f / Em[k]);
    for (int k = 0; k <= G; ++k) tab.push_back(1.0f / Ep[k]);
  } else if (r.kind == 3) {
    tab = {expf(-dt / r.cu.tau_minus), expf(-dt / r.cu.tau_plus), r.cu.a_plus, r.cu.a_minus,
           r.cu.a_star, r.cu.w_max, r.cu.learning_rate};
  } else if (r.kind == 1) {
    tab = {expf(-dt / r.vg.tau_istdp), r.vg.learningrate};
    r.dterm = (double)r.vg.learningrate * (2.0 * (double)r.vg.targetrate * (double)r.vg.tau_istdp);
  } else if (r.kind == 4) {
    tab = {0.0f};   // weight normalization keeps no per-rule constants
  } else {
    tab = {1.0f - (dt / r.ev.decay_term_tau_C), 1.0f - (dt / r.ev.decay_term_tau_D),
           dt * r.ev.synaptic_neurotransmitter_concentration_alpha_C,
           dt * r.ev.model_parameter_alpha_D};
  }
  dfree(r.d_tab);
  n->rules_uploaded = false;
  int rc = dalloc(&r.d_tab, tab.size());
  if (rc) return rc;
  SPK_HIP(hipMemcpy(r.d_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  r.tab_G = n->G;
  return SPK_OK;
}

// Device state of the neurons, inputs, spike lists and counters (every prepared network
// that runs the per-phase kernels).
int alloc_state(spk_network* n) {
  int rc;
  const size_t N = n->N, Nin = n->Nin, L = n->L;
#define ALLOC(p, cnt) if ((rc = dalloc(&(p), (cnt)))) return rc
  ALLOC(n->d_V, N); ALLOC(n->d_v0, N); ALLOC(n->d_vreset, N); ALLOC(n->d_thr, N);
  ALLOC(n->d_c, N); ALLOC(n->d_rp, N); ALLOC(n->d_last_n, N); ALLOC(n->d_last_in, Nin);
  ALLOC(n->d_hist_n, (size_t)n->W * N); ALLOC(n->d_hist_in, (size_t)n->W * Nin);
  ALLOC(n->d_ring, (size_t)n->B * N * L); ALLOC(n->d_g, N * L);
  ALLOC(n->d_decay, L); ALLOC(n->d_erev, L);
  ALLOC(n->d_in_thr, (size_t)std::max(1, n->n_stim) * Nin);
  ALLOC(n->d_ext_cdf, std::max<size_t>(1, n->ext_cdf.size()));
  ALLOC(n->d_lists, (size_t)n->slots * n->G * n->cap);
  ALLOC(n->d_counts, (size_t)n->slots * n->G);
  ALLOC(n->d_slot_time, n->slots); ALLOC(n->d_slot_tidx, n->slots);
  ALLOC(n->d_stats, 4);
  for (size_t r = 0; r < n->rate_pop.size(); ++r) {
    const int pop = n->rate_pop[r];
    const size_t cnt = std::max<size_t>(1, pop == SPK_POP_NEURONS ? N : Nin);
    if (!n->d_rate[pop]) ALLOC(n->d_rate[pop], cnt);
    n->rate_base[r].assign(cnt, 0u);
  }
  if (n->gen) {
    ALLOC(n->d_gen_ids, std::max<size_t>(1, n->gen_ids.size()));
    ALLOC(n->d_gen_times, std::max<size_t>(1, n->gen_times.size()));
  }
#undef ALLOC
  return SPK_OK;
}

int upload_state(spk_network* n) {
  int rc;
  const size_t N = n->N, Nin = n->Nin, L = n->L;
  hipStream_t st = n->stream;
  if ((rc = upload(n->d_v0, n->v0.data(), N, st))) return rc;
  if ((rc = upload(n->d_vreset, n->vreset.data(), N, st))) return rc;
  if ((rc = upload(n->d_thr, n->thr.data(), N, st))) return rc;
  if ((rc = upload(n->d_c, n->c.data(), N, st))) return rc;
  if ((rc = upload(n->d_rp, n->rp.data(), N, st))) return rc;
  if ((rc = upload(n->d_decay, n->decay_g.data(), L, st))) return rc;
  if ((rc = upload(n->d_erev, n->erev.data(), L, st))) return rc;
  for (int pop = 0; pop < 2; ++pop)
    if (n->d_rate[pop])
      SPK_HIP(hipMemsetAsync(n->d_rate[pop], 0, sizeof(uint32_t) * std::max<size_t>(1, pop == SPK_POP_NEURONS ? N : Nin), st));
  if (n->gen) {
    if ((rc = upload(n->d_gen_ids, n->gen_ids.data(), n->gen_ids.size(), st))) return rc;
    if ((rc = upload(n->d_gen_times, n->gen_times.data(), n->gen_times.size(), st))) return rc;
  }
  {
    std::vector<uint32_t> thr(n->rates.size());
    for (size_t i = 0; i < thr.size(); ++i) thr[i] = bernoulli_threshold(n->rates[i], n->dt);
    if ((rc = upload(n->d_in_thr, thr.data(), thr.size(), st))) return rc;
  }
  if ((rc = upload(n->d_ext_cdf, n->ext_cdf.data(), n->ext_cdf.size(), st))) return rc;
  SPK_HIP(hipMemsetAsync(n->d_counts, 0, sizeof(int) * n->slots * n->G, st));
  SPK_HIP(hipMemsetAsync(n->d_stats, 0, sizeof(unsigned long long) * 4, st));
  return SPK_OK;
}

// Spike monitors: device stores of the requested capacity; waits for the stream.
int alloc_monitors(spk_network* n) {
  int rc;
  for (auto& m : n->mons) {
    if ((rc = dalloc(&m.d_ids, (size_t)m.capacity))) return rc;
    if ((rc = dalloc(&m.d_times, (size_t)m.capacity))) return rc;
    if ((rc = dalloc(&m.d_count, 1))) return rc;
    SPK_HIP(hipMemsetAsync(m.d_count, 0, sizeof(unsigned long long), n->stream));
  }
  SPK_HIP(hipStreamSynchronize(n->stream));
  return SPK_OK;
}

// Bitmask of the CSR positions of a network's single spike-driven weight-dependent rule
// (the source side of k_deliver); other networks keep none.
int build_wd_mask(spk_network* n) {
  dfree(n->d_wd_mask);
  if (n->rules.size() != 1 || n->rules[0].kind != 0 || !n->rules[0].ev_ok || n->rules[0].cnt == 0) return SPK_OK;
  const size_t words = (size_t)(n->S + 31) / 32;
  if (int rc = dalloc(&n->d_wd_mask, words)) return rc;
  SPK_HIP(hipMemsetAsync(n->d_wd_mask, 0, sizeof(uint32_t) * words, n->stream));
  const int64_t P = n->rules[0].cnt;
  const int grid = (int)std::min<int64_t>(16384, (P + 255) / 256);
  hipLaunchKernelGGL(spk::k_mark_rule, dim3(grid), dim3(256), 0, n->stream, n->rules[0].d_pos, P, n->d_wd_mask);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// CSR position of caller synapse s (the identity for generated fixed out-degree synapses).
inline int64_t csr_pos(const spk_network* n, int64_t s) { return n->od ? s : n->perm[(size_t)s]; }

int check_time(spk_network* n, float t) {
  if (!(t >= n->t_hi - 0.5f * n->dt))
    return fail(SPK_ERR_ARG, "group start time precedes the latest group simulated "
                             "(spk_reset_state starts the network again at any time)");
  n->t_hi = std::max(n->t_hi, t);
  return SPK_OK;
}

int begin_group(spk_network* n, float t, bool force = false) {
  if (int rc = check_time(n, t)) return rc;
  const int tidx = (int)roundf(t / n->dt);
  if (tidx == n->cur_tidx && !force) return SPK_OK;
  n->cur_tidx = tidx;
  n->cur_slot = (tidx / n->G) % n->slots;
  hipLaunchKernelGGL(spk::k_begin_group, dim3(1), dim3(64), 0, n->stream,
                     n->d_counts + (size_t)n->cur_slot * n->Gprep, n->G, n->d_slot_time,
                     n->d_slot_tidx, n->cur_slot, t, tidx);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

uint32_t* slot_lists(spk_network* n, int slot) {
  return n->d_lists + (size_t)slot * n->Gprep * n->cap;
}
int* slot_counts(spk_network* n, int slot) { return n->d_counts + (size_t)slot * n->Gprep; }

// Statistics arguments of the kernels; the window starts at the first step processed
// after spk_stats_begin.
spk::StatsArgs stats_args(spk_network* n, float start) {
  spk::StatsArgs st{};
  if (!n->st_on) return st;
  if (n->st_pending) {
    n->st_t0 = (int64_t)roundf(start / n->dt);
    n->st_pending = false;
  }
  st.count = n->d_st_cnt; st.nisi = n->d_st_nisi; st.prev = n->d_st_prev;
  st.s1 = n->d_st_s1; st.s2 = n->d_st_s2; st.hist = n->d_st_hist;
  st.t0 = n->st_t0; st.nsteps = n->st_max; st.bin = n->st_bin; st.on = 1;
  return st;
}

// Steps of the statistics window covered once the groups starting at step `first`
// (count `ng`) have run.
void stats_advance(spk_network* n, int64_t first, int64_t ng) {
  if (!n->st_on) return;
  const int64_t end = first + ng * n->G - n->st_t0;
  n->st_steps = std::max(n->st_steps, std::min(end, n->st_max));
}

void fill_lif_args(spk_network* n, float t, int slot, spk::LifArgs& a) {
  a = spk::LifArgs{};
  if (n->lif_uniform) {
    a.uni = 1;
    a.u_v0 = n->v0[0]; a.u_vreset = n->vreset[0]; a.u_thr = n->thr[0]; a.u_c = n->c[0]; a.u_rp = n->rp[0];
  }
  a.V = n->d_V; a.last = n->d_last_n; a.hist = n->d_hist_n; a.ring = n->d_ring; a.gcond = n->d_g;
  a.v0 = n->d_v0; a.vreset = n->d_vreset; a.thr = n->d_thr; a.c = n->d_c; a.rp = n->d_rp;
  a.decay_g = n->d_decay; a.erev = n->d_erev; a.ext_cdf = n->d_ext_cdf;
  a.lists = slot_lists(n, slot); a.counts = slot_counts(n, slot);
  a.ext_events = n->d_stats + 1;
  a.spike_stat = n->d_stats + 2;
  a.N = n->N; a.L = n->L; a.G = n->G; a.B = n->B; a.W = n->W; a.cap = n->cap;
  a.ext_k = (int)n->ext_cdf.size(); a.tidx = (int)roundf(t / n->dt);
  a.t = t; a.dt = n->dt; a.tref = n->tref; a.ibg = n->ibg;
  a.inv_scale = ldexpf(1.0f, -n->frac);
  a.q_ext = n->q_ext; a.seed = n->seed;
  a.st = stats_args(n, t);
  a.rate = n->d_rate[SPK_POP_NEURONS];
  stats_advance(n, a.tidx, 1);
}

int launch_neurons(spk_network* n, float t, int slot) {
  if (n->N == 0) return SPK_OK;
  spk::LifArgs a;
  fill_lif_args(n, t, slot, a);
  const int grid = grid_for(n->N, 256);
  if (n->kind == SPK_SYN_VOLTAGE)
    hipLaunchKernelGGL(spk::k_lif<0>, dim3(grid), dim3(256), 0, n->stream, a);
  else if (n->kind == SPK_SYN_CONDUCTANCE)
    hipLaunchKernelGGL(spk::k_lif<1>, dim3(grid), dim3(256), 0, n->stream, a);
  else
    hipLaunchKernelGGL(spk::k_lif<2>, dim3(grid), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// Scheduled spikes of the selected stimulus: the entries that can fall in the group
// starting at t (a conservative time window; the kernel applies the exact test).
int launch_generator(spk_network* n, float t, int slot) {
  spk::GeneratorArgs a{};
  const float base = t - n->gen_onset;
  const auto b0 = n->gen_times.begin() + n->gen_off[n->stim];
  const auto b1 = n->gen_times.begin() + n->gen_off[n->stim + 1];
  const double lo_t = (double)base - 2.0 * (double)n->dt;
  const double hi_t = (double)base + (double)(n->G + 1) * (double)n->dt;
  a.lo = std::lower_bound(b0, b1, lo_t, [](float x, double v) { return (double)x < v; }) - n->gen_times.begin();
  a.hi = std::upper_bound(b0, b1, hi_t, [](double v, float x) { return v < (double)x; }) - n->gen_times.begin();
  a.last = n->d_last_in; a.hist = n->d_hist_in; a.ids = n->d_gen_ids; a.times = n->d_gen_times;
  a.lists = slot_lists(n, slot); a.counts = slot_counts(n, slot);
  a.spike_stat = n->d_stats + 3;
  a.N = n->N; a.Nin = n->Nin; a.G = n->G; a.W = n->W; a.cap = n->cap; a.tidx = (int)roundf(t / n->dt);
  a.t = t; a.onset = n->gen_onset; a.dt = n->dt;
  a.rate = n->d_rate[SPK_POP_INPUTS];
  hipLaunchKernelGGL(spk::k_generator, dim3(grid_for(n->Nin, 256)), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

void fill_poisson_args(spk_network* n, float t, int slot, spk::PoissonArgs& a) {
  const int tidx = (int)roundf(t / n->dt);
  if (n->pending_arm) {
    n->arm_step = tidx;
    n->pending_arm = false;
  }
  a = spk::PoissonArgs{};
  a.last = n->d_last_in; a.hist = n->d_hist_in;
  a.thr = n->d_in_thr + (size_t)n->stim * n->Nin;
  a.lists = slot_lists(n, slot); a.counts = slot_counts(n, slot);
  a.spike_stat = n->d_stats + 3;
  a.N = n->N; a.Nin = n->Nin; a.G = n->G; a.W = n->W; a.cap = n->cap; a.tidx = tidx;
  a.arm_step = n->arm_step; a.t = t; a.dt = n->dt; a.seed = n->seed;
  a.rate = n->d_rate[SPK_POP_INPUTS];
}

int launch_inputs(spk_network* n, float t, int slot) {
  if (n->Nin == 0) return SPK_OK;
  if (n->gen) return launch_generator(n, t, slot);
  spk::PoissonArgs a;
  fill_poisson_args(n, t, slot, a);
  hipLaunchKernelGGL(spk::k_poisson, dim3(grid_for(n->Nin, 256)), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// Per-neuron traces of a weight-dependent rule (same storage format as its per-synapse ones).
spk::Traces neuron_traces_of(spk_network* n, RuleHost& r) {
  spk::Traces t = traces_of(n, r);
  t.a = r.d_xn_a; t.b = r.d_xn_b; t.a16 = r.d_xn_a16; t.b16 = r.d_xn_b16;
  return t;
}

int copy_traces(spk_network* n, RuleHost& r, bool to_neurons) {
  spk::TraceCopyArgs a{};
  a.syn = traces_of(n, r);
  a.neu = neuron_traces_of(n, r);
  a.pos = r.d_pos; a.pre = r.d_pre; a.tgt = n->d_tgt;
  a.P = r.cnt;
  a.to_neurons = to_neurons ? 1 : 0;
  hipLaunchKernelGGL(spk::k_trace_copy, dim3(grid_for(a.P, 256)), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// Make the per-synapse traces current (after spike-driven groups).
int traces_per_synapse(spk_network* n, RuleHost& r) {
  if (r.syn_tr) return SPK_OK;
  if (int rc = copy_traces(n, r, false)) return rc;
  r.syn_tr = true;
  return SPK_OK;
}

// One group of the weight-dependent rule in its spike-driven form (k_wd_neurons, k_wd_events).
void fill_wd_args(spk_network* n, RuleHost& r, float t, spk::WdEventArgs& a) {
  a = spk::WdEventArgs{};
  a.trn = neuron_traces_of(n, r);
  a.xa = r.d_xcur_a; a.xb = r.d_xcur_b;
  a.wts = weights_of(n);
  a.last_n = n->d_last_n; a.last_in = n->d_last_in; a.tab = r.d_tab;
  a.pos = r.d_pos; a.pre = r.d_pre; a.tgt = n->d_tgt;
  a.pre_ptr = r.d_ev_pre_ptr; a.pre_idx = r.d_ev_pre_idx;
  a.post_ptr = r.d_ev_post_ptr; a.post_pos = r.d_ev_post_pos; a.post_pre = r.d_ev_post_pre;
  a.act = r.d_act; a.cnt = r.d_act_cnt; a.par = r.par;
  a.N = n->N; a.Nin = n->Nin; a.G = n->G; a.nearest = r.nearest;
  a.t = t; a.dt = n->dt;
}

int launch_wd_events(spk_network* n, RuleHost& r, float t) {
  if (!r.neu_tr) {
    if (int rc = copy_traces(n, r, true)) return rc;
    r.neu_tr = true;
  }
  spk::WdEventArgs a;
  fill_wd_args(n, r, t, a);
  hipLaunchKernelGGL(spk::k_wd_neurons, dim3(grid_for((int64_t)n->nu() + n->N, 256)), dim3(256), 0,
                     n->stream, a);
  SPK_HIP(hipGetLastError());
  hipLaunchKernelGGL(spk::k_wd_events, dim3(2048), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  r.par ^= 1;
  r.syn_tr = false;
  return SPK_OK;
}

int launch_rule(spk_network* n, int rid, float t) {
  RuleHost& r = n->rules[rid];
  if (r.cnt == 0) return SPK_OK;
  int rc = refresh_rule_tables(n, r);
  if (rc) return rc;
  if (r.kind == 0 && r.ev_ok && !r.dense_next) return launch_wd_events(n, r, t);
  if (r.kind == 4) {
    spk::WeightNormArgs w{};
    w.wts = weights_of(n); w.pos = r.d_pos; w.w0 = r.d_ta; w.sum_sq = r.d_wn_sumsq;
    w.in_set = r.d_wn_inset; w.ptr = r.d_wn_ptr; w.idx = r.d_wn_idx; w.N = n->N;
    hipLaunchKernelGGL(spk::k_weightnorm, dim3(grid_for(n->N, 256)), dim3(256), 0, n->stream, w);
    SPK_HIP(hipGetLastError());
    return SPK_OK;
  }
  if ((rc = traces_per_synapse(n, r))) return rc;
  r.neu_tr = false;
  r.dense_next = false;
  spk::StdpArgs a{};
  a.pos = r.d_pos; a.pre = r.d_pre; a.tgt = n->d_tgt; a.delay = n->d_delay; a.wts = weights_of(n);
  a.tr = traces_of(n, r); a.last_n = n->d_last_n; a.last_in = n->d_last_in;
  a.hist_n = n->d_hist_n; a.hist_in = n->d_hist_in; a.tab = r.d_tab; a.dterm = r.dterm;
  a.P = r.cnt; a.N = n->N; a.Nin = n->Nin; a.G = n->G; a.W = n->W;
  a.tidx = (int)roundf(t / n->dt); a.nearest = r.nearest; a.t = t; a.dt = n->dt;
  const int grid = grid_for(a.P, 256);
  if (r.kind == 0) hipLaunchKernelGGL(spk::k_stdp_wd, dim3(grid), dim3(256), 0, n->stream, a);
  else if (r.kind == 1) hipLaunchKernelGGL(spk::k_stdp_vogels, dim3(grid), dim3(256), 0, n->stream, a);
  else if (r.kind == 2) hipLaunchKernelGGL(spk::k_stdp_evans, dim3(grid), dim3(256), 0, n->stream, a);
  else hipLaunchKernelGGL(spk::k_stdp_custom, dim3(grid), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

void fill_prop_args(spk_network* n, float t, int slot, spk::PropArgs& a) {
  a = spk::PropArgs{};
  a.lists = slot_lists(n, slot); a.counts = slot_counts(n, slot);
  a.rowptr = n->d_rowptr; a.tgt = n->d_tgt; a.wts = weights_of(n);
  a.scale = n->uniform_scale ? nullptr : n->d_scale;
  a.delay = n->d_delay; a.label = (n->L > 1) ? n->d_label : nullptr;
  a.ring = n->d_ring; a.events = n->d_stats;
  a.scale0 = n->scale0; a.two_k = ldexpf(1.0f, n->frac);
  a.N = n->N; a.L = n->L; a.G = n->G; a.B = n->B; a.cap = n->cap;
  a.tidx = (int)roundf(t / n->dt);
}

int launch_propagate(spk_network* n, float t, int slot) {
  if (n->S == 0) return SPK_OK;
  spk::PropArgs a;
  fill_prop_args(n, t, slot, a);
  hipLaunchKernelGGL(spk::k_propagate, dim3(1024), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// Groups of the per-phase path that run as two launches (k_units, k_deliver): G = 1,
// LIF neurons, Poisson or no inputs, and no rule applied or one weight-dependent rule in
// its spike-driven form.
bool fast_group_ok(const spk_network* n, bool plastic) {
  if (n->G != 1 || n->Gprep != 1 || n->gen || n->N == 0 || n->S == 0 || n->slots < 2) return false;
  if (n->rules.empty() || !plastic) return true;
  if (n->rules.size() != 1) return false;
  const RuleHost& r = n->rules[0];
  return r.kind == 0 && r.ev_ok && r.cnt > 0 && !r.dense_next && n->d_wd_mask != nullptr;
}

constexpr int kDeliverBlocks = 4096;

// One group in two launches.  The current slot's spike counts were cleared by the previous
// group's k_units, or, when `clear_now`, are cleared here by a group-start launch.
int launch_group_fast(spk_network* n, float t, bool plastic, bool clear_now) {
  if (clear_now) {
    if (int rc = begin_group(n, t, true)) return rc;
  } else {
    if (int rc = check_time(n, t)) return rc;
    n->cur_tidx = (int)roundf(t / n->dt);
    n->cur_slot = (n->cur_tidx / n->G) % n->slots;
  }
  const int slot = n->cur_slot;
  const int next = (n->cur_tidx / n->G + 1) % n->slots;
  const bool wd = plastic && !n->rules.empty();
  RuleHost* r = wd ? &n->rules[0] : nullptr;
  spk::UnitsArgs u{};
  if (wd) {
    if (int rc = refresh_rule_tables(n, *r)) return rc;
    if (!r->neu_tr) {
      if (int rc = copy_traces(n, *r, true)) return rc;
      r->neu_tr = true;
    }
    fill_wd_args(n, *r, t, u.wd);
  }
  fill_lif_args(n, t, slot, u.lif);
  if (n->Nin > 0) fill_poisson_args(n, t, slot, u.in);
  u.next_counts = n->d_counts + (size_t)next * n->Gprep;
  u.slot_time = n->d_slot_time;
  u.slot_tidx = n->d_slot_tidx;
  u.slot = slot;
  u.nb_neurons = grid_for(n->N, 256);
  const int blocks = u.nb_neurons + (n->Nin > 0 ? grid_for(n->Nin, 256) : 0);
  const void* fn;
  if (n->kind == SPK_SYN_VOLTAGE) fn = wd ? (const void*)spk::k_units<0, true> : (const void*)spk::k_units<0, false>;
  else if (n->kind == SPK_SYN_CONDUCTANCE) fn = wd ? (const void*)spk::k_units<1, true> : (const void*)spk::k_units<1, false>;
  else fn = wd ? (const void*)spk::k_units<2, true> : (const void*)spk::k_units<2, false>;
  void* uargs[] = {&u};
  SPK_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(256), uargs, 0, n->stream));
  spk::DeliverArgs d{};
  fill_prop_args(n, t, slot, d.p);
  d.row_waves = (int)std::min<int64_t>(64, std::max<int64_t>(1, (n->max_row + spk::kWave - 1) / spk::kWave));
  if (wd) {
    d.wd = u.wd;
    d.mask = n->d_wd_mask;
    d.post_waves = (int)std::min<int64_t>(64, std::max<int64_t>(1, (r->max_post + spk::kWave - 1) / spk::kWave));
    hipLaunchKernelGGL(spk::k_deliver<true>, dim3(kDeliverBlocks), dim3(256), 0, n->stream, d);
    r->par ^= 1;
    r->syn_tr = false;
  } else {
    d.post_waves = 1;
    hipLaunchKernelGGL(spk::k_deliver<false>, dim3(kDeliverBlocks), dim3(256), 0, n->stream, d);
  }
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

int launch_collect(spk_network* n, int mid, int slot0, int nslots) {
  MonitorHost& m = n->mons[mid];
  spk::CollectArgs a{};
  a.lists = n->d_lists; a.counts = n->d_counts; a.slot_time = n->d_slot_time;
  a.ids = m.d_ids; a.times = m.d_times; a.count = m.d_count; a.store_cap = m.capacity;
  a.slot0 = slot0; a.nslots = nslots; a.G = n->Gprep; a.cap = n->cap;
  a.lo = m.pop == SPK_POP_NEURONS ? 0 : n->N;
  a.hi = m.pop == SPK_POP_NEURONS ? n->N : n->nu();
  a.dt = n->dt;
  a.G = n->G;
  a.Gs = n->Gprep;
  hipLaunchKernelGGL(spk::k_collect, dim3(8, nslots * n->G), dim3(256), 0, n->stream, a);
  SPK_HIP(hipGetLastError());
  return SPK_OK;
}

// Drain a monitor's device store into its host vectors, sorted by (time, id).
int drain_monitor(spk_network* n, MonitorHost& m, std::vector<int32_t>& ids,
                  std::vector<float>& times) {
  unsigned long long cnt = 0;
  SPK_HIP(hipMemcpyAsync(&cnt, m.d_count, sizeof(cnt), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  if ((int64_t)cnt > m.capacity)
    return fail(SPK_ERR_CAPACITY, "spike monitor device store overflowed (" +
                                      std::to_string(cnt) + " > " + std::to_string(m.capacity) + ")");
  std::vector<int32_t> bi(cnt);
  std::vector<float> bt(cnt);
  if (cnt) {
    SPK_HIP(hipMemcpyAsync(bi.data(), m.d_ids, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, n->stream));
    SPK_HIP(hipMemcpyAsync(bt.data(), m.d_times, cnt * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  }
  SPK_HIP(hipMemsetAsync(m.d_count, 0, sizeof(unsigned long long), n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  std::vector<size_t> order(cnt);
  for (size_t i = 0; i < cnt; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (bt[a] != bt[b]) return bt[a] < bt[b];
    return bi[a] < bi[b];
  });
  for (size_t i : order) {
    ids.push_back(bi[i]);
    times.push_back(bt[i]);
  }
  return SPK_OK;
}

int ensure_monitor_capacity(spk_network* n, MonitorHost& m, int64_t need) {
  if (m.capacity >= need && m.d_ids) return SPK_OK;
  // keep what is on the device
  std::vector<int32_t> ids;
  std::vector<float> times;
  if (m.d_ids) {
    int rc = drain_monitor(n, m, m.h_ids, m.h_times);
    if (rc) return rc;
  }
  dfree(m.d_ids);
  dfree(m.d_times);
  m.capacity = need;
  int rc = dalloc(&m.d_ids, (size_t)need);
  if (rc) return rc;
  rc = dalloc(&m.d_times, (size_t)need);
  if (rc) return rc;
  if (!m.d_count) {
    rc = dalloc(&m.d_count, 1);
    if (rc) return rc;
  }
  SPK_HIP(hipMemsetAsync(m.d_count, 0, sizeof(unsigned long long), n->stream));
  return SPK_OK;
}

template <typename T>
int download(spk_network* n, T* out, const T* d, size_t cnt) {
  if (cnt == 0) return SPK_OK;
  if (!out) return fail(SPK_ERR_ARG, "null output");
  SPK_HIP(hipMemcpyAsync(out, d, cnt * sizeof(T), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  return SPK_OK;
}


// The persistent kernel instance for a network (synapse kind, single label or not,
// plasticity applied or not).
const void* fused_kernel(const spk_network* n, bool plastic) {
  const bool l1 = n->L == 1;
  // Brunel-type networks (voltage synapses, compact records): kernels holding only their
  // delivery form
  if (!plastic && n->kind == SPK_SYN_VOLTAGE && l1 && n->compact && n->rules.empty())
    return n->lseg_U > 0 ? (const void*)spk::k_fused<0, true, false, spk::kDeliverLdsTable>
                         : (const void*)spk::k_fused<0, true, false, spk::kDeliverCompact>;
#define SPK_FUSED_PICK(K)                                                                  \
  (plastic ? (l1 ? (const void*)spk::k_fused<K, true, true> : (const void*)spk::k_fused<K, false, true>) \
           : (l1 ? (const void*)spk::k_fused<K, true, false> : (const void*)spk::k_fused<K, false, false>))
  if (n->kind == SPK_SYN_VOLTAGE) return SPK_FUSED_PICK(0);
  if (n->kind == SPK_SYN_CONDUCTANCE) return SPK_FUSED_PICK(1);
  return SPK_FUSED_PICK(2);
#undef SPK_FUSED_PICK
}

// Dynamic LDS a persistent workgroup may take: the per-workgroup limit less the static LDS
// of the persistent kernel's two forms, and a margin for the batched form's (the same body
// with its own scan scratch).  Without it a partition sized just under the limit would
// leave the grid no co-resident workgroup.
size_t fused_lds_limit(const spk_network* n) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, n->device) != hipSuccess) return 0;
  size_t st = 0;
  for (bool plastic : {false, true}) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, fused_kernel(n, plastic)) == hipSuccess) st = std::max(st, fa.sharedSizeBytes);
  }
  const size_t lim = (size_t)prop.sharedMemPerBlock;
  return lim > st + 512 ? lim - st - 512 : 0;
}

// The compact form's delivery table of each partition -- a 16-bit local start per source
// unit and the partition's 2-byte targets -- stays in LDS for the whole launch when the
// largest partition's table fits beside the rest (fewer than 2^16 targets).  `seg` is the
// [U][P+1] segment table.
int plan_lds_table(spk_network* n, const uint32_t* seg, int U) {
  n->fused_lds_base = n->fused_lds;
  n->lseg_U = 0;
  n->lrec = 0;
  if (!n->fused_ok || !n->compact || !n->rules.empty() || U <= 0 || U >= (1 << 24)) return SPK_OK;
  const int P = n->P;
  std::vector<int64_t> cnt(P, 0);
  for (int u = 0; u < U; ++u) {
    const uint32_t* r = seg + (size_t)u * (P + 1);
    for (int q = 0; q < P; ++q) cnt[q] += r[q + 1] - r[q];
  }
  const int64_t mx = *std::max_element(cnt.begin(), cnt.end());
  if (mx > 65535) return SPK_OK;
  const size_t full = spk::fused_lds_bytes(n->BL, n->L, n->W, n->G, n->nown_max, n->iown_max, n->xcap, P,
                                           U, (int)mx, n->cring_ok ? 1 : 0);
  if (full > fused_lds_limit(n)) return SPK_OK;
  // and the grid stays co-resident with the larger LDS (both forms of the kernel)
  for (bool plastic : {false, true}) {
    int per_cu = 0;
    SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel(n, plastic), spk::kFusedBlock, full));
    if ((int64_t)per_cu * n->cu_count < n->P) return SPK_OK;
  }
  n->lseg_U = U;
  n->lrec = (int)mx;
  n->fused_lds = full;
  return SPK_OK;
}

// Kernel arguments of the persistent kernel (all but the run position).
int fill_fused_args(spk_network* n, bool plastic, spk::FusedArgs& a) {
  const int G = n->G;
  for (auto& r : n->rules)
    if (int rc = refresh_rule_tables(n, r)) return rc;
  a = spk::FusedArgs{};
  a.nlo = n->d_nlo; a.ilo = n->d_ilo;
  a.rate_n = n->d_rate[SPK_POP_NEURONS]; a.rate_in = n->d_rate[SPK_POP_INPUTS];
  a.V = n->d_V; a.last_n = n->d_last_n; a.hist_n = n->d_hist_n; a.gcond = n->d_g; a.ring = n->d_ring;
  a.v0 = n->d_v0; a.vreset = n->d_vreset; a.thr = n->d_thr; a.c = n->d_c; a.rp = n->d_rp;
  a.decay_g = n->d_decay; a.erev = n->d_erev; a.ext_cdf = n->d_ext_cdf;
  a.last_in = n->d_last_in; a.hist_in = n->d_hist_in;
  a.in_thr = n->d_in_thr + (size_t)n->stim * n->Nin;
  a.seg = n->d_seg; a.syn = n->d_syn;
  a.wc = n->d_wc; a.wfmt = n->wfmt;
  a.scale = n->uniform_scale ? nullptr : n->d_scale;
  if (n->compact && n->rules.empty()) {
    a.syn16 = n->d_syn16;
    a.w_lo = n->cw_lo; a.w_hi = n->cw_hi; a.src_split = n->csplit; a.src_delay = n->cdelay;
    a.lseg_U = n->lseg_U;
    a.lrec = n->lrec;
    a.cring = n->cring_ok ? 1 : 0;
  }
  a.n_rules = plastic ? (int)n->rules.size() : 0;
  for (int r = 0; r < a.n_rules; ++r) {
    RuleHost& rh = n->rules[r];
    if (int rc = traces_per_synapse(n, rh)) return rc;
    rh.neu_tr = false;
    rh.dense_next = false;
  }
  if (a.n_rules > 0 && !n->rules_uploaded) {
    std::vector<spk::FusedRule> fr(a.n_rules);
    for (int r = 0; r < a.n_rules; ++r) {
      RuleHost& rh = n->rules[r];
      spk::StdpArgs& ra = fr[r].args;
      ra = spk::StdpArgs{};
      ra.pos = rh.d_pos; ra.pre = rh.d_pre; ra.tgt = n->d_tgt; ra.delay = n->d_delay; ra.wts = weights_of(n);
      ra.tr = traces_of(n, rh); ra.last_n = n->d_last_n; ra.last_in = n->d_last_in;
      ra.hist_n = n->d_hist_n; ra.hist_in = n->d_hist_in; ra.tab = rh.d_tab; ra.dterm = rh.dterm;
      ra.P = rh.cnt; ra.N = n->N; ra.Nin = n->Nin; ra.G = G; ra.W = n->W;
      ra.nearest = rh.nearest; ra.dt = n->dt;
      fr[r].kind = rh.kind;
      fr[r].part = rh.d_part;
    }
    if (!n->d_rules) {
      if (int rc = dalloc(&n->d_rules, n->rules.size())) return rc;
    }
    SPK_HIP(hipMemcpy(n->d_rules, fr.data(), fr.size() * sizeof(spk::FusedRule), hipMemcpyHostToDevice));
    n->rules_uploaded = true;
  }
  a.rules = n->d_rules;
  a.xlist = n->d_xlist; a.bar = n->d_bar; a.error = n->d_err;
  a.xlist64 = (unsigned long long*)n->d_xlist;
  // random draws one group ahead, made while the workgroup waits at the exchange (passes
  // of 2, 4 or 8 groups are slower: their longer draws no longer hide in the wait)
  a.dahead = 1;
  // short lists (partitions of at most 64 neurons): each quad polls and delivers one
  // source's list; longer ones are staged over the whole workgroup
  a.xquad = n->nown_max + n->iown_max <= 64 ? 1 : 0;
  // one-wavefront partitions publish right after their LIF and draw beside it (Brunel10K
  // 4.19 -> 4.02 ms per simulated second; the batch sets it off: 3 % slower there)
  a.early = a.xquad;
  a.n_mons = (int)n->mons.size();
  for (int m = 0; m < a.n_mons; ++m) {
    const MonitorHost& mh = n->mons[m];
    a.mon[m].ids = mh.d_ids;
    a.mon[m].times = mh.d_times;
    a.mon[m].count = mh.d_count;
    a.mon[m].cap = mh.capacity;
    a.mon[m].lo = mh.pop == SPK_POP_NEURONS ? 0u : (unsigned)n->N;
    a.mon[m].hi = mh.pop == SPK_POP_NEURONS ? (unsigned)n->N : (unsigned)n->nu();
  }
  a.stats = n->d_stats;
  a.q_ext = n->q_ext; a.arm_step = n->Nin > 0 ? n->arm_step : -1; a.seed = n->seed;
  a.spin_limit = (long long)1 << 22;
  a.scale0 = n->scale0; a.two_k = ldexpf(1.0f, n->frac); a.inv_scale = ldexpf(1.0f, -n->frac);
  a.dt = n->dt; a.tref = n->tref; a.ibg = n->ibg;
  a.N = n->N; a.Nin = n->Nin; a.L = n->L; a.G = G; a.B = n->B; a.BL = n->BL; a.W = n->W; a.P = n->P;
  a.xcap = n->xcap; a.xstride = spk::fused_xstride(n->xcap);
  a.nown_max = n->nown_max; a.iown_max = n->iown_max;
  a.ext_k = (int)n->ext_cdf.size();
  a.plastic = plastic ? 1 : 0;
  a.prof = nullptr;
  return SPK_OK;
}

// Barrier generations one launch may use before the level-1 counters (up to
// ceil(P/8) arrivals per generation) could wrap.
constexpr unsigned int kBarGenLimit = 1u << 24;

// Run position of one launch (groups [group0, group0 + ng) of the run that began at
// `start`) and the barrier bookkeeping that goes with it; clears on `st` what must start
// at zero.
int stage_fused(spk_network* n, spk::FusedArgs& a, float start, int64_t group0, int ng, hipStream_t st) {
  if (n->bar_clear || n->bar_gen + (unsigned int)ng > kBarGenLimit) {
    SPK_HIP(hipMemsetAsync(n->d_bar, 0, sizeof(unsigned int) * (spk::kBarGroups + 2) * spk::kBarStride, st));
    SPK_HIP(hipMemsetAsync(n->d_err, 0, sizeof(int), st));
    n->bar_gen = 0;
    n->bar_clear = false;
  }
  if (n->Nin > 0 && n->pending_arm) {
    n->arm_step = (int)roundf(start / n->dt);
    n->pending_arm = false;
  }
  a.arm_step = n->Nin > 0 ? n->arm_step : -1;
  a.st = stats_args(n, start);
  stats_advance(n, (int64_t)roundf(start / n->dt), group0 + ng);
  a.start = start;
  a.group0 = group0;
  a.ngroups = ng;
  a.bar_base = n->bar_gen;
  n->bar_gen += (unsigned int)ng;
  // tags of the launch's groups: xtag + 1 .. xtag + ng, never 0 (the cleared words' tag);
  // each launch starts from cleared blocks, so tags only need to differ within a launch
  // and from the previous launch's
  if ((uint64_t)n->xtag + (uint64_t)ng + 1u >= ((uint64_t)1 << 32)) n->xtag = 0;
  a.xtag_base = n->xtag;
  n->xtag += (unsigned int)ng;
  return SPK_OK;
}

// One cooperative launch of the persistent kernel over `ng` groups starting at group
// index group0 of the run that began at `start`; enqueued only.
int launch_fused(spk_network* n, spk::FusedArgs& a, float start, int64_t group0, int ng,
                 hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  if (int rc = stage_fused(n, a, start, group0, ng, n->stream)) return rc;
  const void* fn = fused_kernel(n, a.plastic != 0);
  const int fi = a.plastic != 0;
  if (n->fused_lds > 65536 && (n->lds_attr[fi] < n->fused_lds || n->lds_fn[fi] != fn)) {
    SPK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)n->fused_lds));
    n->lds_attr[fi] = n->fused_lds;
    n->lds_fn[fi] = fn;
  }
  void* kargs[] = {&a};
  if (ev0) SPK_HIP(hipEventRecord(ev0, n->stream));
  SPK_HIP(hipLaunchCooperativeKernel(fn, dim3(n->P), dim3(spk::kFusedBlock), kargs,
                                     (unsigned int)n->fused_lds, n->stream));
  if (ev1) SPK_HIP(hipEventRecord(ev1, n->stream));
  return SPK_OK;
}

// Report (and clear) a barrier time-out of the persistent kernel; synchronizes.
int check_fused_error(spk_network* n) {
  if (!n->d_err) return SPK_OK;
  int err = 0;
  SPK_HIP(hipMemcpyAsync(&err, n->d_err, sizeof(int), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  if (err) {
    n->bar_clear = true;
    if (err == 2) return fail(SPK_ERR_CAPACITY, "persistent kernel: spikes of a group exceeded the exchange list");
    return fail(SPK_ERR_HIP, "persistent kernel: grid barrier wait exceeded its spin limit");
  }
  return SPK_OK;
}

// The persistent kernel over n_groups groups from `start`, in launches of bounded length;
// monitor records are drained into the host stores after every launch.
int run_fused(spk_network* n, float start, int64_t ngroups, int flags) {
  const bool plastic = (flags & SPK_RUN_PLASTICITY) != 0 && !n->rules.empty();
  const int G = n->G;
  int64_t chunk = 1000;  // bounds the length of one launch
  // every monitor store holds the records of a whole launch (each unit spikes at most once
  // per sub-step); larger stores allow longer launches
  for (auto& m : n->mons) {
    const int64_t pop = std::max<int64_t>(1, m.pop == SPK_POP_NEURONS ? n->N : n->Nin);
    const int64_t per_group = pop * G;
    const int64_t want = std::max<int64_t>(m.capacity, std::min<int64_t>(per_group * 256, (int64_t)1 << 27));
    if (int rc = ensure_monitor_capacity(n, m, std::max<int64_t>(want, per_group))) return rc;
    chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, m.capacity / per_group));
  }
  spk::FusedArgs a;
  if (int rc = fill_fused_args(n, plastic, a)) return rc;
  if (n->profiling) {
    if (!n->d_prof) {
      if (int rc = dalloc(&n->d_prof, spk::kProfPhases)) return rc;
    }
    SPK_HIP(hipMemsetAsync(n->d_prof, 0, sizeof(unsigned long long) * spk::kProfPhases, n->stream));
    a.prof = n->d_prof;
  }
  const int64_t nlaunch = (ngroups + chunk - 1) / chunk;
  while ((int64_t)n->ev_kernel.size() < 2 * nlaunch) {
    hipEvent_t e;
    SPK_HIP(hipEventCreate(&e));
    n->ev_kernel.push_back(e);
  }
  int64_t li = 0;
  for (int64_t s0 = 0; s0 < ngroups; s0 += chunk, ++li) {
    const int ng = (int)std::min<int64_t>(chunk, ngroups - s0);
    if (int rc = launch_fused(n, a, start, s0, ng, n->ev_kernel[2 * li], n->ev_kernel[2 * li + 1])) return rc;
    for (auto& m : n->mons)
      if (int rc = drain_monitor(n, m, m.h_ids, m.h_times)) return rc;
  }
  if (int rc = check_fused_error(n)) return rc;
  double kms = 0.0;
  for (int64_t i = 0; i < nlaunch; ++i) {
    float ms = 0.0f;
    SPK_HIP(hipEventElapsedTime(&ms, n->ev_kernel[2 * i], n->ev_kernel[2 * i + 1]));
    kms += ms;
  }
  n->timing.kernel_ms = kms;
  n->timing.kernel_launches = nlaunch;
  if (n->profiling) {
    unsigned long long ticks[spk::kProfPhases];
    SPK_HIP(hipMemcpy(ticks, n->d_prof, sizeof(ticks), hipMemcpyDeviceToHost));
    int rate_khz = 0;
    SPK_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, n->device));
    for (int i = 0; i < spk::kProfPhases; ++i)
      n->prof_us[i] = rate_khz > 0 ? (double)ticks[i] * 1e3 / (double)rate_khz : 0.0;
  }
  return SPK_OK;
}

}  // namespace

namespace spk {
// error message of the calling thread, shared with the surrogate's entry points (mlp.hip)
void set_last_error(const std::string& msg) { g_error = msg; }
}  // namespace spk

// =====================================================================================
extern "C" {

const char* spk_last_error(void) { return g_error.c_str(); }

int spk_abi_version(void) { return 100; }

int spk_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int spk_create(int n_neurons, int n_inputs, int synapse_kind, int n_labels, float timestep,
               spk_network** out) {
  if (!out) return fail(SPK_ERR_ARG, "out is null");
  *out = nullptr;
  if (n_neurons < 0 || n_inputs < 0) return fail(SPK_ERR_ARG, "negative population size");
  if ((int64_t)n_neurons + n_inputs >= (int64_t)1 << 30) return fail(SPK_ERR_ARG, "too many neurons");
  if (synapse_kind != SPK_SYN_VOLTAGE && synapse_kind != SPK_SYN_CONDUCTANCE && synapse_kind != SPK_SYN_CURRENT)
    return fail(SPK_ERR_ARG, "unknown synapse kind");
  if (n_labels < 1 || n_labels > spk::kMaxLabels)
    return fail(SPK_ERR_ARG, "n_labels must be in [1, 8]");
  if (!(timestep > 0.0f)) return fail(SPK_ERR_ARG, "timestep must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(SPK_ERR_NODEVICE, "no HIP device visible");
  int dev = 0;
  SPK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SPK_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(SPK_ERR_NODEVICE, std::string("device is ") + prop.gcnArchName + ", need gfx950");
  spk_network* n = new spk_network();
  n->N = n_neurons;
  n->Nin = n_inputs;
  n->kind = synapse_kind;
  n->L = n_labels;
  n->dt = timestep;
  n->device = dev;
  n->decay_g.assign(n_labels, 1.0f);
  n->erev.assign(n_labels, 0.0f);
  hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete n;
    return fail(SPK_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  n->own_stream = true;
  (void)hipEventCreate(&n->ev_a);
  (void)hipEventCreate(&n->ev_b);
  *out = n;
  return SPK_OK;
}

int spk_destroy(spk_network* n) {
  if (!n) return SPK_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  free_device(n);
  if (n->ev_a) (void)hipEventDestroy(n->ev_a);
  if (n->ev_b) (void)hipEventDestroy(n->ev_b);
  for (auto e : n->ev_kernel) (void)hipEventDestroy(e);
  if (n->own_stream && n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return SPK_OK;
}

int spk_set_stream(spk_network* n, void* s) {
  if (int rc = check_net(n)) return rc;
  SPK_HIP(hipStreamSynchronize(n->stream));
  if (!s) return SPK_OK;
  if (n->own_stream) (void)hipStreamDestroy(n->stream);
  n->stream = (hipStream_t)s;
  n->own_stream = false;
  return SPK_OK;
}

int spk_set_lif(spk_network* n, const float* v0, const float* vreset, const float* threshold,
                const float* decay_c, const float* resist, float refractory_s,
                float background_current) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (n->N > 0 && (!v0 || !vreset || !threshold || !decay_c || !resist))
    return fail(SPK_ERR_ARG, "null LIF parameter array");
  n->v0.assign(v0, v0 + n->N);
  n->vreset.assign(vreset, vreset + n->N);
  n->thr.assign(threshold, threshold + n->N);
  n->c.assign(decay_c, decay_c + n->N);
  n->rp.assign(resist, resist + n->N);
  n->tref = refractory_s;
  n->ibg = background_current;
  n->lif_set = true;
  // one parameter set for every neuron (bit for bit): the kernels take it as scalars
  auto same = [](const std::vector<float>& v) {
    for (size_t i = 1; i < v.size(); ++i)
      if (std::memcmp(&v[i], &v[0], sizeof(float)) != 0) return false;
    return true;
  };
  n->lif_uniform = n->N > 0 && same(n->v0) && same(n->vreset) && same(n->thr) && same(n->c) && same(n->rp);
  return SPK_OK;
}

int spk_set_poisson_rates(spk_network* n, const float* rates, int n_stimuli) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (n_stimuli < 1 || (n->Nin > 0 && !rates)) return fail(SPK_ERR_ARG, "bad Poisson rates");
  n->rates.assign(rates, rates + (size_t)n_stimuli * n->Nin);
  n->n_stim = n_stimuli;
  n->gen = false;
  for (float r : n->rates)
    if (!(r >= 0.0f) || std::isinf(r)) return fail(SPK_ERR_ARG, "Poisson rates must be finite and >= 0");
  return SPK_OK;
}

int spk_set_generator_stimuli(spk_network* n, const int32_t* counts, int n_stimuli, const int32_t* ids,
                              const float* times) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (n->Nin < 1) return fail(SPK_ERR_ARG, "generator stimuli need an input population");
  if (n_stimuli < 1 || !counts) return fail(SPK_ERR_ARG, "bad generator stimuli");
  int64_t total = 0;
  for (int k = 0; k < n_stimuli; ++k) {
    if (counts[k] < 0) return fail(SPK_ERR_ARG, "negative spike count in a stimulus");
    total += counts[k];
  }
  if (total > 0 && (!ids || !times)) return fail(SPK_ERR_ARG, "null stimulus arrays");
  for (int64_t e = 0; e < total; ++e) {
    if (ids[e] < 0 || ids[e] >= n->Nin) return fail(SPK_ERR_ARG, "stimulus neuron id out of range");
    if (!std::isfinite(times[e])) return fail(SPK_ERR_ARG, "stimulus spike times must be finite");
  }
  n->gen_off.assign(1, 0);
  n->gen_ids.clear();
  n->gen_times.clear();
  int64_t o = 0;
  for (int k = 0; k < n_stimuli; ++k) {
    std::vector<int64_t> idx(counts[k]);
    for (int j = 0; j < counts[k]; ++j) idx[j] = o + j;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return times[x] < times[y]; });
    for (int64_t e : idx) {
      n->gen_ids.push_back(ids[e]);
      n->gen_times.push_back(times[e]);
    }
    o += counts[k];
    n->gen_off.push_back(o);
  }
  n->gen = true;
  n->n_stim = n_stimuli;
  n->stim = 0;
  n->rates.assign((size_t)n_stimuli * n->Nin, 0.0f);
  return SPK_OK;
}

int spk_set_generator_onset(spk_network* n, float onset) {
  if (int rc = check_net(n)) return rc;
  if (!std::isfinite(onset)) return fail(SPK_ERR_ARG, "onset must be finite");
  n->gen_onset = onset;
  return SPK_OK;
}

int spk_select_stimulus(spk_network* n, int s) {
  if (int rc = check_net(n)) return rc;
  if (s < 0 || s >= std::max(1, n->n_stim)) return fail(SPK_ERR_ARG, "stimulus index out of range");
  n->stim = s;
  if (n->prepared) return spk_reset_inputs(n);
  return SPK_OK;
}

int spk_set_seed(spk_network* n, uint64_t seed) {
  if (int rc = check_net(n)) return rc;
  n->seed = seed;
  return SPK_OK;
}

int spk_set_weight_format(spk_network* n, int format) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (format != SPK_WEIGHTS_FP32 && format != SPK_WEIGHTS_BF16 && format != SPK_WEIGHTS_FP16)
    return fail(SPK_ERR_ARG, "weight format must be SPK_WEIGHTS_FP32, _BF16 or _FP16");
  n->wfmt = format;
  return SPK_OK;
}

int spk_set_external_drive(spk_network* n, float rate_hz, float weight) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (!(rate_hz >= 0.0f) || std::isinf(rate_hz)) return fail(SPK_ERR_ARG, "bad external rate");
  n->ext_rate = rate_hz;
  n->ext_w = weight;
  return SPK_OK;
}

int spk_set_labels(spk_network* n, const float* decay_g, const float* reversal) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (!decay_g || !reversal) return fail(SPK_ERR_ARG, "null label arrays");
  n->decay_g.assign(decay_g, decay_g + n->L);
  n->erev.assign(reversal, reversal + n->L);
  return SPK_OK;
}

int spk_set_synapses(spk_network* n, int64_t S, const int32_t* pre, const int32_t* post,
                     const float* weight, const float* scale, const int32_t* delay,
                     const int32_t* label) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (S < 0) return fail(SPK_ERR_ARG, "negative synapse count");
  if (S > 0 && (!pre || !post || !weight || !delay)) return fail(SPK_ERR_ARG, "null synapse array");
  // the first bad synapse names the error, whatever the thread count
  static const char* const why[] = {nullptr, "postsynaptic id out of range", "presynaptic id out of range",
                                    "delay must be in [1, 65535] steps", "label out of range"};
  const int T = host_threads();
  std::vector<int> bad(T, 0);   // chunks are in synapse order: the first nonzero names it
  host_chunks(S, T, [&](int t, int64_t a, int64_t b) {
    for (int64_t s = a; s < b; ++s) {
      int e = 0;
      if (post[s] < 0 || post[s] >= n->N) e = 1;
      else if (pre[s] >= n->N || (pre[s] < 0 && -(int64_t)pre[s] - 1 >= n->Nin)) e = 2;
      else if (delay[s] < 1 || delay[s] > 65535) e = 3;
      else if (label && (label[s] < 0 || label[s] >= n->L)) e = 4;
      if (e) {
        bad[t] = e;
        return;
      }
    }
  });
  for (int t = 0; t < T; ++t)
    if (bad[t]) return fail(SPK_ERR_ARG, why[bad[t]]);
  for (auto* v : {&n->h_pre, &n->h_post, &n->h_delay, &n->h_label}) {
    v->clear();
    v->resize(S);
  }
  for (auto* v : {&n->h_w, &n->h_scale}) {
    v->clear();
    v->resize(S);
  }
  host_chunks(S, T, [&](int, int64_t a, int64_t b) {
    const size_t m = (size_t)(b - a);
    std::memcpy(n->h_pre.data() + a, pre + a, m * sizeof(int32_t));
    std::memcpy(n->h_post.data() + a, post + a, m * sizeof(int32_t));
    std::memcpy(n->h_w.data() + a, weight + a, m * sizeof(float));
    std::memcpy(n->h_delay.data() + a, delay + a, m * sizeof(int32_t));
    if (label) std::memcpy(n->h_label.data() + a, label + a, m * sizeof(int32_t));
    else std::fill(n->h_label.begin() + a, n->h_label.begin() + b, 0);
    if (scale) std::memcpy(n->h_scale.data() + a, scale + a, m * sizeof(float));
    else std::fill(n->h_scale.begin() + a, n->h_scale.begin() + b, 1.0f);
  });
  n->S = S;
  n->syn_set = true;
  return SPK_OK;
}

int spk_set_outdegree_synapses(spk_network* n, const spk_outdegree_params* p) {
  if (int rc = check_net(n)) return rc;
  if (!p) return fail(SPK_ERR_ARG, "null parameters");
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (n->syn_set || !n->rules.empty()) return fail(SPK_ERR_STATE, "synapses or rules already set");
  if (n->L != 1) return fail(SPK_ERR_ARG, "generated synapses carry label 0 of a one-label network");
  if (p->n_exc < 0 || p->n_exc > n->N) return fail(SPK_ERR_ARG, "n_exc out of range");
  if (p->out_degree < 0 || p->input_out_degree < 0) return fail(SPK_ERR_ARG, "negative out-degree");
  if (p->delay_max < 1 || p->delay_max > 65535) return fail(SPK_ERR_ARG, "delay_max must be in [1, 65535] steps");
  if (n->N < 1 && (p->out_degree > 0 || p->input_out_degree > 0)) return fail(SPK_ERR_ARG, "synapses without target neurons");
  const int64_t S = (int64_t)n->N * p->out_degree + (int64_t)n->Nin * p->input_out_degree;
  // the rule's by-target table keeps 32-bit CSR positions
  if (S >= (int64_t)UINT32_MAX) return fail(SPK_ERR_ARG, "2^32 - 1 or more synapses");
  n->odp = *p;
  n->od = true;
  n->S = S;
  n->syn_set = true;
  if (p->ee_stdp) {
    RuleHost r;
    r.kind = 0;
    r.wd = p->stdp;
    r.nearest = p->stdp.nearest_spike_only;
    n->rules.push_back(std::move(r));
  }
  return SPK_OK;
}

static int add_rule(spk_network* n, int kind, const int64_t* syn, int64_t cnt, RuleHost&& r,
                    int* rule_id) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (n->od) return fail(SPK_ERR_STATE, "synapses generated on the device: their rule comes with spk_set_outdegree_synapses");
  if (cnt < 0 || (cnt > 0 && !syn)) return fail(SPK_ERR_ARG, "bad plastic synapse list");
  // a synapse listed twice would be updated by two lanes at once
  std::vector<bool> seen((size_t)n->S, false);
  for (int64_t j = 0; j < cnt; ++j) {
    if (syn[j] < 0 || syn[j] >= n->S) return fail(SPK_ERR_ARG, "plastic synapse index out of range");
    if (seen[(size_t)syn[j]]) return fail(SPK_ERR_ARG, "synapse listed twice in one plastic rule");
    seen[(size_t)syn[j]] = true;
  }
  r.kind = kind;
  r.syn.assign(syn, syn + cnt);
  r.cnt = cnt;
  n->rules.push_back(std::move(r));
  if (rule_id) *rule_id = (int)n->rules.size() - 1;
  return SPK_OK;
}

int spk_add_wd_stdp(spk_network* n, const int64_t* syn, int64_t cnt, const spk_wd_stdp_params* p,
                    int* rule_id) {
  if (!p) return fail(SPK_ERR_ARG, "null parameters");
  RuleHost r;
  r.wd = *p;
  r.nearest = p->nearest_spike_only;
  return add_rule(n, 0, syn, cnt, std::move(r), rule_id);
}

int spk_add_vogels_stdp(spk_network* n, const int64_t* syn, int64_t cnt,
                        const spk_vogels_stdp_params* p, int* rule_id) {
  if (!p) return fail(SPK_ERR_ARG, "null parameters");
  RuleHost r;
  r.vg = *p;
  return add_rule(n, 1, syn, cnt, std::move(r), rule_id);
}

int spk_add_evans_stdp(spk_network* n, const int64_t* syn, int64_t cnt,
                       const spk_evans_stdp_params* p, int* rule_id) {
  if (!p) return fail(SPK_ERR_ARG, "null parameters");
  RuleHost r;
  r.ev = *p;
  return add_rule(n, 2, syn, cnt, std::move(r), rule_id);
}

int spk_add_custom_stdp(spk_network* n, const int64_t* syn, int64_t cnt, const spk_custom_stdp_params* p,
                        int* rule_id) {
  if (!p) return fail(SPK_ERR_ARG, "null parameters");
  if (!(p->w_max != 0.0f)) return fail(SPK_ERR_ARG, "custom STDP needs a nonzero w_max");
  RuleHost r;
  r.cu = *p;
  r.nearest = p->nearest_spike_only;
  return add_rule(n, 3, syn, cnt, std::move(r), rule_id);
}

int spk_add_weightnorm(spk_network* n, const int64_t* syn, int64_t cnt, const float* sum_sq,
                       const uint8_t* in_set, int* rule_id) {
  if (int rc = check_net(n)) return rc;
  if (!sum_sq || !in_set) return fail(SPK_ERR_ARG, "null weight normalization arrays");
  RuleHost r;
  r.wn_sumsq.assign(sum_sq, sum_sq + n->N);
  r.wn_inset.assign(in_set, in_set + n->N);
  return add_rule(n, 4, syn, cnt, std::move(r), rule_id);
}

int spk_weightnorm_set_initial(spk_network* n, int rid, const float* w0) {
  if (int rc = check_prepared(n)) return rc;
  if (rid < 0 || rid >= (int)n->rules.size() || n->rules[rid].kind != 4)
    return fail(SPK_ERR_ARG, "not a weight normalization rule");
  RuleHost& r = n->rules[rid];
  const size_t P = r.syn.size();
  if (P && !w0) return fail(SPK_ERR_ARG, "null initial weights");
  std::vector<float> dev(P);
  for (size_t j = 0; j < P; ++j) dev[j] = w0[r.order[j]];
  if (int rc = upload(r.d_ta, dev.data(), P, n->stream)) return rc;
  SPK_HIP(hipStreamSynchronize(n->stream));
  return SPK_OK;
}

int spk_add_monitor(spk_network* n, int population, int64_t capacity, int* monitor_id) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (population != SPK_POP_NEURONS && population != SPK_POP_INPUTS)
    return fail(SPK_ERR_ARG, "unknown population");
  if (capacity < 1) return fail(SPK_ERR_ARG, "monitor capacity must be >= 1");
  MonitorHost m;
  m.pop = population;
  m.capacity = capacity;
  n->mons.push_back(m);
  if (monitor_id) *monitor_id = (int)n->mons.size() - 1;
  return SPK_OK;
}

int spk_add_rate_monitor(spk_network* n, int population, int* monitor_id) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (population != SPK_POP_NEURONS && population != SPK_POP_INPUTS)
    return fail(SPK_ERR_ARG, "unknown population");
  n->rate_pop.push_back(population);
  n->rate_base.emplace_back();
  if (monitor_id) *monitor_id = (int)n->rate_pop.size() - 1;
  return SPK_OK;
}

// spk_prepare of a network whose fixed out-degree synapses are generated on the device
// (spk_set_outdegree_synapses): the delay range, fraction bits and rule tables come from
// two passes over the same Philox draws; it runs on the per-phase kernels.
static int prepare_outdegree(spk_network* n, int G) {
  const spk_outdegree_params& p = n->odp;
  const int N = n->N, Nin = n->Nin, U = n->nu();
  const int64_t S = n->S;
  const bool plastic = !n->rules.empty();
  hipStream_t st = n->stream;
  n->G = n->Gprep = G;
  n->uniform_scale = false;
  n->scale0 = p.scale_exc;
  {
    hipDeviceProp_t prop;
    SPK_HIP(hipGetDeviceProperties(&prop, n->device));
    n->cu_count = prop.multiProcessorCount;
  }
  n->P = 1;
  n->h_nlo = {0, N};
  n->h_ilo = {0, Nin};
  n->nown_max = N;
  n->iown_max = Nin;
  n->fused_ok = false;
  n->fused_why = "fixed out-degree synapses generated on the device";
  n->cap = std::max(1, U);
  n->slots = 2;
  spk::OutDegreeArgs a{};
  a.seed = p.seed; a.N = N; a.Nin = Nin; a.ne = p.n_exc; a.K = p.out_degree; a.Kin = p.input_out_degree;
  a.dmax = p.delay_max; a.scale_e = p.scale_exc; a.scale_i = p.scale_inh; a.scale_in = p.scale_input;
  a.w_inh = p.weight_inh; a.w_in = p.weight_input;
  a.plastic = plastic ? 1 : 0;
  a.cap_plastic = plastic ? std::fabs(p.stdp.w_max) : 0.0f;
  a.wfmt = n->wfmt;
  int rc;
  // pass 1: delay range, per-neuron input bound, rule counts
  double* d_bound = nullptr;
  unsigned int *d_post_cnt = nullptr, *d_row_cnt = nullptr, *d_cursor = nullptr;
  int* d_range = nullptr;
  auto cleanup = [&]() { dfree(d_bound); dfree(d_post_cnt); dfree(d_row_cnt); dfree(d_cursor); dfree(d_range); };
  if ((rc = dalloc(&d_bound, (size_t)std::max(1, N))) || (rc = dalloc(&d_post_cnt, (size_t)std::max(1, N))) ||
      (rc = dalloc(&d_cursor, (size_t)std::max(1, N))) || (rc = dalloc(&d_row_cnt, (size_t)std::max(1, U))) ||
      (rc = dalloc(&d_range, 2))) {
    cleanup();
    return rc;
  }
  const int range0[2] = {1 << 30, 0};
  SPK_HIP(hipMemsetAsync(d_bound, 0, sizeof(double) * std::max(1, N), st));
  SPK_HIP(hipMemsetAsync(d_post_cnt, 0, sizeof(unsigned int) * std::max(1, N), st));
  SPK_HIP(hipMemsetAsync(d_cursor, 0, sizeof(unsigned int) * std::max(1, N), st));
  if ((rc = upload(d_range, range0, 2, st))) {
    cleanup();
    return rc;
  }
  a.bound = d_bound; a.post_cnt = d_post_cnt; a.row_cnt = d_row_cnt; a.drange = d_range; a.cursor = d_cursor;
  const int grid = std::max(1, std::min(U, 16384));
  if (S > 0) {
    hipLaunchKernelGGL(spk::k_od_count, dim3(grid), dim3(256), 0, st, a);
    SPK_HIP(hipGetLastError());
  }
  std::vector<double> bound(std::max(1, N));
  std::vector<unsigned int> post_cnt(std::max(1, N)), row_cnt(std::max(1, U));
  int range[2] = {0, 0};
  SPK_HIP(hipMemcpyAsync(bound.data(), d_bound, sizeof(double) * bound.size(), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipMemcpyAsync(post_cnt.data(), d_post_cnt, sizeof(unsigned int) * post_cnt.size(), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipMemcpyAsync(row_cnt.data(), d_row_cnt, sizeof(unsigned int) * row_cnt.size(), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipMemcpyAsync(range, d_range, sizeof(range), hipMemcpyDeviceToHost, st));
  SPK_HIP(hipStreamSynchronize(st));
  n->dmin = S ? range[0] : 0;
  n->dmax = S ? range[1] : 0;
  if (S && G > n->dmin) {
    cleanup();
    return fail(SPK_ERR_ARG, "timestep grouping exceeds the minimum delay");
  }
  n->B = n->dmax + 2 * G + 1;           // SpikingSynapses.cu:48
  n->W = (n->dmax + 2 * G) / 32 + 1;
  double mx = 0.0;
  for (int i = 0; i < N; ++i) mx = std::max(mx, bound[i]);
  n->ext_cdf = poisson_count_table((double)n->ext_rate * (double)n->dt);
  if ((int)n->ext_cdf.size() > spk::kMaxExtTable) {
    cleanup();
    return fail(SPK_ERR_ARG, "external drive rate*dt too large for the count table");
  }
  if (!n->ext_cdf.empty()) mx += (double)n->ext_cdf.size() * std::fabs((double)n->ext_w);
  n->frac = choose_frac_bits(mx);
  n->q_ext = n->ext_cdf.empty() ? 0 : (int64_t)std::llrint(std::ldexp((double)n->ext_w, n->frac));
  // rule tables: exclusive scans of the counts
  std::vector<int64_t> pre_ptr((size_t)U + 1, 0), post_ptr((size_t)N + 1, 0);
  for (int u = 0; u < U; ++u) pre_ptr[u + 1] = pre_ptr[u] + (plastic ? row_cnt[u] : 0u);
  for (int i = 0; i < N; ++i) post_ptr[i + 1] = post_ptr[i] + (plastic ? post_cnt[i] : 0u);
  const int64_t PR = pre_ptr[U];
  if (PR >= (int64_t)INT32_MAX) {
    cleanup();
    return fail(SPK_ERR_ARG, "2^31 or more plastic synapses");
  }
  // device arrays
  if ((rc = alloc_state(n))) {
    cleanup();
    return rc;
  }
  const size_t SS = (size_t)std::max<int64_t>(1, S);
#define ALLOC(ptr, cnt) if ((rc = dalloc(&(ptr), (cnt)))) { cleanup(); return rc; }
  ALLOC(n->d_rowptr, (size_t)U + 1); ALLOC(n->d_tgt, SS); ALLOC(n->d_syn, SS);
  ALLOC(n->d_delay, SS); ALLOC(n->d_scale, SS);
  if (n->wfmt != SPK_WEIGHTS_FP32) {
    ALLOC(n->d_wc, SS);
  }
  if (plastic) {
    RuleHost& r = n->rules[0];
    const size_t P = (size_t)std::max<int64_t>(1, PR);
    r.cnt = PR;
    r.order.clear();
    r.part = {0, PR};
    r.ev_ok = PR > 0;
    r.max_post = 0;
    for (int i = 0; i < N; ++i) r.max_post = std::max<int64_t>(r.max_post, post_cnt[i]);
    ALLOC(r.d_pos, P); ALLOC(r.d_pre, P); ALLOC(r.d_ta, P); ALLOC(r.d_tb, P);
    if (n->wfmt != SPK_WEIGHTS_FP32) {
      ALLOC(r.d_ta16, P); ALLOC(r.d_tb16, P);
    }
    ALLOC(r.d_part, 2);
    ALLOC(r.d_ev_pre_ptr, (size_t)U + 1); ALLOC(r.d_ev_post_ptr, (size_t)N + 1);
    ALLOC(r.d_ev_post_pos, P); ALLOC(r.d_ev_post_pre, P);
    if (n->wfmt != SPK_WEIGHTS_FP32) {
      ALLOC(r.d_xn_a16, (size_t)U); ALLOC(r.d_xn_b16, (size_t)N);
    } else {
      ALLOC(r.d_xn_a, (size_t)U); ALLOC(r.d_xn_b, (size_t)N);
    }
    ALLOC(r.d_xcur_a, (size_t)U); ALLOC(r.d_xcur_b, (size_t)N);
    ALLOC(r.d_act, (size_t)U + N); ALLOC(r.d_act_cnt, 4);
  }
#undef ALLOC
  std::vector<int64_t> rowptr((size_t)U + 1);
  for (int u = 0; u <= U; ++u) rowptr[u] = spk::od_row(a, u);
  if ((rc = upload(n->d_rowptr, rowptr.data(), rowptr.size(), st))) {
    cleanup();
    return rc;
  }
  if (plastic) {
    RuleHost& r = n->rules[0];
    if ((rc = upload(r.d_part, r.part.data(), 2, st)) ||
        (rc = upload(r.d_ev_pre_ptr, pre_ptr.data(), pre_ptr.size(), st)) ||
        (rc = upload(r.d_ev_post_ptr, post_ptr.data(), post_ptr.size(), st))) {
      cleanup();
      return rc;
    }
    a.pre_ptr = r.d_ev_pre_ptr; a.post_ptr = r.d_ev_post_ptr;
    a.rpos = r.d_pos; a.rpre = r.d_pre; a.epos = r.d_ev_post_pos; a.epre = r.d_ev_post_pre;
  }
  // pass 2: the synapse arrays and the rule tables
  a.tgt = n->d_tgt; a.syn = n->d_syn; a.delay = n->d_delay; a.scale = n->d_scale;
  a.wc = n->d_wc;
  if (S > 0) {
    hipLaunchKernelGGL(spk::k_od_fill, dim3(grid), dim3(256), 0, st, a);
    SPK_HIP(hipGetLastError());
  }
  n->max_row = S > 0 ? std::max(N > 0 ? p.out_degree : 0, Nin > 0 ? p.input_out_degree : 0) : 0;
  if ((rc = build_wd_mask(n))) {
    cleanup();
    return rc;
  }
  if ((rc = upload_state(n))) {
    cleanup();
    return rc;
  }
  SPK_HIP(hipStreamSynchronize(st));
  cleanup();
  if ((rc = alloc_monitors(n))) return rc;
  n->prepared = true;
  return spk_reset_state(n);
}

int spk_prepare(spk_network* n, int G) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (G < 1 || G > 1024) return fail(SPK_ERR_ARG, "timestep grouping must be in [1, 1024]");
  if (n->N > 0 && !n->lif_set) return fail(SPK_ERR_STATE, "LIF parameters not set");
  if (n->Nin > 0 && n->rates.empty()) {
    n->rates.assign(n->Nin, 0.0f);
    n->n_stim = 1;
  }
  if (n->od) return prepare_outdegree(n, G);
  const int64_t S = n->S;
  n->G = n->Gprep = G;
  n->dmin = S ? 1 << 30 : 0;
  n->dmax = 0;
  const int T = host_threads();
  {
    std::vector<int> lo(T, 1 << 30), hi(T, 0);
    host_chunks(S, T, [&](int t, int64_t a, int64_t b) {
      const int* d = n->h_delay.data();
      int l = 1 << 30, h = 0;   // locals: the per-thread slots share cache lines
      for (int64_t s = a; s < b; ++s) {
        l = std::min(l, d[s]);
        h = std::max(h, d[s]);
      }
      lo[t] = l;
      hi[t] = h;
    });
    for (int t = 0; t < T; ++t)
      if (S) {
        n->dmin = std::min(n->dmin, lo[t]);
        n->dmax = std::max(n->dmax, hi[t]);
      }
  }
  n->B = n->dmax + 2 * G + 1;           // SpikingSynapses.cu:48
  n->W = (n->dmax + 2 * G) / 32 + 1;    // >= dmax + 2G bits of history (SpikingNeurons.cu:33)
  // scale of the synapse weights
  n->uniform_scale = true;
  n->scale0 = S ? n->h_scale[0] : 1.0f;
  for (int64_t s = 0; s < S; ++s)
    if (n->h_scale[s] != n->scale0) { n->uniform_scale = false; break; }
  // fixed-point fraction bits
  hvec<double> cap(S);
  host_chunks(S, T, [&](int, int64_t a, int64_t b) {
    for (int64_t s = a; s < b; ++s) cap[s] = std::fabs((double)n->h_w[s]);
  });
  for (auto& r : n->rules)   // a rule lists a synapse at most once
    host_chunks((int64_t)r.syn.size(), T, [&](int, int64_t a, int64_t b) {
      for (int64_t j = a; j < b; ++j) {
        const int64_t s = r.syn[j];
        double c0 = r.kind == 0 ? std::fabs((double)r.wd.w_max)
                  : (r.kind == 2 ? 1.0 : (r.kind == 3 ? std::fabs((double)r.cu.w_max) : 0.0));
        // normalization keeps a neuron's afferent norm: no weight exceeds sqrt(sum_sq)
        if (r.kind == 4) c0 = std::sqrt(std::fabs((double)r.wn_sumsq[n->h_post[s]]));
        cap[s] = std::max(cap[s], c0);
      }
    });
  std::vector<double> ent((size_t)n->N * n->L, 0.0);
  for (int64_t s = 0; s < S; ++s)
    ent[(size_t)n->h_post[s] * n->L + n->h_label[s]] += cap[s] * std::fabs((double)n->h_scale[s]);
  double mx = 0.0;
  for (double e : ent) mx = std::max(mx, e);
  n->ext_cdf = poisson_count_table((double)n->ext_rate * (double)n->dt);
  if ((int)n->ext_cdf.size() > spk::kMaxExtTable)
    return fail(SPK_ERR_ARG, "external drive rate*dt too large for the count table");
  if (!n->ext_cdf.empty()) mx += (double)n->ext_cdf.size() * std::fabs((double)n->ext_w);
  n->frac = choose_frac_bits(mx);
  n->q_ext = n->ext_cdf.empty() ? 0 : (int64_t)std::llrint(std::ldexp((double)n->ext_w, n->frac));
  // ---- partitions of the persistent kernel: contiguous neuron / input ranges ----
  {
    int dev_cu = 0;
    hipDeviceProp_t prop;
    SPK_HIP(hipGetDeviceProperties(&prop, n->device));
    dev_cu = prop.multiProcessorCount;
    n->cu_count = dev_cu;
    int P = n->req_partitions > 0 ? n->req_partitions : std::min(dev_cu, std::max(1, (n->N + 7) / 8));
    P = std::max(1, std::min(P, std::max(1, n->N)));
    n->P = P;
    n->h_nlo.assign(P + 1, 0);
    n->h_ilo.assign(P + 1, 0);
    for (int q = 0; q <= P; ++q) {
      n->h_nlo[q] = (int)((int64_t)n->N * q / P);
      n->h_ilo[q] = (int)((int64_t)n->Nin * q / P);
    }
    n->nown_max = 0;
    n->iown_max = 0;
    for (int q = 0; q < P; ++q) {
      n->nown_max = std::max(n->nown_max, n->h_nlo[q + 1] - n->h_nlo[q]);
      n->iown_max = std::max(n->iown_max, n->h_ilo[q + 1] - n->h_ilo[q]);
    }
    n->xcap = std::max(1, n->nown_max * max_spikes_per_group(n->tref, n->dt, G) + n->iown_max * G);
  }
  std::vector<int> owner(n->N, 0);
  for (int q = 0; q < n->P; ++q)
    for (int i = n->h_nlo[q]; i < n->h_nlo[q + 1]; ++i) owner[i] = q;
  // ---- CSR by unified presynaptic id; each row sorted by (postsynaptic partition,
  //      delay, caller order) so that a partition's synapses form one segment ----
  const int U = n->nu();
  std::vector<int64_t> rowptr(U + 1, 0);
  auto uid = [&](int32_t pre) { return pre >= 0 ? pre : n->N + (-pre - 1); };
  hvec<int64_t> order(S);
  counting_sort(S, U, T, [&](int64_t s) { return uid(n->h_pre[s]); },
                [&](int64_t s, int64_t k) { order[k] = s; }, rowptr);
  n->max_row = 0;
  for (int u = 0; u < U; ++u) n->max_row = std::max(n->max_row, rowptr[u + 1] - rowptr[u]);
  // ---- eligibility of the persistent kernel ----
  n->fused_ok = true;
  n->fused_why.clear();
  auto no_fused = [&](const char* why) { if (n->fused_ok) { n->fused_ok = false; n->fused_why = why; } };
  if (n->N < 1) no_fused("no LIF neurons");
  if (n->nown_max > spk::kFusedBlock) no_fused("more neurons per partition than threads per workgroup");
  // the draws split work items (input, block) by a float product, exact below 2^21 items
  if (n->iown_max > 65536) no_fused("more than 65536 inputs per partition");
  if (G > 64) no_fused("timestep grouping above 64");
  if (n->dmax > 255) no_fused("delays above 255 steps");
  if (n->L > 15) no_fused("more than 15 labels");
  if ((int64_t)U >= (int64_t)spk::kIdMask) no_fused("more than 2^24 neurons");
  if (S >= ((int64_t)1 << 32)) no_fused("2^32 or more synapses");
  if (n->P > spk::kMaxPartitions) no_fused("too many partitions");
  if ((int)n->mons.size() > spk::kMaxFusedMons) no_fused("more than 4 spike monitors");
  if (n->gen) no_fused("generator inputs");
  for (auto& r : n->rules)
    if (r.kind == 4) no_fused("weight normalization rule");
  n->BL = std::max(n->dmax, G);
  n->fused_lds = spk::fused_lds_bytes(n->BL, n->L, n->W, G, n->nown_max, n->iown_max, n->xcap, n->P);
  if (n->fused_lds > fused_lds_limit(n)) no_fused("input ring slice does not fit in LDS");
  // the (partition, delay) order inside a row serves only the persistent kernel's
  // segments; the per-phase kernels take a row in any order (the ring sums are exact)
  if (n->fused_ok) {
    host_chunks(U, T, [&](int, int64_t u0, int64_t u1) {
      for (int64_t u = u0; u < u1; ++u) {
        auto b = order.begin() + rowptr[u], e = order.begin() + rowptr[u + 1];
        std::stable_sort(b, e, [&](int64_t x, int64_t y) {
          const int ox = owner[n->h_post[x]], oy = owner[n->h_post[y]];
          if (ox != oy) return ox < oy;
          return n->h_delay[x] < n->h_delay[y];
        });
      }
    }, S < ((int64_t)1 << 20) ? INT64_MAX : 0);
  }
  n->perm.clear();
  n->perm.resize(S);   // the order is a permutation: every entry is written below
  hvec<int32_t> tgt(S);
  hvec<float> w(S), sc(S);
  hvec<uint16_t> dl(S);
  hvec<uint8_t> lb(S);
  host_chunks(S, T, [&](int, int64_t a, int64_t b) {
    for (int64_t k = a; k < b; ++k) {
      const int64_t s = order[k];
      n->perm[s] = k;
      tgt[k] = n->h_post[s];
      w[k] = n->h_w[s];
      sc[k] = n->h_scale[s];
      dl[k] = (uint16_t)n->h_delay[s];
      lb[k] = (uint8_t)n->h_label[s];
    }
  });
  std::vector<uint32_t> seg;
  hvec<uint32_t> packed;
  if (n->fused_ok) {
    seg.assign((size_t)U * (n->P + 1), 0);
    host_chunks(U, T, [&](int, int64_t a, int64_t b) {
      for (int64_t u = a; u < b; ++u) {
        int64_t k = rowptr[u];
        for (int q = 0; q <= n->P; ++q) {
          while (k < rowptr[u + 1] && owner[tgt[k]] < q) ++k;
          seg[(size_t)u * (n->P + 1) + q] = (uint32_t)k;
        }
      }
    }, S < ((int64_t)1 << 20) ? INT64_MAX : 0);
    packed.resize(S);
    host_chunks(S, T, [&](int, int64_t a, int64_t b) {
      for (int64_t k = a; k < b; ++k) {
        const uint32_t loc = (uint32_t)(tgt[k] - n->h_nlo[owner[tgt[k]]]);
        packed[k] = loc | ((uint32_t)dl[k] << 20) | ((uint32_t)lb[k] << 28);
      }
    });
    // compact form: one label, uniform scale, no plasticity, every row with one weight and
    // one delay, one delay for all rows, and the row weights w_lo below a split source and
    // w_hi from it on (an excitatory/inhibitory network numbered E first)
    n->compact = false;
    if (n->L == 1 && n->uniform_scale && n->rules.empty() && S > 0 && n->nown_max <= 65536 &&
        n->wfmt == SPK_WEIGHTS_FP32) {
      bool ok = true, have_lo = false, have_hi = false;
      uint32_t blo = 0, bhi = 0;
      int split = U, dconst = -1;
      for (int u = 0; u < U && ok; ++u) {
        const int64_t k0 = rowptr[u], k1 = rowptr[u + 1];
        if (k0 == k1) continue;
        uint32_t wb;
        std::memcpy(&wb, &w[k0], sizeof(float));
        if (dconst < 0) dconst = dl[k0];
        for (int64_t k = k0; k < k1 && ok; ++k) {
          uint32_t b;
          std::memcpy(&b, &w[k], sizeof(float));
          ok = b == wb && (int)dl[k] == dconst;
        }
        if (!have_lo) {
          have_lo = true;
          blo = wb;
        } else if (!have_hi) {
          if (wb != blo) {
            have_hi = true;
            bhi = wb;
            split = u;
          }
        } else if (wb != bhi) {
          ok = false;
        }
      }
      if (ok) {
        n->compact = true;
        std::memcpy(&n->cw_lo, &blo, sizeof(float));
        std::memcpy(&n->cw_hi, have_hi ? &bhi : &blo, sizeof(float));
        n->csplit = split;
        n->cdelay = dconst;
      }
    }
  }
  // count ring of the compact form: a neuron's deliveries of one step from each source type
  // are at most its in-degree from that type, which must stay below 2^16
  n->fused_lds_i64 = n->fused_lds;
  n->cring_ok = false;
  if (n->compact) {
    std::vector<uint32_t> c_lo(n->N, 0u), c_hi(n->N, 0u);
    for (int64_t s = 0; s < S; ++s) {
      const int u = uid(n->h_pre[s]);
      ++(u < n->csplit ? c_lo : c_hi)[n->h_post[s]];
    }
    const uint32_t m_lo = c_lo.empty() ? 0u : *std::max_element(c_lo.begin(), c_lo.end());
    const uint32_t m_hi = c_hi.empty() ? 0u : *std::max_element(c_hi.begin(), c_hi.end());
    n->cring_ok = m_lo < 65536u && m_hi < 65536u;
    if (n->cring_ok)
      n->fused_lds = spk::fused_lds_bytes(n->BL, n->L, n->W, G, n->nown_max, n->iown_max, n->xcap, n->P, 0, 0, 1);
  }
  if (int rc = plan_lds_table(n, seg.data(), U)) return rc;
  // list capacity: every neuron of both populations may spike once per sub-step; two
  // slots, so that a group can clear the next group's counts (k_units)
  n->cap = std::max(1, U);
  n->slots = 2;
  // allocations
  int rc;
  if ((rc = alloc_state(n))) return rc;
#define ALLOC(p, cnt) if ((rc = dalloc(&(p), (cnt)))) return rc
  ALLOC(n->d_rowptr, (size_t)U + 1); ALLOC(n->d_tgt, S); ALLOC(n->d_syn, S);
  ALLOC(n->d_delay, S);
  if (!n->uniform_scale) ALLOC(n->d_scale, S);
  if (n->L > 1) ALLOC(n->d_label, S);
  if (n->fused_ok) {
    ALLOC(n->d_nlo, (size_t)n->P + 1); ALLOC(n->d_ilo, (size_t)n->P + 1);
    ALLOC(n->d_seg, seg.size());
    ALLOC(n->d_xlist, (size_t)2 * 2 * n->P * spk::fused_xstride(n->xcap));   // 64-bit words
    ALLOC(n->d_bar, (size_t)(spk::kBarGroups + 2) * spk::kBarStride); ALLOC(n->d_err, 1);
  }
#undef ALLOC
  hipStream_t st = n->stream;
  if ((rc = upload_state(n))) return rc;
  if ((rc = upload(n->d_rowptr, rowptr.data(), (size_t)U + 1, st))) return rc;
  if ((rc = upload(n->d_tgt, tgt.data(), S, st))) return rc;
  {
    hvec<uint2> rec(S);
    host_chunks(S, T, [&](int, int64_t a, int64_t b) {
      for (int64_t k = a; k < b; ++k) {
        rec[k].x = packed.empty() ? 0u : packed[k];
        std::memcpy(&rec[k].y, &w[k], sizeof(float));
      }
    });
    if ((rc = upload(n->d_syn, rec.data(), S, st))) return rc;
  }
  if (n->wfmt != SPK_WEIGHTS_FP32) {
    // 16-bit weights of the records' weights, compensations zero
    hvec<uint32_t> wc(S);
    host_chunks(S, T, [&](int, int64_t a, int64_t b) {
      for (int64_t k = a; k < b; ++k) wc[k] = spk::enc16(w[k], n->wfmt);
    });
    if ((rc = dalloc(&n->d_wc, S))) return rc;
    if ((rc = upload(n->d_wc, wc.data(), S, st))) return rc;
  }
  if (n->compact) {
    hvec<uint16_t> t16(S);
    host_chunks(S, T, [&](int, int64_t a, int64_t b) {
      for (int64_t k = a; k < b; ++k) t16[k] = (uint16_t)(packed[k] & 0xFFFFFu);
    });
    // padded to whole 16-byte chunks plus one: deliveries read aligned 8-target chunks
    if ((rc = dalloc(&n->d_syn16, (size_t)(S + 7) / 8 * 8 + 8))) return rc;
    if ((rc = upload(n->d_syn16, t16.data(), S, st))) return rc;
  }
  if ((rc = upload(n->d_delay, dl.data(), S, st))) return rc;
  if (n->d_scale && (rc = upload(n->d_scale, sc.data(), S, st))) return rc;
  if (n->d_label && (rc = upload(n->d_label, lb.data(), S, st))) return rc;
  if (n->fused_ok) {
    if ((rc = upload(n->d_nlo, n->h_nlo.data(), (size_t)n->P + 1, st))) return rc;
    if ((rc = upload(n->d_ilo, n->h_ilo.data(), (size_t)n->P + 1, st))) return rc;
    if ((rc = upload(n->d_seg, seg.data(), seg.size(), st))) return rc;
    SPK_HIP(hipMemsetAsync(n->d_xlist, 0, sizeof(uint64_t) * 2 * n->P * spk::fused_xstride(n->xcap), st));
  }
  // rules: synapses grouped by the partition of their postsynaptic neuron
  for (auto& r : n->rules) {
    const size_t P = r.syn.size();
    // persistent kernel: stable counting sort by the partition of the postsynaptic neuron
    // (a workgroup updates the synapses onto its own neurons); per-phase kernels: CSR
    // position order, so that the rule's reads and writes of the synapse arrays stream
    r.part.assign(n->P + 1, 0);
    r.order.resize(P);
    if (n->fused_ok) {
      counting_sort((int64_t)P, n->P, T, [&](int64_t j) { return owner[n->h_post[r.syn[j]]]; },
                    [&](int64_t j, int64_t k) { r.order[k] = j; }, r.part);
    } else {
      for (size_t j = 0; j < P; ++j) r.order[j] = (int64_t)j;
      bool sorted = true;
      for (size_t j = 1; j < P && sorted; ++j) sorted = n->perm[r.syn[j - 1]] <= n->perm[r.syn[j]];
      if (!sorted)
        std::stable_sort(r.order.begin(), r.order.end(),
                         [&](int64_t x, int64_t y) { return n->perm[r.syn[x]] < n->perm[r.syn[y]]; });
      for (int q = 0; q < n->P; ++q) r.part[q + 1] = (int64_t)P;
    }
    if ((rc = dalloc(&r.d_pos, P))) return rc;
    if ((rc = dalloc(&r.d_pre, P))) return rc;
    if ((rc = dalloc(&r.d_ta, P))) return rc;
    if ((rc = dalloc(&r.d_tb, P))) return rc;
    if (n->wfmt != SPK_WEIGHTS_FP32) {
      if ((rc = dalloc(&r.d_ta16, P))) return rc;
      if ((rc = dalloc(&r.d_tb16, P))) return rc;
    }
    if ((rc = dalloc(&r.d_part, (size_t)n->P + 1))) return rc;
    hvec<int64_t> pos(P);
    hvec<int32_t> pre(P);
    host_chunks((int64_t)P, T, [&](int, int64_t a, int64_t b) {
      for (int64_t j = a; j < b; ++j) {
        const int64_t sidx = r.syn[r.order[j]];
        pos[j] = n->perm[sidx];
        pre[j] = uid(n->h_pre[sidx]);
      }
    });
    if ((rc = upload(r.d_pos, pos.data(), P, st))) return rc;
    if ((rc = upload(r.d_pre, pre.data(), P, st))) return rc;
    if ((rc = upload(r.d_part, r.part.data(), (size_t)n->P + 1, st))) return rc;
    // the weight-dependent rule's spike-driven form: its rule indices grouped by
    // presynaptic neuron (already so in CSR order) and by postsynaptic neuron
    r.ev_ok = r.kind == 0 && P > 0 && P < (size_t)INT32_MAX && S < (int64_t)UINT32_MAX;
    if (r.ev_ok) {
      const int U = n->nu();
      std::vector<int64_t> pp, qp;
      bool grouped = true;
      for (size_t j = 1; j < P && grouped; ++j) grouped = pre[j] >= pre[j - 1];
      hvec<uint32_t> qs(P);
      hvec<int32_t> qu(P);
      hvec<int32_t> pi(grouped ? 0 : P);
      counting_sort((int64_t)P, U, T, [&](int64_t j) { return pre[j]; },
                    [&](int64_t j, int64_t k) { if (!grouped) pi[k] = (int32_t)j; }, pp);
      counting_sort((int64_t)P, n->N, T, [&](int64_t j) { return n->h_post[r.syn[r.order[j]]]; },
                    [&](int64_t j, int64_t k) {
                      qs[k] = (uint32_t)pos[j];
                      qu[k] = pre[j];
                    }, qp);
      r.max_post = 0;
      for (size_t v = 0; v + 1 < qp.size(); ++v) r.max_post = std::max(r.max_post, qp[v + 1] - qp[v]);
      if ((rc = dalloc(&r.d_ev_pre_ptr, (size_t)U + 1))) return rc;
      if ((rc = dalloc(&r.d_ev_post_ptr, (size_t)n->N + 1))) return rc;
      if ((rc = dalloc(&r.d_ev_post_pos, P))) return rc;
      if ((rc = dalloc(&r.d_ev_post_pre, P))) return rc;
      if ((rc = upload(r.d_ev_pre_ptr, pp.data(), pp.size(), st))) return rc;
      if ((rc = upload(r.d_ev_post_ptr, qp.data(), qp.size(), st))) return rc;
      if ((rc = upload(r.d_ev_post_pos, qs.data(), P, st))) return rc;
      if ((rc = upload(r.d_ev_post_pre, qu.data(), P, st))) return rc;
      if (!grouped) {
        if ((rc = dalloc(&r.d_ev_pre_idx, P))) return rc;
        if ((rc = upload(r.d_ev_pre_idx, pi.data(), P, st))) return rc;
      }
      if (n->wfmt != SPK_WEIGHTS_FP32) {
        if ((rc = dalloc(&r.d_xn_a16, (size_t)U))) return rc;
        if ((rc = dalloc(&r.d_xn_b16, (size_t)n->N))) return rc;
      } else {
        if ((rc = dalloc(&r.d_xn_a, (size_t)U))) return rc;
        if ((rc = dalloc(&r.d_xn_b, (size_t)n->N))) return rc;
      }
      if ((rc = dalloc(&r.d_xcur_a, (size_t)U))) return rc;
      if ((rc = dalloc(&r.d_xcur_b, (size_t)n->N))) return rc;
      if ((rc = dalloc(&r.d_act, (size_t)U + n->N))) return rc;
      if ((rc = dalloc(&r.d_act_cnt, 4))) return rc;
      SPK_HIP(hipStreamSynchronize(st));   // the host index arrays end with this scope
    }
    if (r.kind == 4) {
      // initial weights (as stored) and the rule indices of each neuron in caller order
      std::vector<int64_t> inv(P);
      for (size_t j = 0; j < P; ++j) inv[r.order[j]] = (int64_t)j;
      std::vector<float> w0(P);
      std::vector<int64_t> qp((size_t)n->N + 1, 0);
      for (size_t j = 0; j < P; ++j) {
        const int64_t sidx = r.syn[r.order[j]];
        const float w = n->h_w[sidx];
        w0[j] = n->wfmt ? spk::dec16(spk::enc16(w, n->wfmt), n->wfmt) : w;
        qp[(size_t)n->h_post[sidx] + 1]++;
      }
      for (int v = 0; v < n->N; ++v) qp[v + 1] += qp[v];
      std::vector<int32_t> qi(std::max<size_t>(1, P));
      std::vector<int64_t> fill(qp.begin(), qp.end() - 1);
      for (size_t c = 0; c < P; ++c) qi[fill[n->h_post[r.syn[c]]]++] = (int32_t)inv[c];
      if ((rc = dalloc(&r.d_wn_sumsq, (size_t)std::max(1, n->N)))) return rc;
      if ((rc = dalloc(&r.d_wn_inset, (size_t)std::max(1, n->N)))) return rc;
      if ((rc = dalloc(&r.d_wn_ptr, (size_t)n->N + 1))) return rc;
      if ((rc = dalloc(&r.d_wn_idx, std::max<size_t>(1, P)))) return rc;
      if ((rc = upload(r.d_wn_sumsq, r.wn_sumsq.data(), (size_t)n->N, st))) return rc;
      if ((rc = upload(r.d_wn_inset, r.wn_inset.data(), (size_t)n->N, st))) return rc;
      if ((rc = upload(r.d_wn_ptr, qp.data(), qp.size(), st))) return rc;
      if ((rc = upload(r.d_wn_idx, qi.data(), P, st))) return rc;
      if ((rc = upload(r.d_ta, w0.data(), P, st))) return rc;
      SPK_HIP(hipStreamSynchronize(st));
    }
  }
  if ((rc = build_wd_mask(n))) return rc;
  if ((rc = alloc_monitors(n))) return rc;
  // the persistent grid must be co-resident: one workgroup per partition
  if (n->fused_ok) {
    int per_cu = 0;
    int per_cu_p = 0;
    SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel(n, false), spk::kFusedBlock, n->fused_lds));
    SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_p, fused_kernel(n, true), spk::kFusedBlock, n->fused_lds));
    per_cu = std::min(per_cu, per_cu_p);
    if ((int64_t)per_cu * n->cu_count < n->P) no_fused("partitions exceed co-resident workgroups");
  }
  n->prepared = true;
  return spk_reset_state(n);
}

int spk_set_partitions(spk_network* n, int partitions) {
  if (int rc = check_net(n)) return rc;
  if (n->prepared) return fail(SPK_ERR_STATE, "network already prepared");
  if (partitions < 0 || partitions > spk::kMaxPartitions) return fail(SPK_ERR_ARG, "bad partition count");
  n->req_partitions = partitions;
  return SPK_OK;
}

int spk_set_timestep_grouping(spk_network* n, int G) {
  if (int rc = check_prepared(n)) return rc;
  if (G < 1 || G > n->Gprep) return fail(SPK_ERR_ARG, "grouping must be in [1, prepared grouping]");
  if (G > n->dmin && n->S > 0) return fail(SPK_ERR_ARG, "grouping larger than the minimum delay");
  n->G = G;
  n->cur_tidx = -1;
  return SPK_OK;
}

int spk_get_info(const spk_network* n, spk_info* info) {
  if (int rc = check_net(n)) return rc;
  if (!info) return fail(SPK_ERR_ARG, "null info");
  info->n_neurons = n->N;
  info->n_inputs = n->Nin;
  info->n_labels = n->L;
  info->synapse_kind = n->kind;
  info->timestep_grouping = n->G;
  info->ring_slots = n->B;
  info->history_words = n->W;
  info->frac_bits = n->frac;
  info->min_delay = n->dmin;
  info->max_delay = n->dmax;
  info->n_synapses = n->S;
  info->n_rules = (int)n->rules.size();
  info->n_monitors = (int)n->mons.size();
  info->partitions = n->P;
  info->fused = n->fused_ok ? 1 : 0;
  info->record_bytes = (n->compact && n->rules.empty()) ? 2 : 8;
  info->lds_bytes = n->fused_ok ? (int)n->fused_lds_base : 0;
  return SPK_OK;
}

int spk_reset_inputs(spk_network* n) {
  if (int rc = check_prepared(n)) return rc;
  if (n->Nin > 0) {
    std::vector<float> ls(n->Nin, -1000.0f);
    if (int rc = upload(n->d_last_in, ls.data(), ls.size(), n->stream)) return rc;
    SPK_HIP(hipMemsetAsync(n->d_hist_in, 0, sizeof(uint32_t) * n->W * n->Nin, n->stream));
  }
  n->pending_arm = true;
  n->arm_step = -1;
  SPK_HIP(hipStreamSynchronize(n->stream));
  return SPK_OK;
}

int spk_reset_state(spk_network* n) {
  if (int rc = check_prepared(n)) return rc;
  hipStream_t st = n->stream;
  if (n->N > 0) {
    if (int rc = upload(n->d_V, n->v0.data(), (size_t)n->N, st)) return rc;
    std::vector<float> ls(n->N, -1000.0f);
    if (int rc = upload(n->d_last_n, ls.data(), ls.size(), st)) return rc;
    SPK_HIP(hipMemsetAsync(n->d_hist_n, 0, sizeof(uint32_t) * n->W * n->N, st));
    SPK_HIP(hipMemsetAsync(n->d_ring, 0, sizeof(int64_t) * n->B * n->N * n->L, st));
    SPK_HIP(hipMemsetAsync(n->d_g, 0, sizeof(float) * n->N * n->L, st));
    SPK_HIP(hipStreamSynchronize(st));
  }
  for (auto& r : n->rules) {
    const size_t P = (size_t)r.cnt;
    if (P && r.kind != 4) {   // (weight normalization keeps its initial weights there)
      SPK_HIP(hipMemsetAsync(r.d_ta, 0, sizeof(float) * P, st));
      SPK_HIP(hipMemsetAsync(r.d_tb, 0, sizeof(float) * P, st));
      if (r.d_ta16) {
        SPK_HIP(hipMemsetAsync(r.d_ta16, 0, sizeof(uint16_t) * P, st));
        SPK_HIP(hipMemsetAsync(r.d_tb16, 0, sizeof(uint16_t) * P, st));
      }
    }
    if (r.ev_ok) {
      const size_t U = (size_t)n->nu(), N = (size_t)n->N;
      if (r.d_xn_a) SPK_HIP(hipMemsetAsync(r.d_xn_a, 0, sizeof(float) * U, st));
      if (r.d_xn_b) SPK_HIP(hipMemsetAsync(r.d_xn_b, 0, sizeof(float) * N, st));
      if (r.d_xn_a16) SPK_HIP(hipMemsetAsync(r.d_xn_a16, 0, sizeof(uint16_t) * U, st));
      if (r.d_xn_b16) SPK_HIP(hipMemsetAsync(r.d_xn_b16, 0, sizeof(uint16_t) * N, st));
      SPK_HIP(hipMemsetAsync(r.d_act_cnt, 0, sizeof(int) * 4, st));
      r.par = 0;
    }
    r.syn_tr = r.neu_tr = true;
  }
  n->cur_tidx = -1;
  n->t_hi = -INFINITY;
  return spk_reset_inputs(n);
}

int spk_update_neurons(spk_network* n, float t) {
  if (int rc = check_listed(n)) return rc;
  if (int rc = begin_group(n, t, true)) return rc;
  return launch_neurons(n, t, n->cur_slot);
}

int spk_update_inputs(spk_network* n, float t) {
  if (int rc = check_listed(n)) return rc;
  if (int rc = begin_group(n, t)) return rc;
  return launch_inputs(n, t, n->cur_slot);
}

int spk_update_plasticity(spk_network* n, int rid, float t) {
  if (int rc = check_listed(n)) return rc;
  if (rid < 0 || rid >= (int)n->rules.size()) return fail(SPK_ERR_ARG, "rule id out of range");
  if (int rc = check_time(n, t)) return rc;
  return launch_rule(n, rid, t);
}

int spk_update_synapses(spk_network* n, float t) {
  if (int rc = check_listed(n)) return rc;
  if (int rc = begin_group(n, t)) return rc;
  return launch_propagate(n, t, n->cur_slot);
}

int spk_collect_spikes(spk_network* n, int mid, float t) {
  if (int rc = check_listed(n)) return rc;
  if (mid < 0 || mid >= (int)n->mons.size()) return fail(SPK_ERR_ARG, "monitor id out of range");
  if (int rc = begin_group(n, t)) return rc;
  return launch_collect(n, mid, n->cur_slot, 1);
}

int spk_run(spk_network* n, float start, int64_t ngroups, int flags) {
  if (int rc = check_prepared(n)) return rc;
  if (ngroups < 0) return fail(SPK_ERR_ARG, "negative group count");
  std::memset(&n->timing, 0, sizeof(n->timing));
  SPK_HIP(hipMemsetAsync(n->d_stats, 0, sizeof(unsigned long long) * 4, n->stream));
  const bool use_fused = n->fused_ok && !(flags & SPK_RUN_FORCE_LAUNCHES) && n->G == n->Gprep;
  SPK_HIP(hipEventRecord(n->ev_a, n->stream));
  if (!use_fused && n->generated) return check_listed(n);
  if (use_fused) {
    if (int rc = check_time(n, start)) return rc;
    if (ngroups > 0) n->t_hi = std::max(n->t_hi, start + (float)((ngroups - 1) * n->G) * n->dt);
    if (int rc = run_fused(n, start, ngroups, flags)) return rc;
  } else {
    // each monitor store must hold 64 groups of its population
    for (auto& m : n->mons) {
      const int64_t pop = m.pop == SPK_POP_NEURONS ? n->N : n->Nin;
      if (int rc = ensure_monitor_capacity(n, m, std::max<int64_t>(m.capacity, pop * n->G * 64))) return rc;
    }
    const bool plastic = (flags & SPK_RUN_PLASTICITY) != 0;
    int since_drain = 0;
    bool cleared = false;   // the previous group cleared this group's spike counts
    for (int64_t s = 0; s < ngroups; ++s) {
      const float t = start + (float)(s * n->G) * n->dt;
      int rc;
      if (fast_group_ok(n, plastic)) {
        if ((rc = launch_group_fast(n, t, plastic, !cleared))) return rc;
        cleared = true;
      } else {
        cleared = false;
        if ((rc = begin_group(n, t, true))) return rc;
        if ((rc = launch_neurons(n, t, n->cur_slot))) return rc;
        if ((rc = launch_inputs(n, t, n->cur_slot))) return rc;
        if (plastic)
          for (int r = 0; r < (int)n->rules.size(); ++r)
            if ((rc = launch_rule(n, r, t))) return rc;
        if ((rc = launch_propagate(n, t, n->cur_slot))) return rc;
      }
      for (int m = 0; m < (int)n->mons.size(); ++m)
        if ((rc = launch_collect(n, m, n->cur_slot, 1))) return rc;
      if (!n->mons.empty() && ++since_drain == 64) {
        since_drain = 0;
        for (auto& m : n->mons)
          if ((rc = drain_monitor(n, m, m.h_ids, m.h_times))) return rc;
      }
    }
    for (auto& m : n->mons)
      if (int rc = drain_monitor(n, m, m.h_ids, m.h_times)) return rc;
  }
  SPK_HIP(hipEventRecord(n->ev_b, n->stream));
  SPK_HIP(hipEventSynchronize(n->ev_b));
  float ms = 0.0f;
  SPK_HIP(hipEventElapsedTime(&ms, n->ev_a, n->ev_b));
  unsigned long long st[4] = {0, 0, 0, 0};
  SPK_HIP(hipMemcpy(st, n->d_stats, sizeof(st), hipMemcpyDeviceToHost));
  n->timing.total_ms = ms;
  n->timing.groups = ngroups;
  n->timing.synaptic_events = (int64_t)st[0];
  n->timing.external_events = (int64_t)st[1];
  n->timing.spikes_neurons = (int64_t)st[2];
  n->timing.spikes_inputs = (int64_t)st[3];
  n->cur_tidx = -1;
  return SPK_OK;
}

int spk_step_group(spk_network* n, float t, int flags) {
  if (int rc = check_prepared(n)) return rc;
  const bool use_fused = n->fused_ok && !(flags & SPK_RUN_FORCE_LAUNCHES) && n->G == n->Gprep;
  if (use_fused) {
    if (int rc = check_time(n, t)) return rc;
    const bool plastic = (flags & SPK_RUN_PLASTICITY) != 0 && !n->rules.empty();
    spk::FusedArgs a;
    if (int rc = fill_fused_args(n, plastic, a)) return rc;
    return launch_fused(n, a, t, 0, 1);
  }
  if (int rc = check_listed(n)) return rc;
  int rc;
  if ((rc = begin_group(n, t, true))) return rc;
  if ((rc = launch_neurons(n, t, n->cur_slot))) return rc;
  if ((rc = launch_inputs(n, t, n->cur_slot))) return rc;
  if (flags & SPK_RUN_PLASTICITY)
    for (int r = 0; r < (int)n->rules.size(); ++r)
      if ((rc = launch_rule(n, r, t))) return rc;
  if ((rc = launch_propagate(n, t, n->cur_slot))) return rc;
  for (int m = 0; m < (int)n->mons.size(); ++m)
    if ((rc = launch_collect(n, m, n->cur_slot, 1))) return rc;
  return SPK_OK;
}

int spk_get_counters(spk_network* n, spk_timing* t) {
  if (int rc = check_prepared(n)) return rc;
  if (!t) return fail(SPK_ERR_ARG, "null counters");
  if (int rc = check_fused_error(n)) return rc;
  unsigned long long st[4] = {0, 0, 0, 0};
  SPK_HIP(hipMemcpy(st, n->d_stats, sizeof(st), hipMemcpyDeviceToHost));
  std::memset(t, 0, sizeof(*t));
  t->synaptic_events = (int64_t)st[0];
  t->external_events = (int64_t)st[1];
  t->spikes_neurons = (int64_t)st[2];
  t->spikes_inputs = (int64_t)st[3];
  return SPK_OK;
}

int spk_stats_begin(spk_network* n, int bin_steps, int64_t max_steps) {
  if (int rc = check_prepared(n)) return rc;
  if (bin_steps < 1 || max_steps < 1) return fail(SPK_ERR_ARG, "bin and window must be positive");
  const int64_t nb = (max_steps + bin_steps - 1) / bin_steps;
  if (nb > ((int64_t)1 << 31)) return fail(SPK_ERR_ARG, "statistics window too long");
  const size_t N = (size_t)std::max(1, n->N);
  if (!n->d_st_cnt) {
    int rc;
    if ((rc = dalloc(&n->d_st_cnt, N)) || (rc = dalloc(&n->d_st_nisi, N)) || (rc = dalloc(&n->d_st_prev, N)) ||
        (rc = dalloc(&n->d_st_s1, N)) || (rc = dalloc(&n->d_st_s2, N)) || (rc = dalloc(&n->d_st_out, 1)))
      return rc;
  }
  if (nb > n->st_nbins) {
    dfree(n->d_st_hist);
    if (int rc = dalloc(&n->d_st_hist, (size_t)nb)) return rc;
    n->st_nbins = nb;
  }
  SPK_HIP(hipMemsetAsync(n->d_st_cnt, 0, N * sizeof(unsigned int), n->stream));
  SPK_HIP(hipMemsetAsync(n->d_st_nisi, 0, N * sizeof(unsigned int), n->stream));
  SPK_HIP(hipMemsetAsync(n->d_st_prev, 0, N * sizeof(float), n->stream));
  SPK_HIP(hipMemsetAsync(n->d_st_s1, 0, N * sizeof(double), n->stream));
  SPK_HIP(hipMemsetAsync(n->d_st_s2, 0, N * sizeof(double), n->stream));
  SPK_HIP(hipMemsetAsync(n->d_st_hist, 0, (size_t)nb * sizeof(unsigned int), n->stream));
  n->st_on = true;
  n->st_pending = true;
  n->st_bin = bin_steps;
  n->st_max = max_steps;
  n->st_steps = 0;
  return SPK_OK;
}

int spk_stats_get(spk_network* n, spk_stats* out) {
  if (int rc = check_prepared(n)) return rc;
  if (!out) return fail(SPK_ERR_ARG, "null output");
  if (!n->d_st_cnt) return fail(SPK_ERR_STATE, "statistics were never started");
  spk::StatsArgs st{};
  st.count = n->d_st_cnt; st.nisi = n->d_st_nisi; st.prev = n->d_st_prev;
  st.s1 = n->d_st_s1; st.s2 = n->d_st_s2; st.hist = n->d_st_hist;
  st.t0 = n->st_t0; st.nsteps = n->st_max; st.bin = n->st_bin; st.on = 1;
  hipLaunchKernelGGL(spk::k_stats_reduce, dim3(1), dim3(spk::kStatsBlock), 0, n->stream, st, n->N,
                     n->st_steps, n->dt, n->d_st_out);
  SPK_HIP(hipGetLastError());
  spk::StatsOut o;
  SPK_HIP(hipMemcpyAsync(&o, n->d_st_out, sizeof(o), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  out->rate_hz = o.rate_hz;
  out->isi_cv = o.isi_cv;
  out->synchrony = o.synchrony;
  out->mean_bin_count = o.mean_bin_count;
  out->spikes = o.spikes;
  out->cv_neurons = o.cv_neurons;
  out->bins = o.bins;
  out->steps = o.steps;
  return SPK_OK;
}

int spk_stats_end(spk_network* n) {
  if (int rc = check_net(n)) return rc;
  n->st_on = false;
  return SPK_OK;
}

int spk_set_profiling(spk_network* n, int on) {
  if (int rc = check_net(n)) return rc;
  n->profiling = on != 0;
  return SPK_OK;
}

int spk_get_phase_profile(const spk_network* n, double* us, int count) {
  if (int rc = check_net(n)) return rc;
  if (!us || count < 0) return fail(SPK_ERR_ARG, "bad output");
  for (int i = 0; i < count && i < spk::kProfPhases; ++i) us[i] = n->prof_us[i];
  return SPK_OK;
}

const char* spk_fused_status(const spk_network* n) {
  if (!n) return "null network";
  return n->fused_ok ? "" : n->fused_why.c_str();
}

int spk_last_timing(const spk_network* n, spk_timing* t) {
  if (int rc = check_net(n)) return rc;
  if (!t) return fail(SPK_ERR_ARG, "null timing");
  *t = n->timing;
  return SPK_OK;
}

int spk_synchronize(spk_network* n) {
  if (int rc = check_net(n)) return rc;
  SPK_HIP(hipStreamSynchronize(n->stream));
  if (n->prepared) return check_fused_error(n);
  return SPK_OK;
}

int spk_monitor_device_count(spk_network* n, int mid, int64_t* count) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->mons.size() || !count) return fail(SPK_ERR_ARG, "bad monitor");
  unsigned long long c = 0;
  SPK_HIP(hipMemcpyAsync(&c, n->mons[mid].d_count, sizeof(c), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  *count = (int64_t)c;
  return check_fused_error(n);
}

int spk_monitor_reset(spk_network* n, int mid) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->mons.size()) return fail(SPK_ERR_ARG, "bad monitor");
  MonitorHost& m = n->mons[mid];
  SPK_HIP(hipMemsetAsync(m.d_count, 0, sizeof(unsigned long long), n->stream));
  m.h_ids.clear();
  m.h_times.clear();
  return SPK_OK;
}

int spk_monitor_drain(spk_network* n, int mid, int32_t* ids, float* times, int64_t cap, int64_t* out_n) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->mons.size() || !out_n) return fail(SPK_ERR_ARG, "bad monitor");
  std::vector<int32_t> bi;
  std::vector<float> bt;
  if (int rc = drain_monitor(n, n->mons[mid], bi, bt)) return rc;
  *out_n = (int64_t)bi.size();
  if ((int64_t)bi.size() > cap) return fail(SPK_ERR_CAPACITY, "output buffer too small");
  if (bi.size()) {
    std::memcpy(ids, bi.data(), bi.size() * sizeof(int32_t));
    std::memcpy(times, bt.data(), bt.size() * sizeof(float));
  }
  return SPK_OK;
}

int spk_monitor_take(spk_network* n, int mid, int32_t* ids, float* times, int64_t cap, int64_t* out_n) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->mons.size() || !out_n) return fail(SPK_ERR_ARG, "bad monitor");
  MonitorHost& m = n->mons[mid];
  *out_n = (int64_t)m.h_ids.size();
  if (!ids) return SPK_OK;
  if ((int64_t)m.h_ids.size() > cap) return fail(SPK_ERR_CAPACITY, "output buffer too small");
  if (!m.h_ids.empty()) {
    std::memcpy(ids, m.h_ids.data(), m.h_ids.size() * sizeof(int32_t));
    std::memcpy(times, m.h_times.data(), m.h_times.size() * sizeof(float));
  }
  m.h_ids.clear();
  m.h_times.clear();
  return SPK_OK;
}

int spk_get_weights(spk_network* n, float* out) {
  if (int rc = check_listed(n)) return rc;
  if (!out && n->S) return fail(SPK_ERR_ARG, "null output");
  if (n->wfmt != SPK_WEIGHTS_FP32) {
    std::vector<uint32_t> wc(n->S);
    if (n->S) SPK_HIP(hipMemcpyAsync(wc.data(), n->d_wc, n->S * sizeof(uint32_t), hipMemcpyDeviceToHost, n->stream));
    SPK_HIP(hipStreamSynchronize(n->stream));
    for (int64_t s = 0; s < n->S; ++s) out[s] = spk::dec16((uint16_t)(wc[csr_pos(n, s)] & 0xFFFFu), n->wfmt);
    return SPK_OK;
  }
  std::vector<uint2> rec(n->S);
  if (n->S) SPK_HIP(hipMemcpyAsync(rec.data(), n->d_syn, n->S * sizeof(uint2), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  for (int64_t s = 0; s < n->S; ++s) std::memcpy(&out[s], &rec[csr_pos(n, s)].y, sizeof(float));
  return SPK_OK;
}

int spk_set_weights(spk_network* n, const float* in) {
  if (int rc = check_listed(n)) return rc;
  if (!in && n->S) return fail(SPK_ERR_ARG, "null input");
  std::vector<uint2> rec(n->S);
  if (n->S) SPK_HIP(hipMemcpyAsync(rec.data(), n->d_syn, n->S * sizeof(uint2), hipMemcpyDeviceToHost, n->stream));
  SPK_HIP(hipStreamSynchronize(n->stream));
  for (int64_t s = 0; s < n->S; ++s) std::memcpy(&rec[csr_pos(n, s)].y, &in[s], sizeof(float));
  if (n->S) SPK_HIP(hipMemcpyAsync(n->d_syn, rec.data(), n->S * sizeof(uint2), hipMemcpyHostToDevice, n->stream));
  if (n->wfmt != SPK_WEIGHTS_FP32 && n->S) {
    // the 16-bit words of the caller's weights; compensations start again from zero
    std::vector<uint32_t> wc(n->S);
    for (int64_t s = 0; s < n->S; ++s) wc[csr_pos(n, s)] = spk::enc16(in[s], n->wfmt);
    SPK_HIP(hipMemcpyAsync(n->d_wc, wc.data(), n->S * sizeof(uint32_t), hipMemcpyHostToDevice, n->stream));
  }
  SPK_HIP(hipStreamSynchronize(n->stream));
  n->compact = false;   // caller weights: deliveries read the full records from now on
  n->cring_ok = false;  // and the LDS ring holds fixed-point sums again
  n->lseg_U = n->lrec = 0;
  if (n->fused_lds_i64) n->fused_lds = n->fused_lds_base = n->fused_lds_i64;
  for (auto& r : n->rules) r.dense_next = true;   // caller weights may lie below 0
  return SPK_OK;
}

int spk_get_membrane(spk_network* n, float* v) {
  if (int rc = check_prepared(n)) return rc;
  return download(n, v, n->d_V, (size_t)n->N);
}

int spk_get_last_spike(spk_network* n, int pop, float* ls) {
  if (int rc = check_prepared(n)) return rc;
  if (pop == SPK_POP_NEURONS) return download(n, ls, n->d_last_n, (size_t)n->N);
  return download(n, ls, n->d_last_in, (size_t)n->Nin);
}

int spk_get_conductances(spk_network* n, float* g) {
  if (int rc = check_prepared(n)) return rc;
  return download(n, g, n->d_g, (size_t)n->N * n->L);
}

int spk_get_ring(spk_network* n, int64_t* ring) {
  if (int rc = check_prepared(n)) return rc;
  return download(n, ring, n->d_ring, (size_t)n->B * n->N * n->L);
}

int spk_get_history(spk_network* n, int pop, uint32_t* words) {
  if (int rc = check_prepared(n)) return rc;
  if (pop == SPK_POP_NEURONS) return download(n, words, n->d_hist_n, (size_t)n->W * n->N);
  return download(n, words, n->d_hist_in, (size_t)n->W * n->Nin);
}

int spk_get_traces(spk_network* n, int rid, float* a, float* b) {
  if (int rc = check_listed(n)) return rc;
  if (rid < 0 || rid >= (int)n->rules.size()) return fail(SPK_ERR_ARG, "rule id out of range");
  RuleHost& r = n->rules[rid];
  const size_t P = (size_t)r.cnt;
  if (int rc = traces_per_synapse(n, r)) return rc;
  std::vector<float> ta(P), tb(P);
  if (r.kind == 4) {   // weight normalization: its initial weights, and zeros
    if (int rc = download(n, ta.data(), r.d_ta, P)) return rc;
  } else if (r.d_ta16) {
    std::vector<uint16_t> a16(P), b16(P);
    if (int rc = download(n, a16.data(), r.d_ta16, P)) return rc;
    if (int rc = download(n, b16.data(), r.d_tb16, P)) return rc;
    for (size_t j = 0; j < P; ++j) {
      ta[j] = spk::dec16(a16[j], SPK_WEIGHTS_FP16);
      tb[j] = spk::dec16(b16[j], SPK_WEIGHTS_FP16);
    }
  } else {
    if (int rc = download(n, ta.data(), r.d_ta, P)) return rc;
    if (int rc = download(n, tb.data(), r.d_tb, P)) return rc;
  }
  for (size_t j = 0; j < P; ++j) {   // device order -> rule order (identity when generated)
    const size_t o = r.order.empty() ? j : (size_t)r.order[j];
    a[o] = ta[j];
    b[o] = tb[j];
  }
  return SPK_OK;
}

int spk_get_rule_size(spk_network* n, int rid, int64_t* count) {
  if (int rc = check_prepared(n)) return rc;
  if (rid < 0 || rid >= (int)n->rules.size()) return fail(SPK_ERR_ARG, "rule id out of range");
  if (!count) return fail(SPK_ERR_ARG, "null output");
  *count = n->rules[rid].cnt;
  return SPK_OK;
}

int spk_rate_monitor_counts(spk_network* n, int mid, int32_t* counts) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->rate_pop.size()) return fail(SPK_ERR_ARG, "rate monitor id out of range");
  const int pop = n->rate_pop[mid];
  const size_t cnt = (size_t)(pop == SPK_POP_NEURONS ? n->N : n->Nin);
  std::vector<uint32_t> cur(std::max<size_t>(1, cnt));
  if (int rc = download(n, cur.data(), n->d_rate[pop], cnt)) return rc;
  for (size_t i = 0; i < cnt; ++i) counts[i] = (int32_t)(cur[i] - n->rate_base[mid][i]);
  return SPK_OK;
}

int spk_rate_monitor_reset(spk_network* n, int mid) {
  if (int rc = check_prepared(n)) return rc;
  if (mid < 0 || mid >= (int)n->rate_pop.size()) return fail(SPK_ERR_ARG, "rate monitor id out of range");
  const int pop = n->rate_pop[mid];
  const size_t cnt = (size_t)(pop == SPK_POP_NEURONS ? n->N : n->Nin);
  return download(n, n->rate_base[mid].data(), n->d_rate[pop], cnt);
}

// ---------------------------------------------------------------------------------------
// Brunel networks generated on the device

int spk_create_brunel(const spk_brunel_params* p, int partitions, spk_network** out) {
  if (!p || !out) return fail(SPK_ERR_ARG, "null argument");
  *out = nullptr;
  if (p->n_exc < 1 || p->n_inh < 0 || p->ce < 0 || p->ci < 0) return fail(SPK_ERR_ARG, "bad population size or in-degree");
  if (p->ci > 0 && p->n_inh < 1) return fail(SPK_ERR_ARG, "inhibitory in-degree without inhibitory neurons");
  if (p->delay_steps < 1 || p->delay_steps > 64) return fail(SPK_ERR_ARG, "delay must be in [1, 64] steps");
  const int N = p->n_exc + p->n_inh;
  const int C = p->ce + p->ci;
  const int64_t S = (int64_t)N * C;
  if (S >= ((int64_t)1 << 32)) return fail(SPK_ERR_ARG, "2^32 or more synapses");
  spk_network* n = nullptr;
  if (int rc = spk_create(N, 0, SPK_SYN_VOLTAGE, 1, p->timestep, &n)) return rc;
  auto bail = [&](int rc) {
    const std::string msg = g_error;
    spk_destroy(n);
    g_error = msg;
    return rc;
  };
  {
    std::vector<float> v0(N, p->v0), vr(N, p->vreset), th(N, p->threshold), c(N, p->decay_c), rp(N, p->resist);
    int rc;
    if ((rc = spk_set_lif(n, v0.data(), vr.data(), th.data(), c.data(), rp.data(), p->refractory_s, 0.0f)) ||
        (rc = spk_set_seed(n, p->seed)) || (rc = spk_set_external_drive(n, p->ext_rate_hz, p->ext_weight)))
      return bail(rc);
  }
  n->req_partitions = partitions;
  const int G = p->delay_steps;
  n->S = S;
  n->G = n->Gprep = G;
  n->dmin = n->dmax = S ? G : 0;
  n->B = n->dmax + 2 * G + 1;
  n->W = (n->dmax + 2 * G) / 32 + 1;
  n->uniform_scale = true;
  n->scale0 = 1.0f;
  // fraction bits: every neuron's per-step input bound, summed in synapse order (E then I)
  double mx = 0.0;
  for (int j = 0; j < p->ce; ++j) mx += std::fabs((double)p->weight_exc);
  for (int j = 0; j < p->ci; ++j) mx += std::fabs((double)p->weight_inh);
  n->ext_cdf = poisson_count_table((double)n->ext_rate * (double)n->dt);
  if ((int)n->ext_cdf.size() > spk::kMaxExtTable) return bail(fail(SPK_ERR_ARG, "external drive too strong"));
  if (!n->ext_cdf.empty()) mx += (double)n->ext_cdf.size() * std::fabs((double)n->ext_w);
  n->frac = choose_frac_bits(mx);
  n->q_ext = n->ext_cdf.empty() ? 0 : (int64_t)std::llrint(std::ldexp((double)n->ext_w, n->frac));
  hipDeviceProp_t prop;
  SPK_HIP(hipGetDeviceProperties(&prop, n->device));
  n->cu_count = prop.multiProcessorCount;
  int P = partitions > 0 ? partitions : std::min(n->cu_count, std::max(1, (N + 7) / 8));
  P = std::max(1, std::min(P, N));
  n->P = P;
  n->h_nlo.assign(P + 1, 0);
  n->h_ilo.assign(P + 1, 0);
  for (int q = 0; q <= P; ++q) n->h_nlo[q] = (int)((int64_t)N * q / P);
  n->nown_max = 0;
  n->iown_max = 0;
  for (int q = 0; q < P; ++q) n->nown_max = std::max(n->nown_max, n->h_nlo[q + 1] - n->h_nlo[q]);
  n->xcap = std::max(1, n->nown_max * max_spikes_per_group(n->tref, n->dt, G));
  n->BL = std::max(n->dmax, G);
  n->cring_ok = p->ce < 65536 && p->ci < 65536;   // counts per step and source type
  n->fused_lds_i64 = spk::fused_lds_bytes(n->BL, 1, n->W, G, n->nown_max, 0, n->xcap, P);
  n->fused_lds = spk::fused_lds_bytes(n->BL, 1, n->W, G, n->nown_max, 0, n->xcap, P, 0, 0, n->cring_ok ? 1 : 0);
  if (n->nown_max > spk::kFusedBlock) return bail(fail(SPK_ERR_ARG, "more neurons per partition than threads per workgroup"));
  if (P > spk::kMaxPartitions) return bail(fail(SPK_ERR_ARG, "too many partitions"));
  if (n->fused_lds > fused_lds_limit(n))
    return bail(fail(SPK_ERR_ARG, "input ring slice does not fit in LDS: use more partitions"));
  n->fused_ok = true;
  n->fused_why.clear();
  n->cap = 1;
  n->slots = 1;
  int rc;
  const size_t NN = (size_t)N;
#define ALLOC(ptr, cnt) if ((rc = dalloc(&(ptr), (cnt)))) return bail(rc)
  ALLOC(n->d_V, NN); ALLOC(n->d_v0, NN); ALLOC(n->d_vreset, NN); ALLOC(n->d_thr, NN);
  ALLOC(n->d_c, NN); ALLOC(n->d_rp, NN); ALLOC(n->d_last_n, NN); ALLOC(n->d_last_in, 1);
  ALLOC(n->d_hist_n, (size_t)n->W * NN); ALLOC(n->d_hist_in, 1);
  ALLOC(n->d_ring, (size_t)n->B * NN); ALLOC(n->d_g, NN);
  ALLOC(n->d_decay, 1); ALLOC(n->d_erev, 1); ALLOC(n->d_in_thr, 1);
  ALLOC(n->d_ext_cdf, std::max<size_t>(1, n->ext_cdf.size()));
  ALLOC(n->d_syn, (size_t)std::max<int64_t>(1, S));
  ALLOC(n->d_syn16, (size_t)(std::max<int64_t>(1, S) + 7) / 8 * 8 + 8);   // whole 16-byte chunks
  ALLOC(n->d_stats, 4);
  ALLOC(n->d_nlo, (size_t)P + 1); ALLOC(n->d_ilo, (size_t)P + 1);
  ALLOC(n->d_seg, (size_t)N * (P + 1));
  ALLOC(n->d_xlist, (size_t)2 * 2 * P * spk::fused_xstride(n->xcap));   // 64-bit words
  ALLOC(n->d_bar, (size_t)(spk::kBarGroups + 2) * spk::kBarStride); ALLOC(n->d_err, 1);
  int* d_owner = nullptr;
  unsigned int *d_count = nullptr, *d_off = nullptr, *d_cursor = nullptr;
  ALLOC(d_owner, NN);
  ALLOC(d_count, NN * P); ALLOC(d_off, NN * P); ALLOC(d_cursor, NN * P);
#undef ALLOC
  hipStream_t st = n->stream;
  auto cleanup = [&]() { dfree(d_owner); dfree(d_count); dfree(d_off); dfree(d_cursor); };
  std::vector<int> owner(N);
  for (int q = 0; q < P; ++q)
    for (int i = n->h_nlo[q]; i < n->h_nlo[q + 1]; ++i) owner[i] = q;
  if ((rc = upload(n->d_v0, n->v0.data(), NN, st)) || (rc = upload(n->d_vreset, n->vreset.data(), NN, st)) ||
      (rc = upload(n->d_thr, n->thr.data(), NN, st)) || (rc = upload(n->d_c, n->c.data(), NN, st)) ||
      (rc = upload(n->d_rp, n->rp.data(), NN, st)) || (rc = upload(n->d_decay, n->decay_g.data(), 1, st)) ||
      (rc = upload(n->d_erev, n->erev.data(), 1, st)) ||
      (rc = upload(n->d_ext_cdf, n->ext_cdf.data(), n->ext_cdf.size(), st)) ||
      (rc = upload(n->d_nlo, n->h_nlo.data(), (size_t)P + 1, st)) ||
      (rc = upload(n->d_ilo, n->h_ilo.data(), (size_t)P + 1, st)) || (rc = upload(d_owner, owner.data(), NN, st))) {
    cleanup();
    return bail(rc);
  }
  SPK_HIP(hipMemsetAsync(n->d_xlist, 0, sizeof(uint64_t) * 2 * P * spk::fused_xstride(n->xcap), st));
  SPK_HIP(hipMemsetAsync(n->d_stats, 0, sizeof(unsigned long long) * 4, st));
  SPK_HIP(hipMemsetAsync(d_count, 0, sizeof(unsigned int) * NN * P, st));
  SPK_HIP(hipMemsetAsync(d_cursor, 0, sizeof(unsigned int) * NN * P, st));
  spk::BrunelGenArgs ga{};
  ga.seed = n->seed; ga.ne = p->n_exc; ga.ni = p->n_inh; ga.ce = p->ce; ga.ci = p->ci; ga.P = P;
  ga.owner = d_owner; ga.nlo = n->d_nlo; ga.count = d_count; ga.off = d_off; ga.cursor = d_cursor;
  ga.syn = n->d_syn; ga.syn16 = n->d_syn16;
  ga.delay_bits = (uint32_t)G << 20; ga.we = p->weight_exc; ga.wi = p->weight_inh;
  // every excitatory source weighs weight_exc, every inhibitory one weight_inh, one delay
  n->compact = S > 0;
  n->cw_lo = p->weight_exc; n->cw_hi = p->weight_inh; n->csplit = p->n_exc; n->cdelay = G;
  const int grid = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (S + 255) / 256));
  if (S > 0) {
    hipLaunchKernelGGL(spk::k_gen_count, dim3(grid), dim3(256), 0, st, ga, S);
    SPK_HIP(hipGetLastError());
  }
  // exclusive scan of the key counts: the segment table seg[u][P+1]
  std::vector<unsigned int> cnt(NN * P), off(NN * P);
  SPK_HIP(hipMemcpyAsync(cnt.data(), d_count, sizeof(unsigned int) * NN * P, hipMemcpyDeviceToHost, st));
  SPK_HIP(hipStreamSynchronize(st));
  std::vector<uint32_t> seg(NN * (P + 1));
  uint64_t run = 0;
  for (size_t u = 0; u < NN; ++u) {
    for (int q = 0; q < P; ++q) {
      off[u * P + q] = (unsigned int)run;
      seg[u * (P + 1) + q] = (uint32_t)run;
      run += cnt[u * P + q];
    }
    seg[u * (P + 1) + P] = (uint32_t)run;
  }
  if ((rc = upload(d_off, off.data(), NN * P, st)) || (rc = upload(n->d_seg, seg.data(), seg.size(), st))) {
    cleanup();
    return bail(rc);
  }
  if ((rc = plan_lds_table(n, seg.data(), N))) {
    cleanup();
    return bail(rc);
  }
  if (S > 0) {
    hipLaunchKernelGGL(spk::k_gen_scatter, dim3(grid), dim3(256), 0, st, ga, S);
    SPK_HIP(hipGetLastError());
  }
  SPK_HIP(hipStreamSynchronize(st));
  cleanup();
  int per_cu = 0;
  SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fused_kernel(n, false), spk::kFusedBlock, n->fused_lds));
  if ((int64_t)per_cu * n->cu_count < P) return bail(fail(SPK_ERR_ARG, "partitions exceed co-resident workgroups"));
  n->prepared = true;
  n->generated = true;
  if ((rc = spk_reset_state(n))) return bail(rc);
  *out = n;
  return SPK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Batches of independent networks in one persistent launch

struct spk_batch {
  std::vector<spk_network*> nets;
  int kind = 0;
  bool l1 = true;
  int nb = 0;                 // workgroups of the launch
  int bs = spk::kFusedBlock;  // threads per workgroup
  size_t lds = 0;
  bool lds_tables = true;     // networks keep their delivery tables in LDS where planned
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  spk::FusedArgs* d_args = nullptr;
  int2* d_tasks = nullptr;
  int* d_off = nullptr;
  size_t tasks_cap = 0;
  double cost_fixed_us = 8.0, cost_neuron_us = 0.004, cost_synapse_us = 2e-5;
  std::vector<double> cost_per_group_us;   // measured per-network costs (empty: the model)
  double last_ms = 0.0;
  double est_ms = 0.0;
  // hipGraph of a run: the argument copies (from pinned staging, read when they execute)
  // and the cooperative launch, captured at the first run and replayed by every later run
  // of the same shape (kernel, task count, device buffers)
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void* g_fn = nullptr;
  size_t g_tasks = 0;
  const int2* g_dtasks = nullptr;
  spk::FusedArgs* h_args = nullptr;
  int2* h_tasks = nullptr;
  size_t h_tasks_cap = 0;
  int* h_off = nullptr;
  int64_t captures = 0, replays = 0;
};

namespace {

void batch_drop_graph(spk_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->g_fn = nullptr;
}

template <int BS>
const void* batch_kernel_bs(const spk_batch* b) {
  if (b->kind == SPK_SYN_VOLTAGE)
    return b->l1 ? (const void*)spk::k_fused_batch<0, true, false, BS> : (const void*)spk::k_fused_batch<0, false, false, BS>;
  if (b->kind == SPK_SYN_CONDUCTANCE)
    return b->l1 ? (const void*)spk::k_fused_batch<1, true, false, BS> : (const void*)spk::k_fused_batch<1, false, false, BS>;
  return b->l1 ? (const void*)spk::k_fused_batch<2, true, false, BS> : (const void*)spk::k_fused_batch<2, false, false, BS>;
}

// Batches of Brunel-type networks (voltage synapses, one label, compact records) run
// kernels without the full-record delivery: compact only when no network keeps its table in
// LDS, else both compact forms.
template <int BS>
const void* batch_kernel_compact(const spk_batch* b, bool tables) {
  return tables ? (const void*)spk::k_fused_batch<0, true, false, BS, spk::kDeliverCompactAny>
                : (const void*)spk::k_fused_batch<0, true, false, BS, spk::kDeliverCompact>;
}

const void* batch_kernel(const spk_batch* b, int bs) {
  bool compact = b->kind == SPK_SYN_VOLTAGE && b->l1, tables = false;
  for (auto* n : b->nets) {
    compact = compact && n->compact && n->rules.empty();
    tables = tables || n->lseg_U > 0;
  }
  tables = tables && bs != spk::kBatchSmallBlock && b->lds_tables;
  if (compact)
    return bs == spk::kBatchSmallBlock ? batch_kernel_compact<spk::kBatchSmallBlock>(b, false)
                                       : batch_kernel_compact<spk::kFusedBlock>(b, tables);
  return bs == spk::kBatchSmallBlock ? batch_kernel_bs<spk::kBatchSmallBlock>(b) : batch_kernel_bs<spk::kFusedBlock>(b);
}

}  // namespace

extern "C" {

int spk_batch_create(spk_network* const* nets, int n_nets, int workgroups, spk_batch** out) {
  if (!nets || n_nets < 1 || !out) return fail(SPK_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  spk_batch* b = new spk_batch();
  for (int k = 0; k < n_nets; ++k) {
    spk_network* n = nets[k];
    if (int rc = check_prepared(n)) { delete b; return rc; }
    if (!n->fused_ok) { delete b; return fail(SPK_ERR_ARG, "network " + std::to_string(k) + " cannot use the persistent kernel: " + n->fused_why); }
    if (k == 0) { b->kind = n->kind; b->l1 = n->L == 1; }
    if (n->kind != b->kind || (n->L == 1) != b->l1) { delete b; return fail(SPK_ERR_ARG, "networks of one batch need the same synapse kind and label count class"); }
    if (n->device != nets[0]->device) { delete b; return fail(SPK_ERR_ARG, "networks on different devices"); }
    b->lds = std::max(b->lds, n->fused_lds);
    b->nets.push_back(n);
  }
  size_t lds_base = 0;
  for (auto* n : b->nets) lds_base = std::max(lds_base, n->fused_lds_base);
  const void* fn = batch_kernel(b, spk::kFusedBlock);
  if (b->lds > 65536) SPK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds));
  int per_cu = 0;
  SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, spk::kFusedBlock, b->lds));
  // smaller workgroups when two of them fit a compute unit (every partition at most
  // kBatchSmallBlock neurons, LDS and registers of two): one partition runs while the
  // other waits at its network's exchange
  int pmax_own = 0;
  for (auto* n : b->nets) pmax_own = std::max(pmax_own, n->nown_max);
  // (without the LDS delivery tables, which would take the second workgroup's room)
  if (pmax_own <= spk::kBatchSmallBlock) {
    const void* fs = batch_kernel(b, spk::kBatchSmallBlock);
    if (lds_base > 65536) SPK_HIP(hipFuncSetAttribute(fs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_base));
    int per_cu_s = 0;
    SPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s, fs, spk::kBatchSmallBlock, lds_base));
    if (per_cu_s >= 2 * per_cu) {
      b->bs = spk::kBatchSmallBlock;
      b->lds = lds_base;
      b->lds_tables = false;
      per_cu = per_cu_s;
    }
  }
  const int cap = per_cu * nets[0]->cu_count;
  b->nb = workgroups > 0 ? std::min(workgroups, cap) : cap;
  int pmax = 0;
  for (auto* n : b->nets) pmax = std::max(pmax, n->P);
  if (b->nb < pmax || b->nb < 1) {
    delete b;
    return fail(SPK_ERR_ARG, "a network has more partitions than the batch has co-resident workgroups");
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return fail(SPK_ERR_HIP, "hipStreamCreate"); }
  (void)hipEventCreate(&b->e0);
  (void)hipEventCreate(&b->e1);
  if (int rc = dalloc(&b->d_args, b->nets.size())) { delete b; return rc; }
  *out = b;
  return SPK_OK;
}

int spk_batch_destroy(spk_batch* b) {
  if (!b) return SPK_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->h_args) (void)hipHostFree(b->h_args);
  if (b->h_tasks) (void)hipHostFree(b->h_tasks);
  if (b->h_off) (void)hipHostFree(b->h_off);
  dfree(b->d_args);
  dfree(b->d_tasks);
  dfree(b->d_off);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return SPK_OK;
}

int spk_batch_set_cost_model(spk_batch* b, double fixed_us, double per_neuron_us, double per_synapse_us) {
  if (!b) return fail(SPK_ERR_ARG, "null batch");
  b->cost_fixed_us = fixed_us;
  b->cost_neuron_us = per_neuron_us;
  b->cost_synapse_us = per_synapse_us;
  return SPK_OK;
}

int spk_batch_set_costs(spk_batch* b, const double* cost_per_group_us) {
  if (!b) return fail(SPK_ERR_ARG, "null batch");
  if (!cost_per_group_us) {
    b->cost_per_group_us.clear();
    return SPK_OK;
  }
  b->cost_per_group_us.assign(cost_per_group_us, cost_per_group_us + b->nets.size());
  return SPK_OK;
}

int spk_batch_workgroups(const spk_batch* b) { return b ? b->nb : 0; }

int spk_batch_threads(const spk_batch* b) { return b ? b->bs : 0; }

int spk_batch_run(spk_batch* b, const float* starts, const int64_t* ngroups, int flags) {
  if (!b || !starts || !ngroups) return fail(SPK_ERR_ARG, "null argument");
  if (flags & SPK_RUN_PLASTICITY) return fail(SPK_ERR_ARG, "batches run without plasticity");
  const int K = (int)b->nets.size();
  for (int k = 0; k < K; ++k) {
    if (ngroups[k] < 0 || ngroups[k] > 0x7FFFFFFF) return fail(SPK_ERR_ARG, "bad group count");
    SPK_HIP(hipStreamSynchronize(b->nets[k]->stream));   // earlier work of every network is done
  }
  // list scheduling in network order: each network takes the P workgroups that are free
  // first; every workgroup's task list is then in network order
  std::vector<double> avail(b->nb, 0.0);
  std::vector<std::vector<int2>> per(b->nb);
  std::vector<int> idx(b->nb);
  double makespan = 0.0;
  for (int k = 0; k < K; ++k) {
    spk_network* n = b->nets[k];
    const double per_group = b->cost_per_group_us.empty()
        ? b->cost_fixed_us + b->cost_neuron_us * n->nown_max + b->cost_synapse_us * (double)n->S / n->P
        : b->cost_per_group_us[k];
    const double cost = (double)ngroups[k] * per_group;
    for (int i = 0; i < b->nb; ++i) idx[i] = i;
    std::partial_sort(idx.begin(), idx.begin() + n->P, idx.end(), [&](int x, int y) {
      return avail[x] != avail[y] ? avail[x] < avail[y] : x < y;
    });
    std::sort(idx.begin(), idx.begin() + n->P);
    double start = 0.0;
    for (int q = 0; q < n->P; ++q) start = std::max(start, avail[idx[q]]);
    for (int q = 0; q < n->P; ++q) {
      avail[idx[q]] = start + cost;
      per[idx[q]].push_back(make_int2(k, q));
    }
    makespan = std::max(makespan, start + cost);
  }
  b->est_ms = makespan * 1e-3;
  std::vector<int2> tasks;
  std::vector<int> off(b->nb + 1, 0);
  for (int i = 0; i < b->nb; ++i) {
    off[i] = (int)tasks.size();
    tasks.insert(tasks.end(), per[i].begin(), per[i].end());
  }
  off[b->nb] = (int)tasks.size();
  if (tasks.size() > b->tasks_cap) {
    dfree(b->d_tasks);
    if (int rc = dalloc(&b->d_tasks, tasks.size())) return rc;
    b->tasks_cap = tasks.size();
  }
  if (!b->d_off) {
    if (int rc = dalloc(&b->d_off, (size_t)b->nb + 1)) return rc;
  }
  std::vector<spk::FusedArgs> args(K);
  for (int k = 0; k < K; ++k) {
    spk_network* n = b->nets[k];
    if (int rc = fill_fused_args(n, false, args[k])) return rc;
    if (!b->lds_tables) args[k].lseg_U = args[k].lrec = 0;
    args[k].early = 0;
    if (int rc = stage_fused(n, args[k], starts[k], 0, (int)ngroups[k], b->stream)) return rc;
    if (n->profiling) {
      if (!n->d_prof) {
        if (int rc = dalloc(&n->d_prof, spk::kProfPhases)) return rc;
      }
      SPK_HIP(hipMemsetAsync(n->d_prof, 0, sizeof(unsigned long long) * spk::kProfPhases, b->stream));
      args[k].prof = n->d_prof;
    }
  }
  // this run's arguments in pinned staging buffers: the copies read them when they execute
  if (!b->h_args) SPK_HIP(hipHostMalloc((void**)&b->h_args, sizeof(spk::FusedArgs) * (size_t)K));
  if (!b->h_off) SPK_HIP(hipHostMalloc((void**)&b->h_off, sizeof(int) * ((size_t)b->nb + 1)));
  if (tasks.size() > b->h_tasks_cap) {
    if (b->h_tasks) SPK_HIP(hipHostFree(b->h_tasks));
    b->h_tasks = nullptr;
    SPK_HIP(hipHostMalloc((void**)&b->h_tasks, sizeof(int2) * tasks.size()));
    b->h_tasks_cap = tasks.size();
  }
  std::memcpy(b->h_args, args.data(), sizeof(spk::FusedArgs) * (size_t)K);
  std::memcpy(b->h_off, off.data(), sizeof(int) * off.size());
  std::memcpy(b->h_tasks, tasks.data(), sizeof(int2) * tasks.size());
  const void* fn = batch_kernel(b, b->bs);
  const spk::FusedArgs* pa = b->d_args;
  const int2* pt = b->d_tasks;
  const int* po = b->d_off;
  void* kargs[] = {(void*)&pa, (void*)&pt, (void*)&po};
  // the run: three copies and the launch (no memset: a captured one is not allowed here)
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(b->d_tasks, b->h_tasks, tasks.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_off, b->h_off, off.size() * sizeof(int), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->d_args, b->h_args, (size_t)K * sizeof(spk::FusedArgs), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipLaunchCooperativeKernel(fn, dim3(b->nb), dim3(b->bs), kargs, (unsigned int)b->lds, b->stream);
    return e;
  };
  if (b->use_graph) {
    if (!b->gexec || b->g_fn != fn || b->g_tasks != tasks.size() || b->g_dtasks != b->d_tasks) {
      batch_drop_graph(b);
      SPK_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t e = enqueue();
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (e != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(SPK_ERR_HIP, std::string("capturing the batch launch: ") +
                                     hipGetErrorString(e != hipSuccess ? e : ec));
      }
      b->graph = g;
      SPK_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
      b->g_fn = fn;
      b->g_tasks = tasks.size();
      b->g_dtasks = b->d_tasks;
      ++b->captures;
    } else {
      ++b->replays;
    }
    SPK_HIP(hipEventRecord(b->e0, b->stream));
    SPK_HIP(hipGraphLaunch(b->gexec, b->stream));
  } else {
    SPK_HIP(hipEventRecord(b->e0, b->stream));
    const hipError_t e = enqueue();
    if (e != hipSuccess) return fail(SPK_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
  }
  SPK_HIP(hipEventRecord(b->e1, b->stream));
  SPK_HIP(hipEventSynchronize(b->e1));
  float ms = 0.0f;
  SPK_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;
  int rate_khz = 0;
  SPK_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, b->nets[0]->device));
  for (auto* n : b->nets) {
    if (!n->profiling || !n->d_prof) continue;
    unsigned long long ticks[spk::kProfPhases];
    SPK_HIP(hipMemcpy(ticks, n->d_prof, sizeof(ticks), hipMemcpyDeviceToHost));
    for (int i = 0; i < spk::kProfPhases; ++i)
      n->prof_us[i] = rate_khz > 0 ? (double)ticks[i] * 1e3 / (double)rate_khz : 0.0;
  }
  for (auto* n : b->nets) {
    int err = 0;
    SPK_HIP(hipMemcpy(&err, n->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
      n->bar_clear = true;
      if (err == 2) return fail(SPK_ERR_CAPACITY, "batched persistent kernel: spikes of a group exceeded the exchange list");
      return fail(SPK_ERR_HIP, "batched persistent kernel: grid barrier wait exceeded its spin limit");
    }
  }
  return SPK_OK;
}

int spk_batch_set_graph(spk_batch* b, int on) {
  if (!b) return fail(SPK_ERR_ARG, "null batch");
  b->use_graph = on != 0;
  if (!b->use_graph) {
    SPK_HIP(hipStreamSynchronize(b->stream));
    batch_drop_graph(b);
  }
  return SPK_OK;
}

int spk_batch_graph_info(const spk_batch* b, int64_t* captures, int64_t* replays) {
  if (!b) return fail(SPK_ERR_ARG, "null batch");
  if (captures) *captures = b->captures;
  if (replays) *replays = b->replays;
  return SPK_OK;
}

int spk_batch_last_ms(const spk_batch* b, double* kernel_ms, double* estimated_ms) {
  if (!b) return fail(SPK_ERR_ARG, "null batch");
  if (kernel_ms) *kernel_ms = b->last_ms;
  if (estimated_ms) *estimated_ms = b->est_ms;
  return SPK_OK;
}

}  // extern "C"
