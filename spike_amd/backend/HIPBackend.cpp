// HIPBackend.cpp -- implementation and registration of the "HIP" Spike backend.
#include "HIPBackend.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace Backend {
namespace HIP {

namespace {

// Errors end the program the way the front-end's print_message_and_exit does
// (Spike/Helpers/TerminalHelpers.hpp:15-20).
[[noreturn]] void die(const std::string& what) {
  printf("\n%s\nExiting...\n", what.c_str());
  fflush(stdout);
  exit(-1);
}

void check(int rc, const char* call) {
  if (rc != SPK_OK) die(std::string("HIP backend: ") + call + " failed: " + spk_last_error());
}

}  // namespace

// One device network per SpikingModel, built from the front-end's host arrays once every
// component has been prepared (SpikingModel::init_backend, SpikingModel.cpp:128-158, runs
// before the first reset_state).
class ModelNetwork {
 public:
  explicit ModelNetwork(::SpikingModel* m) : model(m) {}
  ~ModelNetwork() { release(); }
  // frees the device network (at program exit, before the HIP runtime shuts down)
  void release() {
    if (net) spk_destroy(net);
    net = nullptr;
  }
  spk_network* handle() {
    if (!net) build();
    return net;
  }
  bool built() const { return net != nullptr; }
  void step(float t) {
    spk_network* n = handle();
    if (model->timestep_grouping != grouping) {
      check(spk_set_timestep_grouping(n, model->timestep_grouping), "spk_set_timestep_grouping");
      grouping = model->timestep_grouping;
    }
    check(spk_step_group(n, t, plastic_pending ? SPK_RUN_PLASTICITY : 0), "spk_step_group");
    plastic_pending = false;
  }
  int monitor(const void* front) const {
    auto it = monitor_id.find(front);
    if (it == monitor_id.end()) die("HIP backend: spike monitor is not part of the model");
    return it->second;
  }

  ::SpikingModel* model;
  spk_network* net = nullptr;
  int grouping = 0;
  bool plastic_pending = false;
  std::map<const void*, int> monitor_id;
  std::map<const void*, int> rule_id;

 private:
  void build();
};

namespace {

std::map<::SpikingModel*, std::unique_ptr<ModelNetwork>>& registry() {
  static std::map<::SpikingModel*, std::unique_ptr<ModelNetwork>> r;
  return r;
}

// At program exit: work still queued is waited for, so no kernel outlives the program, and
// the device networks are freed here.  The registry itself is a static constructed before
// the HIP runtime's first call, so its destructor would run after the runtime's own
// teardown; this handler, registered after that first call, runs before it.
void synchronize_all() {
  for (auto& kv : registry())
    if (kv.second && kv.second->built()) spk_synchronize(kv.second->net);
  for (auto& kv : registry())
    if (kv.second) kv.second->release();
}

ModelNetwork* model_network(::SpikingModel* m) {
  if (!m) die("HIP backend: component used before SpikingModel::finalise_model");
  auto& r = registry();
  auto it = r.find(m);
  if (it != r.end()) return it->second.get();
  auto* mn = new ModelNetwork(m);
  r[m].reset(mn);
  return mn;
}

::SpikingModel* model_of(::SpikingNeurons* n) { return n ? n->model : nullptr; }

}  // namespace

ModelNetwork* network_of(::SpikingModel* model) { return model_network(model); }

void ModelNetwork::build() {
  ::SpikingModel* m = model;
  auto* lif = dynamic_cast<::LIFSpikingNeurons*>(m->spiking_neurons);
  if (!lif) die("HIP backend: the neuron population must be LIFSpikingNeurons");
  ::SpikingNeurons* in = m->input_spiking_neurons;
  auto* poisson = dynamic_cast<::PoissonInputSpikingNeurons*>(in);
  auto* generator = dynamic_cast<::GeneratorInputSpikingNeurons*>(in);
  if (in && in->total_number_of_neurons > 0 && !poisson && !generator)
    die("HIP backend: input neurons must be PoissonInputSpikingNeurons or GeneratorInputSpikingNeurons");
  ::SpikingSynapses* syn = m->spiking_synapses;
  auto* cond = dynamic_cast<::ConductanceSpikingSynapses*>(syn);
  auto* volt = dynamic_cast<::VoltageSpikingSynapses*>(syn);
  auto* curr = dynamic_cast<::CurrentSpikingSynapses*>(syn);
  if (syn->total_number_of_synapses > 0 && !cond && !volt && !curr)
    die("HIP backend: synapses must be VoltageSpikingSynapses, ConductanceSpikingSynapses or CurrentSpikingSynapses");

  const int N = lif->total_number_of_neurons;
  const int Nin = in ? in->total_number_of_neurons : 0;
  const int kind = cond ? SPK_SYN_CONDUCTANCE : (curr ? SPK_SYN_CURRENT : SPK_SYN_VOLTAGE);
  const int L = cond ? std::max(1, cond->num_syn_labels) : (curr ? std::max(1, curr->num_syn_labels) : 1);
  const float dt = m->timestep;
  check(spk_create(N, Nin, kind, L, dt, &net), "spk_create");
  // registered once the HIP runtime is up, so that it runs before the runtime's teardown
  static bool exit_hook = false;
  if (!exit_hook) {
    atexit(synchronize_all);
    exit_hook = true;
  }

  // LIF constants as the reference backend derives them on the host
  // (Spike/Backend/CUDA/Neurons/LIFSpikingNeurons.cu:26-46): c = dt/tau, R' = c*R.
  std::vector<float> c(N), rp(N);
  for (int i = 0; i < N; ++i) {
    c[i] = dt / lif->membrane_time_constants_tau_m[i];
    rp[i] = c[i] * lif->membrane_resistances_R[i];
  }
  check(spk_set_lif(net, lif->resting_potentials_v0, lif->after_spike_reset_potentials_vreset,
                    lif->thresholds_for_action_potential_spikes, c.data(), rp.data(),
                    lif->refractory_period_in_seconds, lif->background_current),
        "spk_set_lif");
  if (Nin > 0 && generator) {
    // every stimulus of the front-end (GeneratorInputSpikingNeurons.cpp:28-77), concatenated
    const int ns = generator->total_number_of_input_stimuli;
    std::vector<int32_t> counts, ids;
    std::vector<float> times;
    for (int s = 0; s < ns; ++s) {
      const int k = generator->number_of_spikes_in_stimuli[s];
      counts.push_back(k);
      ids.insert(ids.end(), generator->neuron_id_matrix_for_stimuli[s], generator->neuron_id_matrix_for_stimuli[s] + k);
      times.insert(times.end(), generator->spike_times_matrix_for_stimuli[s],
                   generator->spike_times_matrix_for_stimuli[s] + k);
    }
    if (ns < 1) {   // no stimulus added: inputs that never spike
      counts.push_back(0);
    }
    check(spk_set_generator_stimuli(net, counts.data(), (int)counts.size(), ids.data(), times.data()),
          "spk_set_generator_stimuli");
    check(spk_set_generator_onset(net, generator->stimulus_onset_adjustment), "spk_set_generator_onset");
    if (generator->current_stimulus_index >= 0 && generator->current_stimulus_index < (int)counts.size())
      check(spk_select_stimulus(net, generator->current_stimulus_index), "spk_select_stimulus");
  } else if (Nin > 0) {
    check(spk_set_poisson_rates(net, poisson->rates, 1), "spk_set_poisson_rates");
  }
  // Philox key: the reference seeds its generator with 1
  // (Spike/Backend/CUDA/Helpers/RandomStateManager.hpp:24); SPIKE_HIP_SEED overrides it.
  uint64_t seed = 1;
  if (const char* s = getenv("SPIKE_HIP_SEED")) seed = strtoull(s, nullptr, 10);
  check(spk_set_seed(net, seed), "spk_set_seed");
  // weight and trace storage: FP32 as in the reference; SPIKE_HIP_WEIGHTS=bf16 or fp16
  // selects 16-bit storage with Kahan-compensated plastic updates
  if (const char* f = getenv("SPIKE_HIP_WEIGHTS")) {
    const std::string fs(f);
    const int fmt = fs == "bf16" ? SPK_WEIGHTS_BF16 : (fs == "fp16" ? SPK_WEIGHTS_FP16 : SPK_WEIGHTS_FP32);
    if (fmt == SPK_WEIGHTS_FP32 && fs != "fp32") die("SPIKE_HIP_WEIGHTS must be fp32, bf16 or fp16");
    check(spk_set_weight_format(net, fmt), "spk_set_weight_format");
  }

  if (cond) {
    // decay factors as in Spike/Backend/CUDA/Synapses/ConductanceSpikingSynapses.cu:74-79
    std::vector<float> dg(L, 0.0f), ev(L, 0.0f);
    for (int l = 0; l < L && l < (int)cond->decay_terms_tau_g.size(); ++l) {
      dg[l] = expf(-dt / cond->decay_terms_tau_g[l]);
      ev[l] = cond->reversal_potentials_Vhat[l];
    }
    check(spk_set_labels(net, dg.data(), ev.data()), "spk_set_labels");
  }
  if (curr) {
    // the current traces' decay per label, expf(-dt/tau) as the reference kernel computes
    // it (InlineDeviceFunctions.hpp:66-68), evaluated here on the host once
    std::vector<float> dg(L, 1.0f), ev(L, 0.0f);
    for (int l = 0; l < L && l < (int)curr->decay_terms_tau.size(); ++l) dg[l] = expf(-dt / curr->decay_terms_tau[l]);
    check(spk_set_labels(net, dg.data(), ev.data()), "spk_set_labels");
  }
  const int S = syn->total_number_of_synapses;
  check(spk_set_synapses(net, S, syn->presynaptic_neuron_indices, syn->postsynaptic_neuron_indices,
                         syn->synaptic_efficacies_or_weights, syn->weight_scaling_constants,
                         syn->delays, syn->syn_labels),
        "spk_set_synapses");

  // plasticity rules in the order SpikingModel applies them (SpikingModel.cpp:203-206); the
  // index sets are the front-end's (sorted positions, STDPPlasticity.cpp:13-17)
  for (::STDPPlasticity* rule : m->plasticity_rule_vec) {
    std::vector<int64_t> idx(rule->plastic_synapses.begin(), rule->plastic_synapses.end());
    int rid = -1;
    if (auto* wd = dynamic_cast<::WeightDependentSTDPPlasticity*>(rule)) {
      spk_wd_stdp_params p{};
      p.a_minus = wd->stdp_params->a_minus;
      p.a_plus = wd->stdp_params->a_plus;
      p.tau_minus = wd->stdp_params->tau_minus;
      p.tau_plus = wd->stdp_params->tau_plus;
      p.lambda = wd->stdp_params->lambda;
      p.alpha = wd->stdp_params->alpha;
      p.w_max = wd->stdp_params->w_max;
      p.nearest_spike_only = wd->stdp_params->nearest_spike_only ? 1 : 0;
      check(spk_add_wd_stdp(net, idx.data(), (int64_t)idx.size(), &p, &rid), "spk_add_wd_stdp");
    } else if (auto* vg = dynamic_cast<::InhibitorySTDPPlasticity*>(rule)) {
      spk_vogels_stdp_params p{};
      p.tau_istdp = vg->stdp_params->tau_istdp;
      p.learningrate = vg->stdp_params->learningrate;
      p.targetrate = vg->stdp_params->targetrate;
      p.momentumrate = vg->stdp_params->momentumrate;
      check(spk_add_vogels_stdp(net, idx.data(), (int64_t)idx.size(), &p, &rid), "spk_add_vogels_stdp");
    } else if (auto* ev = dynamic_cast<::EvansSTDPPlasticity*>(rule)) {
      spk_evans_stdp_params p{};
      p.decay_term_tau_D = ev->stdp_params->decay_term_tau_D;
      p.model_parameter_alpha_D = ev->stdp_params->model_parameter_alpha_D;
      p.synaptic_neurotransmitter_concentration_alpha_C =
          ev->stdp_params->synaptic_neurotransmitter_concentration_alpha_C;
      p.decay_term_tau_C = ev->stdp_params->decay_term_tau_C;
      p.learning_rate_rho = ev->stdp_params->learning_rate_rho;
      check(spk_add_evans_stdp(net, idx.data(), (int64_t)idx.size(), &p, &rid), "spk_add_evans_stdp");
    } else if (auto* wn = dynamic_cast<::WeightNormSTDPPlasticity*>(rule)) {
      // the front-end computed the squared afferent norms (and rescaled the weights for a
      // target) in its prepare_backend_early (WeightNormSTDPPlasticity.cpp:32-72)
      const int N = lif->total_number_of_neurons;
      std::vector<float> sq(N, 0.0f);
      std::vector<uint8_t> in(N, 0);
      if (wn->sum_squared_afferent_values && wn->neuron_in_plasticity_set)
        for (int v = 0; v < N; ++v) {
          sq[v] = wn->sum_squared_afferent_values[v];
          in[v] = wn->neuron_in_plasticity_set[v] ? 1 : 0;
        }
      check(spk_add_weightnorm(net, idx.data(), (int64_t)idx.size(), sq.data(), in.data(), &rid),
            "spk_add_weightnorm");
    } else if (auto* cs = dynamic_cast<::CustomSTDPPlasticity*>(rule)) {
      spk_custom_stdp_params p{};
      p.a_minus = cs->stdp_params->a_minus;
      p.a_plus = cs->stdp_params->a_plus;
      p.a_star = cs->stdp_params->a_star;
      p.tau_minus = cs->stdp_params->tau_minus;
      p.tau_plus = cs->stdp_params->tau_plus;
      p.w_max = cs->stdp_params->w_max;
      p.learning_rate = cs->stdp_params->learning_rate;
      p.nearest_spike_only = cs->stdp_params->nearest_spike_only ? 1 : 0;
      check(spk_add_custom_stdp(net, idx.data(), (int64_t)idx.size(), &p, &rid), "spk_add_custom_stdp");
    } else {
      die("HIP backend: unsupported plasticity rule");
    }
    rule_id[rule] = rid;
  }
  for (::ActivityMonitor* am : m->monitors_vec) {
    // SpikingActivityMonitor declares its own `neurons` (SpikingActivityMonitor.hpp:45),
    // hiding ActivityMonitor's, which RateActivityMonitor uses
    auto* sm_ = dynamic_cast<::SpikingActivityMonitor*>(am);
    ::SpikingNeurons* watched = sm_ ? sm_->neurons : am->neurons;
    int pop;
    if (watched == m->spiking_neurons) pop = SPK_POP_NEURONS;
    else if (watched == m->input_spiking_neurons) pop = SPK_POP_INPUTS;
    else die("HIP backend: a monitor must watch the model's neurons or inputs");
    int id = -1;
    if (auto* sm = sm_) {
      check(spk_add_monitor(net, pop, std::max(1, sm->size_of_device_spike_store), &id), "spk_add_monitor");
      monitor_id[sm] = id;
    } else if (auto* rm = dynamic_cast<::RateActivityMonitor*>(am)) {
      check(spk_add_rate_monitor(net, pop, &id), "spk_add_rate_monitor");
      monitor_id[rm] = id;
    } else {
      die("HIP backend: unsupported activity monitor");
    }
  }
  grouping = m->timestep_grouping;
  check(spk_prepare(net, grouping), "spk_prepare");
}

// ---------------------------------------------------------------------------------------
void SpikingNeurons::prepare() {}

void SpikingNeurons::reset_state() {
  // the whole network state is reset by the synapses' reset_state, which
  // SpikingModel::reset_state calls first (SpikingModel.cpp:181-190)
}

void SpikingNeurons::state_update(float, float) {
  // the group runs at the synapses' state_update, after the inputs and plasticity calls
}

void PoissonInputSpikingNeurons::reset_state() {
  // PoissonInputSpikingNeurons::select_stimulus resets the inputs through this call
  ModelNetwork* mn = model_network(model_of(front()));
  if (mn->built()) check(spk_reset_inputs(mn->net), "spk_reset_inputs");
}

void GeneratorInputSpikingNeurons::reset_state() {
  // GeneratorInputSpikingNeurons::select_stimulus resets the inputs through this call
  // (GeneratorInputSpikingNeurons.cu:24-40 copies the selected stimulus to the device)
  auto* g = dynamic_cast<::GeneratorInputSpikingNeurons*>(front());
  ModelNetwork* mn = model_network(model_of(front()));
  if (mn->built() && g) check(spk_select_stimulus(mn->net, g->current_stimulus_index), "spk_select_stimulus");
}

void Synapses::prepare() {}

void Synapses::reset_state() {
  auto* s = dynamic_cast<::SpikingSynapses*>(front());
  if (!s) die("HIP backend: synapses must be SpikingSynapses");
  ModelNetwork* mn = model_network(s->model);
  check(spk_reset_state(mn->handle()), "spk_reset_state");
}

void Synapses::set_neuron_indices_by_sampling_from_normal_distribution(int, int, int, int, int, int*, int*,
                                                                       int, int, int, float, bool) {
  die("HIP backend: Gaussian-sampled connectivity is built by the front-end, not the backend");
}

void Synapses::copy_to_frontend() {
  auto* s = dynamic_cast<::SpikingSynapses*>(front());
  ModelNetwork* mn = s ? model_network(s->model) : nullptr;
  if (mn && mn->built()) check(spk_get_weights(mn->net, front()->synaptic_efficacies_or_weights), "spk_get_weights");
}

void Synapses::copy_to_backend() {
  auto* s = dynamic_cast<::SpikingSynapses*>(front());
  ModelNetwork* mn = s ? model_network(s->model) : nullptr;
  if (mn && mn->built()) check(spk_set_weights(mn->net, front()->synaptic_efficacies_or_weights), "spk_set_weights");
}

void SpikingSynapses::copy_weights_to_host() { copy_to_frontend(); }

void SpikingSynapses::state_update(::SpikingNeurons*, ::SpikingNeurons*, float current_time_in_seconds, float) {
  auto* s = static_cast<::SpikingSynapses*>(front());
  model_network(s->model)->step(current_time_in_seconds);
}

void STDPPlasticity::mark_applied(float) { model_network(front()->model)->plastic_pending = true; }

void WeightDependentSTDPPlasticity::apply_stdp_to_synapse_weights(float t, float) { mark_applied(t); }
void InhibitorySTDPPlasticity::apply_stdp_to_synapse_weights(float t, float) { mark_applied(t); }
void EvansSTDPPlasticity::update_synaptic_efficacies_or_weights(float t, float) { mark_applied(t); }

void CustomSTDPPlasticity::apply_stdp_to_synapse_weights(float t, float) { mark_applied(t); }

void WeightNormSTDPPlasticity::weight_normalization() { mark_applied(0.0f); }

void WeightNormSTDPPlasticity::reset_state() {
  // WeightNormSTDPPlasticity.cu:16-26: the front-end's initial weights become the rule's
  auto* f = dynamic_cast<::WeightNormSTDPPlasticity*>(front());
  ModelNetwork* mn = model_network(front()->model);
  if (!f || !mn->built() || !f->initial_weights) return;
  auto it = mn->rule_id.find(front());
  if (it == mn->rule_id.end()) return;
  check(spk_weightnorm_set_initial(mn->net, it->second, f->initial_weights), "spk_weightnorm_set_initial");
}

void SpikingActivityMonitor::reset_state() {
  ModelNetwork* mn = model_network(front()->model);
  if (mn->built()) check(spk_monitor_reset(mn->net, mn->monitor(front())), "spk_monitor_reset");
}

void RateActivityMonitor::reset_state() {
  ModelNetwork* mn = model_network(front()->model);
  if (mn->built()) check(spk_rate_monitor_reset(mn->net, mn->monitor(front())), "spk_rate_monitor_reset");
}

void RateActivityMonitor::copy_spike_count_to_host() {
  ModelNetwork* mn = model_network(front()->model);
  if (mn->built())
    check(spk_rate_monitor_counts(mn->net, mn->monitor(front()), front()->per_neuron_spike_counts),
          "spk_rate_monitor_counts");
}

void RateActivityMonitor::add_spikes_to_per_neuron_spike_count(float, float) {
  // the counts are kept by the group step where the spikes are made
}

void SpikingActivityMonitor::copy_spikecount_to_front() {
  ModelNetwork* mn = model_network(front()->model);
  int64_t c = 0;
  if (mn->built()) check(spk_monitor_device_count(mn->net, mn->monitor(front()), &c), "spk_monitor_device_count");
  front()->total_number_of_spikes_stored_on_device[0] = (int)c;
}

void SpikingActivityMonitor::copy_spikes_to_front() {
  // records land after those already on the host (SpikingActivityMonitor.cu:41-50), in
  // (time, id) order
  ModelNetwork* mn = model_network(front()->model);
  ::SpikingActivityMonitor* f = front();
  const int64_t want = f->total_number_of_spikes_stored_on_device[0];
  if (!mn->built() || want <= 0) return;
  int64_t got = 0;
  const int off = f->total_number_of_spikes_stored_on_host;
  check(spk_monitor_drain(mn->net, mn->monitor(f), f->neuron_ids_of_stored_spikes_on_host + off,
                          f->spike_times_of_stored_spikes_on_host + off, want, &got),
        "spk_monitor_drain");
  if (got != want) die("HIP backend: spike monitor count changed between count and copy");
}

void SpikingActivityMonitor::collect_spikes_for_timestep(float, float) {
  // the group step records the monitored populations' spikes into the device store
}

// ---------------------------------------------------------------------------------------
// Types this backend does not provide: registered so that selecting "HIP" fails with a
// message instead of silently falling back to the Dummy backend.
#define SPK_HIP_UNSUPPORTED_PREPARE(NAME) \
  void prepare() override { die("HIP backend: " NAME " is not provided by this backend"); }

class ImagePoissonInputSpikingNeurons : public virtual ::Backend::HIP::PoissonInputSpikingNeurons,
                                        public virtual ::Backend::ImagePoissonInputSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(ImagePoissonInputSpikingNeurons);
  SPK_HIP_UNSUPPORTED_PREPARE("ImagePoissonInputSpikingNeurons")
  void copy_rates_to_device() override {}
};

class PatternedPoissonInputSpikingNeurons : public virtual ::Backend::HIP::PoissonInputSpikingNeurons,
                                            public virtual ::Backend::PatternedPoissonInputSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(PatternedPoissonInputSpikingNeurons);
  SPK_HIP_UNSUPPORTED_PREPARE("PatternedPoissonInputSpikingNeurons")
  void copy_rates_to_device() override {}
};

#undef SPK_HIP_UNSUPPORTED_PREPARE

}  // namespace HIP
}  // namespace Backend

SPIKE_EXPORT_BACKEND_TYPE(HIP, SpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, LIFSpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, PoissonInputSpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, SpikingSynapses);
SPIKE_EXPORT_BACKEND_TYPE(HIP, VoltageSpikingSynapses);
SPIKE_EXPORT_BACKEND_TYPE(HIP, ConductanceSpikingSynapses);
SPIKE_EXPORT_BACKEND_TYPE(HIP, WeightDependentSTDPPlasticity);
SPIKE_EXPORT_BACKEND_TYPE(HIP, InhibitorySTDPPlasticity);
SPIKE_EXPORT_BACKEND_TYPE(HIP, EvansSTDPPlasticity);
SPIKE_EXPORT_BACKEND_TYPE(HIP, SpikingActivityMonitor);
SPIKE_EXPORT_BACKEND_TYPE(HIP, RandomStateManager);
SPIKE_EXPORT_BACKEND_TYPE(HIP, GeneratorInputSpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, ImagePoissonInputSpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, PatternedPoissonInputSpikingNeurons);
SPIKE_EXPORT_BACKEND_TYPE(HIP, CurrentSpikingSynapses);
SPIKE_EXPORT_BACKEND_TYPE(HIP, CustomSTDPPlasticity);
SPIKE_EXPORT_BACKEND_TYPE(HIP, WeightNormSTDPPlasticity);
SPIKE_EXPORT_BACKEND_TYPE(HIP, RateActivityMonitor);
