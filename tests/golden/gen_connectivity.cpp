// gen_connectivity.cpp -- writes golden connectivity fixtures with Spike's own front-end.
//
// Built against the reference's front-end and Dummy backend (oracle/ref.mk), it constructs
// small networks through the public API exactly as the Examples do (SpikingModel,
// LIF/Poisson neuron groups, Voltage/Conductance synapse groups, connect_with_sparsity of
// Examples/UtilityFunctions.hpp:51-82, RANDOM / ALL_TO_ALL / ONE_TO_ONE connectivity,
// plasticity attachment) and dumps, for each case, the arrays the front-end holds:
//   before finalise_model: pre/post/weight/scale/delay/label in creation order
//     (Synapses.cpp:102-388, SpikingSynapses.cpp:62-124, glibc rand() after srand(seed));
//   after finalise_model: the same arrays in sorted order plus sort indices,
//     per-neuron efferent start/count (Synapses.cpp:33-81, SpikingSynapses.cpp:12-30),
//     the plastic synapse index sets (STDPPlasticity.cpp:13-17), G and delay bounds.
// The Dummy backend computes nothing, so only construction is pinned by these files.
//
//   oracle/_ref/gen_connectivity tests/golden
#include "Spike/Spike.hpp"
#include "UtilityFunctions.hpp"

#include <sys/stat.h>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::string g_root;

template <typename T>
void write_array(const std::string& dir, const std::string& name, const T* p, size_t n) {
  std::ofstream f(dir + "/" + name + ".bin", std::ios::binary);
  if (n) f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

struct Snapshot {
  std::vector<int> pre, post, delay, label;
  std::vector<float> w, scale;
};

Snapshot take(const SpikingSynapses* s) {
  Snapshot o;
  const int S = s->total_number_of_synapses;
  o.pre.assign(s->presynaptic_neuron_indices, s->presynaptic_neuron_indices + S);
  o.post.assign(s->postsynaptic_neuron_indices, s->postsynaptic_neuron_indices + S);
  o.w.assign(s->synaptic_efficacies_or_weights, s->synaptic_efficacies_or_weights + S);
  o.scale.assign(s->weight_scaling_constants, s->weight_scaling_constants + S);
  o.delay.assign(s->delays, s->delays + S);
  o.label.assign(s->syn_labels, s->syn_labels + S);
  return o;
}

void dump_snapshot(const std::string& dir, const std::string& tag, const Snapshot& s) {
  write_array(dir, tag + "_pre", s.pre.data(), s.pre.size());
  write_array(dir, tag + "_post", s.post.data(), s.post.size());
  write_array(dir, tag + "_w", s.w.data(), s.w.size());
  write_array(dir, tag + "_scale", s.scale.data(), s.scale.size());
  write_array(dir, tag + "_delay", s.delay.data(), s.delay.size());
  write_array(dir, tag + "_label", s.label.data(), s.label.size());
}

struct RuleInfo {
  std::string kind;
  STDPPlasticity* rule;
  std::string params;
};

void dump_case(const std::string& name, SpikingModel* m, const Snapshot& before,
               const std::vector<RuleInfo>& rules, int seed) {
  const std::string dir = g_root + "/" + name;
  mkdir(dir.c_str(), 0755);
  SpikingSynapses* s = m->spiking_synapses;
  LIFSpikingNeurons* n = static_cast<LIFSpikingNeurons*>(m->spiking_neurons);
  SpikingNeurons* in = m->input_spiking_neurons;
  dump_snapshot(dir, "orig", before);
  dump_snapshot(dir, "sorted", take(s));
  const int S = s->total_number_of_synapses;
  write_array(dir, "sort_idx", s->synapse_sort_indices, S);
  write_array(dir, "reversesort_idx", s->synapse_reversesort_indices, S);
  write_array(dir, "neuron_eff_start", n->per_neuron_efferent_synapse_start, n->total_number_of_neurons);
  write_array(dir, "neuron_eff_count", n->per_neuron_efferent_synapse_count, n->total_number_of_neurons);
  write_array(dir, "input_eff_start", in->per_neuron_efferent_synapse_start, in->total_number_of_neurons);
  write_array(dir, "input_eff_count", in->per_neuron_efferent_synapse_count, in->total_number_of_neurons);
  write_array(dir, "lif_v0", n->resting_potentials_v0, n->total_number_of_neurons);
  write_array(dir, "lif_vreset", n->after_spike_reset_potentials_vreset, n->total_number_of_neurons);
  write_array(dir, "lif_thr", n->thresholds_for_action_potential_spikes, n->total_number_of_neurons);
  write_array(dir, "lif_tau", n->membrane_time_constants_tau_m, n->total_number_of_neurons);
  write_array(dir, "lif_R", n->membrane_resistances_R, n->total_number_of_neurons);
  PoissonInputSpikingNeurons* pin = dynamic_cast<PoissonInputSpikingNeurons*>(in);
  if (pin && pin->rates) write_array(dir, "poisson_rates", pin->rates, in->total_number_of_neurons);
  for (size_t r = 0; r < rules.size(); ++r)
    write_array(dir, "plastic_" + std::to_string(r), rules[r].rule->plastic_synapses.data(),
                rules[r].rule->plastic_synapses.size());
  std::ofstream j(dir + "/meta.json");
  j.precision(9);
  j << "{\n";
  j << "  \"seed\": " << seed << ",\n";
  j << "  \"n_neurons\": " << n->total_number_of_neurons << ",\n";
  j << "  \"n_inputs\": " << in->total_number_of_neurons << ",\n";
  j << "  \"n_synapses\": " << S << ",\n";
  j << "  \"timestep\": " << m->timestep << ",\n";
  j << "  \"timestep_grouping\": " << m->timestep_grouping << ",\n";
  j << "  \"min_delay\": " << s->minimum_axonal_delay_in_timesteps << ",\n";
  j << "  \"max_delay\": " << s->maximum_axonal_delay_in_timesteps << ",\n";
  j << "  \"num_syn_labels\": " << s->num_syn_labels << ",\n";
  j << "  \"refractory\": " << n->refractory_period_in_seconds << ",\n";
  j << "  \"background_current\": " << n->background_current << ",\n";
  ConductanceSpikingSynapses* cs = dynamic_cast<ConductanceSpikingSynapses*>(s);
  j << "  \"synapse_kind\": \"" << (cs ? "conductance" : "voltage") << "\",\n";
  if (cs) {
    j << "  \"reversal\": [";
    for (size_t l = 0; l < cs->reversal_potentials_Vhat.size(); ++l)
      j << (l ? ", " : "") << cs->reversal_potentials_Vhat[l];
    j << "],\n  \"tau_g\": [";
    for (size_t l = 0; l < cs->decay_terms_tau_g.size(); ++l)
      j << (l ? ", " : "") << cs->decay_terms_tau_g[l];
    j << "],\n";
  }
  j << "  \"rules\": [";
  for (size_t r = 0; r < rules.size(); ++r)
    j << (r ? ", " : "") << "{\"kind\": \"" << rules[r].kind << "\", \"n\": "
      << rules[r].rule->total_number_of_plastic_synapses << ", " << rules[r].params << "}";
  j << "]\n}\n";
}

void lif_params(lif_spiking_neuron_parameters_struct* p, float cm, float g0, float v0, float vr,
                float thr, float tref, float ibg, int n) {
  p->group_shape[0] = 1;
  p->group_shape[1] = n;
  p->somatic_capacitance_Cm = cm;
  p->somatic_leakage_conductance_g0 = g0;
  p->resting_potential_v0 = v0;
  p->after_spike_reset_potential_vreset = vr;
  p->threshold_for_action_potential_spike = thr;
  p->absolute_refractory_period = tref;
  p->background_current = ibg;
}

// Brunel network in the form of Examples/Brunel10K.cpp:69-231, scaled down.
void case_brunel(const std::string& name, int n_in, int n_e, int n_i, float sparseness, bool plastic) {
  SpikingModel* m = new SpikingModel();
  const float dt = 0.0001f;
  m->SetTimestep(dt);
  const float delay = 1.5f * powf(10.0, -3.0);
  LIFSpikingNeurons* lif = new LIFSpikingNeurons();
  PoissonInputSpikingNeurons* inp = new PoissonInputSpikingNeurons();
  VoltageSpikingSynapses* syn = new VoltageSpikingSynapses(42);
  weightdependent_stdp_plasticity_parameters_struct* wd = new weightdependent_stdp_plasticity_parameters_struct;
  wd->a_plus = 1.0; wd->a_minus = 1.0; wd->tau_plus = 0.02; wd->tau_minus = 0.02;
  wd->lambda = 1.0f * powf(10.0, -2); wd->alpha = 2.02; wd->w_max = 0.3 * powf(10.0, -3);
  WeightDependentSTDPPlasticity* rule = new WeightDependentSTDPPlasticity(
      (SpikingSynapses*)syn, (SpikingNeurons*)lif, (SpikingNeurons*)inp, (stdp_plasticity_parameters_struct*)wd);
  m->spiking_neurons = lif;
  m->input_spiking_neurons = inp;
  m->spiking_synapses = syn;
  if (plastic) m->AddPlasticityRule(rule);
  lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
  lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
  lif_params(E, 200.0f * pow(10.0, -12), 10.0f * pow(10.0, -9), 0.0f, 0.0f, 20.0f * pow(10.0, -3),
             2.0f * pow(10, -3), 0.0f, n_e);
  lif_params(I, 200.0f * pow(10.0, -12), 10.0f * pow(10.0, -9), 0.0f, 0.0f, 20.0f * pow(10.0, -3),
             2.0f * pow(10, -3), 0.0f, n_i);
  poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
  ip->group_shape[0] = 1;
  ip->group_shape[1] = n_in;
  ip->rate = 20.0f;
  const int in_id = m->AddInputNeuronGroup(ip);
  const int e_id = m->AddNeuronGroup(E);
  const int i_id = m->AddNeuronGroup(I);
  voltage_spiking_synapse_parameters_struct* EX = new voltage_spiking_synapse_parameters_struct();
  voltage_spiking_synapse_parameters_struct* IN = new voltage_spiking_synapse_parameters_struct();
  voltage_spiking_synapse_parameters_struct* IP = new voltage_spiking_synapse_parameters_struct();
  for (auto* p : {EX, IN, IP}) {
    p->delay_range[0] = delay;
    p->delay_range[1] = delay;
    p->weight_scaling_constant = 1.0;
  }
  const float wv = 0.1f * powf(10.0, -3.0);
  EX->weight_range[0] = EX->weight_range[1] = wv;
  IN->weight_range[0] = IN->weight_range[1] = -5.0f * wv;
  IP->weight_range[0] = IP->weight_range[1] = wv;
  connect_with_sparsity(in_id, e_id, ip, E, IP, sparseness, m);
  connect_with_sparsity(in_id, i_id, ip, I, IP, sparseness, m);
  connect_with_sparsity(e_id, i_id, E, I, EX, sparseness, m);
  if (plastic) EX->plasticity_vec.push_back(rule);
  connect_with_sparsity(e_id, e_id, E, E, EX, sparseness, m);
  connect_with_sparsity(i_id, e_id, I, E, IN, sparseness, m);
  connect_with_sparsity(i_id, i_id, I, I, IN, sparseness, m);
  Snapshot before = take(syn);
  m->finalise_model();
  std::vector<RuleInfo> rules;
  if (plastic) {
    std::ostringstream ps;
    ps.precision(9);
    ps << "\"a_plus\": " << wd->a_plus << ", \"a_minus\": " << wd->a_minus << ", \"tau_plus\": "
       << wd->tau_plus << ", \"tau_minus\": " << wd->tau_minus << ", \"lambda\": " << wd->lambda
       << ", \"alpha\": " << wd->alpha << ", \"w_max\": " << wd->w_max << ", \"nearest_spike_only\": "
       << (wd->nearest_spike_only ? 1 : 0);
    rules.push_back({"wd", rule, ps.str()});
  }
  dump_case(name, m, before, rules, 42);
}

// Poisson inputs -> LIF, RANDOM connectivity with conductance synapses and random delays
// (the plain small network, scaled down).
void case_plain(const std::string& name, int n_in, int n, float p, int seed) {
  SpikingModel* m = new SpikingModel();
  m->SetTimestep(0.0001f);
  LIFSpikingNeurons* lif = new LIFSpikingNeurons();
  PoissonInputSpikingNeurons* inp = new PoissonInputSpikingNeurons();
  ConductanceSpikingSynapses* syn = new ConductanceSpikingSynapses(seed);
  m->spiking_neurons = lif;
  m->input_spiking_neurons = inp;
  m->spiking_synapses = syn;
  lif_spiking_neuron_parameters_struct* P = new lif_spiking_neuron_parameters_struct();
  lif_params(P, 200e-12f, 10e-9f, -0.070f, -0.070f, -0.050f, 0.002f, 0.0f, n);
  poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
  ip->group_shape[0] = 1;
  ip->group_shape[1] = n_in;
  ip->rate = 20.0f;
  const int in_id = m->AddInputNeuronGroup(ip);
  const int n_id = m->AddNeuronGroup(P);
  conductance_spiking_synapse_parameters_struct* sp = new conductance_spiking_synapse_parameters_struct();
  sp->delay_range[0] = 0.0001f;
  sp->delay_range[1] = 0.0010f;
  sp->weight_range[0] = 0.0f;
  sp->weight_range[1] = 1.0f;
  sp->weight_scaling_constant = 1e-9f;
  sp->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
  sp->random_connectivity_probability = p;
  m->AddSynapseGroup(in_id, n_id, sp);
  Snapshot before = take(syn);
  m->finalise_model();
  dump_case(name, m, before, {}, seed);
}

// Excitatory/inhibitory conductance network with two synapse labels, Vogels inhibitory
// STDP on I->E and Evans STDP on input->E (VogelsAbbottNet.cpp-style parameters).
void case_coba(const std::string& name, int n_in, int n_e, int n_i, float p, int seed) {
  SpikingModel* m = new SpikingModel();
  m->SetTimestep(0.0001f);
  LIFSpikingNeurons* lif = new LIFSpikingNeurons();
  PoissonInputSpikingNeurons* inp = new PoissonInputSpikingNeurons();
  ConductanceSpikingSynapses* syn = new ConductanceSpikingSynapses(seed);
  m->spiking_neurons = lif;
  m->input_spiking_neurons = inp;
  m->spiking_synapses = syn;
  inhibitory_stdp_plasticity_parameters_struct* vp = new inhibitory_stdp_plasticity_parameters_struct;
  vp->tau_istdp = 0.02f;
  vp->learningrate = 0.001f;
  vp->targetrate = 5.0f;
  InhibitorySTDPPlasticity* vog = new InhibitorySTDPPlasticity(
      (SpikingSynapses*)syn, (SpikingNeurons*)lif, (SpikingNeurons*)inp, (stdp_plasticity_parameters_struct*)vp);
  evans_stdp_plasticity_parameters_struct* ep = new evans_stdp_plasticity_parameters_struct;
  EvansSTDPPlasticity* ev = new EvansSTDPPlasticity(
      (SpikingSynapses*)syn, (SpikingNeurons*)lif, (SpikingNeurons*)inp, (stdp_plasticity_parameters_struct*)ep);
  m->AddPlasticityRule(vog);
  m->AddPlasticityRule(ev);
  lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
  lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
  lif_params(E, 200e-12f, 10e-9f, -0.060f, -0.060f, -0.050f, 0.005f, 0.0f, n_e);
  lif_params(I, 200e-12f, 10e-9f, -0.060f, -0.060f, -0.050f, 0.005f, 0.0f, n_i);
  poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
  ip->group_shape[0] = 1;
  ip->group_shape[1] = n_in;
  ip->rate = 10.0f;
  const int in_id = m->AddInputNeuronGroup(ip);
  const int e_id = m->AddNeuronGroup(E);
  const int i_id = m->AddNeuronGroup(I);
  conductance_spiking_synapse_parameters_struct* ex = new conductance_spiking_synapse_parameters_struct();
  conductance_spiking_synapse_parameters_struct* ih = new conductance_spiking_synapse_parameters_struct();
  conductance_spiking_synapse_parameters_struct* xi = new conductance_spiking_synapse_parameters_struct();
  for (auto* c : {ex, ih, xi}) {
    c->delay_range[0] = 0.0008f;
    c->delay_range[1] = 0.0008f;
    c->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
    c->random_connectivity_probability = p;
    c->weight_scaling_constant = 1e-9f;
  }
  ex->weight_range[0] = ex->weight_range[1] = 0.6f;
  ex->reversal_potential_Vhat = 0.0f;
  ex->decay_term_tau_g = 0.005f;
  ih->weight_range[0] = ih->weight_range[1] = 6.7f;
  ih->reversal_potential_Vhat = -0.080f;
  ih->decay_term_tau_g = 0.010f;
  xi->weight_range[0] = 0.2f;
  xi->weight_range[1] = 0.8f;
  xi->reversal_potential_Vhat = 0.0f;
  xi->decay_term_tau_g = 0.005f;
  xi->plasticity_vec.push_back(ev);
  m->AddSynapseGroup(in_id, e_id, xi);
  m->AddSynapseGroup(e_id, e_id, ex);
  m->AddSynapseGroup(e_id, i_id, ex);
  ih->plasticity_vec.push_back(vog);
  m->AddSynapseGroup(i_id, e_id, ih);
  ih->plasticity_vec.clear();
  m->AddSynapseGroup(i_id, i_id, ih);
  Snapshot before = take(syn);
  m->finalise_model();
  std::ostringstream vs, es;
  vs.precision(9);
  es.precision(9);
  vs << "\"tau_istdp\": " << vp->tau_istdp << ", \"learningrate\": " << vp->learningrate
     << ", \"targetrate\": " << vp->targetrate << ", \"momentumrate\": " << vp->momentumrate;
  es << "\"decay_term_tau_D\": " << ep->decay_term_tau_D << ", \"model_parameter_alpha_D\": "
     << ep->model_parameter_alpha_D << ", \"synaptic_neurotransmitter_concentration_alpha_C\": "
     << ep->synaptic_neurotransmitter_concentration_alpha_C << ", \"decay_term_tau_C\": "
     << ep->decay_term_tau_C << ", \"learning_rate_rho\": " << ep->learning_rate_rho;
  dump_case(name, m, before, {{"vogels", vog, vs.str()}, {"evans", ev, es.str()}}, seed);
}

// ALL_TO_ALL from the inputs, ONE_TO_ONE between two neuron groups, random weights and
// delays (voltage synapses).
void case_mixed(const std::string& name, int seed) {
  SpikingModel* m = new SpikingModel();
  m->SetTimestep(0.0001f);
  LIFSpikingNeurons* lif = new LIFSpikingNeurons();
  PoissonInputSpikingNeurons* inp = new PoissonInputSpikingNeurons();
  VoltageSpikingSynapses* syn = new VoltageSpikingSynapses(seed);
  m->spiking_neurons = lif;
  m->input_spiking_neurons = inp;
  m->spiking_synapses = syn;
  lif_spiking_neuron_parameters_struct* A = new lif_spiking_neuron_parameters_struct();
  lif_spiking_neuron_parameters_struct* B = new lif_spiking_neuron_parameters_struct();
  lif_params(A, 200e-12f, 10e-9f, 0.0f, 0.0f, 0.020f, 0.002f, 0.0f, 20);
  lif_params(B, 100e-12f, 10e-9f, 0.0f, 0.005f, 0.015f, 0.002f, 0.0f, 20);
  poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
  ip->group_shape[0] = 1;
  ip->group_shape[1] = 10;
  ip->rate = 50.0f;
  poisson_input_spiking_neuron_parameters_struct* ip2 = new poisson_input_spiking_neuron_parameters_struct();
  ip2->group_shape[0] = 1;
  ip2->group_shape[1] = 6;
  ip2->rate = 30.0f;
  const int in0 = m->AddInputNeuronGroup(ip);
  const int in1 = m->AddInputNeuronGroup(ip2);
  const int a = m->AddNeuronGroup(A);
  const int b = m->AddNeuronGroup(B);
  voltage_spiking_synapse_parameters_struct* all = new voltage_spiking_synapse_parameters_struct();
  all->connectivity_type = CONNECTIVITY_TYPE_ALL_TO_ALL;
  all->weight_range[0] = 0.5e-3f;
  all->weight_range[1] = 2.0e-3f;
  all->delay_range[0] = 0.0002f;
  all->delay_range[1] = 0.0020f;
  voltage_spiking_synapse_parameters_struct* one = new voltage_spiking_synapse_parameters_struct();
  one->connectivity_type = CONNECTIVITY_TYPE_ONE_TO_ONE;
  one->weight_range[0] = 1.0e-3f;
  one->weight_range[1] = 1.0e-3f;
  one->weight_scaling_constant = 2.0f;
  one->delay_range[0] = 0.0003f;
  one->delay_range[1] = 0.0003f;
  m->AddSynapseGroup(in0, a, all);
  m->AddSynapseGroup(in1, b, all);
  m->AddSynapseGroup(a, b, one);
  m->AddSynapseGroup(b, a, all);
  Snapshot before = take(syn);
  m->finalise_model();
  dump_case(name, m, before, {}, seed);
}

}  // namespace

int main(int argc, char** argv) {
  g_root = argc > 1 ? argv[1] : "tests/golden";
  mkdir(g_root.c_str(), 0755);
  case_brunel("brunel_ref_small", 100, 80, 20, 0.1f, true);
  case_plain("plain_small", 100, 100, 0.1f, 42);
  case_coba("coba_small", 40, 80, 20, 0.1f, 7);
  case_mixed("mixed_small", 3);
  return 0;
}
