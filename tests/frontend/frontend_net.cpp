// frontend_net.cpp -- a network built and run through Spike's unchanged C++ front-end.
//
// Linked with the front-end library and the HIP backend (oracle/ref.mk, default backend
// "HIP"), it builds one of a few small networks with the public API the Examples use,
// runs SpikingModel::run, and writes what the parity test needs to replay the same
// network on the CPU oracle:
//   syn_{pre,post,w,scale,delay,label}.bin  the synapse arrays after finalise (sorted order,
//                                           the order the backend receives them)
//   lif_{v0,vreset,thr,tau,R}.bin, rates.bin, decay_g.bin (expf(-dt/tau_g) per label),
//   plastic_<r>.bin, weights_after.bin (weights copied back after the run),
//   conn_*.bin (Synapses::save_connectivity_as_binary after the run), weights_reloaded.bin
//   (those weights loaded back with load_weights_from_binary and copied from the device),
//   <pop>_SpikeIDs.bin / <pop>_SpikeTimes.bin (SpikingActivityMonitor::save_spikes_as_binary),
//   rate_counts.bin (a RateActivityMonitor's per_neuron_spike_counts of the neurons),
//   meta.json.
//
//   frontend_net_hip <case> <out_dir> <simtime_s> [plastic]
//   cases: brunel (voltage synapses, Poisson inputs, optional WD STDP + weight normalization),
//          coba (conductance synapses, two labels, Vogels + Evans + custom STDP),
//          plain (Poisson -> LIF, RANDOM conductance synapses, delays 1-10 steps),
//          current (Poisson -> LIF E/I, CurrentSpikingSynapses with two decay constants),
//          generator (Examples/SimpleExample.cpp's layout: generator inputs -> conductance
//          LIF E/I; stimulus 0 for simtime, then stimulus 1 selected for another simtime;
//          gen_{counts,ids,times}.bin hold the stimuli)
#include "Spike/Spike.hpp"
#include "UtilityFunctions.hpp"

#include <sys/stat.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

std::string g_out;

template <typename T>
void put(const std::string& name, const T* p, size_t n) {
  std::ofstream f(g_out + "/" + name + ".bin", std::ios::binary);
  if (n) f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(T)));
}

void lif(lif_spiking_neuron_parameters_struct* p, float cm, float g0, float v0, float vr, float thr,
         float tref, int n) {
  p->group_shape[0] = 1;
  p->group_shape[1] = n;
  p->somatic_capacitance_Cm = cm;
  p->somatic_leakage_conductance_g0 = g0;
  p->resting_potential_v0 = v0;
  p->after_spike_reset_potential_vreset = vr;
  p->threshold_for_action_potential_spike = thr;
  p->absolute_refractory_period = tref;
  p->background_current = 0.0f;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <brunel|coba|plain> <out_dir> <simtime_s> [plastic]\n", argv[0]);
    return 2;
  }
  const std::string which = argv[1];
  g_out = argv[2];
  const float simtime = (float)atof(argv[3]);
  const bool plastic = argc > 4 && std::string(argv[4]) == "plastic";
  mkdir(g_out.c_str(), 0755);

  SpikingModel* model = new SpikingModel();
  const float dt = 0.0001f;
  model->SetTimestep(dt);
  LIFSpikingNeurons* neurons = new LIFSpikingNeurons();
  PoissonInputSpikingNeurons* poisson = nullptr;
  GeneratorInputSpikingNeurons* generator = nullptr;
  InputSpikingNeurons* inputs;
  if (which == "generator") inputs = generator = new GeneratorInputSpikingNeurons();
  else inputs = poisson = new PoissonInputSpikingNeurons();
  model->spiking_neurons = neurons;
  model->input_spiking_neurons = inputs;
  std::vector<STDPPlasticity*> rules;
  std::vector<std::string> rule_kinds;

  if (which == "brunel") {
    // Examples/Brunel10K.cpp:69-231 at 1/50 of its size, with a stronger external drive
    VoltageSpikingSynapses* syn = new VoltageSpikingSynapses(42);
    model->spiking_synapses = syn;
    weightdependent_stdp_plasticity_parameters_struct* wd = new weightdependent_stdp_plasticity_parameters_struct;
    wd->a_plus = 1.0; wd->a_minus = 1.0; wd->tau_plus = 0.02; wd->tau_minus = 0.02;
    wd->lambda = 1.0e-2f; wd->alpha = 2.02; wd->w_max = 0.3e-3f;
    WeightDependentSTDPPlasticity* rule = new WeightDependentSTDPPlasticity(
        (SpikingSynapses*)syn, (SpikingNeurons*)neurons, (SpikingNeurons*)inputs, (stdp_plasticity_parameters_struct*)wd);
    weightnorm_stdp_plasticity_parameters_struct* np = new weightnorm_stdp_plasticity_parameters_struct;
    WeightNormSTDPPlasticity* norm = new WeightNormSTDPPlasticity(
        (SpikingSynapses*)syn, (SpikingNeurons*)neurons, (SpikingNeurons*)inputs, (plasticity_parameters_struct*)np);
    if (plastic) {
      model->AddPlasticityRule(rule);
      model->AddPlasticityRule(norm);
      rules.push_back(rule);
      rules.push_back(norm);
      rule_kinds.push_back("wd");
      rule_kinds.push_back("weightnorm");
    }
    lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
    lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
    lif(E, 200e-12f, 10e-9f, 0.0f, 0.0f, 20e-3f, 2e-3f, 160);
    lif(I, 200e-12f, 10e-9f, 0.0f, 0.0f, 20e-3f, 2e-3f, 40);
    poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
    ip->group_shape[0] = 1;
    ip->group_shape[1] = 200;
    ip->rate = 400.0f;
    const int in_id = model->AddInputNeuronGroup(ip);
    const int e_id = model->AddNeuronGroup(E);
    const int i_id = model->AddNeuronGroup(I);
    voltage_spiking_synapse_parameters_struct* EX = new voltage_spiking_synapse_parameters_struct();
    voltage_spiking_synapse_parameters_struct* IN = new voltage_spiking_synapse_parameters_struct();
    voltage_spiking_synapse_parameters_struct* IP = new voltage_spiking_synapse_parameters_struct();
    for (auto* p : {EX, IN, IP}) {
      p->delay_range[0] = 1.5e-3f;
      p->delay_range[1] = 1.5e-3f;
      p->weight_scaling_constant = 1.0;
    }
    EX->weight_range[0] = EX->weight_range[1] = 0.2e-3f;
    IN->weight_range[0] = IN->weight_range[1] = -1.0e-3f;
    IP->weight_range[0] = IP->weight_range[1] = 1.0e-3f;
    connect_with_sparsity(in_id, e_id, ip, E, IP, 0.1f, model);
    connect_with_sparsity(in_id, i_id, ip, I, IP, 0.1f, model);
    connect_with_sparsity(e_id, i_id, E, I, EX, 0.1f, model);
    if (plastic) {
      EX->plasticity_vec.push_back(rule);
      EX->plasticity_vec.push_back(norm);
    }
    connect_with_sparsity(e_id, e_id, E, E, EX, 0.1f, model);
    connect_with_sparsity(i_id, e_id, I, E, IN, 0.1f, model);
    connect_with_sparsity(i_id, i_id, I, I, IN, 0.1f, model);
  } else if (which == "coba" || which == "plain") {
    const bool coba = which == "coba";
    ConductanceSpikingSynapses* syn = new ConductanceSpikingSynapses(coba ? 7 : 42);
    model->spiking_synapses = syn;
    lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
    lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
    lif(E, 200e-12f, 10e-9f, -0.060f, -0.060f, -0.050f, coba ? 5e-3f : 2e-3f, coba ? 160 : 200);
    lif(I, 200e-12f, 10e-9f, -0.060f, -0.060f, -0.050f, coba ? 5e-3f : 2e-3f, 40);
    poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
    ip->group_shape[0] = 1;
    ip->group_shape[1] = coba ? 80 : 200;
    ip->rate = coba ? 60.0f : 40.0f;
    const int in_id = model->AddInputNeuronGroup(ip);
    const int e_id = model->AddNeuronGroup(E);
    const int i_id = coba ? model->AddNeuronGroup(I) : -1;
    conductance_spiking_synapse_parameters_struct* xi = new conductance_spiking_synapse_parameters_struct();
    xi->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
    xi->random_connectivity_probability = 0.1f;
    xi->weight_scaling_constant = 1e-9f;
    xi->reversal_potential_Vhat = 0.0f;
    if (coba) {
      inhibitory_stdp_plasticity_parameters_struct* vp = new inhibitory_stdp_plasticity_parameters_struct;
      vp->tau_istdp = 0.02f;
      vp->learningrate = 0.001f;
      vp->targetrate = 5.0f;
      InhibitorySTDPPlasticity* vog = new InhibitorySTDPPlasticity(
          (SpikingSynapses*)syn, (SpikingNeurons*)neurons, (SpikingNeurons*)inputs, (stdp_plasticity_parameters_struct*)vp);
      evans_stdp_plasticity_parameters_struct* ep = new evans_stdp_plasticity_parameters_struct;
      EvansSTDPPlasticity* ev = new EvansSTDPPlasticity(
          (SpikingSynapses*)syn, (SpikingNeurons*)neurons, (SpikingNeurons*)inputs, (stdp_plasticity_parameters_struct*)ep);
      custom_stdp_plasticity_parameters_struct* cp = new custom_stdp_plasticity_parameters_struct;
      cp->a_minus = 0.8f; cp->a_plus = 1.0f; cp->a_star = 0.05f; cp->tau_minus = 0.0337f;
      cp->tau_plus = 0.0168f; cp->w_max = 1.0f; cp->learning_rate = 1e-3f; cp->nearest_spike_only = false;
      CustomSTDPPlasticity* cu = new CustomSTDPPlasticity(
          (SpikingSynapses*)syn, (SpikingNeurons*)neurons, (SpikingNeurons*)inputs, (stdp_plasticity_parameters_struct*)cp);
      if (plastic) {
        model->AddPlasticityRule(vog);
        model->AddPlasticityRule(ev);
        model->AddPlasticityRule(cu);
        rules.push_back(vog);
        rules.push_back(ev);
        rules.push_back(cu);
        rule_kinds.push_back("vogels");
        rule_kinds.push_back("evans");
        rule_kinds.push_back("custom");
      }
      conductance_spiking_synapse_parameters_struct* ex = new conductance_spiking_synapse_parameters_struct();
      conductance_spiking_synapse_parameters_struct* ih = new conductance_spiking_synapse_parameters_struct();
      for (auto* c : {xi, ex, ih}) {
        c->delay_range[0] = 0.0008f;
        c->delay_range[1] = 0.0008f;
        c->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
        c->random_connectivity_probability = 0.1f;
        c->weight_scaling_constant = 1e-9f;
      }
      xi->weight_range[0] = 1.0f;
      xi->weight_range[1] = 3.0f;
      xi->decay_term_tau_g = 0.005f;
      ex->weight_range[0] = ex->weight_range[1] = 0.6f;
      ex->reversal_potential_Vhat = 0.0f;
      ex->decay_term_tau_g = 0.005f;
      ih->weight_range[0] = ih->weight_range[1] = 6.7f;
      ih->reversal_potential_Vhat = -0.080f;
      ih->decay_term_tau_g = 0.010f;
      if (plastic) xi->plasticity_vec.push_back(ev);
      model->AddSynapseGroup(in_id, e_id, xi);
      if (plastic) ex->plasticity_vec.push_back(cu);
      model->AddSynapseGroup(e_id, e_id, ex);
      ex->plasticity_vec.clear();
      model->AddSynapseGroup(e_id, i_id, ex);
      if (plastic) ih->plasticity_vec.push_back(vog);
      model->AddSynapseGroup(i_id, e_id, ih);
      ih->plasticity_vec.clear();
      model->AddSynapseGroup(i_id, i_id, ih);
    } else {
      xi->delay_range[0] = 0.0001f;
      xi->delay_range[1] = 0.0010f;
      xi->weight_range[0] = 0.0f;
      xi->weight_range[1] = 3.0f;
      xi->decay_term_tau_g = 0.001f;
      model->AddSynapseGroup(in_id, e_id, xi);
    }
  } else if (which == "current") {
    // exponential current synapses: input -> E and E -> I with tau 2 ms, I -> E with 5 ms
    CurrentSpikingSynapses* syn = new CurrentSpikingSynapses(42);
    model->spiking_synapses = syn;
    lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
    lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
    lif(E, 200e-12f, 10e-9f, -0.070f, -0.070f, -0.050f, 2e-3f, 150);
    lif(I, 200e-12f, 10e-9f, -0.070f, -0.070f, -0.050f, 2e-3f, 50);
    poisson_input_spiking_neuron_parameters_struct* ip = new poisson_input_spiking_neuron_parameters_struct();
    ip->group_shape[0] = 1;
    ip->group_shape[1] = 100;
    ip->rate = 60.0f;
    const int in_id = model->AddInputNeuronGroup(ip);
    const int e_id = model->AddNeuronGroup(E);
    const int i_id = model->AddNeuronGroup(I);
    current_spiking_synapse_parameters_struct* xe = new current_spiking_synapse_parameters_struct();
    current_spiking_synapse_parameters_struct* ie = new current_spiking_synapse_parameters_struct();
    for (auto* c : {xe, ie}) {
      c->delay_range[0] = 0.0002f;
      c->delay_range[1] = 0.0012f;
      c->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
      c->random_connectivity_probability = 0.1f;
      c->weight_scaling_constant = 1e-10f;
    }
    xe->weight_range[0] = 1.0f;
    xe->weight_range[1] = 4.0f;
    xe->decay_term_tau = 0.002f;
    ie->weight_range[0] = -6.0f;
    ie->weight_range[1] = -2.0f;
    ie->decay_term_tau = 0.005f;
    model->AddSynapseGroup(in_id, e_id, xe);
    model->AddSynapseGroup(e_id, i_id, xe);
    model->AddSynapseGroup(i_id, e_id, ie);
  } else if (which == "generator") {
    // Examples/SimpleExample.cpp:33-150 with more inputs and neurons, RANDOM E -> I
    ConductanceSpikingSynapses* syn = new ConductanceSpikingSynapses(42);
    model->spiking_synapses = syn;
    generator_input_spiking_neuron_parameters_struct* gp = new generator_input_spiking_neuron_parameters_struct();
    gp->group_shape[0] = 1;
    gp->group_shape[1] = 20;
    lif_spiking_neuron_parameters_struct* E = new lif_spiking_neuron_parameters_struct();
    lif_spiking_neuron_parameters_struct* I = new lif_spiking_neuron_parameters_struct();
    lif(E, 500e-12f, 25e-9f, -0.074f, -0.074f, -0.053f, 2e-3f, 100);
    lif(I, 214e-12f, 18e-9f, -0.082f, -0.082f, -0.053f, 2e-3f, 100);
    const int in_id = model->AddInputNeuronGroup(gp);
    const int e_id = model->AddNeuronGroup(E);
    const int i_id = model->AddNeuronGroup(I);
    conductance_spiking_synapse_parameters_struct* ie = new conductance_spiking_synapse_parameters_struct();
    ie->weight_range[0] = 0.5f;
    ie->weight_range[1] = 10.0f;
    ie->weight_scaling_constant = 25e-9f;
    ie->delay_range[0] = ie->delay_range[1] = 8 * dt;
    ie->connectivity_type = CONNECTIVITY_TYPE_ALL_TO_ALL;
    conductance_spiking_synapse_parameters_struct* ei = new conductance_spiking_synapse_parameters_struct();
    ei->weight_range[0] = ei->weight_range[1] = 10.0f;
    ei->weight_scaling_constant = 18e-9f;
    ei->delay_range[0] = 5 * dt;
    ei->delay_range[1] = 3e-3f;
    ei->connectivity_type = CONNECTIVITY_TYPE_RANDOM;
    ei->random_connectivity_probability = 0.05f;
    conductance_spiking_synapse_parameters_struct* iex = new conductance_spiking_synapse_parameters_struct();
    iex->weight_range[0] = -5.0f;
    iex->weight_range[1] = -2.5f;
    iex->weight_scaling_constant = 25e-9f;
    iex->delay_range[0] = 5 * dt;
    iex->delay_range[1] = 3e-3f;
    iex->connectivity_type = CONNECTIVITY_TYPE_ALL_TO_ALL;
    model->AddSynapseGroup(in_id, e_id, ie);
    model->AddSynapseGroup(e_id, i_id, ei);
    model->AddSynapseGroup(i_id, e_id, iex);
    // two stimuli of scheduled spikes over [0, simtime) and [simtime, 2 simtime); an entry
    // repeated within one step in each
    for (int st = 0; st < 2; ++st) {
      std::vector<int> ids;
      std::vector<float> times;
      unsigned int x = 12345u + 77u * (unsigned)st;
      for (int k = 0; k < 120; ++k) {
        x = x * 1664525u + 1013904223u;
        ids.push_back((int)((x >> 8) % 20u));
        x = x * 1664525u + 1013904223u;
        times.push_back(simtime * ((float)st + (float)((x >> 8) & 0xFFFF) / 65536.0f));
      }
      ids.push_back(ids[0]);
      times.push_back(times[0] + 1e-6f);
      generator->add_stimulus(ids, times);
    }
  } else {
    fprintf(stderr, "unknown case %s\n", which.c_str());
    return 2;
  }

  SpikingActivityMonitor* mon_n = new SpikingActivityMonitor(neurons);
  SpikingActivityMonitor* mon_i = new SpikingActivityMonitor(inputs);
  RateActivityMonitor* rate_n = new RateActivityMonitor(neurons);
  model->AddActivityMonitor(mon_n);
  model->AddActivityMonitor(mon_i);
  model->AddActivityMonitor(rate_n);

  model->finalise_model();
  SpikingSynapses* s = model->spiking_synapses;
  const int S = s->total_number_of_synapses;
  put("syn_pre", s->presynaptic_neuron_indices, S);
  put("syn_post", s->postsynaptic_neuron_indices, S);
  put("syn_w", s->synaptic_efficacies_or_weights, S);
  put("syn_scale", s->weight_scaling_constants, S);
  put("syn_delay", s->delays, S);
  put("syn_label", s->syn_labels, S);
  const int N = neurons->total_number_of_neurons;
  put("lif_v0", neurons->resting_potentials_v0, N);
  put("lif_vreset", neurons->after_spike_reset_potentials_vreset, N);
  put("lif_thr", neurons->thresholds_for_action_potential_spikes, N);
  put("lif_tau", neurons->membrane_time_constants_tau_m, N);
  put("lif_R", neurons->membrane_resistances_R, N);
  if (poisson) {
    put("rates", poisson->rates, inputs->total_number_of_neurons);
  } else {
    std::vector<float> zero(inputs->total_number_of_neurons, 0.0f);
    put("rates", zero.data(), zero.size());
    std::vector<int> counts, ids;
    std::vector<float> times;
    for (int st = 0; st < generator->total_number_of_input_stimuli; ++st) {
      const int k = generator->number_of_spikes_in_stimuli[st];
      counts.push_back(k);
      ids.insert(ids.end(), generator->neuron_id_matrix_for_stimuli[st], generator->neuron_id_matrix_for_stimuli[st] + k);
      times.insert(times.end(), generator->spike_times_matrix_for_stimuli[st],
                   generator->spike_times_matrix_for_stimuli[st] + k);
    }
    put("gen_counts", counts.data(), counts.size());
    put("gen_ids", ids.data(), ids.size());
    put("gen_times", times.data(), times.size());
  }
  std::vector<float> decay;
  std::vector<float> rev;
  if (auto* cs = dynamic_cast<ConductanceSpikingSynapses*>(s)) {
    for (size_t l = 0; l < cs->decay_terms_tau_g.size(); ++l) {
      decay.push_back(expf(-dt / cs->decay_terms_tau_g[l]));
      rev.push_back(cs->reversal_potentials_Vhat[l]);
    }
  }
  if (auto* cu = dynamic_cast<CurrentSpikingSynapses*>(s)) {
    for (size_t l = 0; l < cu->decay_terms_tau.size(); ++l) {
      decay.push_back(expf(-dt / cu->decay_terms_tau[l]));
      rev.push_back(0.0f);
    }
  }
  put("decay_g", decay.data(), decay.size());
  put("reversal", rev.data(), rev.size());
  for (size_t r = 0; r < rules.size(); ++r) {
    put("plastic_" + std::to_string(r), rules[r]->plastic_synapses.data(), rules[r]->plastic_synapses.size());
    if (auto* wn = dynamic_cast<WeightNormSTDPPlasticity*>(rules[r])) {
      std::vector<uint8_t> in(N);
      for (int v = 0; v < N; ++v) in[v] = wn->neuron_in_plasticity_set[v] ? 1 : 0;
      put("wn_sumsq_" + std::to_string(r), wn->sum_squared_afferent_values, N);
      put("wn_inset_" + std::to_string(r), in.data(), in.size());
    }
  }

  model->run(simtime, plastic);
  if (generator) {
    generator->select_stimulus(1);
    model->run(simtime, plastic);
  }

  s->backend()->copy_weights_to_host();
  put("weights_after", s->synaptic_efficacies_or_weights, S);
  // the connectivity files (weights copied from the device) and a reload of their weights
  // through Synapses::load_weights_from_binary (copy_to_backend), read back from the device
  s->save_connectivity_as_binary(g_out, "conn_");
  for (int k = 0; k < S; ++k) s->synaptic_efficacies_or_weights[k] = 0.0f;
  s->load_weights_from_binary(g_out + "/conn_SynapticWeights.bin");
  for (int k = 0; k < S; ++k) s->synaptic_efficacies_or_weights[k] = -1.0f;
  s->backend()->copy_weights_to_host();
  put("weights_reloaded", s->synaptic_efficacies_or_weights, S);
  put("rate_counts", rate_n->per_neuron_spike_counts, N);
  mon_n->save_spikes_as_binary(g_out, "neurons_");
  mon_i->save_spikes_as_binary(g_out, "inputs_");

  std::ofstream j(g_out + "/meta.json");
  j.precision(9);
  j << "{\"case\": \"" << which << "\", \"n_neurons\": " << N << ", \"n_inputs\": "
    << inputs->total_number_of_neurons << ", \"n_synapses\": " << S << ", \"timestep\": " << dt
    << ", \"timestep_grouping\": " << model->timestep_grouping << ", \"refractory\": "
    << neurons->refractory_period_in_seconds << ", \"background_current\": " << neurons->background_current
    << ", \"num_syn_labels\": " << s->num_syn_labels << ", \"conductance\": "
    << (dynamic_cast<ConductanceSpikingSynapses*>(s) ? 1 : 0) << ", \"current\": "
    << (dynamic_cast<CurrentSpikingSynapses*>(s) ? 1 : 0) << ", \"simtime\": " << simtime
    << ", \"plastic\": " << (plastic ? 1 : 0) << ", \"generator\": " << (generator ? 1 : 0)
    << ", \"rule_kinds\": [";
  for (size_t r = 0; r < rule_kinds.size(); ++r) j << (r ? ", " : "") << "\"" << rule_kinds[r] << "\"";
  j << "], \"rule_params\": [";
  for (size_t r = 0; r < rules.size(); ++r) {
    j << (r ? ", " : "") << "{";
    if (auto* wd = dynamic_cast<WeightDependentSTDPPlasticity*>(rules[r])) {
      auto* q = wd->stdp_params;
      j << "\"a_minus\": " << q->a_minus << ", \"a_plus\": " << q->a_plus << ", \"tau_minus\": " << q->tau_minus
        << ", \"tau_plus\": " << q->tau_plus << ", \"lambda\": " << q->lambda << ", \"alpha\": " << q->alpha
        << ", \"w_max\": " << q->w_max << ", \"nearest_spike_only\": " << (q->nearest_spike_only ? 1 : 0);
    } else if (auto* vg = dynamic_cast<InhibitorySTDPPlasticity*>(rules[r])) {
      auto* q = vg->stdp_params;
      j << "\"tau_istdp\": " << q->tau_istdp << ", \"learningrate\": " << q->learningrate
        << ", \"targetrate\": " << q->targetrate << ", \"momentumrate\": " << q->momentumrate;
    } else if (auto* cs = dynamic_cast<CustomSTDPPlasticity*>(rules[r])) {
      auto* q = cs->stdp_params;
      j << "\"a_minus\": " << q->a_minus << ", \"a_plus\": " << q->a_plus << ", \"a_star\": " << q->a_star
        << ", \"tau_minus\": " << q->tau_minus << ", \"tau_plus\": " << q->tau_plus << ", \"w_max\": "
        << q->w_max << ", \"learning_rate\": " << q->learning_rate << ", \"nearest_spike_only\": "
        << (q->nearest_spike_only ? 1 : 0);
    } else if (auto* ev = dynamic_cast<EvansSTDPPlasticity*>(rules[r])) {
      auto* q = ev->stdp_params;
      j << "\"decay_term_tau_D\": " << q->decay_term_tau_D << ", \"model_parameter_alpha_D\": "
        << q->model_parameter_alpha_D << ", \"synaptic_neurotransmitter_concentration_alpha_C\": "
        << q->synaptic_neurotransmitter_concentration_alpha_C << ", \"decay_term_tau_C\": " << q->decay_term_tau_C
        << ", \"learning_rate_rho\": " << q->learning_rate_rho;
    }
    j << "}";
  }
  j << "], \"spikes_neurons\": " << mon_n->total_number_of_spikes_stored_on_host
    << ", \"spikes_inputs\": " << mon_i->total_number_of_spikes_stored_on_host << "}\n";
  return 0;
}
