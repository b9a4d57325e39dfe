// HIPBackend.hpp -- the "HIP" backend of Spike's C++ front-end.
//
// Spike selects a backend per component type through string-keyed factory maps
// (Spike/Backend/Backend.hpp:27-48) filled by static registrars
// (SPIKE_EXPORT_BACKEND_TYPE, Spike/Backend/Macros.hpp:54-70); the default name comes from
// SPIKE_DEFAULT_BACKEND (Spike/Backend/Context.hpp:14-21).  The classes below implement the
// abstract Backend::<Type> interfaces under the name "HIP", so a model built with the
// unchanged front-end (SpikingModel, neuron/synapse groups, plasticity rules,
// SpikingActivityMonitor) runs on the gfx950 library behind include/spike_hip.h.
//
// Design: the reference gives every component its own device arrays and reaches siblings
// by dynamic_cast (Spike/Backend/CUDA/Synapses/SpikingSynapses.cu:128-135).  Here all the
// components of one SpikingModel share one network handle (ModelNetwork), built once every
// component has been prepared; the per-group state_update calls of
// SpikingModel::perform_per_step_model_instructions (SpikingModel.cpp:197-212) collapse into
// one spk_step_group at the synapses' call, which is the last compute phase of a group.
// The class layout (virtual bases mirroring the front-end hierarchy) is what the factory
// macros require.
#pragma once

#include "Spike/Spike.hpp"
#include "spike_hip.h"

namespace Backend {
namespace HIP {

class ModelNetwork;

// Shared network of one SpikingModel.  Returns nullptr before the model is finalised.
ModelNetwork* network_of(::SpikingModel* model);

class Neurons : public virtual ::Backend::Neurons {
 public:
  ~Neurons() override = default;
  void prepare() override {}
  void reset_state() override {}
};

class SpikingNeurons : public virtual ::Backend::HIP::Neurons,
                       public virtual ::Backend::SpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(SpikingNeurons);
  SpikingNeurons() = default;
  void prepare() override;
  void reset_state() override;
  void state_update(float current_time_in_seconds, float timestep) override;
 protected:
  ::SpikingNeurons* front() const { return static_cast<::SpikingNeurons*>(_frontend); }
};

class LIFSpikingNeurons : public virtual ::Backend::HIP::SpikingNeurons,
                          public virtual ::Backend::LIFSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(LIFSpikingNeurons);
};

class InputSpikingNeurons : public virtual ::Backend::HIP::SpikingNeurons,
                            public virtual ::Backend::InputSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(InputSpikingNeurons);
  InputSpikingNeurons() = default;
};

class PoissonInputSpikingNeurons : public virtual ::Backend::HIP::InputSpikingNeurons,
                                   public virtual ::Backend::PoissonInputSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(PoissonInputSpikingNeurons);
  PoissonInputSpikingNeurons() = default;
  void reset_state() override;
};

// Scheduled input spikes (Spike/Backend/CUDA/Neurons/GeneratorInputSpikingNeurons.cu): the
// front-end's stimuli are handed to the library when the network is built; selecting a
// stimulus resets the inputs and replays that stimulus.
class GeneratorInputSpikingNeurons : public virtual ::Backend::HIP::InputSpikingNeurons,
                                     public virtual ::Backend::GeneratorInputSpikingNeurons {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(GeneratorInputSpikingNeurons);
  GeneratorInputSpikingNeurons() = default;
  void reset_state() override;
};

class Synapses : public virtual ::Backend::Synapses {
 public:
  ~Synapses() override = default;
  void prepare() override;
  void reset_state() override;
  void set_neuron_indices_by_sampling_from_normal_distribution(
      int original_number_of_synapses, int total_number_of_new_synapses, int postsynaptic_group_id,
      int poststart, int prestart, int* postsynaptic_group_shape, int* presynaptic_group_shape,
      int number_of_new_synapses_per_postsynaptic_neuron, int number_of_postsynaptic_neurons_in_group,
      int max_number_of_connections_per_pair, float standard_deviation_sigma,
      bool presynaptic_group_is_input) override;
  void copy_to_frontend() override;
  void copy_to_backend() override;
 protected:
  ::Synapses* front() const { return static_cast<::Synapses*>(_frontend); }
};

class SpikingSynapses : public virtual ::Backend::HIP::Synapses,
                        public virtual ::Backend::SpikingSynapses {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(SpikingSynapses);
  SpikingSynapses() = default;
  void copy_weights_to_host() override;
  void state_update(::SpikingNeurons* neurons, ::SpikingNeurons* input_neurons,
                    float current_time_in_seconds, float timestep) override;
};

class VoltageSpikingSynapses : public virtual ::Backend::HIP::SpikingSynapses,
                               public virtual ::Backend::VoltageSpikingSynapses {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(VoltageSpikingSynapses);
};

class ConductanceSpikingSynapses : public virtual ::Backend::HIP::SpikingSynapses,
                                   public virtual ::Backend::ConductanceSpikingSynapses {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(ConductanceSpikingSynapses);
};

class CurrentSpikingSynapses : public virtual ::Backend::HIP::SpikingSynapses,
                               public virtual ::Backend::CurrentSpikingSynapses {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(CurrentSpikingSynapses);
};

class Plasticity : public virtual ::Backend::Plasticity {
 public:
  ~Plasticity() override = default;
  void prepare() override {}
  void reset_state() override {}
};

class STDPPlasticity : public virtual ::Backend::HIP::Plasticity,
                       public virtual ::Backend::STDPPlasticity {
 public:
  ~STDPPlasticity() override = default;
 protected:
  void mark_applied(float current_time_in_seconds);
  ::STDPPlasticity* front() const { return static_cast<::STDPPlasticity*>(_frontend); }
};

class WeightDependentSTDPPlasticity : public virtual ::Backend::HIP::STDPPlasticity,
                                      public virtual ::Backend::WeightDependentSTDPPlasticity {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(WeightDependentSTDPPlasticity);
  void apply_stdp_to_synapse_weights(float current_time_in_seconds, float timestep) override;
};

class InhibitorySTDPPlasticity : public virtual ::Backend::HIP::STDPPlasticity,
                                 public virtual ::Backend::InhibitorySTDPPlasticity {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(InhibitorySTDPPlasticity);
  void apply_stdp_to_synapse_weights(float current_time_in_seconds, float timestep) override;
};

class EvansSTDPPlasticity : public virtual ::Backend::HIP::STDPPlasticity,
                            public virtual ::Backend::EvansSTDPPlasticity {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(EvansSTDPPlasticity);
  void update_synaptic_efficacies_or_weights(float current_time_in_seconds, float timestep) override;
};

class CustomSTDPPlasticity : public virtual ::Backend::HIP::STDPPlasticity,
                             public virtual ::Backend::CustomSTDPPlasticity {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(CustomSTDPPlasticity);
  void apply_stdp_to_synapse_weights(float current_time_in_seconds, float timestep) override;
};

class WeightNormSTDPPlasticity : public virtual ::Backend::HIP::STDPPlasticity,
                                 public virtual ::Backend::WeightNormSTDPPlasticity {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(WeightNormSTDPPlasticity);
  void weight_normalization() override;
  void reset_state() override;
};

class ActivityMonitor : public virtual ::Backend::ActivityMonitor {
 public:
  ~ActivityMonitor() override = default;
  void prepare() override {}
  void reset_state() override {}
};

class SpikingActivityMonitor : public virtual ::Backend::HIP::ActivityMonitor,
                               public virtual ::Backend::SpikingActivityMonitor {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(SpikingActivityMonitor);
  void prepare() override {}
  void reset_state() override;
  void copy_spikes_to_front() override;
  void copy_spikecount_to_front() override;
  void collect_spikes_for_timestep(float current_time_in_seconds, float timestep) override;
 private:
  ::SpikingActivityMonitor* front() const { return static_cast<::SpikingActivityMonitor*>(_frontend); }
};

// Per-neuron spike counts (Spike/Backend/CUDA/ActivityMonitor/RateActivityMonitor.cu): the
// library counts spikes where they are made; the front-end's per_neuron_spike_counts are
// filled at final_update.
class RateActivityMonitor : public virtual ::Backend::HIP::ActivityMonitor,
                            public virtual ::Backend::RateActivityMonitor {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(RateActivityMonitor);
  void prepare() override {}
  void reset_state() override;
  void copy_spike_count_to_host() override;
  void add_spikes_to_per_neuron_spike_count(float current_time_in_seconds, float timestep) override;
 private:
  ::RateActivityMonitor* front() const { return static_cast<::RateActivityMonitor*>(_frontend); }
};

class RandomStateManager : public virtual ::Backend::RandomStateManager {
 public:
  SPIKE_MAKE_BACKEND_CONSTRUCTOR(RandomStateManager);
  // Random numbers are counter-based (Philox keyed by the network seed): no state to set up.
  void prepare() override {}
};

}  // namespace HIP
}  // namespace Backend
