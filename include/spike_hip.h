/*
 * spike_hip.h -- C ABI of the MI355X (gfx950) simulation backend for Spike.
 *
 * This is the boundary between Spike's C++ front-end (SpikingModel, neuron groups,
 * synapse groups, plasticity rules, SpikingActivityMonitor) and the hand-written HIP
 * kernels.  The C++ "HIP" backend classes registered in Spike's backend factory
 * (Spike/Backend/Macros.hpp:54-70 pattern) translate every backend virtual call into
 * one of the entry points below.  Everything here is plain C: opaque handles,
 * pointers, sizes and int status codes (0 = success).  Host pointers are host memory;
 * the library owns all device memory and one HIP stream per network.
 *
 * Conventions
 *   - Presynaptic ids follow Spike's encoding: id >= 0 is a LIF neuron, id < 0 is input
 *     neuron (-id - 1) (Spike/Neurons/Neurons.hpp:34-35).
 *   - Synapse arrays are passed in the caller's order; the library keeps its own
 *     delay-sorted CSR internally and returns weights in the caller's order.
 *   - Times are float seconds exactly as the reference passes them to its backends
 *     (t = start + (s*G)*dt, Spike/Models/SpikingModel.cpp:225-226); the timestep index
 *     is roundf(t/dt) as in Spike/Backend/CUDA/Neurons/LIFSpikingNeurons.cu:84.
 *   - Synaptic input is accumulated into the per-delay input ring as exact 64-bit
 *     fixed-point sums (2^-frac_bits units), so results do not depend on the order in
 *     which spikes are delivered.
 */
#ifndef SPIKE_HIP_H
#define SPIKE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spk_network spk_network;

/* status codes */
enum {
  SPK_OK = 0,
  SPK_ERR_ARG = 1,       /* invalid argument or size */
  SPK_ERR_STATE = 2,     /* call out of order (e.g. before spk_prepare) */
  SPK_ERR_HIP = 3,       /* a HIP runtime call failed */
  SPK_ERR_NOMEM = 4,     /* device or host allocation failed */
  SPK_ERR_CAPACITY = 5,  /* a bounded device store overflowed */
  SPK_ERR_NODEVICE = 6   /* no usable gfx950 device */
};

/* synapse models (Spike/Synapses/{Voltage,Conductance}SpikingSynapses.hpp) */
enum { SPK_SYN_VOLTAGE = 0, SPK_SYN_CONDUCTANCE = 1, SPK_SYN_CURRENT = 2 };

/* populations */
enum { SPK_POP_NEURONS = 0, SPK_POP_INPUTS = 1 };

/* storage of synaptic weights and plasticity traces (spk_set_weight_format) */
enum { SPK_WEIGHTS_FP32 = 0, SPK_WEIGHTS_BF16 = 1, SPK_WEIGHTS_FP16 = 2 };

/* run-time switches for spk_run (bit flags) */
enum {
  SPK_RUN_PLASTICITY = 1,   /* apply plasticity rules (SpikingModel::run plasticity_on) */
  SPK_RUN_FORCE_LAUNCHES = 2 /* use the per-phase kernels instead of the fused kernel */
};

/* Weight-dependent STDP parameters.
 * Replaces weightdependent_stdp_plasticity_parameters_struct
 * (Spike/Plasticity/WeightDependentSTDPPlasticity.hpp:28-40). */
typedef struct {
  float a_minus, a_plus, tau_minus, tau_plus, lambda, alpha, w_max;
  int nearest_spike_only;
} spk_wd_stdp_params;

/* Vogels inhibitory STDP parameters.
 * Replaces inhibitory_stdp_plasticity_parameters_struct
 * (Spike/Plasticity/InhibitorySTDPPlasticity.hpp:38-46). */
typedef struct {
  float tau_istdp, learningrate, targetrate, momentumrate;
} spk_vogels_stdp_params;

/* Evans STDP parameters.
 * Replaces evans_stdp_plasticity_parameters_struct
 * (Spike/Plasticity/EvansSTDPPlasticity.hpp:33-40). */
typedef struct {
  float decay_term_tau_D, model_parameter_alpha_D,
        synaptic_neurotransmitter_concentration_alpha_C, decay_term_tau_C,
        learning_rate_rho;
} spk_evans_stdp_params;

/* Custom STDP parameters.
 * Replaces custom_stdp_plasticity_parameters_struct
 * (Spike/Plasticity/CustomSTDPPlasticity.hpp:31-40). */
typedef struct {
  float a_minus, a_plus, a_star, tau_minus, tau_plus, w_max, learning_rate;
  int nearest_spike_only;
} spk_custom_stdp_params;

/* Sizes fixed by spk_prepare. */
typedef struct {
  int n_neurons, n_inputs, n_labels, synapse_kind;
  int timestep_grouping;     /* G: sub-steps per group (Spike's min axonal delay) */
  int ring_slots;            /* B = max_delay + 2G + 1 (SpikingSynapses.cu:48) */
  int history_words;         /* 32-bit words of spike history per neuron */
  int frac_bits;             /* fixed-point fraction bits of the input ring */
  int min_delay, max_delay;  /* in timesteps */
  int64_t n_synapses;
  int n_rules, n_monitors;
  int partitions;            /* workgroups of the persistent kernel (neuron partitions) */
  int fused;                 /* 1 when spk_run uses the persistent kernel */
  int record_bytes;          /* bytes the persistent kernel reads per delivery: 8 for the
                                full records {target|delay|label, weight}; 2 for the compact
                                form (16-bit target; weight and delay per source type) */
  int lds_bytes;             /* LDS per workgroup the persistent kernel needs (0: not used);
                                the optional delivery table it keeps in LDS when that fits
                                is not counted */
} spk_info;

/* Per-run timing of the last spk_run call, from HIP events on the network's stream. */
typedef struct {
  double total_ms;           /* whole run, first launch to last completion */
  double kernel_ms;          /* summed duration of the dominant (fused) kernel */
  int64_t kernel_launches;   /* launches of the dominant kernel */
  int64_t groups;            /* timestep groups simulated */
  int64_t spikes_neurons;    /* LIF spikes emitted */
  int64_t spikes_inputs;     /* Poisson input spikes emitted */
  int64_t synaptic_events;   /* deliveries: sum over spikes of out-degree */
  int64_t external_events;   /* per-neuron external Poisson events */
} spk_timing;

/* ---------------------------------------------------------------------------- */
/* library */

/* Message of the last failed call on this thread. */
const char* spk_last_error(void);
/* ABI version (major*100+minor). */
int spk_abi_version(void);
/* Number of visible HIP devices; 0 when none. Does not fail. */
int spk_device_count(void);

/* ---------------------------------------------------------------------------- */
/* network lifetime
 * Replaces the allocation half of Backend::<Type>::prepare() of every backend type
 * (Spike/Backend/Backend.hpp:15-25) with one network object per SpikingModel. */
int spk_create(int n_neurons, int n_inputs, int synapse_kind, int n_labels,
               float timestep, spk_network** out);
int spk_destroy(spk_network* net);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL keeps the own one. */
int spk_set_stream(spk_network* net, void* hip_stream);

/* LIF neurons.  Replaces Backend::LIFSpikingNeurons::prepare
 * (CUDA LIFSpikingNeurons.cu:26-66, SpikingNeurons.cu:38-42).
 * decay_c[i] = dt/tau_m, resist[i] = decay_c[i]*R (the value the reference stores in
 * membrane_resistances_R on the device); refractory period and background current are
 * scalars as in Spike/Neurons/LIFSpikingNeurons.cpp:24,37. */
int spk_set_lif(spk_network* net, const float* v0, const float* vreset,
                const float* threshold, const float* decay_c, const float* resist,
                float refractory_s, float background_current);

/* Poisson inputs.  Replaces Backend::PoissonInputSpikingNeurons::prepare
 * (CUDA PoissonInputSpikingNeurons.cu:36-47): rates[s*n_inputs + i] in Hz. */
int spk_set_poisson_rates(spk_network* net, const float* rates, int n_stimuli);
/* Generator inputs instead of Poisson inputs: the input population replays scheduled
 * spikes.  Replaces Backend::GeneratorInputSpikingNeurons::prepare / reset_state (CUDA
 * GeneratorInputSpikingNeurons.cu:14-40) and its stimuli (GeneratorInputSpikingNeurons.cpp:
 * 28-77): stimulus s holds counts[s] spikes, (ids, times) concatenated over the stimuli,
 * ids in [0, n_inputs), times in seconds.  In each group starting at t, entry e fires
 * input ids[e] at sub-step g when |(t - onset + g*dt) - times[e]| < 0.5*dt
 * (GeneratorInputSpikingNeurons.cu:106-123); an input spikes at most once per step, and
 * the activated synapses are those of ids[e] (the reference kernel passes the entry index
 * instead).  Before spk_prepare; the network then runs the per-phase kernels. */
int spk_set_generator_stimuli(spk_network* net, const int32_t* counts, int n_stimuli,
                              const int32_t* ids, const float* times);
/* GeneratorInputSpikingNeurons::stimulus_onset_adjustment (default 0). */
int spk_set_generator_onset(spk_network* net, float onset);
/* InputSpikingNeurons::select_stimulus (Spike/Neurons/InputSpikingNeurons.cpp:12-19). */
int spk_select_stimulus(spk_network* net, int stimulus);
/* Philox key of all random streams (replaces RandomStateManager, CUDA
 * RandomStateManager.cu:14-50). */
int spk_set_seed(spk_network* net, uint64_t seed);
/* Per-neuron external Poisson drive (no reference counterpart: the Brunel network's
 * external population nu_ext, generated in registers instead of through synapses).
 * Each event adds `weight` to the neuron's input for that timestep. rate_hz=0 disables. */
int spk_set_external_drive(spk_network* net, float rate_hz, float weight);
/* Storage of synaptic weights and plasticity traces, before spk_prepare.  The reference
 * keeps FP32 weights (synaptic_efficacies_or_weights, Spike/Synapses/Synapses.hpp) and FP32
 * traces; SPK_WEIGHTS_BF16 / _FP16 store the weights as bfloat16 / binary16 words (rounded
 * to nearest even), so a delivery reads 4 B of target record + 2 B of weight, and the
 * traces as binary16 in both modes (bfloat16's spacing is near a trace's 0.5 % decay per
 * 0.1 ms step, which would bias its decay).  Plastic weight updates are
 * computed in FP32 from the stored weight and added with Kahan compensation (a second
 * 16-bit word per synapse holds what the first cannot).  Inputs, conductances and
 * membranes stay FP32; spk_get_weights/spk_get_traces return the stored values. */
int spk_set_weight_format(spk_network* net, int format);

/* Conductance synapse labels.  Replaces ConductanceSpikingSynapses::prepare
 * (CUDA ConductanceSpikingSynapses.cu:19-87): decay_g[l] = expf(-dt/tau_g[l]).
 * Current synapses (SPK_SYN_CURRENT, CurrentSpikingSynapses.cu:13-83) use the same call:
 * decay_g[l] = expf(-dt/tau[l]) of the current trace, reversal ignored; the injected
 * voltage is R' * sum_l I_l, with I_l = decay_g[l] * I_l + input (the reference kernel
 * never allocates the trace array, never copies tau to it and clears a different ring
 * index than it reads, InlineDeviceFunctions.hpp:52-84: none of that is reproduced). */
int spk_set_labels(spk_network* net, const float* decay_g, const float* reversal);

/* Synapses.  Replaces Backend::SpikingSynapses::prepare / Synapses::prepare
 * (CUDA SpikingSynapses.cu:44-117, Synapses.cu:87-102).  delay in timesteps (>=1),
 * label in [0, n_labels).  scale may be NULL (all 1). */
int spk_set_synapses(spk_network* net, int64_t n, const int32_t* pre, const int32_t* post,
                     const float* weight, const float* scale, const int32_t* delay,
                     const int32_t* label);

/* Plasticity rules over synapse indices (caller's order).  Replace
 * Backend::STDPPlasticity::prepare (CUDA STDPPlasticity.cu:13-41) plus the rule's
 * prepare (WeightDependentSTDPPlasticity.cu:28-43, InhibitorySTDPPlasticity.cu:30-48,
 * EvansSTDPPlasticity.cu:30-37).  rule_id receives the index used by
 * spk_update_plasticity. */
int spk_add_wd_stdp(spk_network* net, const int64_t* synapses, int64_t n,
                    const spk_wd_stdp_params* p, int* rule_id);
int spk_add_vogels_stdp(spk_network* net, const int64_t* synapses, int64_t n,
                        const spk_vogels_stdp_params* p, int* rule_id);
int spk_add_evans_stdp(spk_network* net, const int64_t* synapses, int64_t n,
                       const spk_evans_stdp_params* p, int* rule_id);
/* CustomSTDPPlasticity (CustomSTDPPlasticity.cu:13-159): per sub-step, traces decayed by
 * expf(-dt/tau_minus) (post) and expf(-dt/tau_plus) (pre), raised by a_minus / a_plus
 * (or set to them, nearest_spike_only) at the delayed pre and the post spike; a pre spike
 * depresses by lr*(w/w_max)*x_post + lr*a_star, a post spike potentiates by lr*x_pre;
 * weights clamped at 0. */
int spk_add_custom_stdp(spk_network* net, const int64_t* synapses, int64_t n,
                        const spk_custom_stdp_params* p, int* rule_id);
/* WeightNormSTDPPlasticity (WeightNormSTDPPlasticity.cpp:32-72, CUDA
 * WeightNormSTDPPlasticity.cu:68-172): each application rescales the rule's synapses onto
 * every neuron v with in_set[v] by sqrtf(sum_sq[v] + u[v]) / sqrtf(sum_sq[v]), where
 * u[v] = sum over v's synapses, in the caller's order, of d*d + 2*w0*d with d = w - w0
 * (w0: the initial weight), when that divisor is not 1.  sum_sq and in_set ([n_neurons])
 * are the front-end's (it computes them, and may rescale the weights, before the backend
 * is prepared).  The initial weights are the weights at spk_prepare, replaced by
 * spk_weightnorm_set_initial (caller's order) as the reference's reset_state does.  A
 * network with this rule runs the per-phase kernels. */
int spk_add_weightnorm(spk_network* net, const int64_t* synapses, int64_t n, const float* sum_sq,
                       const uint8_t* in_set, int* rule_id);
int spk_weightnorm_set_initial(spk_network* net, int rule_id, const float* initial_weights);

/* Spike monitors.  Replaces Backend::SpikingActivityMonitor::prepare
 * (CUDA SpikingActivityMonitor.cu:22-33): a bounded device store of `capacity`
 * (id, time) records for one population. */
int spk_add_monitor(spk_network* net, int population, int64_t capacity, int* monitor_id);

/* Build device structures.  G = SpikingModel::timestep_grouping (SpikingModel.cpp:98).
 * After this call neurons/synapses/rules/monitors are fixed. */
int spk_prepare(spk_network* net, int timestep_grouping);
/* Change G after prepare (VogelsAbbottNet.cpp:229-231 sets it to 1); must not exceed
 * the value given to spk_prepare. */
int spk_set_timestep_grouping(spk_network* net, int timestep_grouping);
int spk_get_info(const spk_network* net, spk_info* info);
/* Number of neuron partitions (= workgroups) of the persistent kernel; 0 picks
 * min(CU count, ceil(N/8)).  Call before spk_prepare.  (No reference counterpart: it
 * tunes the launch shape, like the reference's threads-per-block context parameter,
 * Spike/Backend/Context.hpp:8-12.) */
int spk_set_partitions(spk_network* net, int partitions);
/* Phase profile of the persistent kernel (no reference counterpart; the reference has
 * only a clock() timer around run(), Examples/Brunel10K.cpp:255-257).  When on, spk_run
 * accumulates workgroup 0's wall-clock time per phase: 0 Poisson inputs and the LIF
 * input read, 9 the LIF recurrence, 1 spike history, statistics and spike append,
 * 2 publication, 3 barrier arrival and the next group's random draws, 4 barrier wait,
 * 5 plasticity, 6 delivery heads, 7 count scan and entry staging, 8 synapse delivery;
 * 10 and 11 are the wall-clock start and end of the
 * last launch's group loop (workgroup 0; in a batch, the network's first workgroup).
 * Values in microseconds summed over the last run (up to 12 entries). */
int spk_set_profiling(spk_network* net, int on);
int spk_get_phase_profile(const spk_network* net, double* us, int count);
/* Empty string when spk_run uses the persistent kernel, else the reason it cannot. */
const char* spk_fused_status(const spk_network* net);

/* reset_state of SpikingModel (SpikingModel.cpp:181-190): V=v0, last spike=-1000,
 * spike history, input ring, conductances and plasticity traces cleared, Poisson
 * inputs disarmed.  Weights are kept. */
int spk_reset_state(spk_network* net);
/* reset of the input population only (PoissonInputSpikingNeurons::select_stimulus). */
int spk_reset_inputs(spk_network* net);

/* ---------------------------------------------------------------------------- */
/* per-phase updates: the backend state_update entry points called once per timestep
 * group by SpikingModel::perform_per_step_model_instructions (SpikingModel.cpp:197-212) */

/* Backend::SpikingNeurons::state_update for LIF (LIFSpikingNeurons.cu:72-195). */
int spk_update_neurons(spk_network* net, float t);
/* Backend::SpikingNeurons::state_update for Poisson (PoissonInputSpikingNeurons.cu:49-130). */
int spk_update_inputs(spk_network* net, float t);
/* apply_stdp_to_synapse_weights / update_synaptic_efficacies_or_weights of one rule
 * (WeightDependentSTDPPlasticity.cu:45-163, InhibitorySTDPPlasticity.cu:50-155,
 * EvansSTDPPlasticity.cu:40-143). */
int spk_update_plasticity(spk_network* net, int rule_id, float t);
/* Backend::SpikingSynapses::state_update (SpikingSynapses.cu:119-202). */
int spk_update_synapses(spk_network* net, float t);
/* Backend::SpikingActivityMonitor::collect_spikes_for_timestep
 * (SpikingActivityMonitor.cu:52-98). */
int spk_collect_spikes(spk_network* net, int monitor_id, float t);

/* Fused run: the whole SpikingModel::run group loop (SpikingModel.cpp:214-236) on the
 * device for n_groups groups starting at start_time.  Monitor records produced during
 * the run are drained into the library's host-side store (spk_monitor_take). */
int spk_run(spk_network* net, float start_time, int64_t n_groups, int flags);
int spk_last_timing(const spk_network* net, spk_timing* timing);
/* One timestep group enqueued on the network's stream without waiting for it: the whole
 * of SpikingModel::perform_per_step_model_instructions (SpikingModel.cpp:197-212) for the
 * group starting at t -- neurons, inputs, plasticity (SPK_RUN_PLASTICITY), propagation
 * and every monitor's collection into its device store.  This is what the backend
 * state_update calls of one group map to; like the reference's kernel launches it returns
 * before the group has run.  Monitor records stay on the device (spk_monitor_device_count,
 * spk_monitor_drain). */
int spk_step_group(spk_network* net, float t, int flags);
/* Cumulative event counters since spk_prepare / the last spk_run (spikes, deliveries,
 * external events; the time fields are zero).  Synchronizes. */
int spk_get_counters(spk_network* net, spk_timing* counters);
int spk_synchronize(spk_network* net);

/* ---------------------------------------------------------------------------- */
/* FP64 spike statistics of the LIF population (no reference counterpart: the reference
 * leaves analysis to scripts over SpikeIDs.bin / SpikeTimes.bin, SpikingActivityMonitor.cpp:
 * 140-153).  Accumulated on the device where spikes are made, over the window of
 * `max_steps` steps that starts at the first step simulated after spk_stats_begin, and
 * reduced in FP64 on the device:
 *   rate_hz        spikes / (N * steps * dt)
 *   isi_cv         mean over neurons with >= 2 in-window intervals of std/mean of them
 *   synchrony      variance / mean of the population spike count per bin of bin_steps
 *   mean_bin_count mean population spike count per bin */
typedef struct {
  double rate_hz, isi_cv, synchrony, mean_bin_count;
  int64_t spikes, cv_neurons, bins, steps;
} spk_stats;
int spk_stats_begin(spk_network* net, int bin_steps, int64_t max_steps);
int spk_stats_get(spk_network* net, spk_stats* stats);
int spk_stats_end(spk_network* net);

/* ---------------------------------------------------------------------------- */
/* read-out */

/* Records currently in the device store of a monitor
 * (SpikingActivityMonitor.cu:35-39).  Overflow is reported as a count > capacity. */
int spk_monitor_device_count(spk_network* net, int monitor_id, int64_t* count);
/* Clear a monitor's device store and host store (SpikingActivityMonitor.cu reset_state,
 * which zeroes the device spike count). */
int spk_monitor_reset(spk_network* net, int monitor_id);
/* Copy device records (SpikingActivityMonitor.cu:41-50) and clear the device store. */
int spk_monitor_drain(spk_network* net, int monitor_id, int32_t* ids, float* times,
                      int64_t cap, int64_t* n);
/* Rate monitor (RateActivityMonitor.cpp:17-34, CUDA RateActivityMonitor.cu:38-74): the
 * number of spikes of every neuron (or input) of a population, counted where the spikes
 * are made.  Added before spk_prepare; counts since prepare or the monitor's last reset
 * (copy_spike_count_to_host / reset_state). */
int spk_add_rate_monitor(spk_network* net, int population, int* monitor_id);
int spk_rate_monitor_counts(spk_network* net, int monitor_id, int32_t* counts);
int spk_rate_monitor_reset(spk_network* net, int monitor_id);
/* Host-side records accumulated by spk_run since the last take; ids are local to the
 * population, sorted by (time, id).  Pass ids=NULL to query the count only. */
int spk_monitor_take(spk_network* net, int monitor_id, int32_t* ids, float* times,
                     int64_t cap, int64_t* n);

/* Synapses::copy_to_frontend / copy_to_backend (CUDA Synapses.cu:104-121). */
int spk_get_weights(spk_network* net, float* weights);
int spk_set_weights(spk_network* net, const float* weights);
/* Neuron state for checks: membrane potentials (LIF) and last spike times. */
int spk_get_membrane(spk_network* net, float* v);
int spk_get_last_spike(spk_network* net, int population, float* last_spike);
/* Per-neuron per-label conductances (conductance synapses only), [n_neurons*n_labels]. */
int spk_get_conductances(spk_network* net, float* g);
/* The input ring: ring[slot*(n_neurons*n_labels) + post*n_labels + label]. */
int spk_get_ring(spk_network* net, int64_t* ring);
/* Spike history bits, words[w*n_pop + i] for the population. */
int spk_get_history(spk_network* net, int population, uint32_t* words);
/* Plasticity traces of a rule in its synapse order (two arrays of n); for a weight
 * normalization rule, its initial weights and zeros. */
int spk_get_traces(spk_network* net, int rule_id, float* trace_a, float* trace_b);
/* Synapse count of a rule (after spk_prepare; a generated rule's size is known only then). */
int spk_get_rule_size(spk_network* net, int rule_id, int64_t* count);

/* ---------------------------------------------------------------------------- */
/* Brunel networks generated on the device (no reference counterpart: the reference builds
 * connectivity on the host, Examples/UtilityFunctions.hpp:51-82 and Synapses.cpp:102-388,
 * which takes minutes for 1e9 synapses).  N = n_exc + n_inh LIF neurons of one parameter
 * set; every neuron receives ce synapses from the excitatory and ci from the inhibitory
 * population, sources drawn with replacement by Philox4x32-10 (key = seed; draw j of
 * neuron i is output j mod 4 of the call with counter (j/4, i, 3)):
 * source = floor(u*n_exc/2^32) or n_exc + floor(u*n_inh/2^32).  All
 * delays are delay_steps (= G).  Such a network runs on the persistent kernel only
 * (spk_run, spk_step_group, statistics, monitors; no plasticity, no weight access). */
typedef struct {
  int n_exc, n_inh, ce, ci, delay_steps;
  float weight_exc, weight_inh;                 /* J and -g*J */
  float v0, vreset, threshold, decay_c, resist, refractory_s;
  float ext_rate_hz, ext_weight;                /* per-neuron external Poisson drive */
  float timestep;
  uint64_t seed;
} spk_brunel_params;
int spk_create_brunel(const spk_brunel_params* params, int partitions, spk_network** out);

/* Fixed out-degree connectivity generated on the device (the fast builder for large
 * networks; no reference counterpart: the reference's CONNECTIVITY_TYPE_RANDOM draws one
 * rand() per (pre, post) pair and keeps 36 B of host arrays per synapse,
 * Synapses.cpp:208-228 and :391-421).  Called on a created network with LIF parameters
 * (and Poisson rates) set, in place of spk_set_synapses and the rule calls; spk_prepare
 * then generates the synapses and the rule tables on the device, with no host copy.
 * Unit u in [0, n_neurons + n_inputs) (neurons, then inputs) has out_degree
 * (input_out_degree) synapses, targets drawn with replacement; synapse j of unit u takes
 * r = Philox4x32-10(ctr = {j, 0, u, 5}, key = seed):
 *   post = floor(r.x * n_neurons / 2^32), delay = 1 + floor(r.y * delay_max / 2^32),
 *   weight = (r.z >> 8) * 2^-24 for excitatory sources (u < n_exc), else weight_inh /
 *   weight_input; scale by the same source classes.
 * Caller synapse s is CSR position s: unit u's synapses are s = row(u) + j with
 * row(u) = u*out_degree for u <= n_neurons, then input_out_degree per input.  With
 * ee_stdp, every synapse from and to the first n_exc neurons belongs to one
 * weight-dependent STDP rule (rule id 0), listed in synapse order.  Such a network runs
 * on the per-phase kernels; weights and traces are readable as for a listed network. */
typedef struct {
  uint64_t seed;
  int n_exc, out_degree, input_out_degree, delay_max;
  float scale_exc, scale_inh, scale_input;
  float weight_inh, weight_input;
  int ee_stdp;
  spk_wd_stdp_params stdp;
} spk_outdegree_params;
int spk_set_outdegree_synapses(spk_network* net, const spk_outdegree_params* params);

/* Batches: many prepared networks (same synapse kind; each able to use the persistent
 * kernel) run in one cooperative launch of `workgroups` workgroups (0 = as many as are
 * co-resident).  Each network keeps its own partitions (spk_set_partitions /
 * spk_create_brunel); a list scheduler assigns every network's partitions to the
 * workgroups that become free first, with the cost model
 *   ngroups * (fixed_us + per_neuron_us * neurons per partition
 *              + per_synapse_us * synapses per partition).
 * No reference counterpart (parameter sweeps run one model at a time in the reference). */
typedef struct spk_batch spk_batch;
int spk_batch_create(spk_network* const* nets, int n_nets, int workgroups, spk_batch** out);
/* A batch run is a hipGraph: the first run captures the argument copies (staged in pinned
 * host memory) and the cooperative launch, and every later run of the same batch replays
 * it with that run's arguments (start times, group counts, seeds, statistics windows).
 * on = 0 runs every launch directly instead.  Default on.  No reference counterpart. */
int spk_batch_set_graph(spk_batch* batch, int on);
/* Graph captures and replays since the batch was created. */
int spk_batch_graph_info(const spk_batch* batch, int64_t* captures, int64_t* replays);
int spk_batch_destroy(spk_batch* batch);
int spk_batch_set_cost_model(spk_batch* batch, double fixed_us, double per_neuron_us,
                             double per_synapse_us);
/* Per-network cost of one group in microseconds (batch order), e.g. measured by a probe
 * run; replaces the model above for scheduling.  NULL returns to the model. */
int spk_batch_set_costs(spk_batch* batch, const double* cost_per_group_us);
int spk_batch_workgroups(const spk_batch* batch);
/* Threads per workgroup of the batch's launches: 512 (two workgroups per compute unit)
 * when every partition has at most 512 neurons and two workgroups' LDS fits a compute
 * unit, else 1024. */
int spk_batch_threads(const spk_batch* batch);
/* Network k runs ngroups[k] groups from starts[k]; returns when all have finished. */
int spk_batch_run(spk_batch* batch, const float* starts, const int64_t* ngroups, int flags);
/* Kernel time of the last run (HIP events) and the scheduler's estimate. */
int spk_batch_last_ms(const spk_batch* batch, double* kernel_ms, double* estimated_ms);

#ifdef __cplusplus
}
#endif

#endif /* SPIKE_HIP_H */
