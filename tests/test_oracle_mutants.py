"""Planted mistakes in the oracle must fail its pins.

Each mutant below is one plausible mistake (a dropped term, a flipped sign, a wrong index,
stream or window) planted in a copy of an oracle source: the C++ restatement
(oracle/spike_oracle.cpp, built here into a temporary library loaded through
SPIKE_ORACLE_LIB) or a Python one (oracle/construction.py, oracle/mlp_oracle.py, loaded
through SPIKE_ORACLE_PYDIR).  The pins named for it are run against the mutant in a
subprocess and must fail; the unmutated copy must pass the same pins (the control).
Everything runs on the CPU."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-pthread", "-shared"]

P = "tests/test_oracle_pinned.py::"
S = "tests/test_oracle_stdp_pinned.py::"
D = "tests/test_outdegree_oracle.py::"

# (name, file, original text, planted text, pins that must fail)
MUTANTS = [
    ("philox_weyl_constant", "spike_oracle.cpp",
     "const uint32_t W0 = 0x9E3779B9u", "const uint32_t W0 = 0x9E3779B8u",
     [P + "test_philox_kat"]),
    ("bernoulli_first_order", "spike_oracle.cpp",
     "double p = -std::expm1(-(double)rate * (double)dt);", "double p = (double)rate * (double)dt;",
     [P + "test_bernoulli_threshold"]),
    ("count_table_shifted", "spike_oracle.cpp",
     "double pmf = std::exp(-lambda), cdf = pmf;", "double pmf = std::exp(-lambda), cdf = 0.0;",
     [P + "test_poisson_count_table_is_the_cdf"]),
    ("frac_bits_off_by_one", "spike_oracle.cpp",
     "int k = 52 - e;", "int k = 53 - e;",
     [P + "test_frac_bits_rule"]),
    ("lif_leak_dropped", "spike_oracle.cpp",
     "Vi = ((cc * (R)v0[i] + ((R)1 - cc) * Vi) + cc * (R)Ibg) + dV;",
     "Vi = ((cc * (R)v0[i] + Vi) + cc * (R)Ibg) + dV;",
     [P + "test_lif_subthreshold_closed_form"]),
    ("lif_background_unscaled", "spike_oracle.cpp",
     "Vi = ((cc * (R)v0[i] + ((R)1 - cc) * Vi) + cc * (R)Ibg) + dV;",
     "Vi = ((cc * (R)v0[i] + ((R)1 - cc) * Vi) + (R)Ibg) + dV;",
     [P + "test_lif_subthreshold_closed_form"]),
    ("refractory_dropped", "spike_oracle.cpp",
     "if (((tg - last_n[i]) - tref) > 0.5f * dt) {", "if ((tg - last_n[i]) > 0.5f * dt) {",
     [P + "test_first_spike_and_refractory_window", P + "test_voltage_injection_step_by_step"]),
    ("voltage_slot_not_cleared", "spike_oracle.cpp",
     "              if (l == 0) a += ext;\n              ring[slot * NL + (size_t)i * L + l] = (Acc)0;\n"
     "              if constexpr (kFixed) total +=",
     "              if (l == 0) a += ext;\n              if constexpr (kFixed) total +=",
     [P + "test_voltage_injection_step_by_step"]),
    ("delay_off_by_one", "spike_oracle.cpp",
     "int slot = (tidx % B + sdelay[s] + g) % B;", "int slot = (tidx % B + sdelay[s] + g + 1) % B;",
     [P + "test_delay_arrival_step", P + "test_voltage_injection_step_by_step"]),
    ("conductance_drop_driving_force", "spike_oracle.cpp",
     "if (kind == 1) total += gl * ((R)Erev[l] - Vi);", "if (kind == 1) total += gl;",
     [P + "test_conductance_injection_step_by_step", P + "test_conductance_at_rest_reversal_is_silent",
      P + "test_conductance_steady_state_is_the_weighted_reversal_mean"]),
    ("conductance_sign_flipped", "spike_oracle.cpp",
     "if (kind == 1) total += gl * ((R)Erev[l] - Vi);", "if (kind == 1) total += gl * (Vi - (R)Erev[l]);",
     [P + "test_conductance_injection_step_by_step",
      P + "test_conductance_steady_state_is_the_weighted_reversal_mean"]),
    ("conductance_rest_not_current_v", "spike_oracle.cpp",
     "if (kind == 1) total += gl * ((R)Erev[l] - Vi);", "if (kind == 1) total += gl * ((R)Erev[l] - (R)v0[i]);",
     [P + "test_conductance_injection_step_by_step",
      P + "test_conductance_steady_state_is_the_weighted_reversal_mean"]),
    ("conductance_no_decay", "spike_oracle.cpp",
     "R gl = (R)decay_g[l] * gcond[(size_t)i * L + l] + in;", "R gl = gcond[(size_t)i * L + l] + in;",
     [P + "test_conductance_injection_step_by_step",
      P + "test_conductance_steady_state_is_the_weighted_reversal_mean"]),
    ("conductance_resistance_dropped", "spike_oracle.cpp",
     "dV = total * (R)Rp[i];", "dV = total;",
     [P + "test_conductance_injection_step_by_step",
      P + "test_conductance_steady_state_is_the_weighted_reversal_mean",
      P + "test_current_synapse_trace_and_injection"]),
    ("drive_wrong_stream", "spike_oracle.cpp",
     "DrawCursor ext_draw(seed, (uint32_t)i, STREAM_EXTERNAL);", "DrawCursor ext_draw(seed, (uint32_t)i, STREAM_INPUT);",
     [P + "test_external_drive_count_is_poisson_inverse_cdf_of_philox"]),
    ("drive_step_shifted", "spike_oracle.cpp",
     "uint32_t u = ext_draw(step);", "uint32_t u = ext_draw(step + 1);",
     [P + "test_external_drive_count_is_poisson_inverse_cdf_of_philox"]),
    ("drive_one_event_per_step", "spike_oracle.cpp",
     "if constexpr (kFixed) ext = (Acc)k * to_acc((R)ext_w, frac);\n            else ext = (Acc)k * (Acc)ext_w;",
     "if constexpr (kFixed) ext = (Acc)(k > 0) * to_acc((R)ext_w, frac);\n            else ext = (Acc)(k > 0) * (Acc)ext_w;",
     [P + "test_external_drive_count_is_poisson_inverse_cdf_of_philox",
      P + "test_external_drive_count_histogram_is_poisson"]),
    ("drive_count_rule_inverted", "spike_oracle.cpp",
     "for (uint32_t cth : ext_cdf) k += (u >= cth) ? 1 : 0;", "for (uint32_t cth : ext_cdf) k += (u < cth) ? 1 : 0;",
     [P + "test_external_drive_count_is_poisson_inverse_cdf_of_philox",
      P + "test_external_drive_count_histogram_is_poisson"]),
    ("poisson_input_not_armed", "spike_oracle.cpp",
     "if (u < thr_s[i] && step != arm_step) {", "if (u < thr_s[i]) {",
     [P + "test_poisson_input_rate_and_isi"]),
    ("poisson_input_complement", "spike_oracle.cpp",
     "if (u < thr_s[i] && step != arm_step) {", "if (u >= thr_s[i] && step != arm_step) {",
     [P + "test_poisson_input_rate_and_isi"]),
    ("generator_modulo_mapping", "spike_oracle.cpp",
     "pre[(size_t)i * C + j] = j < ce ? (int32_t)(((uint64_t)u * (uint64_t)ne) >> 32)",
     "pre[(size_t)i * C + j] = j < ce ? (int32_t)(u % (uint32_t)ne)",
     [P + "test_brunel_connectivity_counts_and_rows"]),
    ("generator_wrong_stream", "spike_oracle.cpp",
     "const uint32_t u = Philox::draw(seed, (int64_t)j, (uint32_t)i, 3u);",
     "const uint32_t u = Philox::draw(seed, (int64_t)j, (uint32_t)i, 2u);",
     [P + "test_brunel_connectivity_counts_and_rows"]),
    ("generator_type_boundary", "spike_oracle.cpp",
     "pre[(size_t)i * C + j] = j < ce ?", "pre[(size_t)i * C + j] = j <= ce ?",
     [P + "test_brunel_connectivity_counts_and_rows"]),
    ("outdegree_delay_from_target_draw", "spike_oracle.cpp",
     "delay[s] = 1 + (int32_t)(((uint64_t)r[1] * (uint64_t)dmax) >> 32);",
     "delay[s] = 1 + (int32_t)(((uint64_t)r[0] * (uint64_t)dmax) >> 32);",
     [D + "test_rows_recomputed_from_philox", D + "test_target_delay_weight_laws"]),
    ("outdegree_delay_from_zero", "spike_oracle.cpp",
     "delay[s] = 1 + (int32_t)(((uint64_t)r[1] * (uint64_t)dmax) >> 32);",
     "delay[s] = (int32_t)(((uint64_t)r[1] * (uint64_t)dmax) >> 32);",
     [D + "test_rows_recomputed_from_philox", D + "test_degrees_layout_and_ranges"]),
    ("outdegree_input_rows_at_neuron_degree", "spike_oracle.cpp",
     "auto row = [&](int u) { return u <= n ? (int64_t)u * k : (int64_t)n * k + (int64_t)(u - n) * k_in; };",
     "auto row = [&](int u) { return u <= n ? (int64_t)u * k : (int64_t)n * k_in + (int64_t)(u - n) * k_in; };",
     [D + "test_rows_recomputed_from_philox"]),
    ("outdegree_class_boundary", "spike_oracle.cpp",
     "        if (u < ne) {\n          weight[s] = (float)(r[2] >> 8) * 0x1p-24f;",
     "        if (u <= ne) {\n          weight[s] = (float)(r[2] >> 8) * 0x1p-24f;",
     [D + "test_rows_recomputed_from_philox"]),
    ("outdegree_wrong_stream", "spike_oracle.cpp",
     "Philox::draws4(seed, (int64_t)j, (uint32_t)u, 5u, r);",
     "Philox::draws4(seed, (int64_t)j, (uint32_t)u, 3u, r);",
     [D + "test_rows_recomputed_from_philox"]),
    ("wd_taus_swapped", "spike_oracle.cpp",
     "Em[k] = std::exp(-((float)k * dt) / tm);", "Em[k] = std::exp(-((float)k * dt) / tp);",
     [S + "test_wd_pair_protocol"]),
    ("wd_alpha_dropped", "spike_oracle.cpp",
     "upd -= (((R)lam * (R)alp) * w) * tmp;", "upd -= ((R)lam * w) * tmp;",
     [S + "test_wd_pair_protocol"]),
    ("wd_no_clamp", "spike_oracle.cpp",
     "R nw = w + upd;\n        if (nw < (R)0) nw = (R)0;", "R nw = w + upd;",
     [S + "test_wd_depression_clamps_at_zero"]),
    ("vogels_target_halved", "spike_oracle.cpp",
     "(2.0 * (double)r.vg[2] * (double)r.vg[0])", "((double)r.vg[2] * (double)r.vg[0])",
     [S + "test_vogels_protocol"]),
    ("vogels_delay_ignored", "spike_oracle.cpp",
     "bool pre = pre_bit(spre[s], step - sdelay[s]);\n          bool post = get_bit(hist_n, N, W, spost[s], step);\n"
     "          if (pre) xpre += (R)1;",
     "bool pre = pre_bit(spre[s], step);\n          bool post = get_bit(hist_n, N, W, spost[s], step);\n"
     "          if (pre) xpre += (R)1;",
     [S + "test_vogels_protocol"]),
    ("evans_potentiation_sign", "spike_oracle.cpp",
     "if (post) upd += ((R)1 - wo) * C;", "if (post) upd += wo * C;",
     [S + "test_evans_protocol"]),
    ("evans_upper_clamp_dropped", "spike_oracle.cpp",
     "if (w > (R)1) w = (R)1;", "",
     [S + "test_evans_clamp_to_unit_interval"]),
    ("custom_a_star_dropped", "spike_oracle.cpp",
     "if (pre) upd -= (((R)lr * (wo / (R)wmax)) * xpost + (R)lr * (R)ast);",
     "if (pre) upd -= ((R)lr * (wo / (R)wmax)) * xpost;",
     [S + "test_custom_protocol"]),
    ("weightnorm_norm_not_rooted", "spike_oracle.cpp",
     "dv[v] = std::sqrt((R)r.wn_sumsq[v] + u[v]) / std::sqrt((R)r.wn_sumsq[v]);",
     "dv[v] = std::sqrt((R)r.wn_sumsq[v] + u[v]) / (R)r.wn_sumsq[v];",
     [S + "test_weight_normalization_keeps_afferent_norms"]),
    ("bf16_truncates", "spike_oracle.cpp",
     "return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);", "return (uint16_t)(u >> 16);",
     ["tests/test_weights16_pinned.py::test_rounding_equals_torch_and_numpy"]),
    ("glibc_rand_lag", "construction.py",
     "v = (r[-31] + r[-3]) & 0xFFFFFFFF", "v = (r[-31] + r[-4]) & 0xFFFFFFFF",
     ["tests/test_construction_pinned.py::test_glibc_rand_known_answers"]),
    ("fixed_indegree_range", "construction.py",
     "pre_v.append(self.rng.rand() % n_pre)", "pre_v.append(self.rng.rand() % (n_pre - 1))",
     ["tests/test_construction_pinned.py::test_creation_order_matches_reference"]),
    ("silu_slope_wrong", "mlp_oracle.py",
     "return s * (1.0 + z * (1.0 - s))", "return s * (1.0 + z * s)",
     ["tests/test_mlp_oracle_pinned.py::test_gradient_equals_finite_differences"]),
]


def _build_cpp(src_text, out):
    d = os.path.dirname(out)
    os.makedirs(d, exist_ok=True)
    src = os.path.join(d, "spike_oracle.cpp")
    with open(src, "w") as f:
        f.write(src_text)
    r = subprocess.run(["g++", *CXXFLAGS, "-o", out, src], capture_output=True, text=True)
    return r.returncode, r.stderr[-2000:]


def _run_pins(pins, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *pins],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, (r.stdout + r.stderr)[-3000:]


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    base = tmp_path_factory.mktemp("mutants")
    sources = {}
    for f in ("spike_oracle.cpp", "construction.py", "mlp_oracle.py"):
        with open(os.path.join(ORACLE, f)) as fh:
            sources[f] = fh.read()
    jobs = {}
    # the control: the unmutated sources through the same route
    ctrl_lib = os.path.join(base, "control", "lib.so")
    jobs["control"] = ("spike_oracle.cpp", sources["spike_oracle.cpp"], ctrl_lib)
    for name, fname, old, new, _ in MUTANTS:
        n = sources[fname].count(old)
        assert n == 1, f"{name}: the original text occurs {n} times in {fname}"
        jobs[name] = (fname, sources[fname].replace(old, new), os.path.join(base, name))
    out = {}
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        builds = {}
        for name, (fname, text, path) in jobs.items():
            if fname == "spike_oracle.cpp":
                lib = path if path.endswith(".so") else os.path.join(path, "lib.so")
                builds[name] = (ex.submit(_build_cpp, text, lib), {"SPIKE_ORACLE_LIB": lib})
            else:
                os.makedirs(path, exist_ok=True)
                for f in ("construction.py", "mlp_oracle.py", "oracle_py.py", "sweep_replay.py"):
                    shutil.copy(os.path.join(ORACLE, f), os.path.join(path, f))
                with open(os.path.join(path, fname), "w") as fh:
                    fh.write(text)
                builds[name] = (None, {"SPIKE_ORACLE_PYDIR": path})
        runs = {}
        for name, (fut, env) in builds.items():
            if fut is not None:
                rc, err = fut.result()
                assert rc == 0, f"{name}: mutant does not compile\n{err}"
            pins = sorted({p for m in MUTANTS for p in m[4]}) if name == "control" else \
                next(m[4] for m in MUTANTS if m[0] == name)
            runs[name] = ex.submit(_run_pins, pins, env)
        for name, fut in runs.items():
            out[name] = fut.result()
    return out


def test_control_passes_every_pin(results):
    rc, log = results["control"]
    assert rc == 0, log


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_pins(results, name):
    rc, log = results[name]
    # pytest exits 1 when tests ran and failed; anything else (0 = all passed, 2+ =
    # collection or usage errors) does not show that the pins caught the mistake
    assert rc == 1, f"mutant {name} survived (rc {rc})\n{log}"
