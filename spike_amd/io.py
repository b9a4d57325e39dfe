"""Spike's on-disk formats, for the Python API.

* Rasters (SpikingActivityMonitor::save_spikes_as_binary / _as_txt,
  SpikingActivityMonitor.cpp:123-153): <prefix>SpikeIDs.bin (int32) and
  <prefix>SpikeTimes.bin (float32), or .txt with one value per line.
* Connectivity (Synapses::save_connectivity_as_binary, Synapses.cpp:466-519, and
  SpikingSynapses::save_connectivity_as_binary, SpikingSynapses.cpp:163-189):
  <prefix>PresynapticIDs.bin, PostsynapticIDs.bin (int32; inputs as -i-1),
  SynapticWeights.bin (float32) and SynapticDelays.bin (int32, in timesteps), one entry per
  synapse in creation order.
* Weights to load (Synapses::load_weights_from_binary / _from_txt, Synapses.cpp:585-614):
  float32 values, or text, in creation order.

The txt connectivity files of the reference are not read: its PresynapticIDs.txt is written
without separators when no group is given (Synapses.cpp:452-455), so the ids cannot be
recovered from it.
"""
from __future__ import annotations

import os
from typing import Dict, Tuple

import numpy as np


def _path(d: str, prefix: str, name: str) -> str:
    return os.path.join(d, prefix + name)


def write_spikes(d: str, ids, times, prefix: str = "", binary: bool = True) -> None:
    ids = np.ascontiguousarray(ids, np.int32)
    times = np.ascontiguousarray(times, np.float32)
    if len(ids) != len(times):
        raise ValueError("ids and times differ in length")
    if binary:
        ids.tofile(_path(d, prefix, "SpikeIDs.bin"))
        times.tofile(_path(d, prefix, "SpikeTimes.bin"))
    else:
        # operator<< of float prints 6 significant digits, as the reference's text files do
        with open(_path(d, prefix, "SpikeIDs.txt"), "w") as f:
            f.writelines(f"{int(i)}\n" for i in ids)
        with open(_path(d, prefix, "SpikeTimes.txt"), "w") as f:
            f.writelines(f"{float(t):.6g}\n" for t in times)


def read_spikes(d: str, prefix: str = "") -> Tuple[np.ndarray, np.ndarray]:
    """(ids int32, times float32) from the binary files, else from the text files."""
    b = _path(d, prefix, "SpikeIDs.bin")
    if os.path.exists(b):
        return (np.fromfile(b, np.int32), np.fromfile(_path(d, prefix, "SpikeTimes.bin"), np.float32))
    ids = np.loadtxt(_path(d, prefix, "SpikeIDs.txt"), dtype=np.int64, ndmin=1).astype(np.int32)
    times = np.loadtxt(_path(d, prefix, "SpikeTimes.txt"), dtype=np.float64, ndmin=1).astype(np.float32)
    return ids, times


def write_connectivity(d: str, pre, post, weight, delay, prefix: str = "") -> None:
    for name, a, t in (("PresynapticIDs", pre, np.int32), ("PostsynapticIDs", post, np.int32),
                       ("SynapticWeights", weight, np.float32), ("SynapticDelays", delay, np.int32)):
        np.ascontiguousarray(a, t).tofile(_path(d, prefix, name + ".bin"))


def read_connectivity(d: str, prefix: str = "") -> Dict[str, np.ndarray]:
    out = dict(pre=np.fromfile(_path(d, prefix, "PresynapticIDs.bin"), np.int32),
               post=np.fromfile(_path(d, prefix, "PostsynapticIDs.bin"), np.int32),
               weight=np.fromfile(_path(d, prefix, "SynapticWeights.bin"), np.float32))
    dl = _path(d, prefix, "SynapticDelays.bin")
    if os.path.exists(dl):
        out["delay"] = np.fromfile(dl, np.int32)
    n = len(out["pre"])
    if any(len(v) != n for v in out.values()):
        raise ValueError("connectivity files differ in length")
    return out


def read_weights(path: str) -> np.ndarray:
    """Weights as load_weights_from_binary (float32 stream) or _from_txt (whitespace-
    separated values) reads them."""
    if path.endswith(".bin"):
        return np.fromfile(path, np.float32)
    with open(path) as f:
        return np.array(f.read().split(), dtype=np.float32)
