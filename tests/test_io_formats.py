"""spike_amd.io against the reference's file layouts (SpikingActivityMonitor.cpp:123-153,
Synapses.cpp:466-614, SpikingSynapses.cpp:163-189): byte layouts written by hand here, read
back, and written files read as raw streams."""
import numpy as np

from spike_amd import io


def test_spikes_binary_layout_and_round_trip(tmp_path):
    ids = np.array([3, 0, 7], np.int32)
    t = np.array([0.1, 0.1003, 2.5], np.float32)
    # the layout the reference writes: raw int32 / float32 streams
    (tmp_path / "x_SpikeIDs.bin").write_bytes(ids.tobytes())
    (tmp_path / "x_SpikeTimes.bin").write_bytes(t.tobytes())
    gi, gt = io.read_spikes(str(tmp_path), "x_")
    assert gi.tolist() == ids.tolist() and gt.tobytes() == t.tobytes()
    io.write_spikes(str(tmp_path), ids, t, prefix="y_")
    assert (tmp_path / "y_SpikeIDs.bin").read_bytes() == ids.tobytes()
    assert (tmp_path / "y_SpikeTimes.bin").read_bytes() == t.tobytes()


def test_spikes_text(tmp_path):
    (tmp_path / "SpikeIDs.txt").write_text("0\n3\n")
    (tmp_path / "SpikeTimes.txt").write_text("0.1\n0.2003\n")
    ids, t = io.read_spikes(str(tmp_path))
    assert ids.tolist() == [0, 3] and np.allclose(t, [0.1, 0.2003])
    io.write_spikes(str(tmp_path), [5], [np.float32(0.123456789)], prefix="o_", binary=False)
    assert (tmp_path / "o_SpikeTimes.txt").read_text() == "0.123457\n"


def test_connectivity_and_weights(tmp_path):
    pre = np.array([-1, 0, 2], np.int32)
    post = np.array([0, 1, 1], np.int32)
    w = np.array([0.5, -1.25, 3.0], np.float32)
    dl = np.array([1, 15, 4], np.int32)
    io.write_connectivity(str(tmp_path), pre, post, w, dl, prefix="c_")
    assert (tmp_path / "c_PresynapticIDs.bin").read_bytes() == pre.tobytes()
    assert (tmp_path / "c_SynapticDelays.bin").read_bytes() == dl.tobytes()
    c = io.read_connectivity(str(tmp_path), "c_")
    assert c["pre"].tolist() == pre.tolist() and c["weight"].tolist() == w.tolist()
    assert io.read_weights(str(tmp_path / "c_SynapticWeights.bin")).tolist() == w.tolist()
    (tmp_path / "w.txt").write_text("0.5\n-1.25 3\n")
    assert io.read_weights(str(tmp_path / "w.txt")).tolist() == [0.5, -1.25, 3.0]
