"""Output of the time loop through the C ABI: RawData frames, the
_VelocityMaxMean.dat monitor, coordination numbers, and grains_gpu_run with
its output times.  The number format is pinned by the reference test's own
output files (tests/test_rawdata.py); here the files written by the library
are compared byte for byte with that format applied to the state the
library reports, and the loop with outputs against a plain run."""
import os

import numpy as np
import pytest

from pacific_amd import GrainsGPU, rawdata, workloads
from tests.helpers import make_gpu, min_image_pairs
from tests.test_xml import write_case


def _expected_line(t, values):
    return rawdata.format_number(t) + "".join(" " + rawdata.format_number(v) for v in values)


@pytest.mark.gpu
def test_coordination_numbers_of_a_squeezed_lattice(gpu):
    # 4^3 spheres on a lattice squeezed so that every lattice neighbour
    # overlaps, walls touching the outer layer: each sphere touches exactly
    # six components (neighbours, and walls where neighbours are missing)
    wl = workloads.dense_pack(side=4)
    g = make_gpu(wl, record=True)
    g.step(1)
    cn = g.coordination_numbers()
    assert cn.shape == (64,)
    assert np.all(cn == 6), cn
    # the particle part against brute force on the positions
    st = g.state()
    pairs = min_image_pairs(st["x"], wl["radius"])
    npart = np.zeros(64, int)
    for i, _ in pairs:
        npart[i - 1] += 1
    side = np.arange(64)
    a, b, c = side % 4, (side // 4) % 4, side // 16
    faces = sum(((u == 0) | (u == 3)).astype(int) for u in (a, b, c))
    assert np.array_equal(npart, 6 - faces)


@pytest.mark.gpu
def test_coordination_numbers_need_records(gpu):
    wl = workloads.dense_pack(side=3)
    g = make_gpu(wl, record=False)
    g.step(1)
    with pytest.raises(Exception) as e:
        g.coordination_numbers()
    assert "record" in str(e.value)


@pytest.mark.gpu
def test_velocity_max_mean_matches_the_state(gpu):
    wl = workloads.dense_pack(side=5, sigma=0.2)
    g = make_gpu(wl, record=False)
    g.step(20)
    vmax, vmean = g.velocity_max_mean()
    v = g.state()["v"]
    mx, s = 0.0, 0.0
    for u in v:  # id order, Norm as sqrt(x*x + y*y + z*z)
        vit = float(np.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]))
        mx = max(mx, vit)
        s += vit
    assert vmax == mx and vmean == s / len(v)


@pytest.mark.gpu
def test_rawdata_text_frames(gpu, tmp_path):
    wl = workloads.dense_pack(side=4, sigma=0.1)
    g = make_gpu(wl, record=True)
    root = str(tmp_path / "run")
    g.write_rawdata(root, truncate=True)
    g.step(10)
    cn = g.coordination_numbers()
    g.write_rawdata(root, coord=cn)
    st, t = g.state(), g.time
    cols = {"position": st["x"], "translational_velocity": st["v"], "angular_velocity": st["w"]}
    for name, arr in cols.items():
        for a, ax in enumerate("xyz"):
            lines = open(f"{root}_{name}_{ax}.dat").read().split("\n")
            assert len(lines) == 3 and lines[2] == ""  # two frames
            assert lines[1] == _expected_line(t, arr[:, a])
    coord = open(f"{root}_coordinationNumber.dat").read().split("\n")
    assert coord[0] == " 0" * 64  # no coordination numbers given: zeros
    assert coord[1] == "".join(f" {c}" for c in cn)
    assert open(f"{root}_particleType.dat").read() == "0 " * 64 + "\n"
    r = rawdata.read_rawdata(root, 64)
    assert np.allclose(r["time"], [0.0, t], rtol=0, atol=1e-15)
    assert np.allclose(r["position_x"][1], st["x"][:, 0], rtol=1e-6, atol=0)


@pytest.mark.gpu
def test_rawdata_binary_frames(gpu, tmp_path):
    wl = workloads.dense_pack(side=3, sigma=0.1)
    g = make_gpu(wl, record=True)
    root = str(tmp_path / "bin")
    g.step(3)
    g.write_rawdata(root, binary=True, truncate=True, coord=g.coordination_numbers())
    s1, t1 = g.state(), g.time
    g.step(4)
    g.write_rawdata(root, binary=True, coord=g.coordination_numbers())
    s2, t2 = g.state(), g.time
    r = rawdata.read_rawdata(root, 27, binary=True)
    assert np.array_equal(r["time"], [t1, t2])
    for k, (key, a) in enumerate([("position", "x"), ("translational_velocity", "v"), ("angular_velocity", "w")]):
        for c, ax in enumerate("xyz"):
            got = r[f"{key}_{ax}"]
            assert np.array_equal(got[0], s1[a][:, c]) and np.array_equal(got[1], s2[a][:, c])
    # the time twice at the head of each z angular velocity frame
    raw = np.fromfile(f"{root}_angular_velocity_z.dat", dtype=np.float64).reshape(2, -1)
    assert raw[0, 0] == raw[0, 1] == t1
    assert os.path.getsize(f"{root}_coordinationNumber.dat") == 2 * 27 * 4
    assert np.array_equal(r["type"], np.zeros(27, np.int32))


OUTPUTS = """    <RestartFile Name="{root}" WritingMode="Text"/>
    <TimeSave Start="0." End="0.005" Every="0.001"/>
    <PostProcessing>
      <Writers>
        <RawData Name="{root}"/>
      </Writers>
    </PostProcessing>
  </Simulation>"""


@pytest.mark.gpu
def test_run_with_outputs_equals_plain_steps(gpu, tmp_path):
    root = str(tmp_path / "out" / "insert")
    os.makedirs(os.path.dirname(root))
    path, _ = write_case(tmp_path, n=200, **{'End="0.1"': 'End="0.005"',
                                             "  </Simulation>": OUTPUTS.format(root=root)})
    g = GrainsGPU.from_xml(path)
    g.run()
    # the same system stepped without outputs, with records in its last step
    h = GrainsGPU.from_xml(path)
    h.step(499)
    h.record_pairs(True)
    h.step(1)
    assert g.time == h.time
    sg, sh = g.state(), h.state()
    for k in ("x", "v", "w", "q"):
        assert np.array_equal(sg[k], sh[k]), k
    # monitor: the initial line and one per output time (the last at the end)
    lines = open(root + "_VelocityMaxMean.dat").read().splitlines()
    assert len(lines) == 6
    assert lines[0] == "0\t0\t0"
    assert lines[-1] == rawdata.velocity_line(h.time, *h.velocity_max_mean())
    times = [float(l.split("\t")[0]) for l in lines]
    assert np.allclose(times, [0, 0.001, 0.002, 0.003, 0.004, 0.005], rtol=0, atol=1e-12)
    # RawData: six frames, the last one the final state and its contacts
    r = rawdata.read_rawdata(root, 200)
    assert r["position_z"].shape == (6, 200)
    last = open(root + "_position_z.dat").read().splitlines()[-1]
    assert last == _expected_line(h.time, sh["x"][:, 2])
    assert np.array_equal(r["coordination"][-1], h.coordination_numbers())
    assert np.array_equal(r["coordination"][0], np.zeros(200))


@pytest.mark.gpu
def test_run_ends_with_a_monitor_line_off_the_output_times(gpu, tmp_path):
    # outputs at 0.002 and 0.004, the run ends at 0.005: the end of the run
    # adds its own monitor line (Grains.cpp:199-211) but no RawData frame
    root = str(tmp_path / "o")
    path, _ = write_case(tmp_path, n=100, **{'End="0.1"': 'End="0.005"',
                                             "  </Simulation>": OUTPUTS.format(root=root).replace(
                                                 'Every="0.001"', 'Every="0.002"')})
    g = GrainsGPU.from_xml(path)
    g.run()
    lines = open(root + "_VelocityMaxMean.dat").read().splitlines()
    times = [float(l.split("\t")[0]) for l in lines]
    assert np.allclose(times, [0, 0.002, 0.004, 0.005], rtol=0, atol=1e-12)
    assert lines[-1] == rawdata.velocity_line(g.time, *g.velocity_max_mean())
    assert rawdata.read_rawdata(root, 100)["position_x"].shape == (3, 100)
