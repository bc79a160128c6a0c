"""The standalone program (pacific_amd/grains_gpu <file.xml>, the
counterpart of `grains <file.xml>`, GRAINS_V1/Main/src/main.cpp:9-76) run as
a child process on an input written by workloads.write_input: exit status,
the monitor and RawData files of its output times, and the final frame
against the same input stepped through the library."""
import os
import subprocess
import sys

import numpy as np
import pytest

from pacific_amd import GrainsGPU, rawdata, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pacific_amd", "grains_gpu")


@pytest.mark.gpu
def test_cli_runs_the_input_with_its_outputs(gpu, tmp_path):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} missing: build it with `python -m pacific_amd.build`")
    path = workloads.write_input(workloads.settling_box(200), str(tmp_path), t_end=0.004, save_every=0.001)
    print("cli input", path, flush=True)
    res = subprocess.run([CLI, path], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "200 particles" in res.stdout
    root = os.path.join(str(tmp_path), "insert")
    lines = open(root + "_VelocityMaxMean.dat").read().splitlines()
    assert len(lines) == 5 and lines[0] == "0\t0\t0"
    r = rawdata.read_rawdata(root, 200)
    assert np.allclose(r["time"], [0, 0.001, 0.002, 0.003, 0.004], rtol=0, atol=1e-12)
    # the last frame is the state of 400 library steps, at 6 digits
    g = GrainsGPU.from_xml(path)
    g.step(400)
    st = g.state()
    last = open(root + "_position_z.dat").read().splitlines()[-1].split()[1:]
    assert last == [rawdata.format_number(v) for v in st["x"][:, 2]]
    assert lines[-1] == rawdata.velocity_line(g.time, *g.velocity_max_mean())


@pytest.mark.gpu
def test_cli_reports_bad_input(gpu, tmp_path):
    bad = tmp_path / "bad.xml"
    bad.write_text("<Grains3D Type=\"Standard\"><Construction></Construction></Grains3D>")
    res = subprocess.run([CLI, str(bad)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert res.returncode == 1 and res.stderr.strip()
