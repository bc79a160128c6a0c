"""Insertion on the device: systems built from insertion windows, and
insertion over time (<Mode Type="OverTime">, Grains.cpp:322-327), where the
system grows by one particle at a step.  A particle inserted over time far
from the others leaves their motion bit for bit as it is without it, and
falls freely from rest (LeapFrog: z_m = z_0 + g dt^2 m^2 / 2)."""
import numpy as np
import pytest

from pacific_amd import GrainsGPU, workloads
from tests.test_insertion import insertion_xml, restate_spheres
from tests.test_xml import write_case

pytestmark = pytest.mark.gpu


def test_window_insertion_initialises_and_runs(tmp_path, gpu):
    windows = [("Box", (0.003, 0.003, 0.003), (0.017, 0.017, 0.025))]
    path = insertion_xml(tmp_path, 150, windows, seed=17)
    g = GrainsGPU.from_xml(path)
    assert g.n == 150
    x, v, _, _ = restate_spheres(150, windows, 17)
    st = g.state()
    assert np.array_equal(st["x"], x) and np.array_equal(st["v"], v)
    g.step(300)
    g.synchronize()
    assert np.all(np.isfinite(g.state()["x"]))


def over_time_case(tmp_path, name, pos, every=1, end="0.1"):
    d = tmp_path / name
    d.mkdir()
    path, _ = write_case(d, n=len(pos))
    np.savetxt(d / "positions.result", pos, fmt="%.17g")
    text = open(path).read().replace('<Mode Type="InitialTime"/>',
                                     f'<Mode Type="OverTime"/><Frequency TryEvery="{every}"/>')
    text = text.replace('<TimeInterval Start="0." End="0.1"/>', f'<TimeInterval Start="0." End="{end}"/>')
    open(path, "w").write(text)
    return path


def test_over_time_growth(tmp_path, gpu):
    # the first lattice layer lowered to 28 um above the floor at the
    # lowest: the first spheres land within the run
    pos = workloads.settling_box(30)["x"] - np.array([0., 0., 0.68e-3])
    far = np.array([[0.01, 0.01, 0.05]])
    a = GrainsGPU.from_xml(over_time_case(tmp_path, "a", np.vstack([pos, far])))
    b = GrainsGPU.from_xml(over_time_case(tmp_path, "b", pos))
    assert a.n == 1 and a.pending_insertions() == 30
    M = 400
    for _ in range(M):
        for g in (a, b):
            g.try_insert()
            g.step(1)
    assert a.n == 31 and b.n == 30 and a.pending_insertions() == 0
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "w", "q"):
        assert np.array_equal(sa[k][:30], sb[k][:30]), k
    # the bed has contacts by now
    b.record_pairs(True)
    b.compute_forces()
    assert len(b.pairs()) > 0
    # the far particle went in before step 31 and fell freely since
    m = M - 30
    dt, gz = 1e-5, -9.81
    dz = sa["x"][30, 2] - far[0, 2]
    assert abs(dz - gz * dt * dt * m * m / 2) < 1e-12 * abs(dz)
    assert abs(sa["v"][30, 2] - gz * dt * m) < 1e-12 * abs(gz * dt * m)


def test_over_time_in_the_time_loop(tmp_path, gpu):
    """grains_gpu_run makes the insertion call before every step: one
    particle every TryEvery steps while the list lasts."""
    wl = workloads.settling_box(40)
    g = GrainsGPU.from_xml(over_time_case(tmp_path, "c", wl["x"], every=5, end="1.e-3"))
    g.run()
    # 100 steps: attempts before steps 1, 6, 11, ... 96
    assert g.n == 20 and g.pending_insertions() == 20


def test_over_time_growth_polyhedra(tmp_path, gpu):
    """The same with tetrahedra at random orientations: the bases of the
    bodies already in carry over the growth bit for bit too."""
    from pacific_amd import shapes
    pos = workloads.settling_box(20)["x"] - np.array([0., 0., 0.68e-3])
    far = np.array([[0.01, 0.01, 0.05]])
    # one system after the other: both draw their rotations from the C
    # library's one random stream, reseeded as each file is read
    systems = []
    for name, p in (("pa", np.vstack([pos, far])), ("pb", pos)):
        path = over_time_case(tmp_path, name, p)
        v, f = shapes.tetrahedron(1.2e-3)
        shapes.write_insert(str(tmp_path / name / "tet.insert"), v, f)
        text = open(path).read().replace('<Sphere Radius="0.001"/>', '<Polyhedron Name="tet.insert"/>')
        text = text.replace('<Mode Type="OverTime"/>', '<Mode Type="OverTime"/><InitialAngularPosition Type="Random"/>')
        text = text.replace('<RandomGeneratorSeed Type="Default"/>', '<RandomGeneratorSeed Type="UserDefined" Value="4"/>')
        open(path, "w").write(text)
        g = GrainsGPU.from_xml(path)
        for _ in range(300):
            g.try_insert()
            g.step(1)
        systems.append(g)
    a, b = systems
    assert a.n == 21 and b.n == 20
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "w", "q"):
        assert np.array_equal(sa[k][:20], sb[k][:20]), k
    assert np.array_equal(a.basis()[:20], b.basis()[:20])
