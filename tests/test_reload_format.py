"""The reload-file layout this library writes, pinned by the reference
test's own reload files (tests/test-grains/spheres-settling-new-grains/
Reference_results, copied as data by tests/golden/rawdata/
copy_reference_results.py).

Those files come from an older writer (records without the byte count and
without the torsor and contact-memory fields of the current
Particle::write2014_binary, Particle.cpp:1183-1193), so they pin the part
of the record both writers share: type, id, tag, activity, the Transform
(origin, basis by rows, transform type), velocity, quaternion with its real
part first, angular velocity -- the order and conventions a plausible slip
(basis by columns, quaternion (x, y, z, w), a missing field) would break.
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

from pacific_amd import reload as rl

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rawdata")


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_reference_binary_record_prefix():
    buf = open(os.path.join(GOLD, "insertA.result.bin"), "rb").read()
    rec = len(buf) // 20
    assert rec * 20 == len(buf) and rec >= rl.PREFIX.size
    ids, vmax = [], []
    for i in range(20):
        r, _ = rl.decode_prefix(buf, i * rec)
        ids.append(r["id"])
        assert r["type"] == 0 and r["tag"] == 0 and r["active"] == 1
        # a unit quaternion whose rotation is the basis, rows first: the
        # quaternion's real part comes first in the record
        assert abs(np.linalg.norm(r["q"]) - 1) < 1e-12
        assert np.abs(quat_matrix(r["q"]) - r["basis"]).max() < 1e-11
        assert np.abs(quat_matrix(r["q"]).T - r["basis"]).max() > 1e-3 or np.abs(r["q"][:3]).max() < 1e-2
        # settled spheres of radius 5 mm in the test's 30 mm wide box:
        # centres at least a radius from the walls, velocities at rest
        assert np.all(r["x"][:2] >= 0.005 - 1e-6) and np.all(r["x"][:2] <= 0.025 + 1e-6)
        assert r["x"][2] >= 0.005 - 1e-6
        assert np.abs(r["v"]).max() < 1e-11 and np.abs(r["w"]).max() < 1e-11
        vmax.append(np.linalg.norm(r["v"]))
    assert sorted(ids) == list(range(20))
    # the largest speed is the one the run's monitor file reports at the
    # time of this file's last save (1.49 s: insertA in the file table)
    rows = rl.read_rftable(os.path.join(GOLD, "insert_RFTable.txt"))
    t_a = [t for t, f in rows if f.endswith("insertA")][-1]
    mon = {float(l.split()[0]): float(l.split()[1])
           for l in open(os.path.join(GOLD, "insert_VelocityMaxMean.dat")) if l.strip()}
    assert t_a == 1.49
    assert abs(max(vmax) / mon[t_a] - 1) < 1e-6


def test_reference_rftable_alternates():
    rows = rl.read_rftable(os.path.join(GOLD, "insert_RFTable.txt"))
    assert len(rows) == 150
    assert [t for t, _ in rows[:3]] == [0.01, 0.02, 0.03]
    assert [f[-1] for _, f in rows[:4]] == ["A", "B", "A", "B"]


def test_reference_matcontact_layout():
    root = ET.parse(os.path.join(GOLD, "insert_MatContact.xml")).getroot()
    mats = {m.text: int(m.get("value")) for m in root.find("Materials")}
    assert mats == {"matP": 1, "matW": 0}
    (model,) = list(root.find("ContactForceModels"))
    names = [p.get("name") for p in model.find("Parameters")]
    assert names == sorted(names)  # a std::map<string, double>
