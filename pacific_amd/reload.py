"""Readers of the __Format2024__ reload files (Grains::saveReload,
Grains.cpp:2543-2584; AllComponents::write, AllComponents.cpp:1299-1415),
as grains_gpu_save_reload writes them: the main file, the particle records
(Particle::write2014 / write2014_binary, Particle.cpp:1167-1193, with the
SecondOrderLeapFrog torsor, SecondOrderLeapFrog.cpp:98-150, and the contact
memory, Component.cpp:776-835) and _RFTable.txt."""
import struct

import numpy as np

# one particle record of write2014_binary, before the integrator's fields
# and the contact memory: GeomType, id, tag (int), activity (unsigned),
# Transform (origin 3 doubles, basis 9 doubles by rows, type unsigned),
# velocity (3 doubles), quaternion (w, then x, y, z), angular velocity
PREFIX = struct.Struct("<3iI3d9dI3d4d3d")
SCHEME_DOUBLES = {"leapfrog": 6, "explicit": 0, "adams_bashforth2": 12}
MEMORY_ENTRY = struct.Struct("<4i9d")


def decode_prefix(buf, offset=0):
    f = PREFIX.unpack_from(buf, offset)
    return {"type": f[0], "id": f[1], "tag": f[2], "active": f[3], "x": np.array(f[4:7]),
            "basis": np.array(f[7:16]).reshape(3, 3), "transform_type": f[16], "v": np.array(f[17:20]),
            "q": np.array([f[21], f[22], f[23], f[20]]), "w": np.array(f[24:27])}, offset + PREFIX.size


def read_main(path):
    """Time, particle-file mode (Text / Hybrid) and count, particle types."""
    tok = open(path).read().split()
    assert tok[0] == "__Format2024__"
    out = {"time": float(tok[2]), "types": int(tok[tok.index("NumberOfParticleTypes") + 1])}
    k = tok.index("NbFiles")
    out["nfiles"], out["mode"], out["n"] = int(tok[k + 1]), tok[k + 2], int(tok[k + 3])
    assert tok[-1] == "#EndTime"
    return out


def read_particles_bin(path, scheme="leapfrog"):
    """Records of a _Particles.bin file: each preceded by its byte count."""
    buf = open(path, "rb").read()
    recs, off = [], 0
    extra = SCHEME_DOUBLES[scheme]
    while off < len(buf):
        (nb,) = struct.unpack_from("<Q", buf, off)
        off += 8
        start = off
        r, off = decode_prefix(buf, off)
        r["scheme"] = np.array(struct.unpack_from(f"<{extra}d", buf, off))
        off += 8 * extra
        (ms,) = struct.unpack_from("<i", buf, off)
        off += 4
        r["memory"] = []
        for _ in range(ms):
            e = MEMORY_ENTRY.unpack_from(buf, off)
            off += MEMORY_ENTRY.size
            r["memory"].append((e[1], np.array(e[4:7]), np.array(e[7:10]), np.array(e[10:13])))
        r["nbytes"] = nb
        assert off - start == nb, (off - start, nb)
        recs.append(r)
    return recs


def read_particles_text(path, scheme="leapfrog"):
    """Records of a text _Particles file, one line each."""
    recs = []
    extra = SCHEME_DOUBLES[scheme]
    for line in open(path):
        t = line.split()
        if not t:
            continue
        v = [float(a) for a in t[4:17]]
        r = {"type": int(t[0]), "id": int(t[1]), "tag": int(t[2]), "active": int(t[3]), "x": np.array(v[0:3]),
             "basis": np.array(v[3:12]).reshape(3, 3), "transform_type": int(t[16])}
        f = [float(a) for a in t[17:27]]
        r["v"], r["q"], r["w"] = np.array(f[0:3]), np.array([f[4], f[5], f[6], f[3]]), np.array(f[7:10])
        k = 27
        r["scheme"] = np.array([float(a) for a in t[k:k + extra]])
        k += extra
        ms = int(t[k])
        k += 1
        r["memory"] = []
        for _ in range(ms):
            e = t[k:k + 13]
            k += 13
            r["memory"].append((int(e[1]), np.array(e[4:7], float), np.array(e[7:10], float),
                                np.array(e[10:13], float)))
        recs.append(r)
    return recs


def read_rftable(path):
    """(time, reload file) lines of <root>_RFTable.txt."""
    out = []
    for line in open(path):
        t = line.split()
        if len(t) == 2:
            out.append((float(t[0]), t[1]))
    return out
