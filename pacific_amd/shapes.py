"""Convex particle shapes: Grains3D polyhedron files and generators.

A polyhedron file (the ``<Polyhedron Name="...">`` format of Grains3D,
read by Polyhedron::create, GRAINS_V1/Grains/Geometry/src/Polyhedron.cpp:
100-160 and readFaces :363-405) holds the dimension (3), the number of
vertices, one vertex per line, the number of faces and one face per line as
its vertex indices.

Generators for the shapes of the polyhedra workload: cube, regular
tetrahedron and regular icosahedron of a given circumscribed radius, centred
on their centre of mass, and uniformly distributed random rotations.
"""
import numpy as np


def read_insert(path):
    """(vertices [nv, 3], faces list of index lists) of a polyhedron file"""
    with open(path) as f:
        lines = [l.split() for l in f if l.strip()]
    toks, li = [], 0
    while len(toks) < 2:
        toks += lines[li]
        li += 1
    nv = int(toks[1])
    need = 2 + 3 * nv + 1
    while len(toks) < need:
        toks += lines[li]
        li += 1
    verts = np.array([float(t) for t in toks[2:2 + 3 * nv]]).reshape(nv, 3)
    nf = int(toks[2 + 3 * nv])
    faces = [[int(t) for t in lines[li + j]] for j in range(nf)]
    return verts, faces


def write_insert(path, verts, faces):
    with open(path, "w") as f:
        f.write("3\n%d\n" % len(verts))
        for v in verts:
            f.write("%.17g %.17g %.17g\n" % tuple(v))
        f.write("\n%d\n" % len(faces))
        for fc in faces:
            f.write(" ".join(str(i) for i in fc) + "\n")


def cube(half):
    """cube of half edge `half` as a polyhedron (vertex order of Box corners)"""
    e = float(half)
    verts = np.array([[-e, -e, -e], [e, -e, -e], [e, -e, e], [-e, -e, e], [-e, e, -e], [e, e, -e], [e, e, e],
                      [-e, e, e]])
    faces = [[4, 5, 6, 7], [5, 1, 2, 6], [1, 0, 3, 2], [0, 4, 7, 3], [4, 0, 1, 5], [3, 7, 6, 2]]
    return verts, faces


def tetrahedron(circumradius):
    """regular tetrahedron, centroid at the origin"""
    v = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]])
    v *= circumradius / np.sqrt(3.0)
    faces = [[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]]
    return v, outward(v, faces)


def icosahedron(circumradius):
    """regular icosahedron, centre at the origin"""
    p = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    v *= circumradius / np.sqrt(1.0 + p * p)
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
             [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
             [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    return v, outward(v, faces)


def outward(verts, faces):
    """faces re-ordered counter-clockwise seen from outside (the polyhedron
    is centred on the origin), so that every edge is walked once in each
    direction -- what the vertex adjacency of Polyhedron::BuildPolyhedron
    (Polyhedron.cpp:324-359) needs to list every neighbour"""
    out = []
    for f in faces:
        a, b, c = (np.asarray(verts[i]) for i in f[:3])
        if np.dot(np.cross(b - a, c - a), a + b + c) < 0:
            f = [f[0]] + list(reversed(f[1:]))
        out.append(list(f))
    return out


def quat_to_matrix(q):
    """rotation matrix (row-major) of unit quaternions q = [x, y, z, w]"""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def random_rotations(n, seed):
    """n rotation matrices from uniformly distributed unit quaternions"""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return quat_to_matrix(q)
