"""Subdivided closed boxes for the simulator tests, and an OBJ writer."""
import numpy as np


def box_mesh(size, n=4, center=(0.0, 0.0, 0.0)):
    """A box of the given edge lengths, each face split into n x n squares of two
    triangles, vertices shared along edges, faces oriented outward."""
    size = np.asarray(size, dtype=np.float64)
    index = {}
    V = []

    def vid(i, j, k):
        key = (i, j, k)
        if key not in index:
            index[key] = len(V)
            V.append((np.array([i, j, k], dtype=np.float64) / n - 0.5) * size + np.asarray(center))
        return index[key]

    F = []
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for side in (0, n):
            for a in range(n):
                for b in range(n):
                    def corner(da, db):
                        c = [0, 0, 0]
                        c[axis], c[u], c[v] = side, a + da, b + db
                        return vid(*c)
                    q = [corner(0, 0), corner(1, 0), corner(1, 1), corner(0, 1)]
                    # (u, v, axis) is right-handed: counter-clockwise in (u, v) faces +axis
                    if side == n:
                        F += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
                    else:
                        F += [[q[0], q[2], q[1]], [q[0], q[3], q[2]]]
    return np.array(V), np.array(F, dtype=np.int64)


def write_obj(path, V, F):
    with open(path, "w") as f:
        for v in V:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        for t in F:
            f.write("f %d %d %d\n" % tuple(np.asarray(t) + 1))
