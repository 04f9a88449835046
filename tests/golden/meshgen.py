"""Seeded hexahedral box meshes as CGNS-style element lists (test input data).

The reference reads one unstructured CGNS zone (src/setup/cell_input.f90:36-93):
3D cells numbered first, then one QUAD_4 section per boundary.  This module
builds the same kind of input for a box: HEXA_8 cells in lexicographic order
(i fastest), CGNS node order, then the sections bottom, south, east, north,
west, top (z-, y-, x+, y+, x-, z+).  Interior vertices may be jittered by
uniform(-a*h, a*h) from a splitmix64 stream so that the C++ generator of the
package can produce bit-identical coordinates.

This file is test infrastructure: it writes the mesh files the reference
driver (oracle/refdrv) reads and feeds the oracle.
"""
import numpy as np

HEXA_8 = 17
QUAD_4 = 7
BC_WALL, BC_LID, BC_SYMMETRY, BC_INLET, BC_OUTLET = 0, 1, 2, 3, 4
SECTION_NAMES = ["bottom", "south", "east", "north", "west", "top"]

_M64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n draws of splitmix64 (Steele, Lea, Flood 2014) as uint64."""
    out = np.empty(n, dtype=np.uint64)
    s = seed & _M64
    for k in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[k] = z ^ (z >> 31)
    return out


def box_mesh(nx, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=20261202):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    nvx = (nx + 1) * (ny + 1) * (nz + 1)
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    X, Y, Z = np.meshgrid(lx * i / nx, ly * j / ny, lz * k / nz, indexing="ij")
    # flatten with i fastest
    x = X.transpose(2, 1, 0).reshape(-1).copy()
    y = Y.transpose(2, 1, 0).reshape(-1).copy()
    z = Z.transpose(2, 1, 0).reshape(-1).copy()
    if jitter > 0.0:
        ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz))
        interior = interior.transpose(2, 1, 0).reshape(-1)
        idx = np.nonzero(interior)[0]
        r = splitmix64(seed, 3 * len(idx)).reshape(-1, 3)
        u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        hx, hy, hz = lx / nx, ly / ny, lz / nz
        x[idx] = x[idx] + (2.0 * u[:, 0] - 1.0) * (jitter * hx)
        y[idx] = y[idx] + (2.0 * u[:, 1] - 1.0) * (jitter * hy)
        z[idx] = z[idx] + (2.0 * u[:, 2] - 1.0) * (jitter * hz)

    def vid(a, b, c):
        return 1 + a + (nx + 1) * (b + (ny + 1) * c)

    ne = nx * ny * nz
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).reshape(-1)
    cj = cj.transpose(2, 1, 0).reshape(-1)
    ck = ck.transpose(2, 1, 0).reshape(-1)
    cells = np.stack([vid(ci, cj, ck), vid(ci + 1, cj, ck), vid(ci + 1, cj + 1, ck), vid(ci, cj + 1, ck),
                      vid(ci, cj, ck + 1), vid(ci + 1, cj, ck + 1), vid(ci + 1, cj + 1, ck + 1),
                      vid(ci, cj + 1, ck + 1)], axis=1).astype(np.int32)

    def grid(n1, n2):
        a, b = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
        return a.T.reshape(-1), b.T.reshape(-1)  # first index fastest

    quads = []
    a, b = grid(nx, ny)   # bottom z=0
    quads.append(np.stack([vid(a, b, 0), vid(a, b + 1, 0), vid(a + 1, b + 1, 0), vid(a + 1, b, 0)], 1))
    a, b = grid(nx, nz)   # south y=0
    quads.append(np.stack([vid(a, 0, b), vid(a + 1, 0, b), vid(a + 1, 0, b + 1), vid(a, 0, b + 1)], 1))
    a, b = grid(ny, nz)   # east x=1
    quads.append(np.stack([vid(nx, a, b), vid(nx, a + 1, b), vid(nx, a + 1, b + 1), vid(nx, a, b + 1)], 1))
    a, b = grid(nx, nz)   # north y=1
    quads.append(np.stack([vid(a + 1, ny, b), vid(a, ny, b), vid(a, ny, b + 1), vid(a + 1, ny, b + 1)], 1))
    a, b = grid(ny, nz)   # west x=0
    quads.append(np.stack([vid(0, a + 1, b), vid(0, a, b), vid(0, a, b + 1), vid(0, a + 1, b + 1)], 1))
    a, b = grid(nx, ny)   # top z=1
    quads.append(np.stack([vid(a, b, nz), vid(a + 1, b, nz), vid(a + 1, b + 1, nz), vid(a, b + 1, nz)], 1))

    nbf = sum(len(q) for q in quads)
    nelem = ne + nbf
    e2vx = np.zeros((nelem, 8), dtype=np.int32)
    e2vx[:ne] = cells
    etype = [HEXA_8]
    esec = [(1, ne)]
    start = ne + 1
    for q in quads:
        e2vx[start - 1:start - 1 + len(q), :4] = q
        etype.append(QUAD_4)
        esec.append((start, start + len(q) - 1))
        start += len(q)
    return dict(nvx=nvx, x=x, y=y, z=z, e2vx=e2vx.reshape(-1), etype=np.array(etype, np.int32),
                esec=np.array(esec, np.int32).reshape(-1), names=["cells"] + SECTION_NAMES,
                ne=ne, nbf=nbf)


def cell_graph(mesh):
    """Interior-face adjacency of the cells (sets of 0-based cell ids): two
    hexahedra are neighbours when they share the four vertices of a face
    (CGNS HEXA_8 face order, src/modules/mod_util.f90:78-85)."""
    ne = mesh["ne"]
    ev = mesh["e2vx"].reshape(-1, 8)[:ne]
    faces = ((0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (4, 5, 6, 7))
    owner = {}
    adj = [set() for _ in range(ne)]
    for e in range(ne):
        for f in faces:
            key = tuple(sorted(int(ev[e, q]) for q in f))
            o = owner.get(key)
            if o is None:
                owner[key] = e
            else:
                adj[e].add(o)
                adj[o].add(e)
    return adj


def rcm(adj):
    """Reverse Cuthill-McKee: each component from its lowest-numbered cell of
    least degree, a breadth-first walk appending each cell's unvisited
    neighbours by increasing (degree, number), the sequence reversed."""
    n = len(adj)
    deg = [len(a) for a in adj]
    seen = [False] * n
    order = []
    for s in sorted(range(n), key=lambda i: (deg[i], i)):
        if seen[s]:
            continue
        seen[s] = True
        head = len(order)
        order.append(s)
        while head < len(order):
            v = order[head]
            head += 1
            for j in sorted((j for j in adj[v] if not seen[j]), key=lambda j: (deg[j], j)):
                seen[j] = True
                order.append(j)
    return order[::-1]


def renumber(mesh, seed):
    """The cell renumbering of the distorted-channel input: a seeded
    permutation (Fisher-Yates from the last position, position i swapping
    with splitmix64 % (i + 1)), then RCM of the shuffled mesh's cells.
    Boundary elements keep their order.  Returns a new mesh dict."""
    ne = mesh["ne"]
    ev = mesh["e2vx"].reshape(-1, 8).copy()
    frm = list(range(ne))
    r = splitmix64(seed, max(ne - 1, 0))
    for t, i in enumerate(range(ne - 1, 0, -1)):
        j = int(r[t] % np.uint64(i + 1))
        frm[i], frm[j] = frm[j], frm[i]
    ev[:ne] = ev[:ne][frm]
    out = dict(mesh)
    out["e2vx"] = ev.reshape(-1).copy()
    order = rcm(cell_graph(out))
    ev2 = ev.copy()
    ev2[:ne] = ev[:ne][order]
    out["e2vx"] = ev2.reshape(-1)
    return out


def cavity_bc(mesh, symmetry=()):
    """Lid-driven cavity binding, src/equations/mod_uvwp.f90:73-78; the
    sections named in symmetry take the symmetry setter (:517-546) instead."""
    kinds = [-1]
    for name in mesh["names"][1:]:
        kinds.append(BC_SYMMETRY if name in symmetry else BC_LID if name == "top" else BC_WALL)
    return np.array(kinds, np.int32)


def channel_bc(mesh):
    """Channel along x (an extension): inlet on west (x-), outlet on east
    (x+), no-slip walls on the other sections."""
    kinds = [-1]
    for name in mesh["names"][1:]:
        kinds.append(BC_INLET if name == "west" else BC_OUTLET if name == "east" else BC_WALL)
    return np.array(kinds, np.int32)


def write_refsim_mesh(mesh, path):
    """Binary mesh file read by oracle/refdrv/drv_cell.f90."""
    with open(path, "wb") as f:
        f.write(np.array([mesh["nvx"], len(mesh["etype"])], np.int32).tobytes())
        for s, name in enumerate(mesh["names"]):
            f.write(name.encode().ljust(32, b" "))
            f.write(np.array([mesh["etype"][s], mesh["esec"][2 * s], mesh["esec"][2 * s + 1]],
                             np.int32).tobytes())
        f.write(mesh["x"].astype("<f8").tobytes())
        f.write(mesh["y"].astype("<f8").tobytes())
        f.write(mesh["z"].astype("<f8").tobytes())
        f.write(mesh["e2vx"].astype("<i4").tobytes())
