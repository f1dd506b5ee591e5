"""Synthetic meshes, Gambit neutral files and initial states for the tests.

Element-local conventions follow the reference:
  * Gambit brick/quad node orders and the reader's permutation into the internal
    shape-node order (src/geometry.cpp:1800-1855): hex internal c2v =
    gambit[0,4,1,5,2,6,3,7] read positionally as c2v(0),c2v(2),c2v(4),c2v(6),
    c2v(1),c2v(3),c2v(5),c2v(7); quad internal c2v(0),c2v(1),c2v(3),c2v(2).
  * Internal shape nodes are tensor ordered, r fastest (src/eles_hexas.cpp:1000).
"""
import math

import numpy as np

HEX_GAMBIT_TO_INTERNAL = [0, 2, 4, 6, 1, 3, 5, 7]   # internal[pos[i]] = gambit[i]
QUAD_GAMBIT_TO_INTERNAL = [0, 1, 3, 2]


def gambit_to_internal(ele_type, conn):
    conn = np.asarray(conn)
    if ele_type == 2:  # linear tets: Gambit order is the internal order
        return conn.copy()
    pos = HEX_GAMBIT_TO_INTERNAL if ele_type == 4 else QUAD_GAMBIT_TO_INTERNAL
    out = np.empty_like(conn)
    for i, p in enumerate(pos):
        out[:, p] = conn[:, i]
    return out


def lattice_coords(n, length, origin=0.0, digits=None):
    """n+1 equispaced coordinates on [origin, origin+length]; optionally rounded to
    `digits` significant digits (as a Gambit file written with %.11e would hold)."""
    x = np.array([origin + length * i / n for i in range(n + 1)])
    if digits is not None:
        x = np.array([float("%.*e" % (digits - 1, v)) for v in x])
    return x


def box_hex(nx, ny, nz, lx, ly, lz, origin=(0.0, 0.0, 0.0), digits=None):
    """Structured hex box in Gambit brick order (x fastest).  Returns (xv, conn_gambit)."""
    xs = lattice_coords(nx, lx, origin[0], digits)
    ys = lattice_coords(ny, ly, origin[1], digits)
    zs = lattice_coords(nz, lz, origin[2], digits)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    xv = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)

    def vid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    conn = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                conn.append([vid(i, j, k), vid(i + 1, j, k), vid(i, j + 1, k), vid(i + 1, j + 1, k),
                             vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i, j + 1, k + 1),
                             vid(i + 1, j + 1, k + 1)])
    return xv, np.array(conn, dtype=np.int32)


def box_quad(nx, ny, lx, ly, origin=(0.0, 0.0), digits=None):
    """Structured quad box in Gambit order (counter-clockwise)."""
    xs = lattice_coords(nx, lx, origin[0], digits)
    ys = lattice_coords(ny, ly, origin[1], digits)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    xv = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1)

    def vid(i, j):
        return i + (nx + 1) * j

    conn = [[vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)]
            for j in range(ny) for i in range(nx)]
    return xv, np.array(conn, dtype=np.int32)


def perturb_interior(xv, amp, seed, periodic_axes=()):
    """Move every vertex not on the box boundary by U(-amp, amp) per coordinate."""
    rng = np.random.default_rng(seed)
    lo, hi = xv.min(axis=0), xv.max(axis=0)
    eps = 1e-9 * (hi - lo).max()
    inner = np.all((xv > lo + eps) & (xv < hi - eps), axis=1)
    out = xv.copy()
    out[inner] += rng.uniform(-amp, amp, size=(inner.sum(), xv.shape[1]))
    return out


def write_gambit(path, xv, conn_gambit, ele_type, bc_groups=()):
    """Gambit neutral file as read by src/geometry.cpp:1544-1881.
    bc_groups: iterable of (name, [(ele_index0, gambit_face_number), ...])."""
    nd = xv.shape[1]
    ne = conn_gambit.shape[0]
    gtype = {4: 4, 2: 6, 1: 2}[ele_type]
    lines = ["        CONTROL INFO 2.3.16", "** GAMBIT NEUTRAL FILE", "graft", "PROGRAM: graft  VERSION: 1",
             "date", "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL",
             "%10d%10d%10d%10d%10d%10d" % (xv.shape[0], ne, 1, len(bc_groups), nd, nd),
             "ENDOFSECTION", "   NODAL COORDINATES 2.3.16"]
    for i, p in enumerate(xv):
        lines.append("%10d" % (i + 1) + "".join(" %24.17e" % v for v in p))
    lines += ["ENDOFSECTION", "      ELEMENTS/CELLS 2.3.16"]
    nv = conn_gambit.shape[1]
    for e in range(ne):
        ids = ["%8d" % (v + 1) for v in conn_gambit[e]]
        lines.append("%8d %2d %2d %s" % (e + 1, gtype, nv, "".join(ids[:7])))
        if nv > 7:   # Gambit wraps element records after 7 nodes
            lines.append("               " + "".join(ids[7:]))
    lines += ["ENDOFSECTION", "       ELEMENT GROUP 2.3.16",
              "GROUP:          1 ELEMENTS: %10d MATERIAL:          2 NFLAGS:          1" % ne,
              "                           fluid", "       0"]
    ids = list(range(1, ne + 1))
    for i in range(0, ne, 10):
        lines.append("".join("%8d" % v for v in ids[i:i + 10]))
    lines.append("ENDOFSECTION")
    for name, faces in bc_groups:
        lines.append(" BOUNDARY CONDITIONS 2.3.16")
        lines.append("%32s%8d%8d%8d%8d" % (name, 1, len(faces), 0, 6))
        for e, f in faces:
            lines.append("%10d%5d%5d" % (e + 1, gtype, f))
        lines.append("ENDOFSECTION")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# internal shape-node position of each Gmsh node of a linear element (the reader's
# permutation, src/geometry.cpp:1987-2050)
GMSH_TO_INTERNAL = {4: [0, 1, 3, 2, 4, 5, 7, 6], 1: [0, 1, 3, 2], 2: [0, 1, 2, 3]}


def write_gmsh(path, xv, c2v_internal, ele_type, bc_groups=(), node_order=None):
    """Gmsh 2.2 ASCII file as read by src/geometry.cpp:1309-1530, 1594-1640,
    1881-2088: a FLUID physical group of the elements and one physical group per
    boundary condition.  bc_groups: iterable of (name, [(element, internal face)]).
    node_order: optional permutation in which the $Nodes lines are listed (the ids
    stay index + 1)."""
    nd = xv.shape[1]
    pos = GMSH_TO_INTERNAL[ele_type]
    fn = face_nodes_of(ele_type)
    cell = {4: 5, 1: 3, 2: 4}[ele_type]
    face = {4: 3, 1: 1, 2: 2}[ele_type]
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "%d" % (1 + len(bc_groups))]
    lines.append('%d %d "FLUID"' % (nd, 1))
    for g, (name, _) in enumerate(bc_groups):
        lines.append('%d %d "%s"' % (nd - 1, g + 2, name))
    lines += ["$EndPhysicalNames", "$Nodes", "%d" % xv.shape[0]]
    order = range(xv.shape[0]) if node_order is None else node_order
    for i in order:
        p = list(xv[i]) + [0.0] * (3 - nd)
        lines.append("%d %.17g %.17g %.17g" % (i + 1, p[0], p[1], p[2]))
    lines += ["$EndNodes", "$Elements"]
    recs = []
    for g, (name, faces) in enumerate(bc_groups):
        for e, f in faces:
            verts = [int(c2v_internal[e][q]) for q in fn[f]]
            if ele_type == 4:  # quadrilateral face: around its perimeter
                verts = [verts[0], verts[1], verts[3], verts[2]]
            recs.append((face, g + 2, verts))
    for e in range(len(c2v_internal)):
        recs.append((cell, 1, [int(c2v_internal[e][pos[q]]) for q in range(len(pos))]))
    lines.append("%d" % len(recs))
    for k, (t, phys, verts) in enumerate(recs):
        lines.append("%d %d 2 %d %d %s" % (k + 1, t, phys, phys, " ".join("%d" % (v + 1) for v in verts)))
    lines.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# Internal local faces as shape-node sets (reference coordinates r,s,t):
# hex 0:t=-1 1:s=-1 2:r=+1 3:s=+1 4:r=-1 5:t=+1 (src/eles_hexas.cpp:266-324);
# quad 0:s=-1 1:r=+1 2:s=+1 3:r=-1 (src/eles_quads.cpp:259-300).
HEX_FACE_NODES = [(0, 1, 2, 3), (0, 1, 4, 5), (1, 3, 5, 7), (2, 3, 6, 7), (0, 2, 4, 6), (4, 5, 6, 7)]
QUAD_FACE_NODES = [(0, 1), (1, 3), (2, 3), (0, 2)]
# tet 0: opposite vertex 0 (slanted face), 1: r=-1, 2: s=-1, 3: t=-1 (src/geometry.cpp:3156-3185)
TET_FACE_NODES = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]
# Gambit face number k -> internal face (src/geometry.cpp read_boundary_gambit)
HEX_GAMBIT_FACE_TO_INTERNAL = {1: 0, 2: 3, 3: 5, 4: 1, 5: 4, 6: 2}
TET_GAMBIT_FACE_TO_INTERNAL = {1: 3, 2: 2, 3: 0, 4: 1}


def internal_face_to_gambit(ele_type, f):
    if ele_type == 4:
        return {v: k for k, v in HEX_GAMBIT_FACE_TO_INTERNAL.items()}[f]
    if ele_type == 2:
        return {v: k for k, v in TET_GAMBIT_FACE_TO_INTERNAL.items()}[f]
    return f + 1


def hex_to_tets(xv, conn_hex_gambit):
    """Split every hex of a Gambit brick connectivity into 6 tets around its
    (0,0,0)-(1,1,1) diagonal, the same split in every hex so that the face
    diagonals match between neighbours; each tet is ordered so that its Jacobian
    is positive.  Returns (n_hex*6, 4) vertex indices."""
    conn = np.asarray(conn_hex_gambit)
    # Gambit brick order: v(i,j,k) at index i + 2j + 4k
    v = lambda c, i, j, k: c[i + 2 * j + 4 * k]
    out = []
    for c in conn:
        d0, d1 = v(c, 0, 0, 0), v(c, 1, 1, 1)
        ring = [v(c, 1, 0, 0), v(c, 1, 1, 0), v(c, 0, 1, 0), v(c, 0, 1, 1), v(c, 0, 0, 1), v(c, 1, 0, 1)]
        for q in range(6):
            t = [d0, ring[q], ring[(q + 1) % 6], d1]
            a, b, cc, d = (xv[i] for i in t)
            if np.linalg.det(np.array([b - a, cc - a, d - a])) < 0:
                t[1], t[2] = t[2], t[1]
            out.append(t)
    return np.array(out, dtype=np.int64)


def face_nodes_of(ele_type):
    return {4: HEX_FACE_NODES, 2: TET_FACE_NODES, 1: QUAD_FACE_NODES}[ele_type]


def boundary_faces(ele_type, c2v_internal):
    """(element, internal face) of every face that no other element shares."""
    fn = face_nodes_of(ele_type)
    seen = {}
    for e, nodes in enumerate(c2v_internal):
        for f, loc in enumerate(fn):
            key = tuple(sorted(int(nodes[i]) for i in loc))
            seen.setdefault(key, []).append((e, f))
    return sorted(v[0] for v in seen.values() if len(v) == 1)


def channel_state(pos, mach, u_bulk=1.0):
    """Plane-channel field between walls at y = 0 and y = 2: parabolic
    u = 1.5 U_b y (2 - y), rho = 1, p = 1/(gamma Mach^2) (non-dimensional units of
    the reference with the bulk velocity as the free-stream velocity)."""
    gamma = 1.4
    y = pos[..., 1]
    u = 1.5 * u_bulk * y * (2.0 - y)
    nd = pos.shape[-1]
    out = np.zeros(pos.shape[:-1] + (nd + 2,))
    p = 1.0 / (gamma * mach ** 2)
    out[..., 0] = 1.0
    out[..., 1] = u
    out[..., nd + 1] = p / (gamma - 1.0) + 0.5 * u * u
    return out


def wall_faces(ele_type, c2v_internal, xv, axis, values, tol=1e-9):
    """(element, internal face) of the boundary faces whose vertices all lie on the
    plane x_axis = v for one of `values`."""
    fn = face_nodes_of(ele_type)
    out = []
    for e, f in boundary_faces(ele_type, c2v_internal):
        coords = xv[[int(c2v_internal[e][i]) for i in fn[f]], axis]
        if any(np.all(np.abs(coords - v) < tol) for v in values):
            out.append((e, f))
    return out


def read_gambit(path):
    """Vertices and element connectivity (Gambit order, 0-based) of a neutral file."""
    with open(path) as fh:
        L = fh.read().split("\n")
    i = 0
    while "NUMNP" not in L[i]:
        i += 1
    nn, ne, _, _, nd = [int(v) for v in L[i + 1].split()[:5]]
    i += 2
    while "NODAL" not in L[i]:
        i += 1
    i += 1
    xv = np.zeros((nn, nd))
    for k in range(nn):
        p = L[i + k].split()
        xv[int(p[0]) - 1] = [float(v) for v in p[1:1 + nd]]
    i += nn
    while "ELEMENTS/CELLS" not in L[i]:
        i += 1
    i += 1
    conn = []
    for _ in range(ne):
        toks = L[i].split()
        i += 1
        nv = int(toks[2])
        nodes = toks[3:]
        while len(nodes) < nv:
            nodes += L[i].split()
            i += 1
        conn.append([int(v) - 1 for v in nodes[:nv]])
    return xv, np.array(conn, dtype=np.int32)


# --------------------------------------------------------------------------
# Initial states at given physical points.  pos: (..., d)

def tgv_state(pos, gamma=1.4, rho=1.0, p0=100.0):
    """Taylor-Green field of src/eles.cpp:560-568 (ic_form 7); p0 = 100 there."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    p = p0 + rho / 16.0 * (np.cos(2.0 * x) + np.cos(2.0 * y)) * (np.cos(2.0 * z) + 2.0)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -1.0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([np.full_like(x, rho), rho * u, rho * v, rho * w, E], axis=-1)


def shu_vortex_state(pos, t=0.0, gamma=1.4, beta=5.0, rc=1.0, uinf=0.5 * math.sqrt(1.4), vinf=0.0,
                     x0=0.0, y0=0.0):
    """Isentropic vortex advected at (uinf, vinf): T = 1 - (g-1) b^2 e^{1-r^2}/(8 g pi^2),
    rho = T^{1/(g-1)}, p = rho^g."""
    x = pos[..., 0] - x0 - uinf * t
    y = pos[..., 1] - y0 - vinf * t
    r2 = (x * x + y * y) / (rc * rc)
    f = np.exp(1.0 - r2)
    T = 1.0 - (gamma - 1.0) * beta * beta * f / (8.0 * gamma * math.pi ** 2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    u = uinf - beta * y * np.exp(0.5 * (1.0 - r2)) / (2.0 * math.pi * rc)
    v = vinf + beta * x * np.exp(0.5 * (1.0 - r2)) / (2.0 * math.pi * rc)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def seeded_perturbation(shape, amp, seed):
    return np.random.default_rng(seed).uniform(-amp, amp, size=shape)
