"""Seeded synthetic meshes and initial states for the explicit residual path.

Meshes are produced in the form the msh 2.2 reader of subfv hands to
``build_mesh`` (src/mesh/mesh_reading_module.F90:1122-1321): vertex
coordinates in file order, 3-D elements as gmsh kinds (4 tet, 5 hex, 6 prism)
with gmsh vertex order, and optional tagged boundary faces.  Two-dimensional
cases are one-layer extrusions with ``boundary_2d`` set, as in the
reference's own test meshes.

Every random number comes from a counter-based generator keyed by
(seed, stream, entity id), so any consumer regenerates identical data.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

DEFAULT_SEED = 20261202
GAMMA = 1.4

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, ids) -> np.ndarray:
    """U[0,1) doubles keyed by (seed, stream, id)."""
    ids = np.asarray(ids, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)))
        h = _splitmix64(ids ^ key)
        h = _splitmix64(h + np.uint64(stream))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Mesh:
    """Raw mesh as read from a msh 2.2 file (1-based vertex ids)."""

    coord: np.ndarray            # (nv, 3) float64
    kind: np.ndarray             # (ne,) int32 gmsh element kind
    elem_vert: np.ndarray        # flat int32, 1-based vertex ids
    bf_nvert: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bf_vert: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    bf_tag: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    boundary_2d: bool = False
    name: str = ""

    @property
    def n_vert(self) -> int:
        return int(self.coord.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.kind.shape[0])


_NV_KIND = {4: 4, 5: 8, 6: 6, 7: 5}


def elem_offsets(kind: np.ndarray) -> np.ndarray:
    nv = np.vectorize(_NV_KIND.get, otypes=[np.int64])(kind) if kind.size else np.zeros(0, np.int64)
    off = np.zeros(kind.size + 1, np.int64)
    np.cumsum(nv, out=off[1:])
    return off


def _jitter_2d(n: int, jitter: float, seed: int, lx: float, ly: float, x0: float, y0: float):
    h_x, h_y = lx / n, ly / n
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    gid = (j * (n + 1) + i).astype(np.uint64)
    x = x0 + i * h_x
    y = y0 + j * h_y
    interior = (i > 0) & (i < n) & (j > 0) & (j < n)
    dx = (2.0 * uniform(seed, 1, gid) - 1.0) * jitter * h_x
    dy = (2.0 * uniform(seed, 2, gid) - 1.0) * jitter * h_y
    x = np.where(interior, x + dx, x)
    y = np.where(interior, y + dy, y)
    return x, y


def quad_mesh(n: int, jitter: float = 0.2, seed: int = DEFAULT_SEED, lx: float = 1.0, ly: float = 1.0,
              x0: float = 0.0, y0: float = 0.0, dz: float | None = None, tags: str | None = None) -> Mesh:
    """n x n jittered quads extruded one layer in z (gmsh hexahedra).

    ``tags='lr'`` tags the x = x0 side faces 1 and the x = x0 + lx side faces 2,
    everything else untagged (default wall); ``tags='lrbt'`` also tags the
    y = y0 side faces 3 and the y = y0 + ly side faces 4."""
    h = lx / n
    if dz is None:
        dz = h
    x, y = _jitter_2d(n, jitter, seed, lx, ly, x0, y0)
    n2 = (n + 1) * (n + 1)
    coord = np.zeros((2 * n2, 3))
    coord[:n2, 0] = x
    coord[:n2, 1] = y
    coord[n2:, 0] = x
    coord[n2:, 1] = y
    coord[n2:, 2] = dz
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()
    a = jj * (n + 1) + ii
    b = a + 1
    c = a + n + 2
    d = a + n + 1
    ev = np.stack([a, b, c, d, a + n2, b + n2, c + n2, d + n2], axis=1) + 1
    kind = np.full(n * n, 5, np.int32)
    mesh = Mesh(coord, kind, ev.astype(np.int32).ravel(), boundary_2d=True, name=f"quad{n}")
    if tags in ("lr", "lrbt"):
        left = np.arange(n) * (n + 1)
        right = left + n
        fl = np.stack([left, left + n2, left + n + 1 + n2, left + n + 1], 1) + 1
        fr = np.stack([right, right + n + 1, right + n + 1 + n2, right + n2], 1) + 1
        groups = [fl, fr]
        if tags == "lrbt":
            bot = np.arange(n)
            top = bot + n * (n + 1)
            groups.append(np.stack([bot, bot + 1, bot + 1 + n2, bot + n2], 1) + 1)
            groups.append(np.stack([top, top + n2, top + 1 + n2, top + 1], 1) + 1)
        faces = np.concatenate(groups).astype(np.int32)
        mesh.bf_nvert = np.full(faces.shape[0], 4, np.int32)
        mesh.bf_vert = faces.ravel()
        mesh.bf_tag = np.concatenate([np.full(n, k + 1) for k in range(len(groups))]).astype(np.int32)
    elif tags is not None:
        raise ValueError(f"unknown tags {tags!r}")
    return mesh


def tri_mesh(n: int, jitter: float = 0.2, seed: int = DEFAULT_SEED, lx: float = 1.0, ly: float = 1.0,
             x0: float = 0.0, y0: float = 0.0, dz: float | None = None) -> Mesh:
    """n x n jittered quads, each split along a seeded random diagonal into two
    triangles, extruded one layer in z (gmsh prisms)."""
    h = lx / n
    if dz is None:
        dz = h
    x, y = _jitter_2d(n, jitter, seed, lx, ly, x0, y0)
    n2 = (n + 1) * (n + 1)
    coord = np.zeros((2 * n2, 3))
    coord[:n2, 0] = x
    coord[:n2, 1] = y
    coord[n2:, 0] = x
    coord[n2:, 1] = y
    coord[n2:, 2] = dz
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()
    a = jj * (n + 1) + ii
    b = a + 1
    c = a + n + 2
    d = a + n + 1
    diag = uniform(seed, 3, (jj * n + ii).astype(np.uint64)) < 0.5
    t1 = np.where(diag[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(diag[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    tris = np.empty((2 * n * n, 3), np.int64)
    tris[0::2] = t1
    tris[1::2] = t2
    ev = np.concatenate([tris, tris + n2], axis=1) + 1
    kind = np.full(2 * n * n, 6, np.int32)
    return Mesh(coord, kind, ev.astype(np.int32).ravel(), boundary_2d=True, name=f"tri{n}")


def tet_mesh(n: int, jitter: float = 0.1, seed: int = DEFAULT_SEED, length: float = 1.0) -> Mesh:
    """n^3 jittered hexahedra on [0, length]^3, each split into 6 tetrahedra
    along its main diagonal (gmsh tets, positively oriented)."""
    h = length / n
    m1 = n + 1
    k, j, i = np.meshgrid(np.arange(m1), np.arange(m1), np.arange(m1), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    k = k.ravel()
    gid = (k * m1 * m1 + j * m1 + i).astype(np.uint64)
    coord = np.stack([i * h, j * h, k * h], 1).astype(np.float64)
    interior = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
    for ax in range(3):
        dd = (2.0 * uniform(seed, 4 + ax, gid) - 1.0) * jitter * h
        coord[:, ax] = np.where(interior, coord[:, ax] + dd, coord[:, ax])
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci = ci.ravel(order="F")
    cj = cj.ravel(order="F")
    ck = ck.ravel(order="F")
    base = ck * m1 * m1 + cj * m1 + ci

    def corner(bx, by, bz):
        return base + bz * m1 * m1 + by * m1 + bx

    c000 = corner(0, 0, 0)
    c111 = corner(1, 1, 1)
    unit = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        e1 = unit[p[0]]
        e2 = tuple(u + w for u, w in zip(unit[p[0]], unit[p[1]]))
        t = np.stack([c000, corner(*e1), corner(*e2), c111], 1)
        tets.append(t)
    tets = np.stack(tets, 1).reshape(-1, 4)
    # orient every tet positively on the unjittered lattice
    x0 = coord[tets[:, 0]]
    v = np.einsum("ij,ij->i", np.cross(coord[tets[:, 1]] - x0, coord[tets[:, 2]] - x0), coord[tets[:, 3]] - x0)
    neg = v < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    kind = np.full(tets.shape[0], 4, np.int32)
    return Mesh(coord, kind, (tets + 1).astype(np.int32).ravel(), boundary_2d=False, name=f"tet{n}")


def hex_mesh(n: int, jitter: float = 0.1, seed: int = DEFAULT_SEED, length: float = 1.0) -> Mesh:
    """n^3 jittered hexahedra on [0, length]^3 (gmsh hexahedra: the z = k
    face counter-clockwise seen from above, then the z = k + 1 face)."""
    h = length / n
    m1 = n + 1
    k, j, i = np.meshgrid(np.arange(m1), np.arange(m1), np.arange(m1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    gid = (k * m1 * m1 + j * m1 + i).astype(np.uint64)
    coord = np.stack([i * h, j * h, k * h], 1).astype(np.float64)
    interior = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
    for ax in range(3):
        dd = (2.0 * uniform(seed, 7 + ax, gid) - 1.0) * jitter * h
        coord[:, ax] = np.where(interior, coord[:, ax] + dd, coord[:, ax])
    ck, cj, ci = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (ck * m1 * m1 + cj * m1 + ci).ravel()
    a, b, c, d = base, base + 1, base + m1 + 1, base + m1
    ev = np.stack([a, b, c, d, a + m1 * m1, b + m1 * m1, c + m1 * m1, d + m1 * m1], 1) + 1
    kind = np.full(n ** 3, 5, np.int32)
    return Mesh(coord, kind, ev.astype(np.int32).ravel(), boundary_2d=False, name=f"hex{n}")


# ----------------------------------------------------------------- states
def primit_to_conserv(w: np.ndarray) -> np.ndarray:
    """Primitive (rho,u,v,w,p) rows to conserved rows, in the reference's
    expression order (src/ns/ns_euler_primitives_module.F90:7-18)."""
    w = np.asarray(w, dtype=np.float64)
    u = np.empty_like(w)
    u[..., 0] = w[..., 0]
    u[..., 1:4] = w[..., 1:4] * w[..., 0:1]
    u[..., 4] = w[..., 4] / (GAMMA - 1) + 0.5 * w[..., 0] * (
        (w[..., 1] * w[..., 1] + w[..., 2] * w[..., 2]) + w[..., 3] * w[..., 3])
    return u


def init_1drp(centroids: np.ndarray, x1drp: float, wl, wr) -> np.ndarray:
    """init_1drp branch of init_sol (src/ns/ns_euler_module.F90:36-43)."""
    left = centroids[:, 0] < x1drp
    ul = primit_to_conserv(np.asarray(wl, np.float64))
    ur = primit_to_conserv(np.asarray(wr, np.float64))
    return np.where(left[:, None], ul[None, :], ur[None, :])


def vortex_state(centroids: np.ndarray, seed: int = DEFAULT_SEED, n_vortex: int = 64,
                 umax: float = 0.3, noise: float = 0.01) -> np.ndarray:
    """rho = 1 (+/- seeded Gaussian noise), p = 1 and a solenoidal velocity
    field from seeded Gaussian vortices, scaled so max|u| <= umax."""
    cid = np.arange(n_vortex, dtype=np.uint64)
    xc = uniform(seed, 10, cid)
    yc = uniform(seed, 11, cid)
    strength = 2.0 * uniform(seed, 12, cid) - 1.0
    radius = 0.03 + 0.05 * uniform(seed, 13, cid)
    x = centroids[:, 0]
    y = centroids[:, 1]
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for q in range(n_vortex):
        dx = x - xc[q]
        dy = y - yc[q]
        g = strength[q] * np.exp(-(dx * dx + dy * dy) / (2.0 * radius[q] ** 2))
        u += -dy / radius[q] * g
        v += dx / radius[q] * g
    sp = np.sqrt(u * u + v * v).max()
    if sp > 0:
        u *= umax / sp
        v *= umax / sp
    ids = np.arange(x.size, dtype=np.uint64)
    u1 = np.maximum(uniform(seed, 14, ids), 1e-300)
    u2 = uniform(seed, 15, ids)
    gauss = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    rho = 1.0 + noise * np.clip(gauss, -3.0, 3.0) / 3.0
    w = np.stack([rho, u, v, np.zeros_like(u), np.ones_like(u)], 1)
    return primit_to_conserv(w)


def blast_state(centroids: np.ndarray, volumes: np.ndarray, h: float, seed: int = DEFAULT_SEED,
                energy: float = 1.0, p0: float = 1e-5) -> np.ndarray:
    """rho = 1, p = p0, plus total energy ``energy`` spread uniformly (per unit
    volume) over the cells whose centroid lies within 2h of a seeded point."""
    pt = 0.25 + 0.5 * uniform(seed, 20, np.arange(3, dtype=np.uint64))
    r = np.sqrt(((centroids - pt[None, :]) ** 2).sum(1))
    sel = r < 2.0 * h
    w = np.zeros((centroids.shape[0], 5))
    w[:, 0] = 1.0
    w[:, 4] = p0
    u = primit_to_conserv(w)
    vdep = volumes[sel].sum()
    u[sel, 4] += energy / vdep
    return u


def sedov_state(centroids: np.ndarray, volumes: np.ndarray, boundary_2d: bool) -> np.ndarray:
    """init_sedov of init_sol (ns_euler_module.F90:64-111, one rank): rho = 1,
    u = 0, E = 1e-16 / (gamma - 1) everywhere, and in the first cell of
    smallest centroid norm (|(x, y)| with boundary_2d) E = 0.984042 / vol in
    2-D, 0.851072 / vol in 3-D."""
    xc = centroids[:, :2] if boundary_2d else centroids
    r = np.sqrt((xc * xc).sum(1))
    imin = int(np.argmin(r))  # first index of the minimum, as the strict < scan
    u = np.zeros((centroids.shape[0], 5))
    u[:, 0] = 1.0
    u[:, 4] = 1e-16 / (GAMMA - 1.0)
    u[imin, 4] = (0.984042 if boundary_2d else 0.851072) / volumes[imin]
    return u


def double_mach_state(centroids: np.ndarray) -> np.ndarray:
    """init_double_mach of init_sol (ns_euler_module.F90:127-138): the post-shock
    state (8, 7.145, -4.125, 0, 116.5) where y >= 1.732 (x - 0.1667), the
    pre-shock state (1.4, 0, 0, 0, 1) below (primitive rho, u, v, w, p)."""
    below = centroids[:, 1] < 1.732 * (centroids[:, 0] - 0.1667)
    w = np.where(below[:, None], np.array([1.4, 0.0, 0.0, 0.0, 1.0])[None, :],
                 np.array([8.0, 7.145, -4.125, 0.0, 116.5])[None, :])
    return primit_to_conserv(w)


def write_msh(mesh: Mesh, path: str, bc_names=()) -> None:
    """Writes the mesh as a gmsh 2.2 ASCII file in the form
    read_mesh_msh_2 reads (mesh_reading_module.F90:989-1324): boundary face
    j-tags become physical groups named bc_names[j-1] (tag 0 faces get an
    unnamed group, i.e. the default wall); coordinates are written with 17
    significant digits so they read back bit-identical."""
    nb = int(mesh.bf_nvert.size)
    phys = {}
    for j, nm in enumerate(bc_names, start=1):
        phys[j] = (100 + j, nm)
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat"]
    lines += ["$PhysicalNames", str(len(phys) + 2), '3 1 "volume"', '2 99 "unnamed_boundary"']
    for j, (tag, nm) in phys.items():
        lines.append(f'2 {tag} "{nm}"')
    lines.append("$EndPhysicalNames")
    lines += ["$Nodes", str(mesh.n_vert)]
    lines += [f"{i + 1} {x:.17g} {y:.17g} {z:.17g}" for i, (x, y, z) in enumerate(mesh.coord)]
    lines.append("$EndNodes")
    items = []
    off = 0
    for f in range(nb):
        nv = int(mesh.bf_nvert[f])
        verts = mesh.bf_vert[off:off + nv]
        off += nv
        t = int(mesh.bf_tag[f])
        ptag = phys[t][0] if t in phys else 99
        items.append((2 if nv == 3 else 3, ptag, verts))
    nvk = {4: 4, 5: 8, 6: 6, 7: 5}
    off = 0
    for e in range(mesh.n_elems):
        k = int(mesh.kind[e])
        verts = mesh.elem_vert[off:off + nvk[k]]
        off += nvk[k]
        items.append((k, 1, verts))
    lines += ["$Elements", str(len(items))]
    for i, (k, ptag, verts) in enumerate(items):
        lines.append(f"{i + 1} {k} 2 {ptag} {ptag} " + " ".join(str(int(v)) for v in verts))
    lines.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


_GMSH_NV = {2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 15: 1, 1: 2}


def read_msh(path: str, bc_names=(), boundary_2d: bool = False) -> Mesh:
    """Reads a gmsh 2.2 ASCII mesh as read_mesh_msh_2 does
    (src/mesh/mesh_reading_module.F90:989-1324): with BC names given, the
    physical group named bc_names[j-1] tags its boundary faces j and every
    other boundary face 0 (the default wall); triangles and quadrangles are
    boundary faces, gmsh kinds 4-7 the 3-D elements, other items are skipped.
    Raises ValueError where the reference stops ("No Physical names inside
    mesh!!!", "No 3D elements in mesh !")."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    bc_tag = {}
    if bc_names:
        while not lines[i].startswith("$PhysicalNames"):
            if lines[i].startswith("$Nodes"):
                raise ValueError("No Physical names inside mesh!!!")
            i += 1
        n = int(lines[i + 1].split()[0])
        for line in lines[i + 2:i + 2 + n]:
            parts = line.split(None, 2)
            tag, name = int(parts[1]), parts[2].strip().strip('"')
            for j, nm in enumerate(bc_names, start=1):
                if nm.strip() == name:
                    bc_tag[j] = tag
    while not lines[i].startswith("$Nodes"):
        i += 1
    nv = int(lines[i + 1].split()[0])
    coord = np.array([[float(x) for x in line.split()[1:4]] for line in lines[i + 2:i + 2 + nv]], np.float64)
    while not lines[i].startswith("$Elements"):
        i += 1
    n_items = int(lines[i + 1].split()[0])
    kinds, ev, bnv, bv, btag = [], [], [], [], []
    for line in lines[i + 2:i + 2 + n_items]:
        f = [int(x) for x in line.split()]
        kind, ntag = f[1], f[2]
        ptag = f[3] if ntag > 0 else 0
        verts = f[3 + ntag:3 + ntag + _GMSH_NV.get(kind, 0)]
        if kind in (2, 3):
            bnv.append(len(verts))
            bv.extend(verts)
            t = 0
            for j in range(1, len(bc_names) + 1):
                if bc_tag.get(j) == ptag:
                    t = j
            btag.append(t)
        elif 4 <= kind <= 7:
            kinds.append(kind)
            ev.extend(verts)
    if not kinds:
        raise ValueError("[-] No 3D elements in mesh !")
    return Mesh(coord, np.array(kinds, np.int32), np.array(ev, np.int32), np.array(bnv, np.int32),
                np.array(bv, np.int32), np.array(btag, np.int32), boundary_2d=boundary_2d,
                name=os.path.basename(path))
