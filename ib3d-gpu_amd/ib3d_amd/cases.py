"""Case generators: stretched staggered grids, Gmsh-2.2 icospheres, ``inputdata``.

The reference ships one case (282x242x242) whose body mesh is missing, so every
case used here is generated.  The file formats are the ones the reference reads:

* ``inputdata``: 27 list-directed values (subReadInput.f90:8-17, key at
  inputdata:12-21).
* ``gridx.dat`` / ``gridy.dat`` / ``gridz.dat``: ``n+1`` node coordinates, one
  per line (subReadInput.f90:55-99).
* ``geometries/sphere.msh``: Gmsh 2.2 text; four header lines, the node count,
  ``id x y z`` lines, two lines, the element count, ``surGeoPoints`` skipped
  lines, then ``id type ntags tag tag n1 n2 n3`` (subReadInput.f90:137-176).

Numbers are written with ``repr`` (shortest round-trip), so every reader
(Fortran list-directed, C ``strtod``, numpy) recovers the same doubles.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field, asdict

import numpy as np

__all__ = [
    "CaseConfig", "CONFIGS", "stretched_axis", "icosphere", "write_case",
    "read_inputdata", "read_grid", "read_gmsh", "read_gmsh_bodies", "search_window", "ellipsoid",
    "place_ellipsoids", "write_result", "read_result", "read_iter_dat", "read_coeff_dat",
]


# --------------------------------------------------------------------------
# grids
# --------------------------------------------------------------------------
def _ratio_for(length: float, h: float, m: int) -> float:
    """Geometric ratio r with h*(r + r^2 + ... + r^m) == length (r >= 1)."""
    if m <= 0:
        if length > 1e-12:
            raise ValueError("no cells left for a stretched side of positive length")
        return 1.0
    if length <= h * m * (1 + 1e-12):
        raise ValueError(f"stretched side too short: length={length} h={h} m={m}")
    lo, hi = 1.0, 4.0
    while h * sum(hi ** k for k in range(1, m + 1)) < length:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        s = h * sum(mid ** k for k in range(1, m + 1))
        if s < length:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def stretched_axis(lo: float, hi: float, ulo: float, uhi: float, h: float, n: int):
    """Node coordinates (n+1 values) of one axis.

    Cells have width ``h`` inside [ulo, uhi] and grow geometrically outward.
    The split of the remaining cells between the two sides is chosen to
    minimise the larger of the two ratios; the ratios are returned so a case
    can state them.
    """
    nu = int(round((uhi - ulo) / h))
    if abs(nu * h - (uhi - ulo)) > 1e-9:
        raise ValueError("uniform box is not a whole number of cells")
    m = n - nu
    if m < 0:
        raise ValueError("more uniform cells than cells")
    best = None
    for ml in range(0, m + 1):
        mr = m - ml
        try:
            rl = _ratio_for(ulo - lo, h, ml)
            rr = _ratio_for(hi - uhi, h, mr)
        except ValueError:
            continue
        key = max(rl, rr)
        if best is None or key < best[0]:
            best = (key, ml, mr, rl, rr)
    if best is None:
        raise ValueError("cannot fit the stretched sides")
    _, ml, mr, rl, rr = best
    left = [ulo]
    w = h
    for _ in range(ml):
        w *= rl
        left.append(left[-1] - w)
    if ml:
        left[-1] = lo
    mid = [ulo + i * h for i in range(1, nu + 1)]
    mid[-1] = uhi
    right = []
    w = h
    x = uhi
    for _ in range(mr):
        w *= rr
        x = x + w
        right.append(x)
    if mr:
        right[-1] = hi
    nodes = np.array(left[::-1] + mid + right, dtype=np.float64)
    assert nodes.size == n + 1
    assert np.all(np.diff(nodes) > 0)
    return nodes, (rl, rr, ml, mr)


# --------------------------------------------------------------------------
# icosphere
# --------------------------------------------------------------------------
def icosphere(level: int, radius: float = 0.5):
    """Icosphere with 20*4**level triangles, wound counter-clockwise seen from
    outside, so (p2-p1)x(p3-p1) points out of the body (``inor`` = 1)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
             (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts) * radius
    f = np.array(faces, dtype=np.int64)
    # orient outward
    p1, p2, p3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(p2 - p1, p3 - p1)
    cen = (p1 + p2 + p3) / 3.0
    flip = np.einsum("ij,ij->i", nrm, cen) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def ellipsoid(center, semi, rot, n_lon: int = 64, n_lat: int = 33):
    """Closed latitude-longitude triangulation of the ellipsoid
    x = center + rot @ (semi * unit sphere point), 2*n_lon*(n_lat-1)
    triangles (64 x 33 gives 4096), wound outward."""
    verts = [np.array([0.0, 0.0, 1.0])]
    for a in range(1, n_lat):
        th = math.pi * a / n_lat
        for b in range(n_lon):
            ph = 2.0 * math.pi * b / n_lon
            verts.append(np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]))
    verts.append(np.array([0.0, 0.0, -1.0]))
    south = len(verts) - 1
    ring = lambda a, b: 1 + (a - 1) * n_lon + (b % n_lon)   # noqa: E731
    faces = []
    for b in range(n_lon):
        faces.append((0, ring(1, b), ring(1, b + 1)))
    for a in range(1, n_lat - 1):
        for b in range(n_lon):
            p, q, r, t = ring(a, b), ring(a, b + 1), ring(a + 1, b), ring(a + 1, b + 1)
            faces.append((p, r, t))
            faces.append((p, t, q))
    for b in range(n_lon):
        faces.append((south, ring(n_lat - 1, b + 1), ring(n_lat - 1, b)))
    unit = np.array(verts)
    v = (unit * np.asarray(semi, dtype=np.float64)) @ np.asarray(rot, dtype=np.float64).T + np.asarray(center)
    f = np.array(faces, dtype=np.int64)
    # outward winding: the normal of each face points away from the centre
    p1, p2, p3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(p2 - p1, p3 - p1)
    flip = np.einsum("ij,ij->i", nrm, (p1 + p2 + p3) / 3.0 - np.asarray(center)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def random_rotation(rng):
    """Uniformly distributed rotation (unit quaternion from four normals)."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def place_ellipsoids(seed: int, count: int, lo, hi, semi_range, gap: float, max_tries: int = 100000):
    """`count` ellipsoids inside the box [lo, hi] (the whole body, bounded by
    its largest semi-axis), semi-axes uniform in `semi_range`, uniform random
    orientations, pairwise clearance >= `gap` between bounding spheres (a
    conservative stand-in for the surface gap)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    bodies = []
    tries = 0
    while len(bodies) < count:
        tries += 1
        if tries > max_tries:
            raise ValueError("could not place the bodies without overlap")
        semi = rng.uniform(semi_range[0], semi_range[1], 3)
        r = float(semi.max())
        if np.any(hi - lo <= 2 * r):
            continue
        c = rng.uniform(lo + r, hi - r)
        if all(np.linalg.norm(c - b["center"]) >= r + b["rmax"] + gap for b in bodies):
            bodies.append({"center": c, "semi": semi, "rot": random_rotation(rng), "rmax": r})
    return bodies


def write_gmsh(path: str, nodes: np.ndarray, tris: np.ndarray, n_points: int = 12, body=None):
    """Gmsh 2.2 ASCII with ``n_points`` point elements (type 15) before the
    triangles (type 2); the reference skips ``surGeoPoints`` element lines.
    ``body`` (1-based, per triangle) goes into the physical tag."""
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n")
        fh.write(f"{len(nodes)}\n")
        for i, (x, y, z) in enumerate(nodes, start=1):
            fh.write(f"{i} {float(x)!r} {float(y)!r} {float(z)!r}\n")
        fh.write("$EndNodes\n$Elements\n")
        fh.write(f"{n_points + len(tris)}\n")
        eid = 1
        for p in range(n_points):
            fh.write(f"{eid} 15 2 0 {p + 1} {p + 1}\n")
            eid += 1
        for q, (a, b, c) in enumerate(tris):
            tag = 0 if body is None else int(body[q])
            fh.write(f"{eid} 2 2 {tag} 1 {a + 1} {b + 1} {c + 1}\n")
            eid += 1
        fh.write("$EndElements\n")


# --------------------------------------------------------------------------
# case description
# --------------------------------------------------------------------------
@dataclass
class CaseConfig:
    name: str
    n: tuple                      # (nx, ny, nz) interior cells
    lo: tuple                     # domain lower corner
    hi: tuple                     # domain upper corner
    ulo: tuple                    # uniform box lower corner
    uhi: tuple                    # uniform box upper corner
    h: float                      # uniform spacing
    level: int                    # icosphere subdivision level
    dt: float                     # time step
    re: float = 100.0
    steps: int = 200
    eps1: float = 1e-5            # max|dpc| stopping tolerance (parity mode)
    radius: float = 0.5
    shift: tuple = (0.0, 0.0, 0.0)
    inor: int = 1
    n_points: int = 12
    margin: int = 3               # search window margin in uniform cells
    extra: dict = field(default_factory=dict)

    @property
    def dt_order(self) -> float:
        return self.dt * 1000.0


CONFIGS = {
    # small, CPU-friendly cases for parity tests
    "tiny": CaseConfig("tiny", (32, 24, 24), (-3.0, -2.5, -2.5), (5.0, 2.5, 2.5),
                       (-0.75, -0.75, -0.75), (1.25, 0.75, 0.75), 0.125, 2, 0.01,
                       steps=5, shift=(0.013, 0.021, -0.017)),
    "small": CaseConfig("small", (48, 32, 32), (-4.0, -3.0, -3.0), (8.0, 3.0, 3.0),
                        (-1.0, -1.0, -1.0), (2.0, 1.0, 1.0), 0.125, 3, 0.01,
                        steps=10, shift=(0.0071, -0.0113, 0.0059)),
    # config 1 of the north star: CPU-runnable fixed sphere
    "cfg1": CaseConfig("cfg1", (128, 64, 64), (-5.0, -5.0, -5.0), (15.0, 5.0, 5.0),
                       (-1.0, -1.0, -1.0), (2.0, 1.0, 1.0), 1.0 / 16.0, 4, 0.005,
                       steps=200),
    # config 2: the headline 256x192x192 fixed sphere
    "cfg2": CaseConfig("cfg2", (256, 192, 192), (-5.0, -5.0, -5.0), (15.0, 5.0, 5.0),
                       (-1.0, -1.0, -1.0), (2.0, 1.0, 1.0), 1.0 / 40.0, 5, 0.002,
                       steps=500),
    # several ellipsoids (small parity case): 2 bodies, 256 triangles each
    "multi": CaseConfig("multi", (48, 32, 32), (-3.0, -2.0, -2.0), (6.0, 2.0, 2.0),
                        (-0.4, -1.0, -1.0), (2.4, 1.0, 1.0), 0.1, 0, 0.01, steps=5,
                        extra={"bodies": 2, "seed": 7, "box": ((0.0, -0.9, -0.9), (2.3, 0.9, 0.9)),
                               "semi": (0.3, 0.5), "gap": 0.2, "n_lon": 16, "n_lat": 9, "re_d": 100.0}),
    # config 5: 8 ellipsoids (seed 2026) in x in [0,6], y,z in [-2,2], 4096
    # triangles each, Re = 300 on the largest diameter, h = 0.02 in the box
    "cfg5": CaseConfig("cfg5", (512, 320, 320), (-4.0, -5.5, -5.5), (20.0, 5.5, 5.5),
                       (0.0, -2.0, -2.0), (6.0, 2.0, 2.0), 0.02, 0, 0.001, steps=500,
                       extra={"bodies": 8, "seed": 2026, "box": ((0.0, -2.0, -2.0), (6.0, 2.0, 2.0)),
                              "semi": (0.3, 0.8), "gap": 0.3, "n_lon": 64, "n_lat": 33, "re_d": 300.0}),
    # config 3: the sphere oscillating transversely, y_c = 0.25 sin(2 pi 0.2 t),
    # 384x256x256 with the uniform h = D/40 band covering the motion envelope
    "cfg3": CaseConfig("cfg3", (384, 256, 256), (-5.0, -5.0, -5.0), (15.0, 5.0, 5.0),
                       (-1.0, -1.25, -1.0), (2.0, 1.25, 1.0), 1.0 / 40.0, 5, 0.002,
                       steps=1000, extra={"a0": 0.25, "freq": 0.2}),
}


def search_window(nodes_1d: np.ndarray, centre: float, radius: float, h: float, margin: int):
    """Fortran (1-based, padded) index range [a, b] of cells whose centre lies
    within ``radius + margin*h`` of ``centre``.  Cell i (2..n+1) spans
    nodes x1(i)..x1(i+1); the file holds x1(2..n+2)."""
    xc = 0.5 * (nodes_1d[:-1] + nodes_1d[1:])        # cell centres, cell i = idx+2
    sel = np.nonzero(np.abs(xc - centre) <= radius + margin * h)[0]
    return int(sel[0]) + 2, int(sel[-1]) + 2


def write_case(cfg: CaseConfig, directory: str, istart: int = 0, a0: float = None, freq: float = None) -> dict:
    """Write inputdata, grid?.dat and geometries/sphere.msh; return metadata.
    ``istart`` = 1 makes the driver restart from a ``result`` file
    (main.f90:15-16)."""
    os.makedirs(os.path.join(directory, "geometries"), exist_ok=True)
    axes = []
    ratios = []
    for d in range(3):
        nodes, r = stretched_axis(cfg.lo[d], cfg.hi[d], cfg.ulo[d], cfg.uhi[d], cfg.h, cfg.n[d])
        axes.append(nodes)
        ratios.append(r)
        with open(os.path.join(directory, "grid" + "xyz"[d] + ".dat"), "w") as fh:
            for val in nodes:
                fh.write(f"{float(val)!r}\n")
    re = cfg.re
    if cfg.extra.get("bodies"):
        ex = cfg.extra
        bodies = place_ellipsoids(ex["seed"], ex["bodies"], ex["box"][0], ex["box"][1], ex["semi"], ex["gap"])
        vs, fs, bid, off = [], [], [], 0
        for k, b in enumerate(bodies):
            v, f = ellipsoid(b["center"], b["semi"], b["rot"], ex["n_lon"], ex["n_lat"])
            vs.append(v)
            fs.append(f + off)
            bid.append(np.full(len(f), k + 1, dtype=np.int32))
            off += len(v)
        v, f = np.concatenate(vs), np.concatenate(fs)
        write_gmsh(os.path.join(directory, "geometries", "sphere.msh"), v, f, cfg.n_points, np.concatenate(bid))
        # Reynolds number on the largest body diameter: Re_D = U D / nu with U = 1
        re = ex["re_d"] / (2.0 * max(float(b["semi"].max()) for b in bodies))
        lo_b = np.min(v, axis=0)
        hi_b = np.max(v, axis=0)
        win = []
        for d in range(3):
            xc = 0.5 * (axes[d][:-1] + axes[d][1:])
            sel = np.nonzero((xc >= lo_b[d] - cfg.margin * cfg.h) & (xc <= hi_b[d] + cfg.margin * cfg.h))[0]
            win.append((int(sel[0]) + 2, int(sel[-1]) + 2))
    else:
        v, f = icosphere(cfg.level, cfg.radius)
        write_gmsh(os.path.join(directory, "geometries", "sphere.msh"), v, f, cfg.n_points)
        win = [search_window(axes[d], cfg.shift[d], cfg.radius, cfg.h, cfg.margin) for d in range(3)]
    nx, ny, nz = cfg.n
    with open(os.path.join(directory, "inputdata"), "w") as fh:
        fh.write(f"      {nx}, {ny}, {nz},\n")
        fh.write(f"      {re!r},\n")
        fh.write(f"      {cfg.steps}, {cfg.eps1!r}, 35000,\n")
        fh.write("      0., 0., 0.,\n")
        amp = cfg.extra.get("a0", 0.0) if a0 is None else a0
        fr = cfg.extra.get("freq", 0.0) if freq is None else freq
        if amp == 0.0 and fr == 0.0:
            fh.write(f"      {cfg.n_points}, 0, 0., 0, 0.,\n")
        else:  # oscillating body: y_c = a0 sin(2 pi freq t)
            fh.write(f"      {cfg.n_points}, {float(amp)!r}, {float(fr)!r}, 0, 0.,\n")
        fh.write(f"      {cfg.shift[0]!r}, {cfg.shift[1]!r}, {cfg.shift[2]!r},\n")
        fh.write(f"      {int(istart)}, {cfg.dt_order!r}, {cfg.inor},\n")
        fh.write(f"      {win[0][0]}, {win[0][1]},\n")
        fh.write(f"      {win[1][0]}, {win[1][1]},\n")
        fh.write(f"      {win[2][0]}, {win[2][1]}\n")
    meta = asdict(cfg)
    meta["ratios"] = ratios
    meta["window"] = win
    meta["n_nodes"] = len(v)
    meta["n_tris"] = len(f)
    return meta


# --------------------------------------------------------------------------
# readers (host side, used by the Python driver and tests)
# --------------------------------------------------------------------------
_INPUT_KEYS = ["nx", "ny", "nz", "re", "itamax", "eps1", "pcItaMax", "u0", "v0", "w0",
               "surGeoPoints", "a0", "freq", "aoa", "piv_pt", "xShift", "yShift", "zShift",
               "istart", "dt_order", "inor", "i_startSearch", "i_endSearch",
               "j_startSearch", "j_endSearch", "k_startSearch", "k_endSearch"]
_INT_KEYS = {"nx", "ny", "nz", "itamax", "pcItaMax", "surGeoPoints", "istart", "inor",
             "i_startSearch", "i_endSearch", "j_startSearch", "j_endSearch",
             "k_startSearch", "k_endSearch"}


def read_inputdata(path: str) -> dict:
    """Parse the 27 list-directed values of ``inputdata``
    (subReadInput.f90:8-17).  Values are separated by commas/blanks; reading
    stops after the 27th value (the key block after it is never read)."""
    toks = []
    with open(path) as fh:
        for line in fh:
            for t in line.replace(",", " ").split():
                toks.append(t)
                if len(toks) == len(_INPUT_KEYS):
                    break
            if len(toks) == len(_INPUT_KEYS):
                break
    if len(toks) < len(_INPUT_KEYS):
        raise ValueError(f"inputdata holds {len(toks)} values, 27 needed")
    out = {}
    for k, t in zip(_INPUT_KEYS, toks):
        if k in _INT_KEYS:
            out[k] = int(float(t))
        else:
            out[k] = float(t.replace("d", "e").replace("D", "e"))
    return out


def read_grid(path: str, count: int) -> np.ndarray:
    """First ``count`` values of a grid file (one per line)."""
    vals = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if s:
                vals.append(float(s.split()[0].replace("d", "e").replace("D", "e")))
            if len(vals) == count:
                break
    if len(vals) < count:
        raise ValueError(f"{path}: {len(vals)} values, {count} needed")
    return np.array(vals, dtype=np.float64)


def read_gmsh(path: str, sur_geo_points: int):
    """Gmsh 2.2 reader with the reference's line accounting
    (subReadInput.f90:143-167): returns nodes (n,3) and 1-based connectivity
    (m,3) of the triangles after the ``sur_geo_points`` skipped elements."""
    with open(path) as fh:
        lines = [ln for ln in fh.read().splitlines()]
    it = iter(lines)
    for _ in range(4):
        next(it)
    nn = int(next(it).split()[0])
    nodes = np.empty((nn, 3), dtype=np.float64)
    for n in range(nn):
        p = next(it).split()
        nodes[n] = [float(p[1]), float(p[2]), float(p[3])]
    next(it)
    next(it)
    ne = int(next(it).split()[0])
    for _ in range(sur_geo_points):
        next(it)
    ne -= sur_geo_points
    conn = np.empty((ne, 3), dtype=np.int64)
    for n in range(ne):
        p = next(it).split()
        conn[n] = [int(p[5]), int(p[6]), int(p[7])]
    return nodes, conn


def read_gmsh_bodies(path: str, sur_geo_points: int) -> np.ndarray:
    """Physical tag of every triangle after the skipped elements (the body id
    of multi-body meshes; the reference reads the connectivity only)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    q = 4
    nn = int(lines[q].split()[0])
    q += 1 + nn + 2
    ne = int(lines[q].split()[0])
    q += 1 + sur_geo_points
    return np.array([int(lines[q + n].split()[3]) for n in range(ne - sur_geo_points)], dtype=np.int32)


# --------------------------------------------------------------------------
# restart file
# --------------------------------------------------------------------------
def write_result(path: str, u, v, w, p, totime: float, ita: int, ita1: int) -> None:
    """The restart file of subWriteOutput.f90:58-72: one record per padded
    cell, i fastest, holding ``u v w p totime ita ita1``.  Values are
    written in shortest round-trip form, so a list-directed read gets the
    same doubles back.  Fields are (nx+2, ny+2, nz+2) arrays indexed [i,j,k]."""
    u, v, w, p = (np.asarray(a, dtype=np.float64) for a in (u, v, w, p))
    uf, vf, wf, pf = (a.reshape(-1, order="F") for a in (u, v, w, p))
    tail = f" {float(totime)!r} {int(ita)} {int(ita1)}\n"
    with open(path, "w") as fh:
        for q in range(uf.size):
            fh.write(f" {float(uf[q])!r} {float(vf[q])!r} {float(wf[q])!r} {float(pf[q])!r}{tail}")


def read_result(path: str, n: tuple):
    """Read a restart file (subLastConditions.f90:24-30); returns
    (u, v, w, p, totime, ita, ita1) with fields shaped (nx+2, ny+2, nz+2).
    Fortran exponent letters (D) are accepted."""
    shape = (n[0] + 2, n[1] + 2, n[2] + 2)
    cnt = shape[0] * shape[1] * shape[2]
    cols = [np.empty(cnt) for _ in range(4)]
    totime, ita, ita1 = 0.0, 0, 0
    with open(path) as fh:
        for q in range(cnt):
            t = fh.readline().replace(",", " ").replace("D", "E").replace("d", "e").split()
            if len(t) < 7:
                raise ValueError(f"{path}: record {q + 1} holds {len(t)} values, 7 needed")
            for c in range(4):
                cols[c][q] = float(t[c])
            totime, ita, ita1 = float(t[4]), int(t[5]), int(t[6])
    fields = [c.reshape(shape, order="F") for c in cols]
    return fields[0], fields[1], fields[2], fields[3], totime, ita, ita1


def read_iter_dat(path: str):
    """Rows of iter.dat (subPcorVcor.f90:60-62, FORMAT(' ',I8,I10,4E15.6)):
    (ita, nIterPcor, divmax, derrStdSt, sor_seconds)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            rows.append((int(line[1:9]), int(line[9:19]), float(line[19:34]), float(line[34:49]),
                         float(line[49:64])))
    return rows


def read_coeff_dat(path: str):
    """Rows of dragcoff.dat / liftcoff.dat (list-directed, three values:
    viscous part, pressure part, totime; subStressCal2.f90:374-379)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            t = line.replace("D", "E").split()
            if t:
                rows.append(tuple(float(x) for x in t[:3]))
    return rows
