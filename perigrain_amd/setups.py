"""Setups: the in-memory form of data/hdf5/all.h5 and seeded generators for
the benchmark workloads.

`Setup` holds exactly what load_particles/load_contact/load_wall read from
all.h5 (src/read/load_hdf5.h:41-201; schema written by exp_dict.py:705-770),
flattened over grains.  `Params` holds the config/main.conf toggles after the
Timeloop/Contact/RectWall apply_config overrides.

The generators build grains as jittered lattices (the reference meshes
shapes with gmsh, which is not available; its bond rule is kept: every node
pair within the horizon delta, exp_dict.py:171-222, minus bonds that leave a
non-convex shape).  Material constants follow material_dict.py:
  2D  c = 6E / (pi delta^3 (1 - nu)),  s0 = sqrt(4 pi Gc / (9 E delta))
      (material_dict.py:116-118, sodalime_similar_to_peridem); this equals
      c = 18K / (pi delta^4 thickness) with K = E (nu = 1/3) and
      thickness = 2 delta.
  3D  K = E / (3 (1 - 2 nu)),  c = 18K / (pi delta^4),
      s0 = sqrt(5 Gc / (9 K delta))  (material_dict.py:194-196).
"""
from dataclasses import dataclass, field, replace
from typing import List

import numpy as np
from scipy.spatial import cKDTree

WALL_SIDES_2D = ("left", "right", "top", "bottom")
WALL_SIDES_3D = ("x_min", "y_min", "z_min", "x_max", "y_max", "z_max")


@dataclass
class Setup:
    dim: int
    node_offset: np.ndarray          # int64 [G+1]
    pos: np.ndarray                  # [N, dim]
    vol: np.ndarray                  # [N]
    disp: np.ndarray
    vel: np.ndarray
    acc: np.ndarray
    extforce: np.ndarray
    conn_offset: np.ndarray          # int64 [G+1]
    conn: np.ndarray                 # uint32 [R, 2] grain-local
    bdry_offset: np.ndarray
    bdry_nodes: np.ndarray
    clamped_offset: np.ndarray
    clamped_nodes: np.ndarray
    delta: np.ndarray
    rho: np.ndarray
    cnot: np.ndarray
    snot: np.ndarray
    movable: np.ndarray
    breakable: np.ndarray
    stoppable: np.ndarray
    torque_axis: np.ndarray
    torque_val: np.ndarray
    contact_radius: float
    normal_stiffness: float
    damping_ratio: float
    friction_coefficient: float
    allow_wall: int
    wall: np.ndarray                 # [6], lrtb() order

    @property
    def n_nodes(self):
        return int(self.node_offset[-1])

    @property
    def n_grains(self):
        return int(len(self.node_offset) - 1)

    def n_slots(self):
        return 2 * int(len(self.conn))

    def grain_nodes(self, g):
        return slice(int(self.node_offset[g]), int(self.node_offset[g + 1]))


@dataclass
class Params:
    """Timeloop/Contact/RectWall toggles (config/main.conf keys)."""
    dt: float = 2e-7
    timesteps: int = 8000
    modulo: int = 100
    enable_fracture: int = 0
    nl_bdry_only: int = 0
    allow_damping: int = 1
    allow_friction: int = 1
    self_contact: int = 1
    gradient_extforce: int = 0
    extforce_maxstep: int = 0
    enable_torque: int = 0
    set_movable_index: int = -1
    set_movable_timestep: int = -1
    set_stoppable_index: int = -1
    set_stoppable_timestep: int = -1
    reset_partzero_y: int = 0
    reset_partzero_y_timestep: int = 0
    wheel_rad: float = 0.0
    wall_speed: np.ndarray = field(default_factory=lambda: np.zeros(6))
    timeloop_variant: int = 0  # 0 run_timeloop, 1 run_timeloop_worm, 2 run_timeloop_compact


# ----------------------------------------------------------------- materials

def material_2d(delta, E=1.34e9, nu=1.0 / 3.0, rho=1200.0, Gc=135.0):
    c = 6 * E / (np.pi * delta ** 3 * (1 - nu))
    s0 = np.sqrt(4 * np.pi * Gc / (9 * E * delta))
    return rho, c, s0


def material_3d(delta, E=1.34e9, nu=0.25, rho=1200.0, Gc=135.0):
    K = E / (3 * (1 - 2 * nu))
    c = 18 * K / (np.pi * delta ** 4)
    s0 = np.sqrt(5 * Gc / (9 * K * delta))
    return rho, c, s0


# -------------------------------------------------------------------- shapes
# Each shape is a signed distance function (negative inside) in the grain's
# own frame, with `radius` its circumradius.

def sdf_disk(p, R):
    return np.linalg.norm(p, axis=-1) - R


def sdf_box(p, half):
    q = np.abs(p) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    inside = np.minimum(np.max(q, axis=-1), 0.0)
    return outside + inside


def sdf_square(p, R):
    a = R / np.sqrt(2.0)
    return sdf_box(p, np.array([a, a]))


def sdf_plus(p, R):
    w = R / 3.0
    L = np.sqrt(R * R - w * w)
    return np.minimum(sdf_box(p, np.array([L, w])), sdf_box(p, np.array([w, L])))


def sdf_annulus(p, R):
    r = np.linalg.norm(p, axis=-1)
    return np.maximum(r - R, 0.5 * R - r)


def sdf_sphere(p, R):
    return np.linalg.norm(p, axis=-1) - R


SHAPES_2D = {"disk": (sdf_disk, np.pi), "square": (sdf_square, 2.0),
             "plus": (sdf_plus, None), "annulus": (sdf_annulus, 0.75 * np.pi)}
CONVEX = {"disk", "square", "sphere"}


def shape_area(name, R):
    if name == "plus":
        w = R / 3.0
        L = np.sqrt(R * R - w * w)
        return 2 * (2 * L) * (2 * w) - (2 * w) ** 2
    return SHAPES_2D[name][1] * R * R


@dataclass
class Grain:
    pos: np.ndarray
    vol: np.ndarray
    conn: np.ndarray
    bdry: np.ndarray
    delta: float
    rho: float
    cnot: float
    snot: float
    radius: float


def lattice_grain(sdf, R, h, dim, rng, contact_radius, convex, jitter=0.05):
    """Jittered lattice nodes inside sdf<0, PD bonds within delta=3h."""
    n = int(np.ceil(R / h)) + 1
    ax = np.arange(-n, n + 1) * h
    grids = np.meshgrid(*([ax] * dim), indexing="ij")
    pts = np.stack([g.reshape(-1) for g in grids], axis=1)
    pts = pts[sdf(pts, R) < 0]
    pts = pts + rng.uniform(-jitter * h, jitter * h, size=pts.shape)
    delta = 3.0 * h
    tree = cKDTree(pts)
    pairs = tree.query_pairs(delta, output_type="ndarray").astype(np.int64).reshape(-1, 2)
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]  # itertools.combinations order
    if len(pairs):
        d = np.sqrt(np.sum((pts[pairs[:, 1]] - pts[pairs[:, 0]]) ** 2, axis=1))
        pairs = pairs[d <= delta]
        if not convex:
            ok = np.ones(len(pairs), dtype=bool)
            for f in np.linspace(0.0, 1.0, 11)[1:-1]:
                q = pts[pairs[:, 0]] * (1 - f) + pts[pairs[:, 1]] * f
                ok &= sdf(q, R) < 0
            pairs = pairs[ok]
    # nonlocal boundary nodes: within the contact radius of the boundary
    bdry = np.nonzero(sdf(pts, R) >= -contact_radius)[0].astype(np.uint32)
    vol = np.full(len(pts), h ** dim)
    return pts, vol, pairs.astype(np.uint32), bdry, delta


def assemble(dim, grains, centers, vels, contact_radius, normal_stiffness, damping_ratio,
             friction_coefficient, gravity, allow_wall, wall, movable=None, breakable=None,
             stoppable=None):
    G = len(grains)
    sizes = [len(g.pos) for g in grains]
    node_offset = np.zeros(G + 1, dtype=np.int64)
    node_offset[1:] = np.cumsum(sizes)
    pos = np.concatenate([g.pos + c for g, c in zip(grains, centers)], axis=0)
    vol = np.concatenate([g.vol for g in grains])
    vel = np.concatenate([np.tile(v, (len(g.pos), 1)) for g, v in zip(grains, vels)], axis=0)
    rho = np.array([g.rho for g in grains])
    grav = np.zeros(dim)
    grav[-1] = -gravity
    ext = np.concatenate([np.tile(g.rho * grav, (len(g.pos), 1)) for g in grains], axis=0)
    conn_offset = np.zeros(G + 1, dtype=np.int64)
    conn_offset[1:] = np.cumsum([len(g.conn) for g in grains])
    conn = (np.concatenate([g.conn for g in grains], axis=0) if G else
            np.zeros((0, 2), np.uint32)).astype(np.uint32)
    bdry_offset = np.zeros(G + 1, dtype=np.int64)
    bdry_offset[1:] = np.cumsum([len(g.bdry) for g in grains])
    bdry = np.concatenate([g.bdry for g in grains]).astype(np.uint32)
    ones = np.ones(G, dtype=np.int32)
    w = np.zeros(6)
    w[:len(wall)] = wall
    return Setup(
        dim=dim, node_offset=node_offset, pos=pos, vol=vol, disp=np.zeros_like(pos),
        vel=vel, acc=np.zeros_like(pos), extforce=ext, conn_offset=conn_offset, conn=conn,
        bdry_offset=bdry_offset, bdry_nodes=bdry, clamped_offset=np.zeros(G + 1, np.int64),
        clamped_nodes=np.zeros(0, np.uint32), delta=np.array([g.delta for g in grains]),
        rho=rho, cnot=np.array([g.cnot for g in grains]),
        snot=np.array([g.snot for g in grains]),
        movable=ones.copy() if movable is None else np.asarray(movable, np.int32),
        breakable=ones.copy() if breakable is None else np.asarray(breakable, np.int32),
        stoppable=ones.copy() if stoppable is None else np.asarray(stoppable, np.int32),
        torque_axis=np.full(G, 2, np.uint32), torque_val=np.zeros(G),
        contact_radius=float(contact_radius), normal_stiffness=float(normal_stiffness),
        damping_ratio=float(damping_ratio), friction_coefficient=float(friction_coefficient),
        allow_wall=int(allow_wall), wall=w)


def make_grain_2d(shape, R, h, rng, contact_radius, mat=None):
    sdf = SHAPES_2D[shape][0]
    pts, vol, conn, bdry, delta = lattice_grain(sdf, R, h, 2, rng, contact_radius,
                                                shape in CONVEX)
    rho, c, s0 = material_2d(delta) if mat is None else mat(delta)
    return Grain(pts, vol, conn, bdry, delta, rho, c, s0, R)


def make_grain_3d(R, h, rng, contact_radius):
    pts, vol, conn, bdry, delta = lattice_grain(sdf_sphere, R, h, 3, rng, contact_radius, True)
    rho, c, s0 = material_3d(delta)
    return Grain(pts, vol, conn, bdry, delta, rho, c, s0, R)


def random_placement(radii, lo, hi, rng, gap, max_tries=20000):
    """Random non-overlapping circles/spheres (circumradius + gap) in [lo, hi]."""
    dim = len(lo)
    C = np.zeros((len(radii), dim))
    Rs = np.zeros(len(radii))
    for k, R in enumerate(radii):
        for _ in range(max_tries):
            c = rng.uniform(np.asarray(lo) + R + gap, np.asarray(hi) - R - gap)
            if k == 0 or np.all(np.linalg.norm(C[:k] - c, axis=1) >= Rs[:k] + R + gap):
                C[k] = c
                Rs[k] = R
                break
        else:
            raise RuntimeError("random placement failed; box too small")
    return [C[k].copy() for k in range(len(radii))]


# --------------------------------------------------------------- workloads

def two_disk_collision(seed=0, R=1e-3, h_div=15, speed=10.0, gap=1e-4):
    """2D two-disk collision: disks of radius R on a jittered lattice
    h=R/15, delta=3h, contact radius delta/5, K_n=c, zeta=0.8, mu=0.3,
    approaching at +-speed along x with `gap` between the two node clouds;
    no walls, no gravity.  dt=1e-8, 5000 steps, fracture on."""
    rng = np.random.default_rng(seed)
    h = R / h_div
    delta = 3 * h
    Rc = delta / 5
    grains = [make_grain_2d("disk", R, h, rng, Rc) for _ in range(2)]
    x0 = float(np.max(grains[0].pos[:, 0]))
    x1 = float(-np.min(grains[1].pos[:, 0]))
    centers = [np.array([-x0 - gap / 2, 0.0]), np.array([x1 + gap / 2, 0.0])]
    vels = [np.array([speed, 0.0]), np.array([-speed, 0.0])]
    s = assemble(2, grains, centers, vels, Rc, grains[0].cnot, 0.8, 0.3, 0.0, 0,
                 [-4 * R, 4 * R, 2 * R, -2 * R])
    p = Params(dt=1e-8, timesteps=5000, modulo=100, enable_fracture=1)
    return s, p


def bulk_settle_2d(seed=0, n_grains=400, box=(40e-3, 80e-3), rmin=0.6e-3, rmax=1.2e-3,
                   nodes_per_grain=300):
    """2D bulk settling: equal mix of disk/square/plus/annulus, radius
    U[rmin,rmax], ~300 lattice nodes per grain (delta=3h per grain), random
    non-overlapping placement, gravity toward 4 walls.  Contact radius and
    wall/contact stiffness come from the median grain (delta/5, its c).
    dt=2e-8, 20000 steps, fracture on."""
    rng = np.random.default_rng(seed)
    shapes = [("disk", "square", "plus", "annulus")[k % 4] for k in range(n_grains)]
    radii = rng.uniform(rmin, rmax, size=n_grains)
    hs = np.array([np.sqrt(shape_area(sh, R) / nodes_per_grain) for sh, R in zip(shapes, radii)])
    dmed = 3 * float(np.median(hs))
    Rc = dmed / 5
    grains = [make_grain_2d(sh, R, h, rng, Rc) for sh, R, h in zip(shapes, radii, hs)]
    _, cmed, _ = material_2d(dmed)
    rng2 = np.random.default_rng(seed + 1)
    centers = random_placement(radii, [0.0, 0.0], list(box), rng2, Rc)
    vels = [np.zeros(2)] * n_grains
    s = assemble(2, grains, centers, vels, Rc, cmed, 0.8, 0.3, 9.81, 1,
                 [0.0, box[0], box[1], 0.0])
    p = Params(dt=2e-8, timesteps=20000, modulo=1000, enable_fracture=1)
    return s, p


def compression_3d(seed=0, n_side=4, R=1e-3, h_div=10, cube=10e-3, wall_speed=-1.0):
    """3D compression: n_side^3 spheres of radius R on jittered cubic lattices
    h=R/10 (~4,200 nodes each), delta=3h, contact radius delta/4, K_n=c,
    zeta=0.8, mu=0.3, fracture on.  Spheres sit on a grid of pitch
    (node-cloud diameter) + Rc/2 with random offsets (so neighbouring grains
    start within
    contact range: node clouds Rc/2 apart), inside a cube of side `cube`;
    gravity -z; the top wall
    (z_max) starts Rc/2 above the highest node and moves at wall_speed.
    dt=1e-8, 10000 steps."""
    rng = np.random.default_rng(seed)
    h = R / h_div
    delta = 3 * h
    Rc = delta / 4
    n = n_side ** 3
    grains = [make_grain_3d(R, h, rng, Rc) for _ in range(n)]
    rng2 = np.random.default_rng(seed + 1)
    # pitch at which the node clouds of two neighbouring spheres are Rc/2
    # apart (bisection on the nearest node pair)
    ta = cKDTree(grains[0].pos)
    other = grains[1 % n].pos
    lo, hi = 0.0, 2 * R + 2 * Rc
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        dmin = ta.query(other + np.array([mid, 0.0, 0.0]))[0].min()
        lo, hi = (mid, hi) if dmin < Rc / 2 else (lo, mid)
    pitch = hi
    span = (n_side - 1) * pitch
    centers = []
    for iz in range(n_side):
        for iy in range(n_side):
            for ix in range(n_side):
                c = np.array([cube / 2 - span / 2 + ix * pitch,
                              cube / 2 - span / 2 + iy * pitch,
                              R + Rc + iz * pitch])
                c[:2] += rng2.uniform(-Rc / 4, Rc / 4, size=2)
                centers.append(c)
    zmax = max(float(np.max(g.pos[:, 2] + c[2])) for g, c in zip(grains, centers))
    wall = [0.0, 0.0, 0.0, cube, cube, zmax + Rc / 2]
    s = assemble(3, grains, centers, [np.zeros(3)] * n, Rc, grains[0].cnot, 0.8, 0.3, 9.81, 1,
                 wall)
    p = Params(dt=1e-8, timesteps=10000, modulo=1000, enable_fracture=1)
    p.wall_speed[5] = wall_speed
    return s, p


def settle_instance_2d(seed):
    """One instance of the batched workload: grains U[20,150], shapes from
    {disk, square, plus, annulus}, radius U[0.5,1.2] mm, 50-400 nodes per
    grain (drawn as 50*8**(u**3)), gravity settle in a box of width
    U[20,60] mm and height twice that (1.5x taller again whenever the random
    placement jams).  dt=2e-8, 5000 steps."""
    rng = np.random.default_rng(seed)
    G = int(rng.integers(20, 151))
    width = float(rng.uniform(20e-3, 60e-3))
    height = 2 * width
    shapes = [("disk", "square", "plus", "annulus")[k] for k in rng.integers(0, 4, size=G)]
    radii = rng.uniform(0.5e-3, 1.2e-3, size=G)
    counts = 50.0 * 8.0 ** (rng.uniform(0, 1, size=G) ** 3)
    hs = np.array([np.sqrt(shape_area(sh, R) / n) for sh, R, n in zip(shapes, radii, counts)])
    dmed = 3 * float(np.median(hs))
    Rc = dmed / 5
    grains = [make_grain_2d(sh, R, h, rng, Rc) for sh, R, h in zip(shapes, radii, hs)]
    _, cmed, _ = material_2d(dmed)
    while True:  # a taller box when the random placement jams
        try:
            centers = random_placement(radii, [0.0, 0.0], [width, height], rng, Rc,
                                       max_tries=2000)
            break
        except RuntimeError:
            height *= 1.5
    s = assemble(2, grains, centers, [np.zeros(2)] * G, Rc, cmed, 0.8, 0.3, 9.81, 1,
                 [0.0, width, height, 0.0])
    p = Params(dt=2e-8, timesteps=5000, modulo=1000, enable_fracture=1)
    return s, p


def batched_settle(n_instances=256, seed=0):
    out = [settle_instance_2d(seed * 100003 + k) for k in range(n_instances)]
    return [s for s, _ in out], [p for _, p in out]


def compression_2d_diff(seed=0, n_grains=64, K=1.5e9, contact_scale=1.0, mu=0.25,
                        nodes_per_grain=250, wall_speed=-0.5):
    """2D compression for the differentiable form: 64 disks and squares of
    radius U[0.8,1.0] mm with ~250 nodes each on an 8x8 grid, top wall moving
    down at 0.5 m/s, gravity on.  The bulk modulus K enters through
    E = K (nu = 1/3) in c and s0; the contact stiffness is contact_scale*c of
    the mean grain.  dt=2e-8, 2000 steps."""
    rng = np.random.default_rng(seed)
    shapes = [("disk", "square")[k % 2] for k in range(n_grains)]
    radii = rng.uniform(0.8e-3, 1.0e-3, size=n_grains)
    hs = np.array([np.sqrt(shape_area(sh, R) / nodes_per_grain) for sh, R in zip(shapes, radii)])
    dmean = 3 * float(np.mean(hs))
    Rc = dmean / 5
    mat = lambda d: material_2d(d, E=K)  # nu = 1/3 gives K = E
    grains = [make_grain_2d(sh, R, h, rng, Rc, mat) for sh, R, h in zip(shapes, radii, hs)]
    _, cmean, _ = mat(dmean)
    side = int(np.ceil(np.sqrt(n_grains)))
    pitch = 2 * 1.0e-3 + Rc / 2
    centers = []
    for k in range(n_grains):
        ix, iy = k % side, k // side
        centers.append(np.array([1.0e-3 + Rc + ix * pitch, 1.0e-3 + Rc + iy * pitch]))
    width = 2 * (1.0e-3 + Rc) + (side - 1) * pitch
    top = max(float(np.max(g.pos[:, 1] + c[1])) for g, c in zip(grains, centers)) + Rc / 2
    s = assemble(2, grains, centers, [np.zeros(2)] * n_grains, Rc, contact_scale * cmean, 0.8, mu,
                 9.81, 1, [0.0, width, top, 0.0])
    p = Params(dt=2e-8, timesteps=2000, modulo=100, enable_fracture=0)
    p.wall_speed[2] = wall_speed
    return s, p


def subset(s: Setup, grains: List[int]) -> Setup:
    """Setup restricted to the given grains (in the given order)."""
    G = len(grains)
    sl = [s.grain_nodes(g) for g in grains]
    node_offset = np.zeros(G + 1, np.int64)
    node_offset[1:] = np.cumsum([x.stop - x.start for x in sl])
    cat = lambda a: np.concatenate([a[x] for x in sl], axis=0)

    def ragged(off, vals):
        parts = [vals[off[g]:off[g + 1]] for g in grains]
        o = np.zeros(G + 1, np.int64)
        o[1:] = np.cumsum([len(p) for p in parts])
        v = np.concatenate(parts, axis=0) if parts else vals[:0]
        return o, v

    co, cv = ragged(s.conn_offset, s.conn)
    bo, bv = ragged(s.bdry_offset, s.bdry_nodes)
    clo, clv = ragged(s.clamped_offset, s.clamped_nodes)
    gi = np.asarray(grains)
    return replace(s, node_offset=node_offset, pos=cat(s.pos), vol=cat(s.vol), disp=cat(s.disp),
                   vel=cat(s.vel), acc=cat(s.acc), extforce=cat(s.extforce), conn_offset=co,
                   conn=cv, bdry_offset=bo, bdry_nodes=bv, clamped_offset=clo, clamped_nodes=clv,
                   delta=s.delta[gi], rho=s.rho[gi], cnot=s.cnot[gi], snot=s.snot[gi],
                   movable=s.movable[gi], breakable=s.breakable[gi], stoppable=s.stoppable[gi],
                   torque_axis=s.torque_axis[gi], torque_val=s.torque_val[gi])
