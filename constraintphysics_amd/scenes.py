"""Seeded synthetic scenes for the benchmark configurations.

C0: unit cubes (density 1, restitution 0, friction 0.5) in an n x n x n grid
    with 1.1 m spacing, U[-0.05, 0.05] jitter and random yaw, on a FIXED
    ground box (-100,-1,-100)+(200,1,200) — the same generator (SplitMix64) as
    the reference fixture harness, so scenes match body for body.
C3: independent scenes with body counts log-uniform in [8, 2048], each in its
    own FIXED box container sized to C1's body density (20k bodies in 60^3 m^3),
    bodies placed without initial overlap, velocities N(0, 1) m/s.
"""
import math

import numpy as np

import constraintphysics_amd as cpa

# ConvexUnionGeometry::box surface order (Geometry.cpp:14-37), corner-origin points.
BOX_FACES = [[0, 3, 2, 1], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [4, 5, 6, 7]]
DEFAULT_MATERIAL = (1.0, 0.3, 0.6, 1.1)  # Material::default_material (ConvexPrimitive.h:18)
C0_MATERIAL = (1.0, 0.0, 0.5, 0.5)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        u = (self.next() >> 11) * (1.0 / 9007199254740992.0)
        return a + (b - a) * u


def box_points(dx, dy, dz):
    """Corner-origin box points in the reference's vertex order, centred on the centre of mass."""
    p = np.array([[0, 0, 0], [dx, 0, 0], [dx, dy, 0], [0, dy, 0], [0, 0, dz], [dx, 0, dz], [dx, dy, dz], [0, dy, dz]], float)
    return p - np.array([dx, dy, dz]) / 2.0


def mass_props_row(mass, tensor):
    """mass, reference tensor, its inverse as the reference stores it (Mat3::inverse)."""
    return np.concatenate([[mass], np.asarray(tensor).ravel(), cpa.mat3_inverse(tensor).ravel()])


_BOX_BODIES = {}


def box_body(dx, dy, dz, density):
    """The exact-hull layout and mass properties of a box body created centred on the origin."""
    key = (dx, dy, dz, density)
    if key not in _BOX_BODIES:
        _BOX_BODIES[key] = cpa.polyhedron_body(box_points(dx, dy, dz), BOX_FACES, density)
    return _BOX_BODIES[key]


def quat_axis_angle(theta, axis):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(theta / 2.0)
    return np.array([math.cos(theta / 2.0), axis[0] * s, axis[1] * s, axis[2] * s])


class SceneBuilder:
    """Collects shapes and bodies for one World; bodies are added scene by scene."""

    def __init__(self, world):
        self.w = world
        self.shape_cache = {}
        self.rows_shape, self.rows_move, self.rows_mp, self.rows_mat, self.rows_scene = [], [], [], [], []
        self.state = []

    def box_shape(self, dx, dy, dz, density):
        """ConvexUnionGeometry::box centred on the origin (Geometry.cpp:14-37) as a body's polyhedron."""
        key = ("box", dx, dy, dz, density)
        if key not in self.shape_cache:
            h = box_body(dx, dy, dz, density)
            self.shape_cache[key] = (self.w.add_hull_exact(h), h["mass_props"])
        return self.shape_cache[key]

    def hull_shape(self, pts, faces, density):
        h = cpa.polyhedron_body(pts, faces, density)
        return self.w.add_hull_exact(h), h["mass_props"]

    def add(self, shape, mp, movement, material, scene, com, q=(1, 0, 0, 0), vel=(0, 0, 0), angv=(0, 0, 0)):
        self.add_many([shape], [mp], [movement], [material], [scene], [com], [q], [vel], [angv])

    def add_many(self, shapes, mps, moves, mats, scenes, coms, qs, vels, angvs):
        n = len(shapes)
        self.rows_shape.append(np.asarray(shapes, np.int32).reshape(n))
        self.rows_move.append(np.asarray(moves, np.int32).reshape(n))
        self.rows_mp.append(np.asarray(mps, float).reshape(n, 19))
        self.rows_mat.append(np.asarray(mats, float).reshape(n, 4))
        self.rows_scene.append(np.asarray(scenes, np.int32).reshape(n))
        self.state.append(np.concatenate([np.asarray(coms, float).reshape(n, 3), np.asarray(qs, float).reshape(n, 4),
                                          np.asarray(vels, float).reshape(n, 3), np.asarray(angvs, float).reshape(n, 3)], axis=1))

    def finish(self):
        import ctypes
        sh = np.ascontiguousarray(np.concatenate(self.rows_shape), np.int32)
        n = len(sh)
        mv = np.ascontiguousarray(np.concatenate(self.rows_move), np.int32)
        mp = np.ascontiguousarray(np.concatenate(self.rows_mp))
        mt = np.ascontiguousarray(np.concatenate(self.rows_mat))
        sc = np.ascontiguousarray(np.concatenate(self.rows_scene), np.int32)
        first = ctypes.c_int()
        cpa._chk(self.w.L.cp_body_add_many(self.w.w, n, sh.ctypes.data, mv.ctypes.data, mp.ctypes.data, mt.ctypes.data, sc.ctypes.data,
                                           ctypes.byref(first)), "cp_body_add_many")
        self.w.n += n
        st = np.ascontiguousarray(np.concatenate(self.state))
        self.w.set_state(st)
        return st


def build_c0(world, n=10, seed=12345, scene=0, builder=None, offset=(0.0, 0.0, 0.0)):
    """C0 pile: n^3 unit cubes on a ground box (same draws as oracle/ref_harness.cpp build_pile)."""
    b = builder or SceneBuilder(world)
    gs, gmp = b.box_shape(200.0, 1.0, 200.0, 1.0)
    ox, oy, oz = offset
    b.add(gs, gmp, 1, C0_MATERIAL, scene, np.array([0.0 + ox, -0.5 + oy, 0.0 + oz]))
    cs, cmp_ = b.box_shape(1.0, 1.0, 1.0, 1.0)
    rng = SplitMix64(seed)
    for j in range(n):
        for i in range(n):
            for k in range(n):
                cx = 1.1 * i + rng.uniform(-0.05, 0.05)
                cy = 0.6 + 1.1 * j + rng.uniform(-0.05, 0.05)
                cz = 1.1 * k + rng.uniform(-0.05, 0.05)
                yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846)
                b.add(cs, cmp_, 0, C0_MATERIAL, scene, np.array([cx + ox, cy + oy, cz + oz]), quat_axis_angle(yaw, (0, 1, 0)))
    if builder is None:
        return b.finish()
    return None


def c0_description(n=10, seed=12345):
    """The build_c0 pile as plain arrays (one scene; for the CPU oracle): the
    same shapes, mass properties, materials and seeded placement."""
    ground = box_body(200.0, 1.0, 200.0, 1.0)
    cube = box_body(1.0, 1.0, 1.0, 1.0)
    hulls, move, mps, state = [ground], [1], [ground["mass_props"]], [np.array([0.0, -0.5, 0.0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0])]
    rng = SplitMix64(seed)
    for j in range(n):
        for i in range(n):
            for k in range(n):
                cx = 1.1 * i + rng.uniform(-0.05, 0.05)
                cy = 0.6 + 1.1 * j + rng.uniform(-0.05, 0.05)
                cz = 1.1 * k + rng.uniform(-0.05, 0.05)
                yaw = rng.uniform(0.0, 2.0 * 3.14159265358979323846)
                hulls.append(cube)
                move.append(0)
                mps.append(cube["mass_props"])
                state.append(np.concatenate([[cx, cy, cz], quat_axis_angle(yaw, (0, 1, 0)), np.zeros(6)]))
    nb = len(hulls)
    return dict(hulls=hulls, movement=np.array(move), mass_props=np.array(mps), material=np.tile(C0_MATERIAL, (nb, 1)),
                state=np.array(state))


# ----------------------------------------------------------------------------- C1 / C3 mix
C1_DENSITY = 20000.0 / 60.0 ** 3


def random_hull(rng, n_min=8, n_max=32):
    """Convex hull of 8..32 seeded points on an ellipsoid with semi-axes U[0.3, 0.8]."""
    from scipy.spatial import ConvexHull
    n = int(rng.integers(n_min, n_max + 1))
    axes = rng.uniform(0.3, 0.8, 3)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = d * axes
    h = ConvexHull(pts)
    used = np.unique(h.simplices)
    remap = -np.ones(n, int)
    remap[used] = np.arange(len(used))
    faces = [list(remap[s]) for s in h.simplices]
    return pts[used], faces


def random_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


class HullLibrary:
    """A seeded library of convex hull shapes (body frame, centre of mass at
    the origin) instanced by many bodies; geometry is independent of any world."""

    def __init__(self, count, seed, density=1.0):
        rng = np.random.default_rng(seed)
        self.items = []
        for _ in range(count):
            pts, faces = random_hull(rng)
            body = cpa.polyhedron_body(pts, faces, density)  # the body createRigidBody makes of these points
            r = float(np.max(np.linalg.norm(body["pts"], axis=1))) + 0.05
            self.items.append(dict(pts=pts, faces=faces, body=body, mp=body["mass_props"], r=r))
        self.r = np.array([it["r"] for it in self.items])
        self.mp = np.array([it["mp"] for it in self.items])


def add_container(b, scene, L, material=DEFAULT_MATERIAL, origin=(0.0, 0.0, 0.0)):
    t = 1.0
    ox, oy, oz = origin
    walls = [((L + 2 * t, t, L + 2 * t), (L / 2, -t / 2, L / 2)), ((L + 2 * t, t, L + 2 * t), (L / 2, L + t / 2, L / 2)),
             ((t, L, L), (-t / 2, L / 2, L / 2)), ((t, L, L), (L + t / 2, L / 2, L / 2)),
             ((L, L, t), (L / 2, L / 2, -t / 2)), ((L, L, t), (L / 2, L / 2, L + t / 2))]
    for (dx, dy, dz), c in walls:
        s, mp = b.box_shape(dx, dy, dz, 1.0)
        b.add(s, mp, 1, material, scene, np.array([c[0] + ox, c[1] + oy, c[2] + oz]))


def place_bodies(rng, radii, L):
    """Positions without overlap: distinct random cells of a lattice of pitch
    2 * max radius, jittered uniformly inside each cell's slack."""
    n = len(radii)
    pitch = 2.0 * float(radii.max())
    m = max(1, int(L // pitch))
    while m ** 3 < n:
        m += 1
    cells = rng.choice(m ** 3, size=n, replace=False)
    ijk = np.stack([cells // (m * m), (cells // m) % m, cells % m], axis=1).astype(float)
    slack = (pitch / 2.0 - radii)[:, None]
    return (ijk + 0.5) * pitch + rng.uniform(-1.0, 1.0, (n, 3)) * slack


def c3_sizes(n_scenes, seed, lo=8, hi=2048):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_scenes))).astype(int)


_LIBS = {}

SPHERE_PI = 3.1415926535  # the reference's PI macro (Quaternion.h:4), used for sphere mass (RigidBody.cpp:556)

# Primitive mix of the mixed-primitive configuration: 40% spheres, 30% cylinders,
# 30% hulls.  Body kinds: 0 hull, 1 sphere, 2 cylinder.
MIX = (("sphere", 0.4), ("cylinder", 0.3), ("hull", 0.3))
KIND = {"hull": 0, "sphere": 1, "cylinder": 2}


def hull_library(seed, count=256):
    key = (seed, count)
    if key not in _LIBS:
        _LIBS[key] = HullLibrary(count, seed ^ 0x5EED)
    return _LIBS[key]


def sphere_library(seed, count=256, density=1.0):
    """Seeded sphere radii U[0.3, 0.7] with mass properties (RigidBody.cpp:552-563)."""
    rng = np.random.default_rng(seed ^ 0x5FE4E)
    r = rng.uniform(0.3, 0.7, count)
    m = (4.0 / 3.0) * SPHERE_PI * r * r * r * density
    k = 2.0 / 5.0 * m * r * r
    mp = np.array([mass_props_row(mi, np.eye(3) * ki) for mi, ki in zip(m, k)])
    return r, mp


def cylinder_library(seed, count=256, density=1.0):
    """Seeded cylinders r ~ U[0.2, 0.5], h ~ U[0.5, 1.5] (height axis +y in the
    body frame): reference geometry, mass properties, bounding radius."""
    rng = np.random.default_rng(seed ^ 0xC711)
    r = rng.uniform(0.2, 0.5, count)
    h = rng.uniform(0.5, 1.5, count)
    geo = [cpa.cylinder_build(ri, hi) for ri, hi in zip(r, h)]
    mp = np.array([mass_props_row(*cpa.cylinder_mass_properties(ri, hi, density)) for ri, hi in zip(r, h)])
    re = r / math.cos(3.1415926535 / 8)  # the cap polygons' encapsulating radius
    return geo, mp, np.sqrt(re * re + (h / 2) ** 2)


def scene_bodies(sizes, seed, s, lib, sph_r, cyl_r=None, mix=MIX):
    """Dynamic bodies of scene s: kind (0 hull, 1 sphere, 2 cylinder), library
    pick, container side, positions, orientations, velocities."""
    n = int(sizes[s])
    rng = np.random.default_rng(seed + s)
    w = np.array([m[1] for m in mix])
    kinds = rng.choice(len(mix), size=n, p=w / w.sum())
    kind = np.array([KIND[mix[k][0]] for k in kinds])
    picks = rng.integers(0, len(lib.items), n)
    if cyl_r is None:
        cyl_r = cylinder_library(seed, len(lib.items))[2]
    radii = np.select([kind == 0, kind == 1], [lib.r[picks], sph_r[picks] + 0.05], cyl_r[picks] + 0.05)
    pitch = 2.0 * float(radii.max())
    L = max((n / C1_DENSITY) ** (1.0 / 3.0), pitch * math.ceil(n ** (1.0 / 3.0)))
    pos = place_bodies(rng, radii, L)
    vel = rng.normal(size=(n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return kind, picks, L, pos, q, vel


CONTAINER_WALLS = 6


def scene_offsets(sizes):
    """First body index of each scene of a batched world (container walls, then the scene's bodies)."""
    cnt = np.asarray(sizes) + CONTAINER_WALLS
    return np.concatenate([[0], np.cumsum(cnt)])


def scene_contacts(contacts, off, s):
    """Contact cache rows of scene s of a batched world, renumbered to a world holding only that scene
    (body ids id - first, MagicID primitive halves body_id * 16 + primitive shifted with them)."""
    a, b = int(off[s]), int(off[s + 1])
    ii, mg, ff = contacts
    sel = (ii[:, 0] > a) & (ii[:, 0] <= b)
    ii, mg, ff = ii[sel].copy(), mg[sel].copy(), ff[sel].copy()
    ii[:, :2] -= a
    shift = np.uint64(16 * a)
    lo = (mg[:, 0] & np.uint64(0xFFFFFFFF)) - shift
    hi = (mg[:, 0] >> np.uint64(32)) - shift
    mg[:, 0] = lo | (hi << np.uint64(32))
    return ii, mg, ff


def scene_snapshot(state, contacts, off, s, inertia=None):
    """State rows (and world inertia rows when given) and contact cache of scene s, renumbered to a
    world holding only that scene."""
    a, b = int(off[s]), int(off[s + 1])
    out = (state[a:b], scene_contacts(contacts, off, s))
    return out + ((inertia[a:b],) if inertia is not None else ())


def container_walls(L, t=1.0):
    return [((L + 2 * t, t, L + 2 * t), (L / 2, -t / 2, L / 2)), ((L + 2 * t, t, L + 2 * t), (L / 2, L + t / 2, L / 2)),
            ((t, L, L), (-t / 2, L / 2, L / 2)), ((t, L, L), (L + t / 2, L / 2, L / 2)),
            ((L, L, t), (L / 2, L / 2, -t / 2)), ((L, L, t), (L / 2, L / 2, L + t / 2))]


def mix_description():
    tot = sum(m[1] for m in MIX)
    parts = ", ".join(f"{100 * m[1] / tot:.0f}% {m[0]}s" for m in MIX)
    return (f"{parts}; hulls: 8-32 points on ellipsoids, semi-axes U[0.3,0.8] (256-shape seeded library); "
            f"spheres r~U[0.3,0.7], cylinders r~U[0.2,0.5] h~U[0.5,1.5] (256-entry seeded libraries); box walls")


def build_batched(world, sizes, seed, scene_indices=None, hull_lib_size=256):
    """Batched scenes of the mixed-primitive distribution: scene_indices picks
    which scenes of the seeded sequence to build (default all), each becoming
    one world scene."""
    b = SceneBuilder(world)
    lib = hull_library(seed, hull_lib_size)
    sph_r, sph_mp = sphere_library(seed, hull_lib_size)
    lib_sid = np.array([world.add_hull_exact(it["body"]) for it in lib.items])
    sph_sid = np.array([world.add_sphere((0.0, 0.0, 0.0), r) for r in sph_r])
    cyl_geo, cyl_mp, cyl_r = cylinder_library(seed, hull_lib_size)
    cyl_sid = np.array([world.add_cylinder(g) for g in cyl_geo])
    idx = range(len(sizes)) if scene_indices is None else scene_indices
    total = 0
    for out_s, s in enumerate(idx):
        kind, picks, L, pos, q, vel = scene_bodies(sizes, seed, s, lib, sph_r, cyl_r)
        for (dx, dy, dz), c in container_walls(L):
            sid, mp = b.box_shape(dx, dy, dz, 1.0)
            b.add(sid, mp, 1, DEFAULT_MATERIAL, out_s, np.array(c))
        n = len(picks)
        shp = np.select([kind == 0, kind == 1], [lib_sid[picks], sph_sid[picks]], cyl_sid[picks])
        mps = np.select([(kind == 0)[:, None], (kind == 1)[:, None]], [lib.mp[picks], sph_mp[picks]], cyl_mp[picks])
        b.add_many(shp, mps, np.zeros(n, np.int32), np.tile(DEFAULT_MATERIAL, (n, 1)), np.full(n, out_s), pos, q, vel, np.zeros((n, 3)))
        total += n
    b.finish()
    return dict(dynamic_bodies=total, mix=mix_description())


def scene_description(sizes, seed, s, hull_lib_size=256, mix=MIX):
    """Scene s as plain arrays (for the CPU oracle): per body the hull in the
    exact-hull layout, movement, mass properties, material, and the state."""
    lib = hull_library(seed, hull_lib_size)
    sph_r, sph_mp = sphere_library(seed, hull_lib_size)
    cyl_geo, cyl_mp, cyl_r = cylinder_library(seed, hull_lib_size)
    kind, picks, L, pos, q, vel = scene_bodies(sizes, seed, s, lib, sph_r, cyl_r, mix)
    hulls, move, mps, mats, state = [], [], [], [], []
    for (dx, dy, dz), c in container_walls(L):
        h = box_body(dx, dy, dz, 1.0)
        hulls.append(h)
        move.append(1)
        mps.append(h["mass_props"])
        mats.append(DEFAULT_MATERIAL)
        state.append(np.concatenate([c, [1, 0, 0, 0], [0, 0, 0], [0, 0, 0]]))
    for i, p in enumerate(picks):
        if kind[i] == 2:
            hulls.append(cyl_geo[p])
            move.append(0)
            mps.append(cyl_mp[p])
            mats.append(DEFAULT_MATERIAL)
            state.append(np.concatenate([pos[i], q[i], vel[i], [0, 0, 0]]))
            continue
        if kind[i] == 1:
            hulls.append(dict(sphere=np.array([0.0, 0.0, 0.0, sph_r[p]])))
            move.append(0)
            mps.append(sph_mp[p])
            mats.append(DEFAULT_MATERIAL)
            state.append(np.concatenate([pos[i], q[i], vel[i], [0, 0, 0]]))
            continue
        hulls.append(lib.items[p]["body"])
        move.append(0)
        mps.append(lib.mp[p])
        mats.append(DEFAULT_MATERIAL)
        state.append(np.concatenate([pos[i], q[i], vel[i], [0, 0, 0]]))
    return dict(hulls=hulls, movement=np.array(move), mass_props=np.array(mps), material=np.array(mats), state=np.array(state))


# ----------------------------------------------------------------------------- holonomic chains on a heightfield
C2_MATERIAL = (1.0, 0.0, 0.5, 0.5)


def heightfield(grid, cell, amp, seed, lattice=8):
    """Points and triangles of a grid x grid heightfield centred on the origin,
    split as generateGridMeshInput splits cells (Geometry.cpp:589-611), heights
    amp * value noise (U[-1, 1] lattice every `lattice` cells, smoothstep
    bilinear)."""
    rng = np.random.default_rng(seed ^ 0x7E22)
    L = grid // lattice + 2
    lat = rng.uniform(-1.0, 1.0, (L, L))
    g = np.arange(grid + 1) / lattice
    ix = np.floor(g).astype(int)
    f = g - ix
    f = f * f * (3 - 2 * f)
    a = lat[ix[:, None], ix[None, :]]
    b = lat[ix[:, None], ix[None, :] + 1]
    c = lat[ix[:, None] + 1, ix[None, :]]
    d = lat[ix[:, None] + 1, ix[None, :] + 1]
    h = (a * (1 - f[None, :]) + b * f[None, :]) * (1 - f[:, None]) + (c * (1 - f[None, :]) + d * f[None, :]) * f[:, None]
    half = 0.5 * grid * cell
    jj, ii = np.meshgrid(np.arange(grid + 1), np.arange(grid + 1), indexing="ij")
    pts = np.stack([-half + cell * ii, amp * h, -half + cell * jj], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(1, grid + 1), np.arange(1, grid + 1), indexing="xy")
    i, j = i.ravel(), j.ravel()
    i1, i2 = i - 1 + (grid + 1) * (j - 1), i + (grid + 1) * (j - 1)
    i3, i4 = i + (grid + 1) * j, i - 1 + (grid + 1) * j
    tris = np.empty((2 * len(i), 3), np.int64)
    tris[0::2] = np.stack([i3, i2, i1], 1)
    tris[1::2] = np.stack([i4, i3, i1], 1)
    return pts, tris


def c2_description(n_chains=512, links=16, grid=256, cell=0.5, amp=2.0, seed=2024, hinge_at=8):
    """Holonomic chains on a heightfield: n_chains chains of `links` cylinder
    links (r 0.15, h 0.8, 0.1 apart, laid along +x) joined end to end by
    ball-sockets, the hinge_at-th joint of each chain a hinge about +z (so a
    16-link chain's system has degree 14 * 3 + 5 = 47), dropped onto a static
    grid x grid heightfield of value noise with amplitude amp.  Body 0 is the
    mesh; joints as (type, id1, id2) and geo14 rows."""
    rng = np.random.default_rng(seed)
    pts, tris = heightfield(grid, cell, amp, seed)
    r, h, gap = 0.15, 0.8, 0.1
    geo = cpa.cylinder_build(r, h)
    m, t = cpa.cylinder_mass_properties(r, h, 1.0)
    mp = mass_props_row(m, t)
    span = grid * cell
    length = links * (h + gap)
    per_row = max(1, int((span - 2.0) // (length + 1.0)))
    rows = (n_chains + per_row - 1) // per_row
    dz = (span - 2.0) / max(rows, 1)
    q = quat_axis_angle(-math.pi / 2.0, (0, 0, 1))
    hulls = [dict(mesh=pts, tris=tris, tri_mat=np.tile(C2_MATERIAL[1:], (len(tris), 1)))]
    move, mps, mats, state = [1], [np.zeros(19)], [np.array(C2_MATERIAL)], [np.array([0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0], float)]
    info, jgeo = [], []
    a = h / 2.0 + gap / 2.0
    for c in range(n_chains):
        row, col = divmod(c, per_row)
        x0 = -span / 2 + 1.0 + col * (length + 1.0)
        z0 = -span / 2 + 1.0 + (row + 0.5) * dz
        y0 = amp + 1.0 + rng.uniform(0.0, 0.5)
        for i in range(links):
            hulls.append(geo)
            move.append(0)
            mps.append(mp)
            mats.append(np.array(C2_MATERIAL))
            v = rng.uniform(-0.5, 0.5, 3)
            state.append(np.concatenate([[x0 + i * (h + gap), y0, z0], q, v, [0, 0, 0]]))
            if i:
                idx = len(hulls) - 1  # body index of this link
                hinge = i == hinge_at
                info.append((5 if hinge else 3, idx, idx + 1))  # ids are index + 1: previous link id = idx, this one idx + 1
                jgeo.append([0, a, 0, 0, -a, 0] + ([0, 0, 1, 0, 0, 1] if hinge else [0] * 6) + [15000.0, 15000.0 if hinge else 0.0])
    return dict(hulls=hulls, movement=np.array(move), mass_props=np.array(mps), material=np.array(mats), state=np.array(state),
                joints=(np.array(info, np.int64), np.array(jgeo, float)))


def build_description(world, desc):
    """Adds a description's bodies (static mesh, spheres, cylinders, exact
    hulls) and joints to a world as one scene and sets the state."""
    for b, sh in enumerate(desc["hulls"]):
        if "mesh" in sh:
            world.add_body(world.add_mesh(sh["mesh"], sh["tris"], sh["tri_mat"]), 1, np.zeros(19), desc["material"][b], 0)
            continue
        if "sphere" in sh:
            sid = world.add_sphere(sh["sphere"][:3], sh["sphere"][3])
        elif "cylinder" in sh:
            sid = world.add_cylinder(sh)
        else:
            sid = world.add_hull_exact(sh)
        world.add_body(sid, int(desc["movement"][b]), desc["mass_props"][b], desc["material"][b], 0)
    if "joints" in desc:
        info, geo = desc["joints"]
        for (kind, id1, id2), g in zip(info, geo):
            world.add_joint(int(kind), int(id1) - 1, int(id2) - 1, g)
    world.set_state(desc["state"])
    return world
