"""MI355X-native per-step contact pipeline of ConstraintPhysics.

Python access to libcp_hip.so through its C ABI (include/cp_abi.h).  The
library is built in-tree by ``python -m constraintphysics_amd.build``; loading
fails loudly when it is missing — there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcp_hip.so")

CP_PGS_COLORED = 0
CP_PGS_REFERENCE_ORDER = 1
CP_PGS_ISLAND_COLORED = 2


CP_ERR_NONFINITE = 6


class CpError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("vel_iters", ctypes.c_int),
        ("pos_iters", ctypes.c_int),
        ("hol_iters", ctypes.c_int),
        ("substeps", ctypes.c_int),
        ("cfm", ctypes.c_double),
        ("contact_hardness", ctypes.c_double),
        ("contact_life", ctypes.c_int),
        ("gyro_ticks", ctypes.c_int),
        ("pgs_mode", ctypes.c_int),
        ("solver_threads", ctypes.c_int),
        ("diagnostics", ctypes.c_int),
        ("defect_fixes", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("pgs_class_bodies", ctypes.c_int * 3),
        ("pgs_class_threads", ctypes.c_int * 3),
        ("graph", ctypes.c_int),
        ("fused", ctypes.c_int),
        ("sleeping", ctypes.c_int),
        ("sleep_tick_threshold", ctypes.c_int),
        ("sleep_vel_coeff", ctypes.c_double),
        ("sleep_ang_vel_eps", ctypes.c_double),
        ("sleep_accel_coeff", ctypes.c_double),
        ("sleep_delay", ctypes.c_double),
    ]


_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int


def lib():
    """Load libcp_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CpError(f"{LIB_PATH} is missing: build it with `python -m constraintphysics_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "cp_config_default": (None, [_P]),
        "cp_status_string": (ctypes.c_char_p, [_I]),
        "cp_world_create": (_I, [_P, _P]),
        "cp_world_destroy": (_I, [_P]),
        "cp_world_set_config": (_I, [_P, _P]),
        "cp_shape_add_hull_exact": (_I, [_P, _I, _P, _I] + [_P] * 9 + [_I, _P, _I, _P, _P]),
        "cp_shape_add_polyhedron": (_I, [_P, _I, _P, _I, _P, _P, _P]),
        "cp_shape_add_sphere": (_I, [_P, _P, ctypes.c_double, _P]),
        "cp_shape_add_cylinder": (_I, [_P, _P, _P, _P, _P, _P]),
        "cp_cylinder_build": (_I, [_P, ctypes.c_double, ctypes.c_double, _P, _P, _P, _P]),
        "cp_polyhedron_mass_properties": (_I, [_I, _P, _I, _P, _P, ctypes.c_double, _P]),
        "cp_polyhedron_body": (_I, [_I, _P, _I, _P, _P, ctypes.c_double, _P, _P, _P]),
        "cp_polyhedron_build": (_I, [_I, _P, _I, _P, _P] + [_P] * 8 + [_I, _P, _P, _P, _P]),
        "cp_body_add_many": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
        "cp_body_add": (_I, [_P, _I, _I, _P, _P, _I, _P]),
        "cp_world_body_count": (_I, [_P, _P]),
        "cp_world_set_state": (_I, [_P, _P, _P]),
        "cp_world_get_state": (_I, [_P, _P]),
        "cp_world_get_aabbs": (_I, [_P, _P]),
        "cp_world_get_inertia": (_I, [_P, _P]),
        "cp_world_set_contacts": (_I, [_P, _I, _P, _P, _P]),
        "cp_world_get_contact_count": (_I, [_P, _P]),
        "cp_world_get_contacts": (_I, [_P, _I, _P, _P, _P, _P]),
        "cp_world_set_node_order": (_I, [_P, _I, _P]),
        "cp_reference_node_order": (_I, [_I, _P]),
        "cp_world_step": (_I, [_P, _I]),
        "cp_broadphase": (_I, [_P, _I, _P, _P]),
        "cp_narrowphase": (_I, [_P, _I, _P, _P, _P, _P, _P]),
        "cp_narrowphase_noalloc": (_I, [_P, _I, _P, _P, _P, _P, _P]),
        "cp_world_last_islands": (_I, [_P, _I, _P, _I, _P, _P, _P]),
        "cp_world_last_timing": (_I, [_P, _P]),
        "cp_world_last_residual": (_I, [_P, _P]),
        "cp_world_last_scene_contacts": (_I, [_P, _I, _P, _P]),
        "cp_world_scene_flags": (_I, [_P, _I, _P, _P]),
        "cp_world_get_sleep_state": (_I, [_P, _P]),
        "cp_world_wake_bodies": (_I, [_P, _I, _P]),
        "cp_world_set_sleep_disabled": (_I, [_P, _I, _P, _I]),
        "cp_world_raycast": (_I, [_P, _I, _P, _P, _P, _P, _P]),
        "cp_world_save_snapshot": (_I, [_P, ctypes.c_char_p]),
        "cp_world_load_snapshot": (_I, [_P, ctypes.c_char_p]),
        "cp_world_trajectory_open": (_I, [_P, ctypes.c_char_p]),
        "cp_world_trajectory_close": (_I, [_P]),
        "cp_trajectory_frames": (_I, [ctypes.c_char_p, _I, _P]),
        "cp_trajectory_read": (_I, [ctypes.c_char_p, _I, _I, _P]),
        "cp_world_scene_snapshot": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
        "cp_world_admit_scene_resume": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
        "cp_shape_add_mesh": (_I, [_P, _I, _P, _I, _P, _P, _P]),
        "cp_narrowphase_multi": (_I, [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P]),
        "cp_world_serving_init": (_I, [_P, _I]),
        "cp_world_admit_scene": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
        "cp_world_retire_scene": (_I, [_P, _I]),
        "cp_world_serving_commit": (_I, [_P]),
        "cp_world_get_scene_state": (_I, [_P, _I, _P, _P]),
        "cp_world_get_scene_states": (_I, [_P, _I, _P, _I, _P, _P, _P]),
        "cp_world_serving_info": (_I, [_P, _P]),
        "cp_joint_add": (_I, [_P, _I, _I, _I, _P, _P]),
        "cp_world_joint_count": (_I, [_P, _P]),
        "cp_joint_set_drive": (_I, [_P, _I, _I, ctypes.c_double, ctypes.c_double]),
        "cp_joint_set_cfm": (_I, [_P, _I, ctypes.c_double]),
        "cp_world_set_joint_state": (_I, [_P, _P]),
        "cp_world_get_joint_state": (_I, [_P, _P]),
        "cp_world_get_joint_rows": (_I, [_P, _I, _P, _P, _P, _P, _P]),
        "cp_world_set_holonomic_systems": (_I, [_P, _I, _P, _P]),
        "cp_world_holonomic_system_count": (_I, [_P, _P]),
        "cp_world_get_holonomic_system": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return [ln.split("(")[0].split()[-1].lstrip("*") for ln in _abi_decls()]


def _abi_decls():
    path = os.path.join(os.path.dirname(HERE), "include", "cp_abi.h")
    out = []
    buf = ""
    for line in open(path):
        s = line.strip()
        if buf or s.startswith(("int cp_", "void cp_", "const char* cp_")):
            buf += " " + s
            if s.endswith(";"):
                out.append(buf.strip())
                buf = ""
    return out


def _chk(st, what=""):
    if st != 0:
        msg = lib().cp_status_string(st).decode()
        raise CpError(f"{what}: {msg} ({st})")


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def default_config(**kw):
    c = Config()
    lib().cp_config_default(ctypes.byref(c))
    for k, v in kw.items():
        _set_field(c, k, v)
    return c


def _set_field(c, k, v):
    if k == "gravity":
        c.gravity = (ctypes.c_double * 3)(*v)
    elif k in ("pgs_class_bodies", "pgs_class_threads"):
        setattr(c, k, (ctypes.c_int * 3)(*[int(x) for x in v]))
    else:
        setattr(c, k, v)


def reference_node_order(n):
    out = np.zeros(n, np.int64)
    _chk(lib().cp_reference_node_order(n, out.ctypes.data), "cp_reference_node_order")
    return out


def polyhedron_mass_properties(pts, faces, density):
    pts = _a(pts, np.float64)
    off = np.zeros(len(faces) + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in faces])
    fv = _a(np.concatenate([np.asarray(f) for f in faces]), np.int64)
    out = np.zeros(13)
    _chk(lib().cp_polyhedron_mass_properties(len(pts), pts.ctypes.data, len(faces), off.ctypes.data, fv.ctypes.data, float(density),
                                             out.ctypes.data), "cp_polyhedron_mass_properties")
    return out[0], out[1:10].reshape(3, 3), out[10:13]


def trajectory_frames(path, n_bodies):
    """Frames in a recorded trajectory file of n_bodies bodies (7 doubles per body per frame)."""
    n = ctypes.c_int()
    _chk(lib().cp_trajectory_frames(os.fsencode(path), int(n_bodies), ctypes.byref(n)), "cp_trajectory_frames")
    return n.value


def trajectory_read(path, n_bodies, frame):
    """Frame `frame` of a recorded trajectory: [n_bodies, 7] = com xyz, orientation r i j k."""
    out = np.zeros((int(n_bodies), 7))
    _chk(lib().cp_trajectory_read(os.fsencode(path), int(n_bodies), int(frame), out.ctypes.data), "cp_trajectory_read")
    return out


def polyhedron_body(pts, faces, density):
    """The polyhedron of a body created from Polyhedron(points, faces) with
    the given density (RigidBody ctor, RigidBody.cpp:13-38): the exact-hull
    layout (Gauss map of the points as given, body-frame points shifted to
    the centre of mass) plus ``mass_props`` (mass, reference tensor, its
    Mat3::inverse) and ``com`` (centre of mass in the input frame)."""
    h = polyhedron_build(pts, faces)
    pts = _a(pts, np.float64)
    off = np.zeros(len(faces) + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in faces])
    fv = _a(np.concatenate([np.asarray(f) for f in faces]), np.int64)
    body = np.zeros_like(pts)
    mp = np.zeros(19)
    com = np.zeros(3)
    _chk(lib().cp_polyhedron_body(len(pts), pts.ctypes.data, len(faces), off.ctypes.data, fv.ctypes.data, float(density),
                                  body.ctypes.data, mp.ctypes.data, com.ctypes.data), "cp_polyhedron_body")
    h["pts"] = body
    # reference_geometry's Gauss normals: the identity rotation applied (Mat3 * Vec3, Mat3.h:70-74)
    g = np.asarray(h["gauss_normal"], np.float64)
    x, y, z = g[:, 0], g[:, 1], g[:, 2]
    h["gauss_normal"] = np.stack([(1.0 * x + 0.0 * y) + 0.0 * z, (0.0 * x + 1.0 * y) + 0.0 * z, (0.0 * x + 0.0 * y) + 1.0 * z], axis=1)
    h["mass_props"] = mp
    h["com"] = com
    return h


def mat3_inverse(m):
    """Mat3::inverse (Mat3.h:57-68): cofactors divided by the determinant, in
    the reference's operation order (the inverse tensor a body stores)."""
    (a, b, c), (d, e, f), (g, h, i) = [[float(x) for x in row] for row in np.asarray(m, float).reshape(3, 3)]
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    o = [(e * i - f * h), (c * h - b * i), (b * f - c * e), (f * g - d * i), (a * i - c * g), (c * d - a * f), (d * h - e * g),
         (b * g - a * h), (a * e - b * d)]
    return np.array([x / det for x in o]).reshape(3, 3)


def cylinder_build(radius, height, center=(0.0, 0.0, 0.0)):
    """Reference geometry of Cylinder(center, radius, height) with height axis +y."""
    c = _a(center, np.float64)
    out = dict(cylinder=np.zeros(8), gauss=np.zeros((14, 3)), top=np.zeros((8, 3)), bot=np.zeros((8, 3)))
    _chk(lib().cp_cylinder_build(c.ctypes.data, float(radius), float(height), out["cylinder"].ctypes.data, out["gauss"].ctypes.data,
                                 out["top"].ctypes.data, out["bot"].ctypes.data), "cp_cylinder_build")
    return out


def cylinder_mass_properties(radius, height, density=1.0):
    """Mass and inertia tensor about the centre of a solid cylinder with height
    axis +y (RigidBody.cpp:603-631)."""
    m = 3.1415926535 * radius * radius * height * density
    t = np.zeros((3, 3))
    t[0, 0] = t[2, 2] = m * (height * height / 12 + radius * radius / 4)
    t[1, 1] = m * radius * radius / 2
    return m, t


def polyhedron_build(pts, faces):
    """Host-built polyhedron (surfaces, winding, Gauss map) in the exact-hull layout."""
    pts = _a(pts, np.float64)
    nf = len(faces)
    off = np.zeros(nf + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in faces])
    fv = _a(np.concatenate([np.asarray(f) for f in faces]), np.int64)
    fvo = np.zeros_like(fv)
    sid, ncalc, ndir, gred = (np.zeros(nf, np.int64) for _ in range(4))
    gn = np.zeros((nf, 3))
    gint = np.zeros((nf, 3), np.int64)
    gval = np.zeros((nf, 3))
    cap = int(off[-1])
    arcs = np.zeros((cap, 2), np.int64)
    edges = np.zeros((cap, 2), np.int64)
    na, ne = ctypes.c_int(), ctypes.c_int()
    _chk(lib().cp_polyhedron_build(len(pts), pts.ctypes.data, nf, off.ctypes.data, fv.ctypes.data, fvo.ctypes.data, sid.ctypes.data,
                                   ncalc.ctypes.data, ndir.ctypes.data, gn.ctypes.data, gred.ctypes.data, gint.ctypes.data,
                                   gval.ctypes.data, cap, arcs.ctypes.data, ctypes.byref(na), edges.ctypes.data, ctypes.byref(ne)),
         "cp_polyhedron_build")
    return dict(pts=pts, face_off=off, face_verts=fvo, surface_id=sid, normal_calc=ncalc, normal_dir=ndir, gauss_normal=gn,
                gauss_red=gred, gauss_ints=gint, gauss_vals=gval, arcs=arcs[:na.value], edges=edges[:ne.value])


class World:
    """A device-resident world (one or many scenes)."""

    def __init__(self, config=None, **kw):
        self.L = lib()
        self.cfg = config if config is not None else default_config(**kw)
        w = ctypes.c_void_p()
        _chk(self.L.cp_world_create(ctypes.byref(self.cfg), ctypes.byref(w)), "cp_world_create")
        self.w = w
        self.n = 0

    def close(self):
        if self.w:
            self.L.cp_world_destroy(self.w)
            self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, **kw):
        for k, v in kw.items():
            _set_field(self.cfg, k, v)
        _chk(self.L.cp_world_set_config(self.w, ctypes.byref(self.cfg)), "cp_world_set_config")

    # ---- shapes / bodies
    def add_hull_exact(self, h):
        arr = [_a(h["pts"], np.float64), _a(h["face_off"], np.int64), _a(h["face_verts"], np.int64), _a(h["surface_id"], np.int64),
               _a(h["normal_calc"], np.int64), _a(h["normal_dir"], np.int64), _a(h["gauss_normal"], np.float64),
               _a(h["gauss_red"], np.int64), _a(h["gauss_ints"], np.int64), _a(h["gauss_vals"], np.float64), _a(h["arcs"], np.int64),
               _a(h["edges"], np.int64)]
        sid = ctypes.c_int()
        _chk(self.L.cp_shape_add_hull_exact(self.w, len(arr[0]), arr[0].ctypes.data, len(arr[1]) - 1, *[a.ctypes.data for a in arr[1:10]],
                                            len(arr[10]), arr[10].ctypes.data, len(arr[11]), arr[11].ctypes.data, ctypes.byref(sid)),
             "cp_shape_add_hull_exact")
        return sid.value

    def add_sphere(self, center, radius):
        c = _a(center, np.float64)
        sid = ctypes.c_int()
        _chk(self.L.cp_shape_add_sphere(self.w, c.ctypes.data, float(radius), ctypes.byref(sid)), "cp_shape_add_sphere")
        return sid.value

    def add_cylinder(self, cyl):
        """cyl: dict(cylinder=[centre xyz, axis xyz, r, h], gauss=[14, 3], top=[8, 3], bot=[8, 3])."""
        arr = [_a(cyl["cylinder"], np.float64), _a(cyl["gauss"], np.float64), _a(cyl["top"], np.float64), _a(cyl["bot"], np.float64)]
        sid = ctypes.c_int()
        _chk(self.L.cp_shape_add_cylinder(self.w, *[a.ctypes.data for a in arr], ctypes.byref(sid)), "cp_shape_add_cylinder")
        return sid.value

    def add_mesh(self, points, tris, tri_mat):
        """Static triangle mesh: points [P,3], triangles [T,3] (point indices),
        per-triangle (restitution, kinetic mu, static mu) [T,3]."""
        p = _a(points, np.float64).reshape(-1, 3)
        t = _a(tris, np.int64).reshape(-1, 3)
        m = _a(tri_mat, np.float64).reshape(-1, 3)
        sid = ctypes.c_int()
        _chk(self.L.cp_shape_add_mesh(self.w, len(p), p.ctypes.data, len(t), t.ctypes.data, m.ctypes.data, ctypes.byref(sid)),
             "cp_shape_add_mesh")
        return sid.value

    def narrowphase_multi(self, pairs, cap=None):
        """Every manifold of the given (id_lo, id_hi) pairs: list of
        (pair index, normal+depth[4], points[n,4], magic[n,2])."""
        pairs = _a(pairs, np.int64).reshape(-1, 2)
        cap = cap or max(64, 4 * len(pairs))
        idx = np.zeros(cap, np.int64)
        counts = np.zeros(cap, np.int32)
        heads = np.zeros((cap, 4))
        pts = np.zeros((cap, 4, 4))
        mg = np.zeros((cap, 4, 2), np.uint64)
        n = ctypes.c_int()
        _chk(self.L.cp_narrowphase_multi(self.w, len(pairs), pairs.ctypes.data, cap, idx.ctypes.data, counts.ctypes.data, heads.ctypes.data,
                                         pts.ctypes.data, mg.ctypes.data, ctypes.byref(n)), "cp_narrowphase_multi")
        return [(int(idx[k]), heads[k], pts[k, :counts[k]], mg[k, :counts[k]]) for k in range(n.value)]

    def add_polyhedron(self, pts, faces):
        pts = _a(pts, np.float64)
        off = np.zeros(len(faces) + 1, np.int64)
        off[1:] = np.cumsum([len(f) for f in faces])
        fv = _a(np.concatenate([np.asarray(f) for f in faces]), np.int64)
        sid = ctypes.c_int()
        _chk(self.L.cp_shape_add_polyhedron(self.w, len(pts), pts.ctypes.data, len(faces), off.ctypes.data, fv.ctypes.data,
                                            ctypes.byref(sid)), "cp_shape_add_polyhedron")
        return sid.value

    def add_body(self, shape, movement, mass_props, material, scene=0):
        mp = _a(mass_props, np.float64)
        mt = _a(material, np.float64)
        bi = ctypes.c_int()
        _chk(self.L.cp_body_add(self.w, shape, movement, mp.ctypes.data, mt.ctypes.data, scene, ctypes.byref(bi)), "cp_body_add")
        self.n += 1
        return bi.value

    # ---- state
    def set_state(self, state, inertia=None):
        st = _a(state, np.float64)
        assert st.shape == (self.n, 13)
        if inertia is not None:
            it = _a(inertia, np.float64)
            assert it.shape == (self.n, 18)
            _chk(self.L.cp_world_set_state(self.w, st.ctypes.data, it.ctypes.data), "cp_world_set_state")
        else:
            _chk(self.L.cp_world_set_state(self.w, st.ctypes.data, None), "cp_world_set_state")

    def state(self):
        s = np.zeros((self.n, 13))
        _chk(self.L.cp_world_get_state(self.w, s.ctypes.data), "cp_world_get_state")
        return s

    def inertia(self):
        """World inertia tensor and inverse per body [n, 18] (exact restore with set_state)."""
        a = np.zeros((self.n, 18))
        _chk(self.L.cp_world_get_inertia(self.w, a.ctypes.data), "cp_world_get_inertia")
        return a

    def aabbs(self):
        a = np.zeros((self.n, 6))
        _chk(self.L.cp_world_get_aabbs(self.w, a.ctypes.data), "cp_world_get_aabbs")
        return a

    def set_contacts(self, ints, magic, f):
        ii, mg, ff = _a(ints, np.int64), _a(magic, np.uint64), _a(f, np.float64)
        _chk(self.L.cp_world_set_contacts(self.w, len(ii), ii.ctypes.data, mg.ctypes.data, ff.ctypes.data), "cp_world_set_contacts")

    def contacts(self):
        n = ctypes.c_int()
        _chk(self.L.cp_world_get_contact_count(self.w, ctypes.byref(n)), "cp_world_get_contact_count")
        k = n.value
        ii = np.zeros((k, 6), np.int64)
        mg = np.zeros((k, 2), np.uint64)
        ff = np.zeros((k, 21))
        _chk(self.L.cp_world_get_contacts(self.w, k, ii.ctypes.data, mg.ctypes.data, ff.ctypes.data, ctypes.byref(n)),
             "cp_world_get_contacts")
        return ii[:n.value], mg[:n.value], ff[:n.value]

    def set_node_order(self, ids):
        ids = _a(ids, np.int64)
        _chk(self.L.cp_world_set_node_order(self.w, len(ids), ids.ctypes.data), "cp_world_set_node_order")

    def step(self, n=1):
        """Advances n steps.  Returns False when a dynamic body's pose or
        velocity went NaN or infinite (CP_ERR_NONFINITE: the steps completed;
        scene_flags() names the scenes), else True."""
        st = self.L.cp_world_step(self.w, n)
        if st == CP_ERR_NONFINITE:
            return False
        _chk(st, "cp_world_step")
        return True

    def sleep_state(self):
        """Per body: asleep, sleep_ready_counter, non_sleepy_tick_count, history length."""
        out = np.zeros((self.n, 4))
        _chk(self.L.cp_world_get_sleep_state(self.w, out.ctypes.data), "cp_world_get_sleep_state")
        return out

    def wake_bodies(self, idx):
        """RigidBody::alertWakingAction for the listed bodies (their islands wake at the next step)."""
        i = _a(idx, np.int32)
        _chk(self.L.cp_world_wake_bodies(self.w, len(i), i.ctypes.data), "cp_world_wake_bodies")

    def set_sleep_disabled(self, idx, disabled=True):
        i = _a(idx, np.int32)
        _chk(self.L.cp_world_set_sleep_disabled(self.w, len(i), i.ctypes.data, int(disabled)), "cp_world_set_sleep_disabled")

    def scene_flags(self):
        """Per scene 1 when one of its dynamic bodies has gone non-finite since its state was set."""
        n = ctypes.c_int()
        self.L.cp_world_scene_flags(self.w, 0, None, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), np.int32)
        _chk(self.L.cp_world_scene_flags(self.w, n.value, out.ctypes.data, ctypes.byref(n)), "cp_world_scene_flags")
        return out[:n.value]

    def broadphase(self):
        cap = max(64, self.n * 16)
        for _ in range(4):
            out = np.zeros((cap, 2), np.int64)
            n = ctypes.c_int()
            st = self.L.cp_broadphase(self.w, cap, out.ctypes.data, ctypes.byref(n))
            if st == 4:
                cap = n.value
                continue
            _chk(st, "cp_broadphase")
            return out[:n.value]
        raise CpError("cp_broadphase: capacity")

    def narrowphase(self, pairs):
        pairs = _a(pairs, np.int64).reshape(-1, 2)
        n = len(pairs)
        counts = np.zeros(n, np.int32)
        heads = np.zeros((n, 4))
        pts = np.zeros((n, 4, 4))
        mg = np.zeros((n, 4, 2), np.uint64)
        _chk(self.L.cp_narrowphase(self.w, n, pairs.ctypes.data, counts.ctypes.data, heads.ctypes.data, pts.ctypes.data, mg.ctypes.data),
             "cp_narrowphase")
        return counts, heads, pts, mg

    def last_islands(self):
        nr, ni = ctypes.c_int(), ctypes.c_int()
        st = self.L.cp_world_last_islands(self.w, 0, None, 0, None, ctypes.byref(nr), ctypes.byref(ni))
        rows = np.zeros((max(nr.value, 1), 2), np.int64)
        offs = np.zeros(ni.value + 1, np.int64)
        _chk(self.L.cp_world_last_islands(self.w, nr.value, rows.ctypes.data, ni.value + 1, offs.ctypes.data, ctypes.byref(nr),
                                          ctypes.byref(ni)), "cp_world_last_islands")
        return rows[:nr.value], offs

    def last_scene_contacts(self):
        """First post-narrowphase cache index of each scene's contacts (n_scenes + 1)."""
        n = ctypes.c_int()
        self.L.cp_world_last_scene_contacts(self.w, 0, None, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), np.int64)
        _chk(self.L.cp_world_last_scene_contacts(self.w, n.value, out.ctypes.data, ctypes.byref(n)), "cp_world_last_scene_contacts")
        return out[:n.value]

    def last_residual(self):
        r = np.zeros(4)
        _chk(self.L.cp_world_last_residual(self.w, r.ctypes.data), "cp_world_last_residual")
        return r

    def add_joint(self, kind, b1, b2, geo14):
        """Joint of kind 1 distance, 3 ball-socket, 4 sliding hinge, 5 hinge,
        6 weld, 7 slider, 8 cone limit, 9 twist limit, 10 motor, 11 piston
        drive, 12 piston slide limit or 13 spring between body indices b1, b2; geo14 =
        attach points in the body frames, two body-frame axes (hinge or slide
        axes; a distance joint's target distance first), position and
        rotation correction strengths.  A cone limit holds its maximum angle,
        a twist limit and a motor their minimum and maximum angles, in the
        first attach-point slots; a slide limit its minimum and maximum
        positions in the second axis slot.  A piston is its slide limit, then
        its drive.  A spring holds stiffness, damping and resting length (a
        negative length: the current distance) in the first axis slot."""
        g = _a(geo14, np.float64).reshape(14)
        k = ctypes.c_int()
        _chk(self.L.cp_joint_add(self.w, int(kind), int(b1), int(b2), g.ctypes.data, ctypes.byref(k)), "cp_joint_add")
        return k.value

    def set_drive(self, k, mode, max_=0.0, target=0.0):
        """Powered mode of motor or piston drive k: 0 off, 1 constant torque /
        force max_, 2 target velocity, 3 target position (max_, target)."""
        _chk(self.L.cp_joint_set_drive(self.w, int(k), int(mode), float(max_), float(target)), "cp_joint_set_drive")

    def set_joint_cfm(self, k, cfm):
        """Custom constraint force mixing of joint k's rows
        (setConstraintUseCustomCFM); NaN returns it to the world's CFM."""
        _chk(self.L.cp_joint_set_cfm(self.w, int(k), float(cfm)), "cp_joint_set_cfm")

    def joint_count(self):
        n = ctypes.c_int()
        _chk(self.L.cp_world_joint_count(self.w, ctypes.byref(n)), "cp_world_joint_count")
        return n.value

    def set_joint_state(self, state12):
        """Per joint: 6 accumulated impulses (zero padded), basis u, w (a
        limit: is_inactive and the twist's tracked angle in place of u)."""
        s = _a(state12, np.float64)
        assert s.shape == (self.joint_count(), 12)
        _chk(self.L.cp_world_set_joint_state(self.w, s.ctypes.data), "cp_world_set_joint_state")

    def joint_state(self):
        s = np.zeros((self.joint_count(), 12))
        _chk(self.L.cp_world_get_joint_state(self.w, s.ctypes.data), "cp_world_get_joint_state")
        return s

    def joint_rows(self, k):
        a = [np.zeros(36) for _ in range(4)]
        n = ctypes.c_int()
        _chk(self.L.cp_world_get_joint_rows(self.w, int(k), *[x.ctypes.data for x in a], ctypes.byref(n)), "cp_world_get_joint_rows")
        t = n.value
        return a[0][:6 * t].reshape(t, 6), a[1][:6 * t].reshape(t, 6), a[2][:6 * t].reshape(6, t), a[3][:6 * t].reshape(6, t)

    def set_holonomic_systems(self, off, joints):
        off, joints = _a(off, np.int64), _a(joints, np.int64)
        _chk(self.L.cp_world_set_holonomic_systems(self.w, len(off) - 1, off.ctypes.data, joints.ctypes.data),
             "cp_world_set_holonomic_systems")

    def holonomic_systems(self):
        """Per system (og, order, table, vloc, factor buffer)."""
        n = ctypes.c_int()
        _chk(self.L.cp_world_holonomic_system_count(self.w, ctypes.byref(n)), "cp_world_holonomic_system_count")
        out = []
        for k in range(n.value):
            sz = np.zeros(4, np.int64)
            _chk(self.L.cp_world_get_holonomic_system(self.w, k, None, None, None, None, None, sz.ctypes.data), "cp_world_get_holonomic_system")
            og, o, t, v = (np.zeros(sz[0], np.int64), np.zeros(sz[0], np.int64), np.zeros(sz[1], np.int64), np.zeros(sz[0], np.int64))
            b = np.zeros(max(int(sz[3]), 1))
            _chk(self.L.cp_world_get_holonomic_system(self.w, k, og.ctypes.data, o.ctypes.data, t.ctypes.data, v.ctypes.data, b.ctypes.data,
                                                      sz.ctypes.data), "cp_world_get_holonomic_system")
            out.append((og, o, t, v, b[:sz[3]]))
        return out

    # ---- serving (continuous batching)
    def serving_init(self, capacity):
        """Turns this (empty) world into a serving world of `capacity` body slots."""
        _chk(self.L.cp_world_serving_init(self.w, int(capacity)), "cp_world_serving_init")
        self.n = int(capacity)

    def admit_scene(self, shape_ids, movement, mass_props, material, state):
        """Stages a scene into a free body range; returns its handle, or None
        when no free range is large enough (the request waits)."""
        sid = _a(shape_ids, np.int32)
        mv = _a(movement, np.int32)
        mp = _a(mass_props, np.float64).reshape(-1, 19)
        mt = _a(material, np.float64).reshape(-1, 4)
        st = _a(state, np.float64).reshape(-1, 13)
        h = ctypes.c_int()
        rc = self.L.cp_world_admit_scene(self.w, len(sid), sid.ctypes.data, mv.ctypes.data, mp.ctypes.data, mt.ctypes.data, st.ctypes.data,
                                         ctypes.byref(h))
        if rc == 4:
            return None
        _chk(rc, "cp_world_admit_scene")
        return h.value

    def retire_scene(self, handle):
        _chk(self.L.cp_world_retire_scene(self.w, int(handle)), "cp_world_retire_scene")

    def scene_states(self, handles, rows):
        """Results of several admitted scenes in one call: a list of (rows, finite)
        per handle; rows = an upper bound on their total body count."""
        hs = _a(list(handles), np.int32)
        if len(hs) == 0:
            return []
        out = np.zeros((int(rows), 13))
        cnt = np.zeros(len(hs), np.int32)
        st = np.zeros(len(hs), np.int32)
        _chk(self.L.cp_world_get_scene_states(self.w, len(hs), hs.ctypes.data, int(rows), out.ctypes.data, cnt.ctypes.data, st.ctypes.data),
             "cp_world_get_scene_states")
        off = np.concatenate([[0], np.cumsum(cnt)])
        return [(out[off[k]:off[k + 1]], st[k] != CP_ERR_NONFINITE) for k in range(len(hs))]

    def scene_state(self, handle, with_status=False):
        """State rows of an admitted scene; with_status: (rows, finite) where
        finite is False when the scene holds a body whose pose or velocity went
        NaN or infinite (CP_ERR_NONFINITE)."""
        n = ctypes.c_int()
        _chk(self.L.cp_world_get_scene_state(self.w, int(handle), None, ctypes.byref(n)), "cp_world_get_scene_state")
        out = np.zeros((n.value, 13))
        st = self.L.cp_world_get_scene_state(self.w, int(handle), out.ctypes.data, ctypes.byref(n))
        if st != CP_ERR_NONFINITE:
            _chk(st, "cp_world_get_scene_state")
        elif not with_status:
            raise CpError(f"cp_world_get_scene_state: scene {handle} holds a non-finite body state ({st})")
        return (out, st != CP_ERR_NONFINITE) if with_status else out

    def raycast(self, origins, dirs, scenes=None):
        """First body hit per ray: (body index [n] (-1 none), hit [n, 7] = point, normal, distance)."""
        o = _a(origins, np.float64).reshape(-1, 3)
        d = _a(dirs, np.float64).reshape(-1, 3)
        n = len(o)
        sc = None if scenes is None else _a(scenes, np.int32)
        hb = np.zeros(n, np.int32)
        out = np.zeros((n, 7))
        _chk(self.L.cp_world_raycast(self.w, n, o.ctypes.data, d.ctypes.data, None if sc is None else sc.ctypes.data, hb.ctypes.data,
                                     out.ctypes.data), "cp_world_raycast")
        return hb, out

    def save_snapshot(self, path):
        _chk(self.L.cp_world_save_snapshot(self.w, os.fsencode(path)), "cp_world_save_snapshot")

    def load_snapshot(self, path):
        _chk(self.L.cp_world_load_snapshot(self.w, os.fsencode(path)), "cp_world_load_snapshot")

    def trajectory_open(self, path):
        _chk(self.L.cp_world_trajectory_open(self.w, os.fsencode(path)), "cp_world_trajectory_open")

    def trajectory_close(self):
        _chk(self.L.cp_world_trajectory_close(self.w), "cp_world_trajectory_close")

    def scene_snapshot(self, handle):
        """(state13 [n,13], inertia18 [n,18], contacts (ints6, magic2, f21) with scene-relative ids)."""
        nb, nc = ctypes.c_int(), ctypes.c_int()
        _chk(self.L.cp_world_scene_snapshot(self.w, int(handle), 0, None, None, None, None, None, ctypes.byref(nb), ctypes.byref(nc)),
             "cp_world_scene_snapshot")
        st, ine = np.zeros((nb.value, 13)), np.zeros((nb.value, 18))
        ii, mg, ff = np.zeros((nc.value, 6), np.int64), np.zeros((nc.value, 2), np.uint64), np.zeros((nc.value, 21))
        s = self.L.cp_world_scene_snapshot(self.w, int(handle), nc.value, st.ctypes.data, ine.ctypes.data, ii.ctypes.data, mg.ctypes.data,
                                           ff.ctypes.data, ctypes.byref(nb), ctypes.byref(nc))
        if s != CP_ERR_NONFINITE:
            _chk(s, "cp_world_scene_snapshot")
        return st, ine, (ii, mg, ff)

    def admit_scene_resume(self, shape_ids, movement, mass_props, material, snapshot):
        """Admits a scene evicted with scene_snapshot, resuming its inertia and contacts; None when no range holds it."""
        st, ine, (ii, mg, ff) = snapshot
        sid, mv = _a(shape_ids, np.int32), _a(movement, np.int32)
        mp, mt, st, ine = _a(mass_props, np.float64), _a(material, np.float64), _a(st, np.float64), _a(ine, np.float64)
        ii, mg, ff = _a(ii, np.int64), _a(mg, np.uint64), _a(ff, np.float64)
        h = ctypes.c_int()
        r = self.L.cp_world_admit_scene_resume(self.w, len(sid), sid.ctypes.data, mv.ctypes.data, mp.ctypes.data, mt.ctypes.data,
                                               st.ctypes.data, ine.ctypes.data, len(ii), ii.ctypes.data, mg.ctypes.data, ff.ctypes.data,
                                               ctypes.byref(h))
        if r == 4:  # CP_ERR_CAPACITY: no free range holds it now
            return None
        _chk(r, "cp_world_admit_scene_resume")
        return h.value

    def serving_info(self):
        """(capacity, free bodies, largest free range, admitted scenes)."""
        info = np.zeros(4, np.int32)
        _chk(self.L.cp_world_serving_info(self.w, info.ctypes.data), "cp_world_serving_info")
        return tuple(int(x) for x in info)

    def last_timing(self):
        ms = np.zeros(8)
        _chk(self.L.cp_world_last_timing(self.w, ms.ctypes.data), "cp_world_last_timing")
        return ms
