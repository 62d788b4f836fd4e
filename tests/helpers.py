"""Small in-code scenes mirroring the reference's test scenes
(tests/accel/bvh_test_scene.hpp:9-53) and array helpers."""
import ctypes as C

import numpy as np

from yaoray_amd import native as N
from yaoray_amd.api import PRIMITIVE_DTYPE, VERTEX_DTYPE, Scene


class MeshArrays:
    def __init__(self):
        self.vertices = []
        self.indices = []
        self.primitives = []

    def add_test_triangle(self, x_offset, z=0.0, material=0):
        # AddBvhTestTriangle: (x-0.25,-0.25) (x+0.25,-0.25) (x,0.25) at z, normal +z.
        first_vertex = len(self.vertices)
        first_index = len(self.indices)
        for p in ((x_offset - 0.25, -0.25, z), (x_offset + 0.25, -0.25, z), (x_offset, 0.25, z)):
            self.vertices.append((p, (0.0, 0.0, 1.0)))
        self.indices += [first_vertex, first_vertex + 1, first_vertex + 2]
        self.primitives.append((first_index, 3, material, 1))
        return self

    def add_triangles(self, tris, material=0):
        """One primitive holding every triangle of `tris` (n, 3, 3)."""
        first_index = len(self.indices)
        for tri in np.asarray(tris, np.float32):
            base = len(self.vertices)
            for p in tri:
                self.vertices.append((tuple(p), (0.0, 0.0, 0.0)))
            self.indices += [base, base + 1, base + 2]
        self.primitives.append((first_index, 3 * len(tris), material, 0))
        return self

    def arrays(self):
        v = np.zeros(len(self.vertices), VERTEX_DTYPE)
        for k, (p, n) in enumerate(self.vertices):
            v[k]["position"] = p
            v[k]["normal"] = n
            v[k]["tangent_handedness"] = 1.0
        i = np.array(self.indices, np.uint32)
        p = np.zeros(len(self.primitives), PRIMITIVE_DTYPE)
        for k, (fi, ic, m, hn) in enumerate(self.primitives):
            p[k]["first_index"] = fi
            p[k]["index_count"] = ic
            p[k]["material_index"] = m
            p[k]["has_normals"] = hn
        return v, i, p

    def scene(self, instance_primitive=None, instance_matrices=None):
        v, i, p = self.arrays()
        return Scene.from_arrays(v, i, p, instance_primitive, instance_matrices)


def clustered():
    m = MeshArrays()
    for x in (0.0, 0.05, 0.10, 0.15, 5.0, 5.05, 5.10, 5.15, 10.0, 10.05, 10.10, 10.15):
        m.add_test_triangle(x)
    return m


def rays(origins, directions, t_min=1e-5, t_max=np.inf):
    o = np.zeros((len(origins), 4), np.float32)
    d = np.zeros((len(origins), 4), np.float32)
    o[:, :3] = origins
    o[:, 3] = t_min
    d[:, :3] = directions
    d[:, 3] = t_max
    return o, d


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype == np.float32:
        a = a.view(np.uint32)
    if b.dtype == np.float32:
        b = b.view(np.uint32)
    return np.array_equal(a, b)


def trs(translate, axis, angle_deg, scale):
    """Column-major 4x4 T * R(axis, angle) * S as float32[16] (computed in float64)."""
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    a = np.radians(angle_deg)
    c, s, t = np.cos(a), np.sin(a), 1 - np.cos(a)
    x, y, z = axis
    r = np.array([[t * x * x + c, t * x * y - s * z, t * x * z + s * y],
                  [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
                  [t * x * z - s * y, t * y * z + s * x, t * z * z + c]])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(np.broadcast_to(np.asarray(scale, np.float64), 3))
    m[:3, 3] = translate
    return m.T.reshape(16).astype(np.float32)  # column-major


def base_path_scene():
    """MakeBasePathJob (tests/backends/cpu/cpu_path_tracer_test_support.hpp:40-66):
    one diffuse triangle facing a pinhole camera, constant environment."""
    m = MeshArrays()
    m.vertices = [((-1.5, -1.0, 0.0), (0.0, 0.0, 1.0)), ((1.5, -1.0, 0.0), (0.0, 0.0, 1.0)),
                  ((0.0, 1.25, 0.0), (0.0, 0.0, 1.0))]
    m.indices = [0, 1, 2]
    m.primitives = [(0, 3, 0, 1)]
    s = m.scene()
    s.set_camera((0.0, 0.0, 3.0), (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.8)
    s.set_materials([{"kind": 0, "reflectance": (0.8, 0.8, 0.8)}])
    s.set_constant_environment((0.02, 0.03, 0.04), 1.0)
    s.build_accel()
    return s


# ---- learned guiding ---------------------------------------------------------------


def guide_desc(log2_table=12, rmin=4, rmax=64, bounds=((0, 0, 0), (1, 1, 1)), capacity=1 << 17, lr=1e-2):
    return N.GuideDesc((C.c_float * 3)(*bounds[0]), (C.c_float * 3)(*bounds[1]), log2_table, rmin, rmax, lr, 0.5,
                       capacity, 1)


def guide_params(d, seed, grid_scale=0.5):
    rng = np.random.default_rng(seed)
    mlp = np.zeros(N.GUIDE_MLP_PARAMS, np.float32)
    dims = [(32, 64), (64, 64), (64, 64), (64, 48)]
    off = 0
    for k, (i, o) in enumerate(dims):
        a = np.sqrt(6.0 / i) * (0.5 if k == 3 else 1.0)
        w = rng.uniform(-a, a, (i, o)).astype(np.float32)
        if k == 3:
            w[:, 40:] = 0.0
        mlp[off:off + i * o] = w.ravel()
        off += i * o
    grid = rng.uniform(-grid_scale, grid_scale, 8 * (1 << d.log2_table_size) * 2).astype(np.float32)
    return mlp, grid


def guide_samples(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.95, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, (n, 1))
    return np.hstack([p, nrm, wo, wi, w]).astype(np.float32)


# Ground + one material per icosphere variant of Scene.instanced(variants=8):
# every BSDF branch of the in-scope kinds, including an absorbing smooth
# dielectric (Beer-Lambert inside the medium), rough dielectrics, a delta
# mirror and smooth / rough thin dielectrics.
ZOO_MATERIALS = [
    {"kind": 0, "reflectance": (0.5, 0.5, 0.5)},
    {"kind": 0, "reflectance": (0.7, 0.4, 0.3)},
    {"kind": 1, "reflectance": (0.9, 0.7, 0.4), "uroughness": 0.4},
    {"kind": 1, "reflectance": (0.95, 0.95, 0.95), "uroughness": 0.0},
    {"kind": 2, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5, "absorption_color": (0.6, 0.8, 0.9),
     "absorption_distance": 0.5},
    {"kind": 2, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5, "uroughness": 0.3},
    {"kind": 2, "reflectance": (0.9, 0.85, 0.8), "ior": 1.33, "uroughness": 0.55},
    {"kind": 3, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5},
    {"kind": 3, "reflectance": (0.9, 0.95, 1.0), "ior": 1.5, "uroughness": 0.35},
]


def material_zoo_scene(instances=150, subdivisions=2, env_width=64, env_height=32, seed=4):
    """Two-level environment-lit scene whose instances carry every BSDF branch."""
    s = Scene.instanced(instances=instances, subdivisions=subdivisions, variants=8, env_width=env_width,
                        env_height=env_height, seed=seed)
    s.set_materials(ZOO_MATERIALS)
    s.build_accel()
    return s


def cornell_materials(scene):
    """The scene's materials as set_materials dicts."""
    import ctypes as C

    from yaoray_amd import native as N
    sd = scene.shading_desc()
    arr = (N.Material * sd.material_count).from_address(sd.materials)
    out = []
    for m in arr:
        out.append({f: (list(getattr(m, f)) if isinstance(getattr(m, f), C.Array) else getattr(m, f))
                    for f, _ in N.Material._fields_ if f != "reserved"})
    return out


# Spheres over every sphere path: absorbing glass (a closed medium), mirror, rough conductor, diffuse,
# one with flipped normals and one carrying the emissive material (a sphere's
# emission is not added by the reference).
def sphere_cornell():
    """Cornell box (flat BVH) plus analytic spheres; materials 8.. are added."""
    s = Scene.cornell(1)
    mats = cornell_materials(s) + [
        {"kind": 2, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5, "absorption_color": (0.7, 0.9, 0.8),
         "absorption_distance": 0.6},
        {"kind": 1, "reflectance": (0.95, 0.95, 0.95), "uroughness": 0.0},
        {"kind": 1, "reflectance": (0.9, 0.7, 0.4), "uroughness": 0.4},
    ]
    s.set_materials(mats)
    s.set_spheres([((-0.45, 0.4, 0.2), 0.4, 8), ((0.5, 0.35, -0.3), 0.35, 9), ((0.45, 1.2, 0.4), 0.2, 10),
                   ((-0.3, 1.4, -0.4), 0.25, 1, True), ((0.0, 0.15, 0.6), 0.15, 2), ((0.6, 1.7, -0.6), 0.12, 7)])
    s.build_accel()
    return s


def sphere_zoo(instances=150):
    """The two-level material-zoo scene plus spheres in front of its camera (TLAS spheres)."""
    s = Scene.instanced(instances=instances, subdivisions=2, variants=8, env_width=64, env_height=32, seed=4)
    s.set_materials(ZOO_MATERIALS)
    s.set_spheres([((-4.0, 2.0, 20.0), 2.0, 7), ((3.0, 1.5, 25.0), 1.5, 3), ((0.0, 3.0, 10.0), 3.0, 0),
                   ((6.0, 1.0, 30.0), 1.0, 1, True), ((-6.0, 1.5, 18.0), 1.2, 2), ((5.0, 1.2, 16.0), 1.2, 5),
                   ((-1.0, 1.0, 22.0), 0.9, 4)])
    s.build_accel()
    return s


# ---- analytic lights ---------------------------------------------------------------

def lit_sphere_scene(lights):
    """MakeLitSphereJob (tests/backends/cpu/cpu_path_tracer_test_support.hpp:59-84):
    a diffuse 0.8 sphere of radius 0.5 at the origin, camera at z = 3, no
    environment, lit by `lights` (Scene.set_analytic_lights dicts)."""
    s = MeshArrays().scene()
    s.set_camera((0.0, 0.0, 3.0), (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1.04719758)
    s.set_materials([{"kind": 0, "reflectance": (0.8, 0.8, 0.8)}])
    s.set_constant_environment((0.0, 0.0, 0.0), 1.0)
    s.set_spheres([((0.0, 0.0, 0.0), 0.5, 0)])
    s.set_analytic_lights(lights)
    s.build_accel()
    return s


CORNELL_LIGHTS = [
    {"kind": "point", "position": (0.3, 1.6, 0.5), "intensity": (1.5, 1.4, 1.2)},
    {"kind": "spot", "position": (-0.5, 1.9, -0.3), "direction": (0.1, -1.0, 0.05), "intensity": (3.0, 2.5, 2.0),
     "cone_angle": float(np.cos(np.radians(35.0))), "cone_cos_inner": float(np.cos(np.radians(25.0)))},
    {"kind": "distant", "direction": (0.1, -0.45, -1.0), "intensity": (0.6, 0.65, 0.7)},
]


def cornell_lights_scene(spheres=True):
    """The sphere Cornell box (or the plain one) with a point, a spot and a
    distant light (the distant one shines in through the open front)."""
    s = sphere_cornell() if spheres else Scene.cornell(1)
    s.set_analytic_lights(CORNELL_LIGHTS)
    s.build_accel()
    return s


def many_point_lights_scene(count=12, seed=3):
    """Cornell box with `count` point lights and no others kinds: more than 8
    local lights, so light selection descends the light tree."""
    rng = np.random.default_rng(seed)
    s = Scene.cornell(1)
    s.set_analytic_lights([{"kind": "point", "position": tuple(rng.uniform((-0.9, 0.2, -0.9), (0.9, 1.9, 0.9))),
                            "intensity": tuple(rng.uniform(0.05, 0.6, 3))} for _ in range(count)])
    s.build_accel()
    return s


# ---- textures, alpha cut-outs, normal maps ----------------------------------------


class TexturedMesh:
    """Quads (two triangles each) with per-vertex uvs and tangents, one
    primitive per quad."""

    def __init__(self):
        self.v = []
        self.i = []
        self.p = []
        self.areas = []

    def quad(self, corners, uvs=None, material=0, normals=None, tangent=None, handedness=(1.0, 1.0, 1.0, 1.0)):
        c = np.asarray(corners, np.float32)
        base, first = len(self.v), len(self.i)
        for k in range(4):
            rec = np.zeros((), VERTEX_DTYPE)
            rec["position"] = c[k]
            if normals is not None:
                rec["normal"] = normals[k] if np.ndim(normals) == 2 else normals
            if uvs is not None:
                rec["uv"] = uvs[k]
            if tangent is not None:
                rec["tangent"] = tangent[k] if np.ndim(tangent) == 2 else tangent
            rec["tangent_handedness"] = handedness[k]
            self.v.append(rec)
        self.i += [base, base + 1, base + 2, base, base + 2, base + 3]
        self.p.append((first, 6, material, normals is not None, uvs is not None, tangent is not None))
        a = 0.5 * (np.linalg.norm(np.cross(c[1] - c[0], c[2] - c[0])) + np.linalg.norm(np.cross(c[2] - c[0], c[3] - c[0])))
        self.areas.append(float(a))
        return len(self.p) - 1

    def arrays(self):
        v = np.array(self.v, VERTEX_DTYPE)
        p = np.zeros(len(self.p), PRIMITIVE_DTYPE)
        for k, (fi, ic, m, hn, hu, ht) in enumerate(self.p):
            p[k]["first_index"], p[k]["index_count"], p[k]["material_index"] = fi, ic, m
            p[k]["has_normals"], p[k]["has_uvs"], p[k]["has_tangents"] = hn, hu, ht
        return v, np.array(self.i, np.uint32), p


def _checker(h, w, seed, lo=0.15, hi=0.95):
    rng = np.random.default_rng(seed)
    t = rng.uniform(lo, hi, (h, w, 4)).astype(np.float32)
    t[..., 3] = 1.0
    return t


def _normal_texels(h, w, seed):
    """Encoded tangent-space normals ((n + 1) / 2) leaning up to ~40 degrees."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-0.6, 0.6, (h, w, 2))
    n = np.concatenate([xy, np.ones((h, w, 1))], axis=2)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    t = np.ones((h, w, 4), np.float32)
    t[..., :3] = (n + 1.0) * 0.5
    return t


def textured_scene(two_level=False, filters=None):
    """Textured planes under an area light: an EWA-filtered repeat ground seen
    at a grazing angle (anisotropic footprints), trilinear clamp, nearest
    mirrored and bilinear repeat walls (uv from -1 to 2), a normal-mapped
    quad with tangents, an alpha cut-out quad (alpha texture), a quad with
    constant alpha 0.3 (never visible), and a thin-dielectric quad whose
    textured reflectance tints the shadows behind it. `two_level`: every
    quad is an instance, and the normal-mapped and cut-out quads appear a
    second time under a mirroring, rotating transform. `filters`: the four
    image textures' filters (default EWA, trilinear, nearest, bilinear)."""
    f = filters or (N.FILTER_EWA, N.FILTER_TRILINEAR, N.FILTER_NEAREST, N.FILTER_BILINEAR)
    m = TexturedMesh()
    up = (0.0, 1.0, 0.0)
    uvq = [(-1.0, -1.0), (2.0, -1.0), (2.0, 2.0), (-1.0, 2.0)]
    m.quad([(-4, 0, 3), (4, 0, 3), (4, 0, -6), (-4, 0, -6)], [(-2, 1.5), (2, 1.5), (2, -3), (-2, -3)], 0)
    m.quad([(-3, 0, -2.5), (-0.5, 0, -2.5), (-0.5, 2, -2.5), (-3, 2, -2.5)], uvq, 1)
    m.quad([(-0.5, 0, -2.5), (1.5, 0, -2.5), (1.5, 2, -2.5), (-0.5, 2, -2.5)], uvq, 2)
    nm = m.quad([(1.5, 0, -2.4), (3, 0, -2.4), (3, 2, -2.4), (1.5, 2, -2.4)], [(0, 0), (1, 0), (1, 1), (0, 1)], 3,
                normals=(0.0, 0.0, 1.0), tangent=(1.0, 0.0, 0.0), handedness=(1.0, 1.0, -1.0, 1.0))
    cut = m.quad([(-2, 0.2, 0.5), (0, 0.2, 0.5), (0, 1.6, 0.5), (-2, 1.6, 0.5)], [(0, 0), (1, 0), (1, 1), (0, 1)], 4)
    m.quad([(0.5, 0, 1.0), (2, 0, 1.0), (2, 1, 1.0), (0.5, 1, 1.0)], [(0, 0), (1, 0), (1, 1), (0, 1)], 5)
    m.quad([(0.3, 0.3, -0.5), (1.3, 0.3, -0.5), (1.3, 1.3, -0.5), (0.3, 1.3, -0.5)], [(0, 0), (1, 0), (1, 1), (0, 1)], 6)
    light = m.quad([(-1, 3, -1), (1, 3, -1), (1, 3, 1), (-1, 3, 1)], None, 7)
    v, i, p = m.arrays()
    if two_level:
        ip = list(range(len(p))) + [nm, cut]
        mats = [np.eye(4, dtype=np.float32).reshape(16)] * len(p)
        mats += [trs((4.2, 0.0, -3.5), (0.0, 1.0, 0.0), 35.0, (-1.0, 1.0, 1.0)),
                 trs((-1.0, 0.1, 1.8), (0.3, 1.0, 0.0), -25.0, (1.0, 1.2, -0.8))]
        s = Scene.from_arrays(v, i, p, np.array(ip, np.int32), np.array(mats, np.float32))
    else:
        s = Scene.from_arrays(v, i, p)
    t = [s.add_texture(_checker(8, 8, 1), f[0], N.WRAP_REPEAT),
         s.add_texture(_checker(3, 5, 2), f[1], N.WRAP_CLAMP),
         s.add_texture(_checker(6, 4, 3), f[2], N.WRAP_MIRRORED),
         s.add_texture(_checker(3, 3, 4), f[3], N.WRAP_REPEAT, N.WRAP_MIRRORED)]
    alpha = np.ones((4, 4, 4), np.float32)
    alpha[..., 0] = np.indices((4, 4)).sum(0) % 2
    t.append(s.add_texture(alpha, N.FILTER_BILINEAR, N.WRAP_REPEAT))
    t.append(s.add_texture(_normal_texels(8, 8, 5), N.FILTER_BILINEAR, N.WRAP_REPEAT))
    s.set_materials([
        {"kind": 0, "reflectance": (1.0, 1.0, 1.0), "reflectance_texture": t[0]},
        {"kind": 0, "reflectance": (0.9, 0.8, 0.9), "reflectance_texture": t[1]},
        {"kind": 1, "reflectance": (1.0, 0.9, 0.8), "uroughness": 0.3, "reflectance_texture": t[2]},
        {"kind": 0, "reflectance": (0.8, 0.8, 0.8), "reflectance_texture": t[3], "normal_map": t[5],
         "normal_scale": 0.8},
        {"kind": 0, "reflectance": (0.7, 0.3, 0.2), "alpha_texture": t[4]},
        {"kind": 0, "reflectance": (0.2, 0.7, 0.3), "alpha": 0.3},
        {"kind": 3, "reflectance": (1.0, 1.0, 1.0), "ior": 1.5, "reflectance_texture": t[3]},
        {"kind": 0, "reflectance": (0.0, 0.0, 0.0), "emission": (6.0, 6.0, 6.0)},
    ])
    s.set_emissives([(light, (6.0, 6.0, 6.0), m.areas[light])])
    s.set_constant_environment((0.10, 0.12, 0.15), 1.0)
    fwd = np.array([0.0, -0.4, -1.0])
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    s.set_camera((0.3, 2.2, 5.0), fwd, right, np.cross(right, fwd), 0.9)
    s.build_accel()
    return s


def alpha_stack_scene(layers=70, alpha=0.3):
    """`layers` parallel quads of constant alpha `alpha` between the camera and
    a lit diffuse wall, the light off to the side on the camera's side of the
    stack (part of the wall sees it through the stack): with more than 64 layers
    every camera ray and shadow ray through the stack exhausts the alpha skips
    (a miss for the path, blocked for the shadow)."""
    m = TexturedMesh()
    m.quad([(-2, -2, -4), (2, -2, -4), (2, 2, -4), (-2, 2, -4)], None, 0)
    for k in range(layers):
        z = -3.0 + 2.5 * k / layers
        m.quad([(-1, -1, z), (1, -1, z), (1, 1, z), (-1, 1, z)], [(0, 0), (1, 0), (1, 1), (0, 1)], 1)
    light = m.quad([(1.5, -0.5, 1.0), (2.5, -0.5, 1.0), (2.5, 0.5, 1.0), (1.5, 0.5, 1.0)][::-1], None, 2)
    v, i, p = m.arrays()
    s = Scene.from_arrays(v, i, p)
    s.set_materials([{"kind": 0, "reflectance": (0.8, 0.7, 0.6)}, {"kind": 0, "reflectance": (0.9, 0.1, 0.1),
                                                                    "alpha": alpha},
                     {"kind": 0, "reflectance": (0.0, 0.0, 0.0), "emission": (4.0, 4.0, 4.0)}])
    s.set_emissives([(light, (4.0, 4.0, 4.0), m.areas[light])])
    s.set_constant_environment((0.05, 0.05, 0.05), 1.0)
    s.set_camera((0.0, 0.0, 3.0), (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.9)
    s.build_accel()
    return s
