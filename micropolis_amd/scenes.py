"""Seeded synthetic workloads (host side; the inputs of tests and benchmarks).

Each generator returns a ``Workload``: renderer options (reyes.options names)
and one or more views, each a projection plus objects (patch arrays,
modelview, colour).  The modelview of an object is
``inverse(camera) * object_transform`` as Scene::draw computes it
(src/micropolis/Scene.cpp:190).

  c1_sphere   64 bicubic Bezier patches tiling a unit sphere (8x8 in u,v),
              interior control points jittered N(0, 0.03); 1024x768, BREADTH
  c2_torus    4096 patches on a 64x64 torus (R=1, r=0.35), value-noise
              displacement 0.02; 1920x1080, LOCAL, limit 16, N 16
  c3_batch    64 scenes of log-uniform [100, 20000] random patches; 640x480
  c4_gregory  16384 Gregory patches on 32 stacked, overlapping planes;
              1920x1080, BALANCED, limit 8, N 8
  c5_views    the c2 torus seen from 32 cameras at 960x540
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .renderer import BEZIER, GREGORY, Projection

# Bezier -> Gregory control point mapping (tools/bez2mscene_greg.py:111):
# corners, e+, e-, f+, f- taken from the 16 Bezier points.
BEZ_TO_GREGORY = [0, 3, 15, 12, 1, 7, 14, 8, 4, 2, 11, 13, 5, 6, 10, 9, 5, 6, 10, 9]


@dataclass
class SceneObject:
    handle: int
    patches: np.ndarray          # (P, 16|20, 3) float32
    patch_type: int
    matrix: np.ndarray           # 4x4 modelview (row-major numpy, column vectors)
    color: np.ndarray            # (4,) float32


@dataclass
class View:
    projection: Projection
    objects: List[SceneObject] = field(default_factory=list)


@dataclass
class Workload:
    name: str
    options: dict
    views: List[View]

    @property
    def patch_count(self):
        return sum(o.patches.shape[0] for v in self.views for o in v.objects)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """World-to-eye matrix (camera looks down -z)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def _bernstein(t):
    t = np.asarray(t, np.float64)
    s = 1.0 - t
    return np.stack([s ** 3, 3 * s * s * t, 3 * s * t * t, t ** 3], axis=-1)


_T4 = np.array([0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0])
_BINV = np.linalg.inv(_bernstein(_T4))


def fit_bezier(f, u0, u1, v0, v1):
    """Bicubic Bezier interpolating f at the 4x4 parameter grid; control
    point k = 4*i + j with i along u (kernels/utility.h:41-72)."""
    us = u0 + (u1 - u0) * _T4
    vs = v0 + (v1 - v0) * _T4
    S = np.array([[f(u, v) for v in vs] for u in us])           # (4,4,3)
    P = np.einsum("ia,abc,jb->ijc", _BINV, S, _BINV)             # (4,4,3)
    return P.reshape(16, 3)


def _bezier_eval(P, u, v):
    a = _bernstein(u)
    b = _bernstein(v)
    return np.einsum("i,j,ijc->c", a, b, P.reshape(4, 4, 3))


def _bezier_normal(P, u, v, h=1e-4):
    pu = (_bezier_eval(P, u + h, v) - _bezier_eval(P, u - h, v)) / (2 * h)
    pv = (_bezier_eval(P, u, v + h) - _bezier_eval(P, u, v - h)) / (2 * h)
    return np.cross(pv, pu)       # the orientation shade uses (reyes.cl:177)


def orient(P, outward):
    """Transpose the control grid if cross(dP/dv, dP/du) points inward."""
    n = _bezier_normal(P, 0.5, 0.5)
    if np.dot(n, outward) < 0:
        P = P.reshape(4, 4, 3).transpose(1, 0, 2).reshape(16, 3)
    return P


def _obj(handle, patches, ptype, mv, color):
    return SceneObject(handle, np.ascontiguousarray(patches, dtype=np.float32), ptype,
                       np.asarray(mv, np.float32), np.asarray(color, np.float32))


# ---------------------------------------------------------------------------

def sphere_patches(seed=1, jitter=0.03):
    rng = np.random.default_rng(seed)
    pats = []
    for a in range(8):           # phi sectors
        for b in range(8):       # theta bands
            def f(u, v):
                return np.array([np.sin(v) * np.cos(u), np.cos(v), np.sin(v) * np.sin(u)])
            u0, u1 = 2 * np.pi * a / 8, 2 * np.pi * (a + 1) / 8
            v0, v1 = np.pi * b / 8, np.pi * (b + 1) / 8
            P = fit_bezier(f, u0, u1, v0, v1)
            P = orient(P, f(0.5 * (u0 + u1), 0.5 * (v0 + v1)))
            for k in (5, 6, 9, 10):
                P[k] += rng.normal(0.0, jitter, 3)
            pats.append(P)
    return np.array(pats, np.float32)


def c1_sphere(seed=1):
    view = look_at((0, 0, 4), (0, 0, 0))
    opts = dict(window_size=(1024, 768), bound_n_split_method="BREADTH", bound_n_split_limit=8.0,
                reyes_patch_size=8, reyes_patches_per_pass=1024)
    proj = Projection(60.0, 0.01, (1024, 768))
    o = _obj(1, sphere_patches(seed), BEZIER, view, (1.0, 0.6, 0.3, 1.0))
    return Workload("c1_sphere_64p_1024x768_breadth", opts, [View(proj, [o])])


def torus_patches(n=64, R=1.0, r=0.35):
    pats = []
    for a in range(n):
        for b in range(n):
            def f(u, v):
                return np.array([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u),
                                 r * np.sin(v)])
            u0, u1 = 2 * np.pi * a / n, 2 * np.pi * (a + 1) / n
            v0, v1 = 2 * np.pi * b / n, 2 * np.pi * (b + 1) / n
            P = fit_bezier(f, u0, u1, v0, v1)
            um, vm = 0.5 * (u0 + u1), 0.5 * (v0 + v1)
            outward = f(um, vm) - np.array([R * np.cos(um), R * np.sin(um), 0.0])
            pats.append(orient(P, outward))
    return np.array(pats, np.float32)


# Camera of the 1080p torus frame: close to the tube, looking across the hole
# at the far side of the ring.  30.4% of the 4096 patches lie in front of the
# eye with a projected control-point box wider than 200 pixels, and the DRAW
# ranges of the LOCAL split (limit 16) have depths 1..10.
TORUS_EYE = (0.084, 1.388, -0.414)
TORUS_TARGET = (-1.609, 0.233, -0.518)
TORUS_UP = (0.0, 0.0, 1.0)


def c2_torus(width=1920, height=1080):
    view = look_at(TORUS_EYE, TORUS_TARGET, TORUS_UP)
    opts = dict(window_size=(width, height), bound_n_split_method="LOCAL", bound_n_split_limit=16.0,
                reyes_patch_size=16, displacement=1, displacement_amplitude=0.02,
                displacement_frequency=4.0, displacement_seed=7, displacement_octaves=4,
                local_bns_work_groups=0)
    proj = Projection(60.0, 0.01, (width, height))
    o = _obj(1, torus_patches(), BEZIER, view, (0.85, 0.75, 0.55, 1.0))
    return Workload("c2_torus_4096p_%dx%d_local_disp" % (width, height), opts, [View(proj, [o])])


def c5_views(n_views=32, width=960, height=540, seed=5):
    rng = np.random.default_rng(seed)
    pats = torus_patches()
    views = []
    for k in range(n_views):
        ang = 2 * np.pi * k / n_views
        dist = rng.uniform(2.2, 3.2)
        elev = rng.uniform(0.3, 1.2)
        eye = (dist * np.cos(ang), dist * np.sin(ang), elev)
        mv = look_at(eye, (0, 0, 0), (0, 0, 1))
        o = _obj(1, pats, BEZIER, mv, (0.85, 0.75, 0.55, 1.0))
        views.append(View(Projection(60.0, 0.01, (width, height)), [o]))
    opts = dict(window_size=(width, height), bound_n_split_method="LOCAL", bound_n_split_limit=16.0,
                reyes_patch_size=16, displacement=1, displacement_amplitude=0.02,
                displacement_frequency=4.0, displacement_seed=7, displacement_octaves=4,
                local_bns_work_groups=0)
    return Workload("c5_torus_32views_%dx%d" % (width, height), opts, views)


def random_patches(rng, n, centre_radius):
    centres = rng.uniform(-centre_radius, centre_radius, (n, 1, 3))
    return (centres + rng.uniform(-1.0, 1.0, (n, 16, 3))).astype(np.float32)


def c3_batch(n_scenes=64, seed=3, width=640, height=480):
    rng = np.random.default_rng(seed)
    views = []
    for s in range(n_scenes):
        n = int(np.floor(np.exp(rng.uniform(np.log(100), np.log(20000)))))
        radius = 2.0 * n ** (1.0 / 3.0)
        pats = random_patches(rng, n, radius)
        # camera just outside the cloud: the nearest patches fill the screen
        dist = radius * np.sqrt(3.0) + 1.0 + rng.uniform(0.0, 3.0)
        ang = rng.uniform(0, 2 * np.pi)
        eye = (dist * np.cos(ang), 0.3 * dist, dist * np.sin(ang))
        mv = look_at(eye, (0, 0, 0))
        col = (rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), 1.0)
        views.append(View(Projection(60.0, 0.01, (width, height)), [_obj(s + 1, pats, BEZIER, mv, col)]))
    opts = dict(window_size=(width, height), bound_n_split_method="LOCAL", bound_n_split_limit=8.0,
                reyes_patch_size=8, local_bns_work_groups=0, backface_culling=1)
    return Workload("c3_batch_64scenes_%dx%d" % (width, height), opts, views)


def gregory_planes(rng, n_planes=32, nu=32, nv=16, width=4.0, height=2.0, spacing=0.04):
    pats = []
    for k in range(n_planes):
        off = rng.uniform(-0.8, 0.8, 2)
        z = -k * spacing
        xs = np.linspace(-width / 2, width / 2, nu + 1) + off[0]
        ys = np.linspace(-height / 2, height / 2, nv + 1) + off[1]
        grid = np.zeros((nu + 1, nv + 1, 3))
        grid[..., 0] = xs[:, None]
        grid[..., 1] = ys[None, :]
        grid[..., 2] = z
        grid[1:-1, 1:-1, :2] += rng.uniform(-0.3, 0.3, (nu - 1, nv - 1, 2)) * np.array(
            [width / nu, height / nv])
        for i in range(nu):
            for j in range(nv):
                c00, c10 = grid[i, j], grid[i + 1, j]
                c01, c11 = grid[i, j + 1], grid[i + 1, j + 1]
                P = np.zeros((4, 4, 3))
                for a in range(4):
                    for b in range(4):
                        s, t = a / 3.0, b / 3.0
                        P[a, b] = (1 - s) * (1 - t) * c00 + s * (1 - t) * c10 + (1 - s) * t * c01 + s * t * c11
                P[1:3, 1:3, 2] += rng.normal(0.0, 0.01, (2, 2))
                P = orient(P.reshape(16, 3), np.array([0.0, 0.0, 1.0]))
                G = P[BEZ_TO_GREGORY].copy()
                G[16:20] += rng.normal(0.0, 0.005, (4, 3))   # f- differs from f+
                pats.append(G)
    return np.array(pats, np.float32)


def c4_gregory(seed=4, width=1920, height=1080, planes=32, quads=512):
    """32 stacked planes x 512 perturbed quads as Gregory patches (16384);
    `planes` and `quads` (32 x 16 grid scaled down) shrink it for tests."""
    rng = np.random.default_rng(seed)
    nu = 32 if quads >= 512 else max(2, int(round(np.sqrt(quads * 2))))
    nv = max(1, quads // nu)
    pats = gregory_planes(rng, n_planes=planes, nu=nu, nv=nv)
    mv = look_at((0.0, 0.0, 3.0), (0.0, 0.0, -0.6))
    opts = dict(window_size=(width, height), bound_n_split_method="BALANCED", bound_n_split_limit=8.0,
                reyes_patch_size=8, local_bns_work_groups=0)
    proj = Projection(60.0, 0.01, (width, height))
    o = _obj(1, pats, GREGORY, mv, (0.6, 0.8, 0.9, 1.0))
    return Workload("c4_gregory_%dp_%dx%d_balanced" % (len(pats), width, height), opts, [View(proj, [o])])


def big_patch(limit=2.83, width=1920, height=1080):
    """One bicubic patch, a plane tilted away from the eye (z -0.9 .. -2.2)
    filling a third of a 90-degree view: LOCAL split with bound_n_split_limit
    `limit` and N 8 gives 154,634 DRAW ranges of that one patch at 2.83 (split
    depths 15-19) and 90,654 at 4.0, the case a per-patch rank by counting
    would pay for with the square of the count."""
    P = np.zeros((16, 3))
    for i in range(4):
        for j in range(4):
            P[4 * i + j] = (-1.2 + 2.4 * i / 3.0, -0.8 + 1.3 * j / 3.0, -0.9 - 1.3 * j / 3.0)
    P = orient(P, np.array([0.0, 0.3, 1.0]))[None].astype(np.float32)
    o = _obj(1, P, BEZIER, np.eye(4), (0.7, 0.7, 0.7, 1.0))
    opts = dict(window_size=(width, height), bound_n_split_method="LOCAL", bound_n_split_limit=float(limit),
                reyes_patch_size=8, local_bns_work_groups=0)
    return Workload("big_patch_%dx%d_limit%g" % (width, height, limit), opts,
                    [View(Projection(90.0, 0.01, (width, height)), [o])])


WORKLOADS = {"c1": c1_sphere, "c2": c2_torus, "c3": c3_batch, "c4": c4_gregory, "c5": c5_views}
