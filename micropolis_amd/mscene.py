"""Reader for the reference's `.mscene` scene files, and the frame they draw.

A `.mscene` file is a Cap'n Proto message of the schema in
src/micropolis/mscene.capnp:4-74, written in the packed stream encoding
(read by `capnp::PackedFdMessageReader`, src/micropolis/Scene.cpp:89-100).
No Cap'n Proto library is installed here, so this module decodes the format
from its published specification:

* packed encoding: each 8-byte word is preceded by a tag byte whose bit i
  says whether byte i is non-zero (and present); tag 0x00 is followed by a
  count of further all-zero words, tag 0xff by a count of further words
  copied verbatim;
* stream framing: (segment count - 1) and the segment sizes in words as
  32-bit integers, padded to a word, then the segments;
* pointers: struct (kind 0: word offset, data and pointer section sizes),
  list (kind 1: offset, element size code, count; code 7 = composite with a
  tag word), far (kind 2: landing pad in another segment, single or double).

`load(path)` returns a `SceneFile`; `draw_workload(scene, ...)` turns it into
the frame `Scene::draw` renders (Scene.cpp:179-197): one object per
scene object, modelview = inverse(camera transform) * object transform,
projection from the camera's fovy and near plane and the window size.
"""
import struct
from dataclasses import dataclass, field

import numpy as np


class FormatError(ValueError):
    pass


# ---------------------------------------------------------------------------
# packed stream -> segments

def unpack(data):
    """Packed Cap'n Proto bytes -> unpacked bytes."""
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        tag = data[i]
        i += 1
        word = bytearray(8)
        for b in range(8):
            if tag & (1 << b):
                if i >= n:
                    raise FormatError("truncated packed word")
                word[b] = data[i]
                i += 1
        out += word
        if tag == 0x00:
            if i >= n:
                raise FormatError("truncated zero run")
            out += bytes(8 * data[i])
            i += 1
        elif tag == 0xFF:
            if i >= n:
                raise FormatError("truncated literal run")
            cnt = 8 * data[i]
            i += 1
            if i + cnt > n:
                raise FormatError("truncated literal words")
            out += data[i:i + cnt]
            i += cnt
    return bytes(out)


def segments(raw):
    if len(raw) < 8:
        raise FormatError("message too short")
    nseg = struct.unpack_from("<I", raw, 0)[0] + 1
    sizes = struct.unpack_from("<%dI" % nseg, raw, 4)
    off = 4 + 4 * nseg
    off += (-off) % 8
    segs = []
    for s in sizes:
        if off + 8 * s > len(raw):
            raise FormatError("segment extends past the message")
        segs.append(raw[off:off + 8 * s])
        off += 8 * s
    return segs


# ---------------------------------------------------------------------------
# pointers

class Message:
    def __init__(self, segs):
        self.segs = segs

    def word(self, seg, idx):
        s = self.segs[seg]
        if idx < 0 or 8 * idx + 8 > len(s):
            raise FormatError("pointer outside its segment")
        return struct.unpack_from("<Q", s, 8 * idx)[0]

    def resolve(self, seg, idx):
        """Follow the pointer at (seg, idx) -> (seg, target word, tag word)."""
        p = self.word(seg, idx)
        if p == 0:
            return None
        kind = p & 3
        if kind == 2:
            double = (p >> 2) & 1
            pad = (p >> 3) & 0x1FFFFFFF
            tseg = p >> 32
            if not double:
                return self.resolve(tseg, pad)
            far = self.word(tseg, pad)
            tag = self.word(tseg, pad + 1)
            return (far >> 32, (far >> 3) & 0x1FFFFFFF, tag)
        off = (p >> 2) & 0x3FFFFFFF
        if off & 0x20000000:
            off -= 0x40000000
        return (seg, idx + 1 + off, p)

    def struct(self, seg, idx):
        r = self.resolve(seg, idx)
        if r is None:
            return None
        s, start, p = r
        if p & 3 != 0:
            raise FormatError("expected a struct pointer")
        return Struct(self, s, start, (p >> 32) & 0xFFFF, p >> 48)

    def list_of(self, seg, idx):
        """(segment, start word, element size code, count, struct sizes)."""
        r = self.resolve(seg, idx)
        if r is None:
            return None
        s, start, p = r
        if p & 3 != 1:
            raise FormatError("expected a list pointer")
        code = (p >> 32) & 7
        count = p >> 35
        if code == 7:
            tag = self.word(s, start)
            n = (tag >> 2) & 0x3FFFFFFF
            return (s, start + 1, 7, n, ((tag >> 32) & 0xFFFF, tag >> 48))
        return (s, start, code, count, None)


class Struct:
    def __init__(self, msg, seg, start, data_words, ptr_words):
        self.msg, self.seg, self.start = msg, seg, start
        self.dw, self.pw = data_words, ptr_words

    def f32(self, slot):
        """Float32 field at 32-bit slot `slot` (0 when outside the data section)."""
        if 4 * slot + 4 > 8 * self.dw:
            return 0.0
        return struct.unpack_from("<f", self.msg.segs[self.seg], 8 * self.start + 4 * slot)[0]

    def u16(self, slot):
        if 2 * slot + 2 > 8 * self.dw:
            return 0
        return struct.unpack_from("<H", self.msg.segs[self.seg], 8 * self.start + 2 * slot)[0]

    def _ptr(self, k):
        return (self.seg, self.start + self.dw + k) if k < self.pw else None

    def struct(self, k):
        loc = self._ptr(k)
        return self.msg.struct(*loc) if loc else None

    def text(self, k):
        loc = self._ptr(k)
        lst = self.msg.list_of(*loc) if loc else None
        if lst is None:
            return ""
        s, start, code, count, _ = lst
        if code != 2:
            raise FormatError("text is not a byte list")
        b = self.msg.segs[s][8 * start:8 * start + count]
        return b.rstrip(b"\0").decode("utf-8")

    def f32_list(self, k):
        loc = self._ptr(k)
        lst = self.msg.list_of(*loc) if loc else None
        if lst is None:
            return np.zeros(0, np.float32)
        s, start, code, count, _ = lst
        if code != 4:
            raise FormatError("expected a list of 32-bit values")
        return np.frombuffer(self.msg.segs[s], np.float32, count, 8 * start).copy()

    def structs(self, k):
        loc = self._ptr(k)
        lst = self.msg.list_of(*loc) if loc else None
        if lst is None:
            return []
        s, start, code, count, sizes = lst
        if code != 7:
            raise FormatError("expected a list of structs")
        dw, pw = sizes
        return [Struct(self.msg, s, start + i * (dw + pw), dw, pw) for i in range(count)]


# ---------------------------------------------------------------------------
# the scene schema (field slots follow the schema's declaration order)

@dataclass
class Transform:
    translation: tuple
    rotation: tuple   # (r, i, j, k)

    def matrix(self):
        """glm::translate(mat4(1), t) * glm::mat4_cast(quat(r, i, j, k))
        (Scene.cpp:67-73), row-major float64."""
        w, x, y, z = (float(c) for c in self.rotation)
        xx, yy, zz = x * x, y * y, z * z
        xz, xy, yz = x * z, x * y, y * z
        wx, wy, wz = w * x, w * y, w * z
        m = np.eye(4)
        m[0, 0] = 1 - 2 * (yy + zz); m[1, 0] = 2 * (xy + wz); m[2, 0] = 2 * (xz - wy)
        m[0, 1] = 2 * (xy - wz); m[1, 1] = 1 - 2 * (xx + zz); m[2, 1] = 2 * (yz + wx)
        m[0, 2] = 2 * (xz + wy); m[1, 2] = 2 * (yz - wx); m[2, 2] = 1 - 2 * (xx + yy)
        m[:3, 3] = [float(c) for c in self.translation]
        return m


@dataclass
class Camera:
    name: str
    transform: Transform
    near: float
    far: float
    fovy: float


@dataclass
class Light:
    name: str
    transform: Transform
    color: tuple
    intensity: float
    type: int


@dataclass
class Mesh:
    name: str
    type: int                 # 0 bezier (16 points per patch), 1 gregory (20)
    positions: np.ndarray     # (n_points, 3) float32

    @property
    def patches(self):
        per = 20 if self.type == 1 else 16
        return self.positions.reshape(-1, per, 3)


@dataclass
class SceneObject:
    name: str
    transform: Transform
    meshname: str
    color: tuple


@dataclass
class SceneFile:
    meshes: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    cameras: list = field(default_factory=list)

    def mesh(self, name):
        for m in self.meshes:
            if m.name == name:
                return m
        raise KeyError(name)


def _vec3(s):
    return (s.f32(0), s.f32(1), s.f32(2)) if s is not None else (0.0, 0.0, 0.0)


def _transform(s):
    if s is None:
        return Transform((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 0.0))
    q = s.struct(1)
    rot = (q.f32(0), q.f32(1), q.f32(2), q.f32(3)) if q is not None else (0.0, 0.0, 0.0, 0.0)
    return Transform(_vec3(s.struct(0)), rot)


def parse(data):
    """Packed `.mscene` bytes -> SceneFile."""
    msg = Message(segments(unpack(data)))
    root = msg.struct(0, 0)
    if root is None:
        raise FormatError("empty message")
    sc = SceneFile()
    for m in root.structs(0):                      # meshes @0
        pos = m.f32_list(1)                        # positions @2 (second pointer)
        if pos.size % 3:
            raise FormatError("mesh positions are not xyz triples")
        mesh = Mesh(m.text(0), m.u16(0), pos.reshape(-1, 3))
        per = 20 if mesh.type == 1 else 16
        if mesh.positions.shape[0] % per:
            raise FormatError("mesh %r: %d points is not a whole number of patches" % (mesh.name, mesh.positions.shape[0]))
        sc.meshes.append(mesh)
    for l in root.structs(1):                      # lights @1
        sc.lights.append(Light(l.text(0), _transform(l.struct(1)), _vec3(l.struct(2)), l.f32(0), l.u16(2)))
    for o in root.structs(2):                      # objects @2
        sc.objects.append(SceneObject(o.text(0), _transform(o.struct(1)), o.text(2), _vec3(o.struct(3))))
    for c in root.structs(3):                      # cameras @3
        sc.cameras.append(Camera(c.text(0), _transform(c.struct(1)), c.f32(0), c.f32(1), c.f32(2)))
    return sc


def load(path):
    with open(path, "rb") as f:
        return parse(f.read())


# ---------------------------------------------------------------------------
# Scene::draw -> a frame of the library's scene description

def draw_workload(scene, window=(1024, 768), camera=0, options=None, name="mscene"):
    """The frame Scene::draw renders (Scene.cpp:179-197) as a scenes.Workload.

    The camera inverse is computed in float64 and rounded to float32 (the
    reference uses glm::inverse in float32; both the library and the oracle
    receive the same rounded modelview)."""
    from . import scenes
    from .renderer import Projection
    cam = scene.cameras[camera]
    inv_cam = np.linalg.inv(cam.transform.matrix())
    proj = Projection(cam.fovy, cam.near, window)
    objs = []
    handles = {}
    for o in scene.objects:
        mesh = scene.mesh(o.meshname)
        h = handles.setdefault(mesh.name, len(handles) + 1)
        mv = (inv_cam @ o.transform.matrix()).astype(np.float32)
        color = np.array(list(o.color) + [1.0], np.float32)
        objs.append(scenes.SceneObject(h, mesh.patches.astype(np.float32), mesh.type, mv, color))
    opts = dict(window_size=window)
    opts.update(options or {})
    return scenes.Workload(name, opts, [scenes.View(proj, objs)])


def draw(renderer, scene, camera=0):
    """Scene::draw (Scene.cpp:179-197) on a Renderer: prepare, load each mesh
    once (handle = mesh), draw every object with inverse(camera) * object
    transform and the camera's projection, finish."""
    from .renderer import Projection
    cam = scene.cameras[camera]
    inv_cam = np.linalg.inv(cam.transform.matrix())
    window = (renderer.cfg.window_width, renderer.cfg.window_height)
    proj = Projection(cam.fovy, cam.near, window)
    handles = {m.name: i + 1 for i, m in enumerate(scene.meshes)}
    renderer.prepare()
    for o in scene.objects:
        mesh = scene.mesh(o.meshname)
        h = handles[mesh.name]
        if not renderer.are_patches_loaded(h):
            renderer.load_patches(h, mesh.patches.astype(np.float32), mesh.type)
        mv = (inv_cam @ o.transform.matrix()).astype(np.float32)
        renderer.draw_patches(h, mv, proj, np.array(list(o.color) + [1.0], np.float32))
    renderer.finish()
