"""`.mscene` reader (micropolis_amd.mscene) against hand-built messages in the
published Cap'n Proto encoding and against invariants of the reference's own
scene files (tests/golden/mscene)."""
import os
import struct

import numpy as np
import pytest

from micropolis_amd import mscene, scenes

HERE = os.path.join(os.path.dirname(__file__), "golden", "mscene")


def pack(raw):
    """Independent packer (Cap'n Proto packed encoding) for test messages."""
    out = bytearray()
    words = [raw[i:i + 8] for i in range(0, len(raw), 8)]
    i = 0
    while i < len(words):
        w = words[i]
        tag = sum(1 << b for b in range(8) if w[b])
        out.append(tag)
        out += bytes(b for b in w if b)
        i += 1
        if tag == 0:
            n = 0
            while i < len(words) and n < 255 and not any(words[i]):
                n += 1
                i += 1
            out.append(n)
        elif tag == 0xFF:
            n = 0
            while i + n < len(words) and n < 255 and sum(1 for b in words[i + n] if b) >= 7:
                n += 1
            out.append(n)
            for k in range(n):
                out += words[i + k]
            i += n
    return bytes(out)


def sptr(off, dw, pw):
    return struct.pack("<Q", ((off & 0x3FFFFFFF) << 2) | (dw << 32) | (pw << 48))


def lptr(off, code, count):
    return struct.pack("<Q", 1 | ((off & 0x3FFFFFFF) << 2) | (code << 32) | (count << 35))


def farptr(seg, pad, double):
    return struct.pack("<Q", 2 | (double << 2) | (pad << 3) | (seg << 32))


def test_unpack_zero_and_literal_runs():
    raw = bytes(8) * 5 + bytes(range(1, 9)) * 3 + b"\x01" + bytes(7)
    assert mscene.unpack(pack(raw)) == raw


def test_mesh_through_a_double_far_pointer():
    """Scene -> meshes list -> one Mesh whose positions list lives in a
    second segment, reached through a double-far landing pad."""
    seg0 = bytearray()
    seg0 += sptr(0, 0, 4)                       # 0: root -> Scene at word 1
    seg0 += lptr(3, 7, 3)                       # 1: Scene.meshes -> tag at word 5
    seg0 += bytes(24)                           # 2-4: lights, objects, cameras = null
    seg0 += struct.pack("<Q", (1 << 2) | (1 << 32) | (2 << 48))  # 5: tag: 1 element, 1 data word, 2 ptrs
    seg0 += struct.pack("<HHI", 1, 0, 0)        # 6: Mesh data: type = gregory
    seg0 += lptr(1, 2, 4)                       # 7: name -> word 9 (4 bytes "abc\\0")
    seg0 += farptr(1, 0, 1)                     # 8: positions -> double-far pad at seg 1 word 0
    seg0 += b"abc\0" + bytes(4)                 # 9
    pts = np.arange(60, dtype=np.float32) * 0.5  # 20 xyz points = one Gregory patch
    seg1 = bytearray()
    seg1 += farptr(1, 2, 0)                     # pad word 0: content starts at seg 1 word 2
    seg1 += lptr(0, 4, 60)                      # pad word 1: tag: 60 four-byte elements
    seg1 += pts.tobytes()                       # words 2..31
    header = struct.pack("<III", 1, len(seg0) // 8, len(seg1) // 8) + bytes(4)
    sc = mscene.parse(pack(header + bytes(seg0) + bytes(seg1)))
    assert len(sc.meshes) == 1 and sc.lights == [] and sc.cameras == []
    m = sc.meshes[0]
    assert m.name == "abc" and m.type == 1
    np.testing.assert_array_equal(m.patches.reshape(-1), pts)


def test_teapot_is_the_utah_teapot():
    sc = mscene.load(os.path.join(HERE, "teapot.mscene"))
    assert len(sc.meshes) == 1 and len(sc.objects) == 1 and len(sc.cameras) == 1
    p = sc.meshes[0].patches
    assert p.shape == (32, 16, 3) and sc.meshes[0].type == 0
    # Newell's data: the rim patch starts at (1.4, 0, 2.4), (1.4, -0.784, 2.4)
    np.testing.assert_allclose(p[0, :2], [[1.4, 0.0, 2.4], [1.4, -0.784, 2.4]], rtol=0, atol=1e-6)
    cam = sc.cameras[0]
    assert cam.fovy == pytest.approx(60.0) and cam.near == pytest.approx(0.01)
    R = cam.transform.matrix()[:3, :3]
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)


def test_gregory_teapot_is_the_bezier_teapot_remapped():
    """tools/bez2mscene_greg.py:111 maps Bezier points to the Gregory layout;
    the two reference files agree exactly under that map."""
    b = mscene.load(os.path.join(HERE, "teapot.mscene")).meshes[0].patches
    g = mscene.load(os.path.join(HERE, "teapot_greg.mscene")).meshes[0]
    assert g.type == 1 and g.patches.shape == (32, 20, 3)
    np.testing.assert_array_equal(g.patches, b[:, scenes.BEZ_TO_GREGORY])


def test_teapots_objects_and_draw_workload():
    sc = mscene.load(os.path.join(HERE, "teapots.mscene"))
    assert len(sc.objects) == 20 and all(o.meshname == sc.meshes[0].name for o in sc.objects)
    w = mscene.draw_workload(sc, window=(320, 240))
    objs = w.views[0].objects
    assert len(objs) == 20 and len({o.handle for o in objs}) == 1
    for o, so in zip(objs, sc.objects):
        # modelview = inverse(camera) * object: camera-space origin of the object
        want = np.linalg.inv(sc.cameras[0].transform.matrix()) @ np.append(so.transform.translation, 1.0)
        np.testing.assert_allclose(o.matrix[:, 3], want, rtol=1e-5, atol=1e-5)
        assert o.color[3] == 1.0


def test_oracle_renders_the_teapot(oracle):
    sc = mscene.load(os.path.join(HERE, "teapot.mscene"))
    w = mscene.draw_workload(sc, window=(320, 240), options=dict(bound_n_split_method="LOCAL"))
    out = oracle.render(w)[0]
    cov = out["ids"] != 0xFFFFFFFF
    assert 0.1 < cov.mean() < 0.4
    ys, xs = np.nonzero(cov)
    assert abs(xs.mean() - 160) < 40 and abs(ys.mean() - 120) < 40
