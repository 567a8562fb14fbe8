"""The pbrt scene-file reader (pbrt-v2_amd/host/pbrt_scene.{h,cpp}) driven through
tests/cpp/libmirror_driver.so.

CPU tests pin the reader without a device:
- world-space vertices against closed forms: dyadic Translate / Scale give exact FP32
  results, and Rotate 90 about z is checked to within an ulp of the rotated axes;
- the reference's parameter typing (integers read as floats, "bool" strings, prefix-matched
  type names) and its directive rules: Include, attribute / transform stacks,
  ObjectBegin / ObjectInstance, animated transforms, non-triangle shapes, heightfields;
- error behaviour: syntax errors are fatal, what pbrt reports with Error() and ignores is
  recorded as a message;
- the reference's own sample scenes when the reference tree is present, against an
  independent count of their "integer indices".
PFM output and an image comparison with exrdiff's statistics are pinned on synthetic images.
The GPU test renders a scene file through RenderScene and compares the image with one built
from the oracle's hits and occlusion on the same camera rays.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "cpp" / "libmirror_driver.so"
REF_SCENES = Path("/root/reference/scenes")
F = np.float32


def _driver():
    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(DRIVER))
    lib.mirror_parse.restype = ctypes.c_int64
    lib.mirror_render_scene.restype = ctypes.c_int
    lib.mirror_last_error.restype = ctypes.c_char_p
    return lib


def parse(path):
    """Returns (v9 (n, 9) float32, info dict, messages list); raises RuntimeError on Severe."""
    lib = _driver()
    info = ctypes.create_string_buffer(1 << 20)
    n = lib.mirror_parse(str(path).encode(), None, ctypes.c_uint64(0), info,
                         ctypes.c_uint64(1 << 20))
    if n < 0:
        raise RuntimeError(lib.mirror_last_error().decode())
    v9 = np.zeros((max(n, 1), 9), F)
    n2 = lib.mirror_parse(str(path).encode(), ctypes.c_void_p(v9.ctypes.data), ctypes.c_uint64(n),
                          info, ctypes.c_uint64(1 << 20))
    assert n2 == n
    d, msgs = {}, []
    for line in info.value.decode().splitlines():
        k, _, v = line.partition("=")
        if k == "message":
            msgs.append(v)
        else:
            d[k] = v
    return v9[:n], d, msgs


def floats(s):
    return np.array([float(x) for x in s.split(",")], F)


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


SCENE = """# options block
LookAt 0 0 -3  0 0 0  0 1 0
Camera "perspective" "float fov" [50]
Film "image" "integer xresolution" [64] "integer yresolution" [32]
Sampler "stratified" "integer xsamples" [2] "integer ysamples" [1]
Accelerator "ray-bvh" "integer chunkXSize" [8] "string node" ["ia"] "bool sort" "true"
  "integer height" [2 3]
SurfaceIntegrator "path" "integer maxdepth" [2]
Renderer "breadth-first"
WorldBegin
LightSource "point" "point from" [1 2 3] "color I" [2 2 2]
AttributeBegin
  Translate 1 2 3
  Scale 2 0.5 4
  Shape "trianglemesh" "point P" [0 0 0  1 0 0  0 1 0  1 1 1]
        "integer indices" [0 1 2  1 3 2]
AttributeEnd
Shape "trianglemesh" "point P" [ 5 5 5  6 5 5  5 6 5 ] "integer indices" [0 1 2 1]
      "float uv" [0 0 1 0 1 1]
Shape "sphere" "float radius" 1
ObjectBegin "thing"
  Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]
ObjectEnd
ObjectInstance "thing"
TransformBegin
  Rotate 90 0 0 1
  Shape "trianglemesh" "point P" [1 0 0  0 1 0  0 0 1] "integer indices" [0 1 2]
TransformEnd
Include "part.pbrt"
Shape "heightfield" "integer nu" 3 "integer nv" 2 "float Pz" [0 1 2 3 4 5]
ActiveTransform EndTime
Translate 0 0 1
ActiveTransform All
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]
WorldEnd
"""

PART = """# included relative to the main file's directory
Shape "trianglemesh" "integerindices" [0 2 1] "point P" [-1 -1 -1  -2 -1 -1  -1 -2 -1]
"""


def test_scene_geometry_and_options(tmp_path):
    write(tmp_path, "part.pbrt", PART)
    v9, info, msgs = parse(write(tmp_path, "main.pbrt", SCENE))
    print("[scene] messages:", *msgs, sep="\n  ")
    tri = v9.reshape(-1, 3, 3)
    # Translate 1 2 3 then Scale 2 .5 4: p' = (2x + 1, .5y + 2, 4z + 3), exact in FP32
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 1]], F)
    W = P * np.array([2, 0.5, 4], F) + np.array([1, 2, 3], F)
    assert np.array_equal(tri[0], W[[0, 1, 2]])
    assert np.array_equal(tri[1], W[[1, 3, 2]])
    # identity transform; the 4th index (a partial triangle) is dropped as nvi / 3 does
    assert np.array_equal(tri[2], np.array([[5, 5, 5], [6, 5, 5], [5, 6, 5]], F))
    # Rotate 90 about z: x -> y, y -> -x (sinf / cosf of (float)pi/180*90)
    rot = np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1]], F)
    np.testing.assert_allclose(tri[3], rot, atol=1e-7, rtol=0)
    assert tri[3][2].tolist() == [0.0, 0.0, 1.0]
    # Include, with the prefix-matched type name "integerindices"
    assert np.array_equal(tri[4], np.array([[-1, -1, -1], [-1, -2, -1], [-2, -1, -1]], F))
    # heightfield 3 x 2: vertices (x / 2, y, Pz), two triangles per cell
    hp = np.array([[x / 2, y, 3 * y + x] for y in range(2) for x in range(3)], F)
    cells = [[0, 1, 4], [0, 4, 3], [1, 2, 5], [1, 5, 4]]
    assert np.array_equal(tri[5:9], hp[cells])
    assert len(tri) == 9  # sphere, instance shape, ObjectInstance and animated shape skipped
    assert int(info["skipped"]) == 4
    assert int(info["meshes"]) == 5
    # options
    assert info["accelerator"] == "ray-bvh"
    p = info["params"]
    assert "integer chunkXSize=8;" in p and "string node=ia;" in p and "bool sort=true;" in p
    assert "height" not in p  # two values: not a single-valued parameter
    assert info["integrator"] == "path" and int(info["max_depth"]) == 2
    assert info["renderer"] == "breadth-first"
    assert info["film"] == "64,32" and int(info["spp"]) == 2
    assert float(info["fov"]) == 50.0
    # LookAt from (0,0,-3) toward +z: identity rotation, eye as translation
    c2w = floats(info["camera_to_world"]).reshape(3, 4)
    assert np.array_equal(c2w, np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, -3]], F))
    assert np.array_equal(floats(info["screen"]), np.array([-2, 2, -1, 1], F))  # aspect 2
    assert np.array_equal(floats(info["light"]), np.array([1, 2, 3], F))
    assert abs(float(info["intensity"]) - 2.0) < 1e-6
    assert any("animated" in m for m in msgs)
    assert any("ObjectInstance" in m for m in msgs)


def _f32_normalize(v):
    # Normalize (geometry.h:511): v * (1 / sqrtf((x*x + y*y) + z*z)) in float
    l2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    inv = F(1) / np.sqrt(F(l2))
    return np.array([v[0] * inv, v[1] * inv, v[2] * inv], F)


def _cross(a, b, dtype):
    a, b = a.astype(dtype), b.astype(dtype)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]]).astype(F)


def test_scene_lookat_cross_in_double(tmp_path):
    # LookAt (transform.cpp:228-254) builds left = Normalize(Cross(Normalize(up), dir)) and
    # newUp = Cross(dir, left); pbrt's Cross computes in double and rounds once
    # (geometry.h:481-488).  For this LookAt an FP32 cross product gives a different frame.
    eye, look, up = (np.array(x, F) for x in ((1.4, -2.3, -4.6), (-4.8, 3.1, 4.1), (0.2, 0.5, 0.1)))
    text = f"""LookAt {' '.join(map(str, [1.4, -2.3, -4.6, -4.8, 3.1, 4.1, 0.2, 0.5, 0.1]))}
Camera "perspective"
WorldBegin
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]
WorldEnd
"""
    _, info, _ = parse(write(tmp_path, "l.pbrt", text))
    c2w = floats(info["camera_to_world"]).reshape(3, 4)

    def frame(dtype):
        d = _f32_normalize((look - eye).astype(F))
        left = _f32_normalize(_cross(_f32_normalize(up), d, dtype))
        return np.stack([left, _cross(d, left, dtype), d], 1)

    assert np.array_equal(c2w[:, :3], frame(np.float64))
    assert not np.array_equal(frame(np.float64), frame(np.float32))
    assert np.array_equal(c2w[:, 3], eye)


def test_scene_defaults_and_light_frames(tmp_path):
    # no options: pbrt's defaults; the light follows CoordSysTransform "camera"
    text = """LookAt 1 2 3  1 2 4  0 1 0
Camera "perspective"
WorldBegin
AttributeBegin
CoordSysTransform "camera"
LightSource "point" "color I" [.7 .7 .7] "rgb scale" [2 2 2]
AttributeEnd
AttributeBegin
Translate 0 0 1
LightSource "point" "point from" [5 5 5]
AttributeEnd
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]
WorldEnd
"""
    v9, info, msgs = parse(write(tmp_path, "d.pbrt", text))
    assert len(v9) == 1
    assert info["accelerator"] == "bvh" and info["integrator"] == "directlighting"
    assert info["sampler"] == "lowdiscrepancy" and int(info["spp"]) == 4
    assert info["film"] == "640,480" and float(info["fov"]) == 90.0
    assert np.array_equal(floats(info["light"]), np.array([1, 2, 3], F))  # at the camera
    assert abs(float(info["intensity"]) - 1.4) < 1e-6
    assert any("more than one point light" in m for m in msgs)


def test_scene_errors_and_reports(tmp_path):
    with pytest.raises(RuntimeError, match="Parsing error"):
        parse(write(tmp_path, "a.pbrt", 'WorldBegin\nShapee "trianglemesh"\nWorldEnd\n'))
    with pytest.raises(RuntimeError, match="Parsing error"):
        parse(write(tmp_path, "b.pbrt", "WorldBegin\nTranslate 1 2\nWorldEnd\n"))
    with pytest.raises(RuntimeError, match="Unable to open"):
        parse(write(tmp_path, "c.pbrt", 'WorldBegin\nInclude "missing.pbrt"\nWorldEnd\n'))
    with pytest.raises(RuntimeError, match="Unable to open"):
        parse(tmp_path / "no_such_file.pbrt")
    text = """WorldBegin
Film "image" "integer xresolution" [10]
AttributeEnd
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 3]
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2] "float uv" [0 0 1]
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2] "bool x" [1]
Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1] "integer indices" [0 1 2]
Shape "nurbs" "integer nu" 2 "integer nv" 2
WorldEnd
"""
    v9, info, msgs = parse(write(tmp_path, "e.pbrt", text))
    joined = "\n".join(msgs)
    assert "Options cannot be set inside world block" in joined
    assert "Unmatched pbrtAttributeEnd" in joined
    assert "out of-bounds vertex index 3" in joined
    assert "Discarding" in joined  # too few uvs: the mesh stays, without uvs
    assert "Expected string parameter value" in joined
    assert "Excess values" in joined  # 8 numbers for "point P"
    assert "nurbs" in joined
    assert info["film"] == "640,480"  # the Film inside the world block was ignored
    # dropped: index 3 of three points, and index 2 of the two points left after the excess
    # value; kept: the mesh whose uvs were discarded and the one with the ignored "bool x"
    assert len(v9) == 2


def _mesh_text(P, idx, levels, shape="loopsubdiv"):
    pts = " ".join(f"{x:.9g}" for x in np.asarray(P, np.float64).reshape(-1))
    ind = " ".join(str(int(i)) for i in np.asarray(idx).reshape(-1))
    return (f'WorldBegin\nShape "{shape}" "integer nlevels" [{levels}] "point P" [{pts}] '
            f'"integer indices" [{ind}]\nWorldEnd\n')


def test_loop_subdivision(tmp_path):
    """Loop subdivision (LoopSubdiv) pinned by what the scheme fixes:
    - each level splits every face in 4, and adds one vertex per edge (Euler count);
    - the limit surface of a planar control mesh stays in its plane, exactly;
    - on a regular planar grid, affine functions are reproduced, so interior vertices keep
      their place in the plane (to rounding);
    - a closed regular octahedron subdivides to a surface symmetric under its own symmetries:
      the limit vertices come in orbits of equal radius;
    - nlevels 0 gives the limit positions of the control mesh itself."""
    # octahedron: closed, every vertex of valence 4
    P = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], F)
    idx = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4],
                    [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    for levels in (0, 1, 2, 3):
        v9, info, msgs = parse(write(tmp_path, f"oct{levels}.pbrt", _mesh_text(P, idx, levels)))
        assert len(v9) == 8 * 4 ** levels, msgs
        verts = np.unique(v9.reshape(-1, 3), axis=0).astype(np.float64)
        nv = 6 + sum(12 * 4 ** l for l in range(levels))  # V' = V + E, E = 3F/2 per level
        assert len(verts) == nv, (levels, len(verts), nv)
        # the octahedron's symmetries map the limit vertex set onto itself (to rounding: the
        # one-ring sums start at different neighbours)
        for T in (lambda v: v * [-1, 1, 1], lambda v: v[:, [1, 0, 2]], lambda v: v[:, [2, 0, 1]]):
            d = np.abs(T(verts)[:, None, :] - verts[None, :, :]).max(-1).min(1)
            assert d.max() < 1e-5, (levels, d.max())
        if levels == 0:
            # limit rule alone (gamma(4) = 1/8): (1 - 4/8) v + 1/8 (sum of the 4 neighbours),
            # which is v / 2 here, exactly in FP32
            want = np.unique(P.astype(np.float64) * 0.5, axis=0)
            assert np.array_equal(verts, want)
        else:
            r = np.sqrt((verts ** 2).sum(1))
            assert r.min() > 0.4 and r.max() <= 0.5 + 1e-6  # inside the level-0 limit hull
    # regular tetrahedron, valence 3 (gamma = 1/5): the limit of v is 0.4 v + 0.2 (sum of the
    # other three) = 0.2 v, since the four vertices sum to 0
    T = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], F)
    v9, _, _ = parse(write(tmp_path, "tet.pbrt", _mesh_text(T, [[0, 1, 2], [0, 3, 1], [0, 2, 3],
                                                                 [1, 3, 2]], 0)))
    np.testing.assert_allclose(np.unique(v9.reshape(-1, 3), axis=0),
                               np.unique(T * F(0.2), axis=0), atol=1e-6)
    # a lone triangle: every vertex on the boundary (rule 1/5): 0.6 v + 0.2 (a + b)
    v9, _, _ = parse(write(tmp_path, "tri.pbrt", _mesh_text([[0, 0, 0], [1, 0, 0], [0, 1, 0]],
                                                            [[0, 1, 2]], 0)))
    np.testing.assert_allclose(v9.reshape(3, 3), [[0.2, 0.2, 0], [0.6, 0.2, 0], [0.2, 0.6, 0]],
                               atol=1e-6)
    # planar grid 5 x 5 (with boundary), dyadic coordinates
    g = 5
    P = np.array([[x, y, 0] for y in range(g) for x in range(g)], F)
    tri = []
    for y in range(g - 1):
        for x in range(g - 1):
            a, b, c, d = y * g + x, y * g + x + 1, (y + 1) * g + x + 1, (y + 1) * g + x
            tri += [[a, b, c], [a, c, d]]
    v9, info, msgs = parse(write(tmp_path, "grid.pbrt", _mesh_text(P, tri, 2)))
    assert len(v9) == len(tri) * 16
    q = v9.reshape(-1, 3)
    assert np.all(q[:, 2] == 0)  # planar stays planar, exactly
    inner = (q[:, 0] > 1.01) & (q[:, 0] < 2.99) & (q[:, 1] > 1.01) & (q[:, 1] < 2.99)
    # the interior of a regular planar grid reproduces the parametrisation: the limit points
    # lie on the 1/4 lattice (the level-2 vertices), to rounding
    np.testing.assert_allclose(q[inner, :2] * 4, np.round(q[inner, :2] * 4), atol=2e-5)
    # out-of-range control index: reported and skipped
    v9, info, msgs = parse(write(tmp_path, "bad.pbrt", _mesh_text(P, [[0, 1, 99]], 1)))
    assert len(v9) == 0 and any("out of-bounds" in m for m in msgs)


def _index_count(path, seen):
    """Triangles of every "trianglemesh" in a scene file and its Includes, counted from the
    text alone (independent of the reader): len("integer indices") // 3."""
    text = path.read_text()
    n = 0
    for m in re.finditer(r'"integer\s+indices"\s*\[([^\]]*)\]', text):
        n += len(m.group(1).split()) // 3
    for inc in re.findall(r'^\s*Include\s+"([^"]+)"', text, flags=re.M):
        n += _index_count(REF_SCENES / inc, seen)
    return n


@pytest.mark.skipif(not REF_SCENES.is_dir(), reason="reference scene files not present")
@pytest.mark.parametrize("name", ["bunny.pbrt", "prt-teapot.pbrt", "killeroo-simple.pbrt"])
def test_reference_sample_scenes(name):
    """The reference's own sample scenes parse; triangle counts match an independent count of
    their index lists (these two have no non-triangle or skipped meshes)."""
    v9, info, msgs = parse(REF_SCENES / name)
    print(f"[scene] {name}: {len(v9)} triangles, accelerator {info['accelerator']} "
          f"{info['params']}, messages {msgs}")
    assert np.all(np.isfinite(v9))
    if name == "killeroo-simple.pbrt":
        # two quads, and the killeroo twice: its control faces x 4^nlevels (nlevels 1)
        text = (REF_SCENES / "geometry" / "killeroo.pbrt").read_text()
        nf = len(re.search(r'"integer indices"\s*\[([^\]]*)\]', text).group(1).split()) // 3
        lv = int(re.search(r'"integer nlevels"\s*\[\s*(\d+)', text).group(1))
        assert len(v9) == 4 + 2 * nf * 4 ** lv
        assert int(info["skipped"]) == 1  # the sphere of the area light
        return
    assert len(v9) == _index_count(REF_SCENES / name, set())
    assert info["accelerator"] == "rayhierarchy"
    if name == "bunny.pbrt":
        # the bunny's vertices sit under "Translate 0 -.033 0": y' = fl(y + (-0.033f))
        head = (REF_SCENES / "geometry" / "bunny.pbrt").read_text()
        P = re.search(r'"point P"\s*\[([^\]]*)\]', head).group(1).split()
        I = re.search(r'"integer indices"\s*\[([^\]]*)\]', head).group(1).split()
        p = np.array(P, dtype=np.float64).astype(F).reshape(-1, 3)
        idx = np.array(I[:3], dtype=np.int64)
        want = p[idx] + np.array([0, F(-0.033), 0], F)
        assert np.array_equal(v9[2].reshape(3, 3), want)  # after the 2-triangle floor
        assert np.array_equal(floats(info["light"]), floats(info["camera_to_world"])[[3, 7, 11]])
    else:
        assert "integer chunkXSize=100;" in info["params"] and "string node=ia;" in info["params"]
        assert info["integrator"] == "path" and int(info["max_depth"]) == 8


def _pfm_bytes(img, little):
    """A PFM file written by numpy alone: header, then rows bottom first."""
    H, W, _ = img.shape
    scale = b"-1" if little else b"1"
    body = img[::-1].astype("<f4" if little else ">f4").tobytes()
    return b"PF\n%d %d\n" % (W, H) + scale + b"\n" + body


def test_pfm_write_and_read(tmp_path):
    lib = _driver()
    rng = np.random.default_rng(11)
    W, H = 7, 5
    img = rng.random((H, W, 3)).astype(F)
    back = np.zeros_like(img)
    path = tmp_path / "a.pfm"
    assert lib.mirror_pfm_roundtrip(str(path).encode(), W, H, ctypes.c_void_p(img.ctypes.data),
                                    ctypes.c_void_p(back.ctypes.data)) == 0
    assert np.array_equal(back, img)
    # WritePFM's bytes are the format itself (little-endian, bottom row first)
    assert path.read_bytes() == _pfm_bytes(img, little=True)
    # ReadPFM takes either byte order
    for little in (True, False):
        q = tmp_path / f"b{int(little)}.pfm"
        q.write_bytes(_pfm_bytes(img, little))
        out = np.zeros_like(img)
        x, y = ctypes.c_int(), ctypes.c_int()
        assert lib.mirror_read_pfm(str(q).encode(), ctypes.byref(x), ctypes.byref(y),
                                   ctypes.c_void_p(out.ctypes.data), ctypes.c_uint64(out.size)) == 0
        assert (x.value, y.value) == (W, H) and np.array_equal(out, img)
    (tmp_path / "bad.pfm").write_bytes(b"P6\n2 2\n255\n")
    out = np.zeros(12, F)
    assert lib.mirror_read_pfm(str(tmp_path / "bad.pfm").encode(), ctypes.byref(x),
                               ctypes.byref(y), ctypes.c_void_p(out.ctypes.data),
                               ctypes.c_uint64(12)) == 1


def test_image_diff_statistics():
    """exrdiff's statistics on images with planted differences: a sample off by 1% is small,
    one off by 10% is small and big, samples 0 in both images are skipped."""
    lib = _driver()
    W, H = 4, 3
    a = np.ones((H, W, 3), F)
    a[0, 0] = 0  # a pixel black in both images
    b = a.copy()
    b[1, 1, 0] = F(1.01)
    b[2, 3, 2] = F(1.10)
    out = np.zeros(7, np.float64)

    def diff(x, y, tol):
        assert lib.mirror_diff_images(ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data),
                                      W, H, ctypes.c_float(tol),
                                      ctypes.c_void_p(out.ctypes.data)) == 0
        return out.copy()

    d = diff(a, b, 0.0)
    n = 3 * W * H
    assert (d[0], d[1]) == (2, 1)
    assert d[2] == pytest.approx(float(a.sum(dtype=np.float64)) / n)
    assert d[3] == pytest.approx(float(b.sum(dtype=np.float64)) / n)
    assert d[4] == pytest.approx(100 * (d[2] - d[3]) / min(d[2], d[3]))
    assert d[5] == pytest.approx(float(((a.astype(np.float64) - b) ** 2).sum()) / n)
    assert d[6] == 1  # tol 0: any difference counts
    assert diff(a, b, 1.0)[6] == 0  # average delta ~0.3% is within 1%
    assert diff(a, b, 0.1)[6] == 1
    assert diff(a, a, 0.0)[6] == 0


# ------------------------------------------------------------------------------------ GPU
def _camera_rays(W, H, fov, c2w, screen):
    """numpy twin of the renderer loop's perspective camera (one centred sample per pixel),
    in the loop's FP32 operation order."""
    t = F(np.tan(np.float64(fov) * (np.pi / 180.0) / 2.0))
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    fx = xs.astype(F) + F(0.5)
    fy = ys.astype(F) + F(0.5)
    sx = screen[0] + (fx / F(W)) * (screen[1] - screen[0])
    sy = screen[3] - (fy / F(H)) * (screen[3] - screen[2])
    x, y, z = sx * t, sy * t, np.ones_like(sx)
    inv = F(1) / np.sqrt((x * x + y * y) + z * z)
    d = [x * inv, y * inv, z * inv]
    M = c2w
    w = [(M[r, 0] * d[0] + M[r, 1] * d[1]) + M[r, 2] * d[2] for r in range(3)]
    n = W * H
    return {"ox": np.full(n, M[0, 3], F), "oy": np.full(n, M[1, 3], F),
            "oz": np.full(n, M[2, 3], F), "dx": w[0].reshape(-1), "dy": w[1].reshape(-1),
            "dz": w[2].reshape(-1), "tmin": np.zeros(n, F), "tmax": np.full(n, np.inf, F),
            "valid": np.ones(n, np.uint8), "width": W, "height": H}


@pytest.mark.gpu
@pytest.mark.parametrize("accel,accel_params,kw", [
    ("rayhierarchy", "", {}),
    ("ray-bvh", "", {}),
    # the Accelerator line of prt-teapot.pbrt and sibenik.pbrt (200x8-ray blocks here)
    ("rayhierarchy", '"integer chunkXSize" [100] "integer chunkYSize" [4] "integer height" [4] '
     '"string node" ["ia"]', {"cx": 100, "cy": 4, "height": 4, "node": "ia"}),
], ids=["rayhierarchy", "ray-bvh", "rayhierarchy-100x4-ia"])
def test_scene_file_render_matches_oracle(scenes, orc, tmp_path, accel, accel_params, kw):
    """A torus scene written as a pbrt file: the reader returns its vertices bit for bit, and
    RenderScene (MakeAccelerator from the file's Accelerator, the breadth-first loop) gives
    the direct-lighting image of the oracle's hits and occlusion on the same camera rays."""
    v9 = scenes.torus_scene()
    pts = " ".join(f"{x:.9g}" for x in v9.reshape(-1))
    idx = " ".join(str(i) for i in range(3 * len(v9)))
    W, H = 256, 128
    text = f"""LookAt 0 0 -3  0 0 0  0 1 0
Camera "perspective" "float fov" [60]
Film "image" "integer xresolution" [{W}] "integer yresolution" [{H}]
Sampler "stratified" "integer xsamples" [1] "integer ysamples" [1]
Accelerator "{accel}" {accel_params}
Renderer "breadth-first"
SurfaceIntegrator "directlighting"
WorldBegin
LightSource "point" "point from" [2 2 -2] "color I" [10 10 10]
Shape "trianglemesh" "point P" [{pts}] "integer indices" [{idx}]
WorldEnd
"""
    path = write(tmp_path, "torus.pbrt", text)
    vp, info, msgs = parse(path)
    assert np.array_equal(vp, v9.reshape(-1, 9)), "vertices changed on the way through the file"
    lib = _driver()
    rgb = np.zeros(3 * W * H, F)
    stats = np.zeros(6, np.int64)
    print(f"[scene] render {accel} {W}x{H}", flush=True)
    pfm = tmp_path / "torus.pfm"
    rc = lib.mirror_render_scene(str(path).encode(), b"", ctypes.c_void_p(rgb.ctypes.data),
                                 ctypes.c_uint64(rgb.size), ctypes.c_void_p(stats.ctypes.data),
                                 str(pfm).encode())
    if rc != 0:
        raise RuntimeError(lib.mirror_last_error().decode())
    img = rgb.reshape(H, W, 3)[..., 0]
    assert stats[0] == 1 and stats[3] == W * H

    c2w = floats(info["camera_to_world"]).reshape(3, 4)
    r = _camera_rays(W, H, 60.0, c2w, floats(info["screen"]))
    light = tuple(floats(info["light"]).tolist())
    intensity = F(float(info["intensity"]))
    p = orc.params(accel, **kw)
    if accel == "ray-bvh":
        nodes, roots, order = orc.raybvh_handoff(v9, max_prims=1, bvh_height=3)
        trace = lambda q, m: orc.trace_raybvh(v9, nodes, roots, order, q, p, m)
    else:
        trace = lambda q, m: orc.trace(v9, q, p, m)
    h = trace(r, orc.CLOSEST_PRIMARY)
    sh = scenes.shadow_rays_np(r, h["id"], h["t"], light)
    occ = trace(sh, orc.OCCLUDED)["occ"]
    tri = v9.reshape(-1, 9)
    n = np.cross(tri[:, 3:6] - tri[:, 0:3], tri[:, 6:9] - tri[:, 0:3]).astype(F)
    nl = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    n = (n * (F(1) / nl)[:, None]).astype(F)
    nn = n[np.maximum(h["id"], 0)]
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1)
    flip = ((nn[:, 0] * d[:, 0] + nn[:, 1] * d[:, 1]) + nn[:, 2] * d[:, 2]) > 0
    nn = np.where(flip[:, None], -nn, nn)
    wi = np.stack([sh["dx"], sh["dy"], sh["dz"]], 1)
    c = np.maximum(F(0), (nn[:, 0] * wi[:, 0] + nn[:, 1] * wi[:, 1]) + nn[:, 2] * wi[:, 2])
    with np.errstate(divide="ignore", invalid="ignore"):  # misses: masked below
        L = ((F(0.5) * F(0.31830988618379067154)) * intensity) * c / (sh["tmax"] * sh["tmax"])
    lit = (h["id"] >= 0) & (occ == 0)
    ref = np.where(lit, L, F(0)).reshape(H, W)
    assert np.count_nonzero(ref) > 2000
    np.testing.assert_allclose(img, ref, rtol=1e-5, atol=1e-7)
    # the PFM written by RenderScene holds the same image, and exrdiff's test passes on it
    back = np.zeros(3 * W * H, F)
    x, y = ctypes.c_int(), ctypes.c_int()
    assert lib.mirror_read_pfm(str(pfm).encode(), ctypes.byref(x), ctypes.byref(y),
                               ctypes.c_void_p(back.ctypes.data), ctypes.c_uint64(back.size)) == 0
    assert np.array_equal(back, rgb)
    ref3 = np.repeat(ref.reshape(-1), 3).astype(F)
    out = np.zeros(7, np.float64)
    assert lib.mirror_diff_images(ctypes.c_void_p(back.ctypes.data),
                                  ctypes.c_void_p(ref3.ctypes.data), W, H, ctypes.c_float(0),
                                  ctypes.c_void_p(out.ctypes.data)) == 0
    assert out[6] == 0, f"images differ: {out}"
