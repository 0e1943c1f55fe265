"""Texture maps and vertex normals in the oracle and the scene loader.

The filter is pinned by what the CUDA Programming Guide's linear filtering
fixes (texel centres return the texel, weights carry 8 fractional bits,
wrap/clamp addressing, bottom-up rows, R in the low byte); the integrator's
texture plumbing by exact invariants (a white texture leaves every path
bit-identical, a black one equals a black material) and by reading the
textured debug frames back through the hit positions."""
import numpy as np
import pytest

from rtlib_amd import scenefile, scenes

NEE, RIS, MIS, RR = 1, 2, 128, 256


def _rgba(r, g, b, a=255):
    return np.uint32(r) | (np.uint32(g) << 8) | (np.uint32(b) << 16) | (np.uint32(a) << 24)


def test_texel_centres_return_the_texel(orc):
    rng = np.random.default_rng(5)
    W, H = 8, 4
    tx = rng.integers(0, 2**32, W * H, dtype=np.uint64).astype(np.uint32)
    for mode in (0, 1):
        for j in range(H):
            for i in range(W):
                got = orc.tex_sample(tx, [0, W, H, mode], (i + 0.5) / W, (j + 0.5) / H)
                c = tx[j * W + i]
                exp = np.array([(c >> s) & 0xFF for s in (0, 8, 16)], np.float32) / np.float32(255)
                assert np.array_equal(got, exp), (mode, i, j)


def test_midpoints_average_four_texels(orc):
    rng = np.random.default_rng(6)
    W, H = 4, 4
    tx = rng.integers(0, 2**32, W * H, dtype=np.uint64).astype(np.uint32)
    t = np.array([[(c >> s) & 0xFF for s in (0, 8, 16)] for c in tx], np.float64).reshape(H, W, 3) / 255.0
    for j in range(H - 1):
        for i in range(W - 1):
            got = orc.tex_sample(tx, [0, W, H, 0], (i + 1) / W, (j + 1) / H)
            exp = t[j:j + 2, i:i + 2].mean(axis=(0, 1))
            assert np.allclose(got, exp, atol=2e-7)


def test_weights_have_eight_fractional_bits(orc):
    # 2 x 1 ramp 0 -> 255 along x: between the centres the result is the
    # weight itself, round(frac * 256) / 256, not frac
    tx = np.array([_rgba(0, 0, 0), _rgba(255, 0, 0)], np.uint32)
    for u, frac in [(0.3, 0.1), (0.4, 0.3), (0.6, 0.7), (0.55, 0.6)]:
        got = float(orc.tex_sample(tx, [0, 2, 1, 0], u, 0.5)[0])
        x = np.float32(np.float32(u) * np.float32(2.0) - np.float32(0.5))
        f = float(x - np.floor(x))
        assert abs(f - frac) < 1e-6
        assert got == np.rint(f * 256.0) / 256.0
        assert got != np.float32(f)  # the quantisation is visible at these points
    # exactly representable weights pass through
    for k in (0, 1, 77, 128, 255):
        u = (0.5 + k / 256.0) / 2.0
        assert float(orc.tex_sample(tx, [0, 2, 1, 0], u, 0.5)[0]) == k / 256.0


def test_wrap_and_clamp_addressing(orc):
    rng = np.random.default_rng(7)
    W, H = 4, 2
    tx = rng.integers(0, 2**32, W * H, dtype=np.uint64).astype(np.uint32)
    s = lambda u, v, m: orc.tex_sample(tx, [0, W, H, m], u, v)
    for u, v in [(0.3, 0.7), (0.61, 0.2), (0.05, 0.9)]:
        # wrap: periodic in u and v
        assert np.array_equal(s(u, v, 1), s(u + 1.0, v, 1))
        assert np.array_equal(s(u, v, 1), s(u - 2.0, v - 1.0, 1))
    # clamp: outside [0, 1] is the edge texel row / column
    t = np.array([[(c >> sh) & 0xFF for sh in (0, 8, 16)] for c in tx], np.float32).reshape(H, W, 3) / np.float32(255)
    assert np.array_equal(s(-3.0, 0.25, 0), t[0, 0]) and np.array_equal(s(7.5, 0.25, 0), t[0, W - 1])
    assert np.array_equal(s(0.125, 5.0, 0), t[H - 1, 0])
    # at u = 0 wrap blends the last and first columns half and half; clamp does not blend
    assert np.allclose(s(0.0, 0.25, 1), 0.5 * (t[0, W - 1] + t[0, 0]), atol=1e-7)
    assert np.array_equal(s(0.0, 0.25, 0), t[0, 0])
    # non-finite coordinates read as 0
    assert np.array_equal(s(np.nan, 0.25, 1), s(0.0, 0.25, 1))


def test_rows_bottom_up_and_channel_order(orc):
    tx = np.array([_rgba(1, 2, 3, 4), _rgba(250, 251, 252, 5)], np.uint32)  # row 0, row 1 of a 1 x 2 texture
    lo = orc.tex_sample(tx, [0, 1, 2, 0], 0.5, 0.25)
    hi = orc.tex_sample(tx, [0, 1, 2, 0], 0.5, 0.75)
    assert np.array_equal(lo, np.array([1, 2, 3], np.float32) / np.float32(255))
    assert np.array_equal(hi, np.array([250, 251, 252], np.float32) / np.float32(255))


def _render(orc, sc, lm, spp, flags, **kw):
    cam = sc["camera"]
    w, h = cam["width"], cam["height"]
    uvw = scenes.camera_uvw(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], w / h)
    acc, _ = orc.render(orc.Scene(sc, light_mesh=lm), uvw, w, h, spp, flags, **kw)
    return acc


def _with_constant_textures(sc, c):
    """Every slot of every triangle on one 1 x 1 texture of value c."""
    s = dict(sc)
    nt = sc["idx"].shape[0]
    s["texcoord"] = np.random.default_rng(1).random((nt, 3, 2)).astype(np.float32) * 3 - 1
    s["tex_id"] = np.zeros((nt, 4), np.uint32)
    s["texels"] = np.array([_rgba(c, c, c)], np.uint32)
    s["tex"] = np.array([[0, 1, 1, 1]], np.uint32)
    return s


@pytest.mark.parametrize("flags", [NEE | MIS | RR, NEE | RIS])
def test_white_texture_leaves_every_path_unchanged(orc, flags):
    # a white texel is exactly 1 and every bilinear weight set sums to 1
    # exactly: the render must be bit-identical to the untextured one
    sc, lm = scenes.cornell_materials(24, 24)
    kw = dict(max_depth=6, rr_depth=2, num_candidates=4, seed=3)
    a = _render(orc, sc, lm, 4, flags, **kw)
    b = _render(orc, _with_constant_textures(sc, 255), lm, 4, flags, **kw)
    assert np.array_equal(a, b)


def test_black_diffuse_texture_equals_black_material(orc):
    sc, lm = scenes.cornell_materials(24, 24)
    t = _with_constant_textures(sc, 0)
    t["tex_id"] = np.full_like(t["tex_id"], 0xFFFFFFFF)
    lamb = sc["mat"] == scenes.MAT_LAMBERT
    t["tex_id"][lamb, 0] = 0
    black = dict(sc)
    black["albedo"] = sc["albedo"].copy()
    black["albedo"][lamb, :3] = 0.0
    kw = dict(max_depth=6, rr_depth=2, seed=4)
    assert np.array_equal(_render(orc, t, lm, 4, NEE | MIS | RR, **kw), _render(orc, black, lm, 4, NEE | MIS | RR, **kw))


def test_textured_debug_frames_read_back_through_positions(orc):
    # planar texture coordinates are an affine function of the hit point, so
    # the diffuse / specular / emission frames must equal the material colour
    # times the texture fetched at the uv recomputed from the position frame
    sc, lm = scenes.textured_cornell(40, 32)
    cam = sc["camera"]
    uvw = scenes.camera_uvw(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], 40 / 32)
    S = orc.Scene(sc, light_mesh=lm)
    pos, tri = orc.debug_frame(S, uvw, 40, 32, 8, seed=2)
    q = sc["verts"][sc["idx"].astype(np.int64)].astype(np.float64)
    fn = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
    ax = np.argmax(np.abs(fn), axis=1)
    checked = {2: 0, 3: 0, 4: 0}
    for ft in (2, 3, 4):
        val, tri2 = orc.debug_frame(S, uvw, 40, 32, ft, seed=2)
        assert np.array_equal(tri, tri2)
        for y in range(32):
            for x in range(40):
                t = int(tri[y, x])
                if t < 0:
                    continue
                a, b = [(1, 2), (0, 2), (0, 1)][ax[t]]
                p = pos[y, x].astype(np.float64)
                u, v = p[a] * 0.8 + 0.1, p[b] * 0.8 + 0.1
                mt = int(sc["mat"][t])
                light = sc["emission"][t].sum() > 0
                if ft == 2 and mt != scenes.MAT_GLASS and not light:
                    base, slot = sc["albedo"][t, :3], 0
                elif ft == 3 and mt == scenes.MAT_PHONG:
                    base, slot = sc["spec"][t, :3], 1
                elif ft == 4 and light:
                    base, slot = sc["emission"][t], 2
                else:
                    continue
                tid = int(sc["tex_id"][t, slot])
                f = np.ones(3) if tid == 0xFFFFFFFF else orc.tex_sample(sc["texels"], sc["tex"][tid], u, v)
                assert np.allclose(val[y, x], base * f, rtol=2e-3, atol=2e-3), (ft, x, y, t)
                checked[ft] += 1
    assert min(checked.values()) > 10, checked


def test_glass_with_face_normals_as_vertex_normals_is_unchanged(orc):
    # vertex normals equal to the face normal interpolate back to it (up to
    # rounding, which paths trapped in the glass box amplify): the image is
    # closer to the face-normal render than an independent one
    sc, lm = scenes.cornell_materials(24, 24)
    q = sc["verts"][sc["idx"].astype(np.int64)].astype(np.float64)
    fn = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    s2 = dict(sc)
    s2["vnormal"] = np.repeat(fn[:, None, :], 3, axis=1).astype(np.float32)
    kw = dict(max_depth=6, rr_depth=2, seed=5)
    a = _render(orc, sc, lm, 16, NEE | MIS | RR, **kw) / 16
    b = _render(orc, s2, lm, 16, NEE | MIS | RR, **kw) / 16
    c = _render(orc, sc, lm, 16, NEE | MIS | RR, **dict(kw, seed=6)) / 16
    err = lambda x, y: float(np.mean((x - y) ** 2 / (y * y + 1e-2)))
    assert err(b, a) < 0.6 * err(c, a) and abs(b.mean() - a.mean()) < 0.02 * a.mean()
    # smooth normals on the glass box change its image
    s3 = dict(sc)
    t3, _ = scenes.textured_cornell(24, 24)
    s3["vnormal"] = t3["vnormal"]
    d = _render(orc, s3, lm, 16, NEE | MIS | RR, **kw) / 16
    assert err(d, a) > 0.2 * err(c, a)


def test_scene_loader_textures_normals_and_slots(tmp_path):
    from PIL import Image

    img = np.zeros((2, 3, 4), np.uint8)  # 2 rows x 3 columns, top row first
    img[0] = [[10, 20, 30, 255], [40, 50, 60, 255], [70, 80, 90, 255]]
    img[1] = [[110, 120, 130, 255], [140, 150, 160, 200], [170, 180, 190, 255]]
    Image.fromarray(img, "RGBA").save(tmp_path / "d.png")
    Image.fromarray(np.full((1, 1), 77, np.uint8), "L").save(tmp_path / "s.png")
    Image.fromarray(np.full((2, 2, 3), 9, np.uint8), "RGB").save(tmp_path / "e.png")
    (tmp_path / "m.mtl").write_text(
        "newmtl tex\nKd 1 1 1\nKs 0.5 0.5 0.5\nNs 10\nmap_Kd d.png\nmap_Ks s.png\nmap_Ke e.png\nillum 2\n"
        "newmtl glass\nKs 1 1 1\nNi 1.5\nillum 7\n")
    (tmp_path / "m.obj").write_text(
        "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0.25 0.5\nvt 0.75 0.5\nvn 0 0 2\n"
        "usemtl tex\nf 1/1 2/2 3\nusemtl glass\nf 1//1 3 4\n")
    tx = scenefile.load_texture(tmp_path / "d.png")
    assert tx.shape == (2, 3) and tx[0, 0] == _rgba(110, 120, 130) and tx[1, 2] == _rgba(70, 80, 90)
    assert tx[0, 1] == _rgba(140, 150, 160, 200)
    assert scenefile.load_texture(tmp_path / "s.png")[0, 0] == _rgba(77, 77, 77)
    sj = {"ObjModels": {"Assets": {"M": "m.obj"}},
          "World": {"Geometries": {"G": {"Type": "ObjModel", "Base": "M"}},
                    "GeometryASs": {"GAS": {"Type": "GeometryAS", "Geometries": ["G"]}},
                    "Instances": {"I": {"Type": "Instance", "ASType": "Geometry", "Base": "GAS",
                                        "Transform": [2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 4, 0]}},
                    "InstanceASs": {"Root": {"Type": "InstanceAS", "Instances": ["I"]}}},
          "Config": {"Width": 8, "Height": 8}}
    sc, _, _ = scenefile.load_scene(sj, base_dir=str(tmp_path))
    tex = sc["tex"]
    names = {int(r[1]) * 10 + int(r[2]): k for k, r in enumerate(tex)}  # by size: 3x2 d, 1x1 s, 2x2 e
    d, s, e = names[32], names[11], names[22]
    assert np.all(tex[:, 3] == 1)  # LoadFromPath textures wrap
    t0 = sc["tex_id"][0]
    assert t0.tolist() == [d, e, s, e]  # the reference's binding: specular <- emitTex, emission <- specTex
    assert sc["tex_id"][1].tolist() == [0xFFFFFFFF] * 4
    sc2, _, _ = scenefile.load_scene(sj, base_dir=str(tmp_path), texture_slots="named")
    n = {int(r[1]) * 10 + int(r[2]): k for k, r in enumerate(sc2["tex"])}
    assert sc2["tex_id"][0].tolist() == [n[32], n[11], n[22], n[22]]
    # texture coordinates: given corners, else (0.5, 0.5)
    assert np.allclose(sc["texcoord"][0], [[0.25, 0.5], [0.75, 0.5], [0.5, 0.5]])
    # normals in world space (inverse transpose of diag(2, 1, 4)); missing -> (0, 1, 0)
    vn = sc["vnormal"]
    assert np.allclose(vn[1, 0], [0, 0, 0.5]) and np.allclose(vn[1, 1], [0, 1, 0]) and np.allclose(vn[1, 2], [0, 1, 0])
    sc3, _, _ = scenefile.load_scene(sj, base_dir=str(tmp_path), missing_normal="face")
    assert np.allclose(sc3["vnormal"][1, 1] / np.linalg.norm(sc3["vnormal"][1, 1]), [0, 0, 1])
    texels = sc["texels"]
    r = tex[d]
    assert texels[r[0]:r[0] + 6].reshape(2, 3)[1, 0] == _rgba(10, 20, 30)
