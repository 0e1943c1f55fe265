"""Procedural scenes of the benchmark workloads (seeded, nothing read from disk).

Every scene is a dict of numpy arrays in the layout rtpg_scene_upload takes:
  verts (nv,3) f32, idx (nt,3) u32, albedo (nt,4) f32 (rgb + roughness/ior),
  mat (nt,) u32, emission (nt,3) f32, spec (nt,4) f32 or None,
plus a camera dict (eye, lookat, up, fovy, width, height).
"""
from __future__ import annotations

import math

import numpy as np

MAT_LAMBERT, MAT_GGX, MAT_PHONG, MAT_GLASS = 0, 1, 2, 3


def camera_uvw(eye, lookat, up, fovy_deg, aspect):
    """Camera::GetUVW (Core/Include/RTLib/Core/Camera.h:78-92)."""
    eye = np.asarray(eye, np.float32)
    w = np.asarray(lookat, np.float32) - eye
    u = np.cross(np.asarray(up, np.float32), w)
    u = u / np.linalg.norm(u)
    v = np.cross(w, u)
    v = v / np.linalg.norm(v)
    vlen = np.linalg.norm(w) * math.tan(math.pi * fovy_deg / 360.0)
    ulen = vlen * aspect
    return eye, (u * ulen).astype(np.float32), (v * vlen).astype(np.float32), w.astype(np.float32)


class _Builder:
    def __init__(self):
        self.v, self.i, self.alb, self.mat, self.em = [], [], [], [], []
        self.nv = 0

    def tris(self, verts, idx, albedo, mat=MAT_LAMBERT, emission=(0, 0, 0), param=0.0):
        verts = np.asarray(verts, np.float32).reshape(-1, 3)
        idx = np.asarray(idx, np.int64).reshape(-1, 3) + self.nv
        nt = idx.shape[0]
        self.v.append(verts)
        self.i.append(idx)
        a = np.zeros((nt, 4), np.float32)
        a[:, :3] = np.broadcast_to(np.asarray(albedo, np.float32), (nt, 3)) if np.ndim(albedo) == 1 else albedo
        a[:, 3] = param
        self.alb.append(a)
        self.mat.append(np.full(nt, mat, np.uint32) if np.ndim(mat) == 0 else np.asarray(mat, np.uint32))
        self.em.append(np.broadcast_to(np.asarray(emission, np.float32), (nt, 3)).copy())
        self.nv += verts.shape[0]

    def quad(self, p0, p1, p2, p3, albedo, **kw):
        """Quad p0 p1 p2 p3 (counter-clockwise seen from the side its normal faces)."""
        self.tris([p0, p1, p2, p3], [[0, 1, 2], [0, 2, 3]], albedo, **kw)

    def box(self, center, size, angle_deg, albedo):
        cx, cy, cz = center
        sx, sy, sz = size
        a = math.radians(angle_deg)
        ca, sa = math.cos(a), math.sin(a)
        corners = []
        for dy in (0.0, sy):
            for dx, dz in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
                x, z = dx * sx, dz * sz
                corners.append((cx + ca * x - sa * z, cy + dy, cz + sa * x + ca * z))
        c = np.asarray(corners, np.float32)
        faces = [[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]]
        idx = []
        for f in faces:
            idx += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
        self.tris(c, idx, albedo)

    def finish(self, camera, spec=None):
        s = {
            "verts": np.ascontiguousarray(np.concatenate(self.v), np.float32),
            "idx": np.ascontiguousarray(np.concatenate(self.i), np.uint32),
            "albedo": np.ascontiguousarray(np.concatenate(self.alb), np.float32),
            "mat": np.ascontiguousarray(np.concatenate(self.mat), np.uint32),
            "emission": np.ascontiguousarray(np.concatenate(self.em), np.float32),
            "spec": spec,
            "camera": camera,
        }
        return s


RED = (0.63, 0.065, 0.05)
GREEN = (0.14, 0.45, 0.09)
WHITE = (0.725, 0.71, 0.68)
CORNELL_LE = (17.0, 12.0, 4.0)


def cornell(width=256, height=256, fovy=40.0):
    """Cornell box: 5 wall quads + 2 rotated boxes (34 triangles) + a 2-triangle
    ceiling light with radiance (17, 12, 4).  Room x in [-1,1], y in [0,2],
    z in [-1,1], open towards +z where the camera is."""
    b = _Builder()
    b.quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), WHITE)      # floor, +y
    b.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), WHITE)      # ceiling, -y
    b.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), WHITE)    # back, +z
    b.quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), RED)      # left, +x
    b.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), GREEN)        # right, -x
    b.box((0.33, 0.0, 0.3), (0.6, 0.6, 0.6), -17.0, WHITE)
    b.box((-0.33, 0.0, -0.35), (0.6, 1.2, 0.6), 17.0, WHITE)
    b.quad((-0.25, 1.98, -0.2), (0.25, 1.98, -0.2), (0.25, 1.98, 0.2), (-0.25, 1.98, 0.2), (0, 0, 0),
           emission=CORNELL_LE)                                         # light, -y
    cam = {"eye": (0.0, 1.0, 3.8), "lookat": (0.0, 1.0, 0.0), "up": (0, 1, 0), "fovy": fovy,
           "width": width, "height": height}
    return b.finish(cam)


def cornell_materials(width=64, height=64, fovy=40.0):
    """Cornell box with the reference's other materials and light layout: the
    tall box normalised Phong (diffuse 0.3, specular 0.5, shininess 30), the
    short box glass (ior 1.5), and two light meshes (the ceiling quad and a
    smaller quad on the back wall, light_mesh ids 0 (4 triangles) and 1 (2
    triangles)) for the uniform mesh-then-triangle selection.  Returns
    (scene, light_mesh)."""
    b = _Builder()
    b.quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1), WHITE)
    b.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), WHITE)
    b.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), WHITE)
    b.quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), RED)
    b.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), GREEN)
    n_walls = sum(len(i) for i in b.i)
    b.box((0.33, 0.0, 0.3), (0.6, 0.6, 0.6), -17.0, WHITE)
    n_short = sum(len(i) for i in b.i)
    b.box((-0.33, 0.0, -0.35), (0.6, 1.2, 0.6), 17.0, (0.3, 0.3, 0.3))
    n_tall = sum(len(i) for i in b.i)
    # ceiling light as two quads (4 triangles, mesh 0); back-wall light 2 triangles (mesh 1)
    b.quad((-0.25, 1.98, -0.2), (0.0, 1.98, -0.2), (0.0, 1.98, 0.2), (-0.25, 1.98, 0.2), (0, 0, 0),
           emission=CORNELL_LE)
    b.quad((0.0, 1.98, -0.2), (0.25, 1.98, -0.2), (0.25, 1.98, 0.2), (0.0, 1.98, 0.2), (0, 0, 0),
           emission=CORNELL_LE)
    b.quad((0.5, 1.2, -0.99), (0.8, 1.2, -0.99), (0.8, 1.5, -0.99), (0.5, 1.5, -0.99), (0, 0, 0),
           emission=(4.0, 6.0, 9.0))
    cam = {"eye": (0.0, 1.0, 3.8), "lookat": (0.0, 1.0, 0.0), "up": (0, 1, 0), "fovy": fovy,
           "width": width, "height": height}
    nt = sum(len(i) for i in b.i)
    spec = np.zeros((nt, 4), np.float32)
    s = b.finish(cam, spec=spec)
    s["mat"][n_walls:n_short] = MAT_GLASS
    s["albedo"][n_walls:n_short] = (1.0, 1.0, 1.0, 1.5)
    s["mat"][n_short:n_tall] = MAT_PHONG
    spec[n_short:n_tall] = (0.5, 0.5, 0.5, 30.0)
    light_mesh = np.zeros(nt, np.uint32)
    light_mesh[nt - 2:] = 1
    return s, light_mesh


def _rgba(r, g, b, a=255):
    return np.uint32(r) | (np.uint32(g) << 8) | (np.uint32(b) << 16) | (np.uint32(a) << 24)


def textured_cornell(width=64, height=64, fovy=40.0):
    """cornell_materials with texture maps and vertex normals: planar texture
    coordinates per face (outside [0, 1] on purpose), an 8 x 8 checker
    (wrap) on the floor, ceiling and back wall and as the Phong box's
    specular map, a 16 x 4 colour ramp (clamp) on the side walls and as the
    Phong box's diffuse map, a 4 x 4 emission map (clamp) on both light
    meshes (on hit and when light-sampled), and on the glass box vertex
    normals pointing away from its centre.  Returns (scene, light_mesh); the
    scene holds the arrays of rtpg_scene_attributes."""
    s, lm = cornell_materials(width, height, fovy)
    nt = s["idx"].shape[0]
    q = s["verts"][s["idx"].astype(np.int64)].astype(np.float64)  # nt x 3 x 3
    fn = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    ax = np.argmax(np.abs(fn), axis=1)
    uv = np.empty((nt, 3, 2), np.float64)
    for t in range(nt):
        a, b = [(1, 2), (0, 2), (0, 1)][ax[t]]
        uv[t] = q[t][:, [a, b]] * 0.8 + 0.1
    checker = np.empty((8, 8), np.uint32)
    for j in range(8):
        for i in range(8):
            checker[j, i] = _rgba(230, 225, 215) if (i + j) % 2 == 0 else _rgba(60, 90, 200)
    ramp = np.empty((4, 16), np.uint32)
    for j in range(4):
        for i in range(16):
            ramp[j, i] = _rgba(17 * i, 255 - 17 * i, 60 * j + 30)
    glow = np.empty((4, 4), np.uint32)
    for j in range(4):
        for i in range(4):
            glow[j, i] = _rgba(255 - 40 * i, 200 + 15 * j, 128 + 30 * ((i + j) % 3))
    texels = np.concatenate([checker.ravel(), ramp.ravel(), glow.ravel()])
    tex = np.array([[0, 8, 8, 1], [64, 16, 4, 0], [128, 4, 4, 0]], np.uint32)
    none = 0xFFFFFFFF
    tid = np.full((nt, 4), none, np.uint32)
    mat, em = s["mat"], s["emission"]
    n_walls = 10
    side = (ax[:n_walls] == 0)
    tid[:n_walls, 0] = np.where(side, 1, 0)
    tid[mat == MAT_PHONG, 0] = 1
    tid[mat == MAT_PHONG, 1] = 0
    lights = em.sum(axis=1) > 0
    tid[lights, 2] = 2
    tid[lights, 3] = 2
    vn = np.repeat(fn[:, None, :], 3, axis=1)
    glass = mat == MAT_GLASS
    centre = q[glass].reshape(-1, 3).mean(axis=0)
    d = q[glass] - centre
    vn[glass] = d / np.linalg.norm(d, axis=2, keepdims=True)
    s["texcoord"] = uv.astype(np.float32)
    s["tex_id"] = tid
    s["vnormal"] = vn.astype(np.float32)
    s["texels"] = texels
    s["tex"] = tex
    return s, lm


def _value_noise(nx, nz, octaves, seed):
    """fBm of bilinearly interpolated value noise on an nx x nz grid in [0,1]^2."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, nx, dtype=np.float64)
    zs = np.linspace(0.0, 1.0, nz, dtype=np.float64)
    h = np.zeros((nz, nx), np.float64)
    amp, freq = 1.0, 4
    for _ in range(octaves):
        lat = rng.random((freq + 2, 2 * freq + 2))
        fx = xs * 2 * freq
        fz = zs * freq
        ix = np.floor(fx).astype(np.int64)
        iz = np.floor(fz).astype(np.int64)
        tx = (fx - ix)[None, :]
        tz = (fz - iz)[:, None]
        tx = tx * tx * (3 - 2 * tx)
        tz = tz * tz * (3 - 2 * tz)
        a = lat[iz[:, None], ix[None, :]]
        bb = lat[iz[:, None], ix[None, :] + 1]
        c = lat[iz[:, None] + 1, ix[None, :]]
        d = lat[iz[:, None] + 1, ix[None, :] + 1]
        h += amp * ((a * (1 - tx) + bb * tx) * (1 - tz) + (c * (1 - tx) + d * tx) * tz)
        amp *= 0.5
        freq *= 2
    return h


def guided_room(width=1280, height=720, grid=(1024, 512), seed_geom=7, seed_mat=1, fovy=60.0):
    """Room with a half-open dividing wall; a small ceiling light (0.5% of the
    ceiling area) hidden behind the wall; the floor is a grid heightfield of
    8-octave fBm (about 1M triangles at 1024 x 512)."""
    X, Y, Z = 8.0, 3.0, 4.0
    b = _Builder()
    nx, nz = grid
    h = _value_noise(nx, nz, 8, seed_geom)
    h = 0.15 * (h - h.min()) / max(h.max() - h.min(), 1e-9)
    gx = np.linspace(0.0, X, nx, dtype=np.float32)
    gz = np.linspace(0.0, Z, nz, dtype=np.float32)
    vx, vz = np.meshgrid(gx, gz)
    verts = np.stack([vx, h.astype(np.float32), vz], -1).reshape(-1, 3)
    ii = np.arange(nz - 1)[:, None] * nx + np.arange(nx - 1)[None, :]
    a, bq, c, d = ii, ii + 1, ii + nx + 1, ii + nx
    # normals +y: (a, d, c) and (a, c, b) with x right, z forward
    t1 = np.stack([a, d, c], -1).reshape(-1, 3)
    t2 = np.stack([a, c, bq], -1).reshape(-1, 3)
    fidx = np.empty((t1.shape[0] * 2, 3), np.int64)
    fidx[0::2] = t1
    fidx[1::2] = t2
    b.tris(verts, fidx, WHITE)
    # walls and ceiling (normals into the room)
    b.quad((0, Y, 0), (X, Y, 0), (X, Y, Z), (0, Y, Z), WHITE)        # ceiling
    b.quad((0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0), WHITE)        # z = 0 wall
    b.quad((0, 0, Z), (0, Y, Z), (X, Y, Z), (X, 0, Z), WHITE)        # z = Z wall
    b.quad((0, 0, 0), (0, Y, 0), (0, Y, Z), (0, 0, Z), WHITE)        # x = 0 wall
    b.quad((X, 0, 0), (X, 0, Z), (X, Y, Z), (X, Y, 0), WHITE)        # x = X wall
    # dividing wall at x = 5: solid for z in [0, 2.6], open doorway above z = 2.6
    # (both faces, 0.1 thick)
    W0, W1, ZD = 5.0, 5.1, 2.6
    b.quad((W0, 0, 0), (W0, 0, ZD), (W0, Y, ZD), (W0, Y, 0), WHITE)
    b.quad((W1, 0, 0), (W1, Y, 0), (W1, Y, ZD), (W1, 0, ZD), WHITE)
    b.quad((W0, 0, ZD), (W1, 0, ZD), (W1, Y, ZD), (W0, Y, ZD), WHITE)
    # lintel over the doorway (upper third)
    b.quad((W0, 2.2, ZD), (W0, 2.2, Z), (W0, Y, Z), (W0, Y, ZD), WHITE)
    b.quad((W1, 2.2, ZD), (W1, Y, ZD), (W1, Y, Z), (W1, 2.2, Z), WHITE)
    b.quad((W0, 2.2, ZD), (W1, 2.2, ZD), (W1, 2.2, Z), (W0, 2.2, Z), WHITE)
    # light: 0.5% of the ceiling (0.16 m^2), in the far corner behind the wall
    s = math.sqrt(0.005 * X * Z) / 2.0
    lx, lz = 7.3, 0.8
    yl = Y - 0.01
    b.quad((lx - s, yl, lz - s), (lx + s, yl, lz - s), (lx + s, yl, lz + s), (lx - s, yl, lz + s), (0, 0, 0),
           emission=(900.0, 800.0, 650.0))
    sc = b.finish({"eye": (0.6, 1.5, 3.5), "lookat": (5.0, 1.0, 1.8), "up": (0, 1, 0), "fovy": fovy,
                   "width": width, "height": height})
    # Materials: albedo U[0.2, 0.8] per triangle; 20% GGX with roughness U[0.05, 0.5].
    rng = np.random.default_rng(seed_mat)
    nt = sc["idx"].shape[0]
    light = sc["emission"].sum(1) > 0
    alb = rng.uniform(0.2, 0.8, size=(nt, 3)).astype(np.float32)
    ggx = rng.random(nt) < 0.2
    rough = rng.uniform(0.05, 0.5, size=nt).astype(np.float32)
    sc["albedo"][:, :3] = np.where(light[:, None], 0.0, alb)
    sc["mat"] = np.where(ggx & ~light, MAT_GGX, MAT_LAMBERT).astype(np.uint32)
    sc["albedo"][:, 3] = np.where(ggx, rough, 0.0)
    return sc


def _icosphere(subdiv):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b_, c in faces:
            ab, bc, ca = mid(a, b_), mid(b_, c), mid(c, a)
            nf += [(a, ab, ca), (b_, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.asarray(verts, np.float32), np.asarray(faces, np.int64)


def many_lights(width=1920, height=1080, n_spheres=2000, subdiv=3, n_emissive=4096, seed=3, fovy=60.0):
    """2000 icospheres of 1280 triangles (2.56M), centres U[0,100]^3, radii
    U[0.5,3]; 4096 triangles emissive with radiance log-uniform in [0.1, 100]."""
    rng = np.random.default_rng(seed)
    sv, sf = _icosphere(subdiv)
    centres = rng.uniform(0.0, 100.0, size=(n_spheres, 3)).astype(np.float32)
    radii = rng.uniform(0.5, 3.0, size=n_spheres).astype(np.float32)
    verts = (sv[None, :, :] * radii[:, None, None] + centres[:, None, :]).reshape(-1, 3)
    idx = (sf[None, :, :] + (np.arange(n_spheres) * sv.shape[0])[:, None, None]).reshape(-1, 3)
    b = _Builder()
    nt = idx.shape[0]
    b.tris(verts, idx, WHITE)
    sc = b.finish({"eye": (50.0, 50.0, -40.0), "lookat": (50.0, 50.0, 50.0), "up": (0, 1, 0), "fovy": fovy,
                   "width": width, "height": height})
    sc["albedo"][:, :3] = rng.uniform(0.2, 0.8, size=(nt, 3)).astype(np.float32)
    em_idx = rng.choice(nt, size=n_emissive, replace=False)
    le = np.exp(rng.uniform(math.log(0.1), math.log(100.0), size=n_emissive)).astype(np.float32)
    sc["emission"][em_idx] = le[:, None]
    sc["albedo"][em_idx, :3] = 0.0
    return sc


def furnace_sphere(albedo=0.5, mat=MAT_LAMBERT, roughness=0.0, subdiv=3, width=32, height=32):
    """A unit icosphere under a uniform white background, for furnace tests."""
    sv, sf = _icosphere(subdiv)
    b = _Builder()
    b.tris(sv, sf, (albedo, albedo, albedo), mat=mat, param=roughness)
    return b.finish({"eye": (0.0, 0.0, 4.0), "lookat": (0.0, 0.0, 0.0), "up": (0, 1, 0), "fovy": 45.0,
                     "width": width, "height": height})
