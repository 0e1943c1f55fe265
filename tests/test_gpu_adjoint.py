"""Path-replay adjoint on the GPU: against the FP64 oracle gradient on the
same samples, against central finite differences (h = 1e-3) of GPU renders
with the same random streams, and through the torch.autograd wrapper."""
import numpy as np
import pytest

from rtlib_amd import scenes
from rtlib_amd.tracer import FLAG_MIS, FLAG_NEE

pytestmark = pytest.mark.gpu

W = H = 32
FLAGS = FLAG_NEE | FLAG_MIS


def _setup(ctx):
    sc = scenes.cornell(W, H)
    ctx.upload_scene(sc)
    return sc


def test_adjoint_matches_oracle_fp64(ctx, orc):
    sc = _setup(ctx)
    rng = np.random.default_rng(0)
    dldi = rng.uniform(-1, 1, (H, W, 3)).astype(np.float32)
    ga, gl = ctx.adjoint(dldi, spp=8, flags=FLAGS, max_depth=6, seed=4)
    cam = sc["camera"]
    uvw = scenes.camera_uvw(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], 1.0)
    oa, ol = orc.adjoint(orc.Scene(sc), uvw, W, H, 8, FLAGS, dldi, 2, max_depth=6, seed=4)
    scale = np.abs(oa).max()
    # identical paths up to float rounding of a few diverging samples
    assert np.max(np.abs(ga - oa)) <= 2e-3 * scale
    assert np.allclose(gl, ol, rtol=2e-3, atol=2e-3 * np.abs(ol).max())


def test_adjoint_matches_central_differences(ctx):
    # L2 loss against the target of the adjoint workload: the image rendered
    # with albedo + N(0, 0.1), clipped to [0, 1].  Central differences (h =
    # 1e-3) of the loss of GPU renders with the same random streams, on 64
    # random (triangle, channel) parameters, each within 1e-2 relative.  At
    # this h the float32 image's rounding limits the difference quotient of the
    # smallest gradients to a few 1e-3 (the FP64 oracle gradient shows the
    # same spread against its own differences; h = 1e-2 agrees to 2e-3).
    sc = _setup(ctx)
    rng = np.random.default_rng(1)
    alb = sc["albedo"][:, :3].copy()

    def render(a, spp=8, seed=7):
        ctx.set_albedo(a)
        ctx.clear()
        ctx.render(spp=spp, flags=FLAGS, max_depth=6, seed=seed)
        return ctx.image().astype(np.float64)

    target = render(np.clip(alb + rng.normal(0.0, 0.1, alb.shape), 0.0, 1.0).astype(np.float32), spp=64, seed=99)
    img = render(alb)
    dldi = (2.0 * (img - target)).astype(np.float32)
    ga, _ = ctx.adjoint(dldi, spp=8, flags=FLAGS, max_depth=6, seed=7)

    def loss(a):
        return float(((render(a) - target) ** 2).sum())

    nt = alb.shape[0]
    lambert = [t for t in range(nt) if sc["mat"][t] == 0 and sc["emission"][t].sum() == 0]
    picks = [(int(t), int(c)) for t, c in zip(rng.choice(lambert, 64), rng.integers(0, 3, 64))]
    h = 1e-3
    rel = []
    for t, c in picks:
        ap, am = alb.copy(), alb.copy()
        ap[t, c] += h
        am[t, c] -= h
        fd = (loss(ap) - loss(am)) / (2 * h)
        # a channel no path scatters from has fd = grad = 0 exactly
        err = 0.0 if fd == 0.0 and ga[t, c] == 0.0 else abs(fd - ga[t, c]) / abs(fd)
        rel.append((float(err), t, c, fd, float(ga[t, c])))
    ctx.set_albedo(alb)
    rel.sort(reverse=True)
    print("adjoint vs central differences, worst relative errors:", rel[:4], flush=True)
    assert rel[0][0] < 1e-2, rel[:4]


def test_torch_autograd_wrapper(ctx):
    torch = pytest.importorskip("torch")
    from rtlib_amd.adjoint import render_differentiable

    sc = _setup(ctx)
    alb = torch.tensor(sc["albedo"][:, :3], dtype=torch.float32, requires_grad=True)
    le = torch.tensor([17.0, 12.0, 4.0], requires_grad=True)
    target = torch.zeros(H, W, 3)
    img = render_differentiable(ctx, alb, le, spp=4, seed=2)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    # d loss / d Le . Le = 2 * loss for a loss quadratic in a linear function of Le
    assert abs(float((le.grad * le).sum()) - 2 * float(loss)) <= 2e-3 * float(loss)
    assert alb.grad.shape == alb.shape and torch.isfinite(alb.grad).all() and alb.grad.abs().sum() > 0


def test_adjoint_with_textures_matches_oracle(ctx, orc):
    # diffuse maps on the Lambertian walls and an emission map on the light:
    # the gradient is still with respect to the per-triangle albedo and the
    # light's radiance (the texture is a fixed factor)
    sc = scenes.cornell(W, H)
    nt = sc["idx"].shape[0]
    q = sc["verts"][sc["idx"].astype(np.int64)].astype(np.float64)
    fn = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
    ax = np.argmax(np.abs(fn), axis=1)
    uv = np.stack([q[t][:, [(1, 2), (0, 2), (0, 1)][ax[t]]] for t in range(nt)]) * 0.7
    rng = np.random.default_rng(3)
    sc["texcoord"] = uv.astype(np.float32)
    sc["tex_id"] = np.full((nt, 4), 0xFFFFFFFF, np.uint32)
    light = sc["emission"].sum(axis=1) > 0
    sc["tex_id"][~light, 0] = 0
    sc["tex_id"][light, 2] = 1
    sc["tex_id"][light, 3] = 1
    sc["texels"] = rng.integers(0, 2**32, 64 + 4, dtype=np.uint64).astype(np.uint32)
    sc["tex"] = np.array([[0, 8, 8, 1], [64, 2, 2, 0]], np.uint32)
    ctx.upload_scene(sc)
    dldi = rng.uniform(-1, 1, (H, W, 3)).astype(np.float32)
    ga, gl = ctx.adjoint(dldi, spp=8, flags=FLAGS, max_depth=6, seed=4)
    cam = sc["camera"]
    uvw = scenes.camera_uvw(cam["eye"], cam["lookat"], cam["up"], cam["fovy"], 1.0)
    oa, ol = orc.adjoint(orc.Scene(sc), uvw, W, H, 8, FLAGS, dldi, 2, max_depth=6, seed=4)
    scale = np.abs(oa).max()
    assert np.max(np.abs(ga - oa)) <= 2e-3 * scale
    assert np.allclose(gl, ol, rtol=2e-3, atol=2e-3 * np.abs(ol).max())
