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
    sc = _setup(ctx)
    rng = np.random.default_rng(1)
    target = rng.uniform(0, 0.3, (H, W, 3)).astype(np.float32)
    alb = sc["albedo"][:, :3].copy()

    def render(a):
        ctx.set_albedo(a)
        ctx.clear()
        ctx.render(spp=8, flags=FLAGS, max_depth=6, seed=7)
        return ctx.image().astype(np.float64)

    img = render(alb)
    dldi = (2.0 * (img - target)).astype(np.float32)
    ga, _ = ctx.adjoint(dldi, spp=8, flags=FLAGS, max_depth=6, seed=7)

    def loss(a):
        return float(((render(a) - target) ** 2).sum())

    nt = alb.shape[0]
    lambert = [t for t in range(nt) if sc["mat"][t] == 0 and sc["emission"][t].sum() == 0]
    picks = [(int(t), int(c)) for t, c in zip(rng.choice(lambert, 64), rng.integers(0, 3, 64))]
    h = 1e-3
    scale = np.abs(ga).max()
    bad = []
    for t, c in picks:
        ap, am = alb.copy(), alb.copy()
        ap[t, c] += h
        am[t, c] -= h
        fd = (loss(ap) - loss(am)) / (2 * h)
        if abs(fd - ga[t, c]) > 1e-2 * abs(ga[t, c]) + 1e-3 * scale:
            bad.append((t, c, fd, float(ga[t, c])))
    ctx.set_albedo(alb)
    assert not bad, bad


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
