"""Learned guide on the GPU against the CPU oracle (oracle/mlp.py): forward
pass and gradient of the bf16 matrix-core kernels against the oracle's
bf16-emulating float64 restatement, the vMF sampler against its own density
and closed-form moments, and training against a known target distribution."""
import numpy as np
import pytest

from oracle import mlp as om
from rtlib_amd.tracer import MLPGuide

pytestmark = pytest.mark.gpu

LO, HI = np.array([-1.0, -0.5, -1.0]), np.array([1.0, 1.5, 0.5])


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _guide(ctx, seed=0, **kw):
    cfg = dict(log2_table=12, base_res=4.0, max_res=64.0)
    cfg.update(kw)
    g = MLPGuide(ctx, LO, HI, seed=seed, **cfg)
    enc = om.Encoding(LO, HI, cfg["log2_table"], cfg["base_res"], cfg["max_res"])
    return g, enc


def _randomise(g, enc, rng):
    # weights at their initial scale, table entries large enough to matter
    w, t = g.export()
    t = rng.uniform(-0.5, 0.5, t.shape).astype(np.float32)
    g.import_(w, t)
    return w, t


def _samples(rng, n):
    return dict(pos=rng.uniform(LO, HI, (n, 3)).astype(np.float32), wo=_unit(rng, n), nrm=_unit(rng, n),
                rough=rng.uniform(0, 0.5, n).astype(np.float32), omega=_unit(rng, n),
                weight=rng.uniform(0.1, 2.0, n).astype(np.float32),
                pdf_bsdf=rng.uniform(0.05, 0.3, n).astype(np.float32))


def test_levels_match_oracle(ctx):
    g, enc = _guide(ctx)
    info = g.info()
    assert np.array_equal(info["res"], enc.res.astype(np.float32))
    assert info["dense_mask"] == sum(1 << l for l in range(16) if enc.dense[l])


def test_forward_matches_oracle(ctx):
    rng = np.random.default_rng(0)
    g, enc = _guide(ctx)
    w, t = _randomise(g, enc, rng)
    s = _samples(rng, 300)
    raw, mix = g.forward(s["pos"], s["wo"], s["nrm"], s["rough"])
    ref = om.Net(enc, w, t, bf16_emulation=True).forward(s["pos"], s["wo"], s["nrm"], s["rough"])["raw"][:, :41]
    scale = np.abs(ref).max()
    # identical bf16 operands; only fp32 summation order differs
    assert np.abs(raw - ref).max() <= 2e-3 * scale, np.abs(raw - ref).max() / scale
    hd = om.head(ref)
    pi = mix[:, 0:40:5]
    mu = np.stack([mix[:, 1:40:5], mix[:, 2:40:5], mix[:, 3:40:5]], -1)
    kappa = mix[:, 4:40:5]
    assert np.allclose(pi, hd["pi"], atol=2e-3)
    assert np.allclose(kappa, hd["kappa"], rtol=1e-2)
    assert np.allclose(mix[:, 40], hd["c"], atol=1e-3)
    assert np.mean(np.sum(mu * hd["mu"], -1)) > 0.999


def test_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    g, enc = _guide(ctx)
    w, t = _randomise(g, enc, rng)
    s = _samples(rng, 1000)
    loss, gw, gt = g.gradient(s)
    ol, ogw, ogt = om.Net(enc, w, t, bf16_emulation=True).loss_grad(s)
    assert abs(loss - ol) <= 1e-3 * abs(ol)
    ew = np.linalg.norm(gw - ogw) / np.linalg.norm(ogw)
    et = np.linalg.norm(gt - ogt.ravel()) / np.linalg.norm(ogt)
    assert ew < 2e-2 and et < 2e-2, (ew, et)
    # every layer carries gradient
    for a, b in [(om.W0, om.W1), (om.W1, om.W2), (om.W2, om.W3), (om.W3, om.N_WEIGHTS)]:
        assert np.abs(gw[a:b]).max() > 0


def test_sampler_density_and_moments(ctx):
    rng = np.random.default_rng(2)
    g, enc = _guide(ctx)
    _randomise(g, enc, rng)
    s = _samples(rng, 1)
    _, mix = g.forward(s["pos"], s["wo"], s["nrm"], s["rough"])
    n = 1 << 16
    out = g.sample(np.repeat(mix, n, 0), seed=7)
    d, q = out[:, :3].astype(np.float64), out[:, 3].astype(np.float64)
    pi = mix[0, 0:40:5].astype(np.float64)
    mu = np.stack([mix[0, 1:40:5], mix[0, 2:40:5], mix[0, 3:40:5]], -1).astype(np.float64)
    kappa = mix[0, 4:40:5].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    qo = om.mix_pdf(pi[None], mu[None], kappa[None], d)
    assert np.allclose(q, qo, rtol=1e-3)
    # E[g / q] = integral of g = 1 for a normalised broad g
    gref = om.vmf_pdf(np.array([0.0, 0.0, 1.0]), 1.0, d)
    r = gref / q
    assert abs(r.mean() - 1.0) < 4 * r.std() / np.sqrt(n)
    # mean direction of the mixture: sum_j pi_j (coth k_j - 1/k_j) mu_j
    want = np.sum((pi * (1.0 / np.tanh(kappa) - 1.0 / kappa))[:, None] * mu, 0)
    assert np.all(np.abs(d.mean(0) - want) < 4 * d.std(0) / np.sqrt(n))


def test_training_fits_target(ctx):
    # Target: at each position the radiance-weighted distribution is a vMF
    # (kappa 20) towards a point light L; samples drawn uniformly on the
    # sphere, so weight = target density / (1 / 4 pi), BSDF density uniform.
    rng = np.random.default_rng(3)
    g, enc = _guide(ctx, lr=1e-2, batch=1 << 14)
    L = np.array([0.3, 1.2, -0.2])
    n = 1 << 14

    def batch():
        p = rng.uniform(LO, HI, (n, 3))
        mu = L - p
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        om_ = _unit(rng, n).astype(np.float64)
        wt = om.vmf_pdf(mu, 20.0, om_) * 4 * np.pi
        nrm = _unit(rng, n)
        return dict(pos=p.astype(np.float32), wo=nrm, nrm=nrm, rough=np.zeros(n, np.float32),
                    omega=om_.astype(np.float32), weight=wt.astype(np.float32),
                    pdf_bsdf=np.full(n, 1 / (4 * np.pi), np.float32)), mu

    losses = []
    for it in range(300):
        s, _ = batch()
        losses.append(g.train(s, 1)[0])
    assert np.mean(losses[-20:]) < np.mean(losses[:5]) - 1.0
    s, mu = batch()
    _, mix = g.forward(s["pos"][:2000], s["wo"][:2000], s["nrm"][:2000], s["rough"][:2000])
    pi = mix[:, 0:40:5]
    m = np.stack([mix[:, 1:40:5], mix[:, 2:40:5], mix[:, 3:40:5]], -1)
    kappa = mix[:, 4:40:5]
    # density of the learned mixture at the true mode vs. the target's
    q_mode = om.mix_pdf(pi, m, kappa, mu[:2000])
    assert np.median(q_mode) > 0.5 * om.vmf_norm(20.0)
    # the guide is preferred over the (uniform) BSDF
    assert np.median(mix[:, 40]) > 0.8
    assert g.info()["steps"] == 300


def test_spatial_cache_is_the_network_at_cell_centres(ctx):
    # cache entry (cell, face) = the network at the cell centre with normal =
    # face direction, wo = normal, roughness 0
    rng = np.random.default_rng(9)
    g, enc = _guide(ctx, seed=2, cache_res=4)
    _randomise(g, enc, rng)
    mix, c = g.cache()
    res = 4
    assert mix.shape == (res ** 3 * 6, 48) and c.shape == (res ** 3 * 6,)
    lo32, hi32 = LO.astype(np.float32), HI.astype(np.float32)
    ext = np.float32(np.max(hi32 - lo32))
    lo = np.float32(0.5) * (lo32 + hi32) - np.float32(0.5) * ext
    step = (np.float32(1) / (np.float32(1) / ext)) / np.float32(res)
    e = np.arange(res ** 3 * 6)
    cell, face = e // 6, e % 6
    xyz = np.stack([cell % res, (cell // res) % res, cell // (res * res)], -1).astype(np.float32)
    pos = lo + (xyz + np.float32(0.5)) * step
    nrm = np.zeros((e.size, 3), np.float32)
    nrm[np.arange(e.size), face // 2] = np.where(face % 2 == 0, 1.0, -1.0)
    _, fwd = g.forward(pos, nrm, nrm, np.zeros(e.size, np.float32))
    pi, mu, kappa = fwd[:, 0:40:5], np.stack([fwd[:, 1:40:5], fwd[:, 2:40:5], fwd[:, 3:40:5]], -1), fwd[:, 4:40:5]
    assert np.allclose(mix[:, 32:40], pi, rtol=1e-3, atol=1e-5)
    assert np.allclose(mix[:, :32].reshape(-1, 8, 4)[..., :3], mu, atol=1e-3)
    assert np.allclose(mix[:, :32].reshape(-1, 8, 4)[..., 3], kappa, rtol=1e-3)
    assert np.allclose(c, fwd[:, 40], rtol=1e-3)
    # A_j = pi_j kappa_j / (2 pi (1 - exp(-2 kappa_j)))
    a = pi * kappa / (2 * np.pi * -np.expm1(-2 * kappa))
    assert np.allclose(mix[:, 40:48], a, rtol=2e-3)
    g.close()
