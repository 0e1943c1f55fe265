"""Pins the learned guide's CPU oracle (oracle/mlp.py): its analytic gradient
against central finite differences of its own FP64 loss, the vMF density and
sampler against closed forms, and the hash-grid encoding against brute-force
properties.  A dropped term, a wrong sign or a transposed weight in the
backward pass fails the finite-difference check."""
import numpy as np
import pytest

from oracle import mlp as om


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _setup(seed=0, n=40):
    rng = np.random.default_rng(seed)
    enc = om.Encoding([0, 0, 0], [2.0, 1.0, 1.5], log2_table=8, base_res=2.0, max_res=16.0)
    w = np.zeros(om.N_WEIGHTS)
    for lo, hi, fan in [(om.W0, om.W1, 40), (om.W1, om.W2, 64), (om.W2, om.W3, 64), (om.W3, om.N_WEIGHTS, 64)]:
        w[lo:hi] = rng.uniform(-1, 1, hi - lo) * np.sqrt(6.0 / fan)
    table = rng.uniform(-0.5, 0.5, (om.LEVELS, enc.T, 2))
    s = dict(pos=rng.uniform([0, 0, 0], [2.0, 1.0, 1.5], (n, 3)), wo=_unit(rng, n), nrm=_unit(rng, n),
             rough=rng.uniform(0, 0.5, n), omega=_unit(rng, n), weight=rng.uniform(0.1, 2.0, n),
             pdf_bsdf=rng.uniform(0.05, 0.3, n))
    return enc, w, table, s, rng


def test_encoding_properties():
    enc = om.Encoding([0, 0, 0], [1, 1, 1], log2_table=8, base_res=2.0, max_res=16.0)
    assert enc.dense[0] and not enc.dense[-1]
    # dense levels index every grid point exactly once
    for l in np.nonzero(enc.dense)[0]:
        r1 = int(enc.res[l]) + 1
        g = np.stack(np.meshgrid(*[np.arange(r1)] * 3, indexing="ij"), -1).reshape(-1, 3)
        idx = enc.index(l, g[:, 0], g[:, 1], g[:, 2])
        assert len(np.unique(idx)) == r1 ** 3 and idx.max() < enc.T
    # partition of unity: a constant table encodes to the constant
    pos = np.random.default_rng(1).uniform(0, 1, (100, 3))
    cs = enc.corners(pos)
    for idx, wt in cs:
        assert np.allclose(wt.sum(1), 1.0)
    table = np.ones((om.LEVELS, enc.T, 2)) * np.array([0.25, -2.0])
    f, _ = enc.encode(table, pos, cs)
    assert np.allclose(f[:, 0::2], 0.25) and np.allclose(f[:, 1::2], -2.0)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 30.0, 1000.0])
def test_vmf_density_integrates_to_one(kappa):
    # integrate over the cosine to the mean axis (the density is axially symmetric)
    t = np.linspace(-1.0, 1.0, 2_000_001)
    mu = np.array([0.0, 0.0, 1.0])
    w = np.stack([np.sqrt(1 - t * t), np.zeros_like(t), t], 1)
    f = om.vmf_pdf(mu, kappa, w) * 2.0 * np.pi
    assert abs(np.trapezoid(f, t) - 1.0) < 1e-6


def test_vmf_sampler_mean_cosine():
    rng = np.random.default_rng(3)
    for kappa in (0.3, 5.0, 200.0):
        mu = _unit(rng, 1)[0]
        n = 200_000
        d = om.vmf_sample(np.broadcast_to(mu, (n, 3)), kappa, rng.uniform(size=n), rng.uniform(size=n))
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
        c = d @ mu
        want = 1.0 / np.tanh(kappa) - 1.0 / kappa  # E[mu . w] of vMF on S^2
        assert abs(c.mean() - want) < 4 * c.std() / np.sqrt(n) + 1e-9


def test_mixture_pdf_single_lobe():
    rng = np.random.default_rng(4)
    mu = _unit(rng, 8)[None]
    kappa = rng.uniform(0.1, 50, (1, 8))
    pi = np.eye(8)[3][None]
    w = _unit(rng, 1)
    assert np.isclose(om.mix_pdf(pi, mu, kappa, w)[0], om.vmf_pdf(mu[0, 3], kappa[0, 3], w[0]))


def test_gradient_matches_central_differences():
    enc, w, table, s, rng = _setup()
    net = om.Net(enc, w, table)
    loss, gw, gt = net.loss_grad(s)
    h = 1e-6
    bad = []
    picks = np.concatenate([rng.integers(a, b, 6) for a, b in
                            [(om.W0, om.W0 + 64 * 40), (om.W1, om.W2), (om.W2, om.W3), (om.W3, om.W3 + 41 * 64)]])
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (om.Net(enc, wp, table).loss(s) - om.Net(enc, wm, table).loss(s)) / (2 * h)
        if abs(fd - gw[k]) > 1e-5 * abs(gw[k]) + 1e-9:
            bad.append(("w", int(k), fd, gw[k]))
    touched = np.argwhere(np.abs(gt) > 0)
    for l, i, f in touched[rng.choice(len(touched), 24, replace=False)]:
        tp, tm = table.copy(), table.copy()
        tp[l, i, f] += h
        tm[l, i, f] -= h
        fd = (om.Net(enc, w, tp).loss(s) - om.Net(enc, w, tm).loss(s)) / (2 * h)
        if abs(fd - gt[l, i, f]) > 1e-5 * abs(gt[l, i, f]) + 1e-9:
            bad.append(("t", (int(l), int(i), int(f)), fd, gt[l, i, f]))
    assert not bad, bad
    assert np.isfinite(loss) and np.abs(gw).max() > 0


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.14159, 0.0], np.float32)
    r = om.bf16(x)
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[2] == 1.0 and r[1] == 1.0 and r[3] == np.float32(1.0078125)
    assert r[4] == np.float32(-3.140625) and r[5] == 0.0
