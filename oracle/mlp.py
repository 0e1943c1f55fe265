"""CPU oracle of the learned directional guide (TEST INFRASTRUCTURE ONLY).

A float64 numpy restatement of the network the library evaluates on the
matrix cores: multiresolution hash-grid encoding (16 levels x 2 features,
instant-NGP spatial hash, dense indexing where (res+1)^3 fits the table),
inputs [32 features, wo, n, roughness, 1, 0 x 8], three 64-wide ReLU layers
without biases, 41 outputs -> 8 vMF lobes (softmax weight, normalised mean,
kappa = exp(clamp(s, -3, 8))) and c = 0.05 + 0.9 sigmoid(t).  Loss per
sample: -w log(c q(omega) + (1 - c) p_bsdf).

The reference (RTLib) has no learned guide; this oracle is pinned by central
finite differences of its own loss (tests/test_oracle_mlp.py), by closed
forms of the vMF density and by brute-force integration, not by reference
vectors.  With bf16=True it rounds weights, table entries, inputs,
activations and back-propagated output gradients to bf16 (round to nearest
even) where the GPU kernels do, for direct comparison with them.

Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

LEVELS = 16
HIDDEN = 64
LOBES = 8
N_IN = 48
N_OUT = 48
N_USED = 41
W0, W1, W2, W3 = 0, 64 * 48, 64 * 48 + 64 * 64, 64 * 48 + 2 * 64 * 64
N_WEIGHTS = W3 + 48 * 64
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -3.0, 8.0
SEL_MIN, SEL_SPAN = 0.05, 0.9
P1, P2 = np.uint32(2654435761), np.uint32(805459861)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def split_weights(w):
    w = np.asarray(w, np.float64)
    return (w[W0:W1].reshape(64, 48), w[W1:W2].reshape(64, 64), w[W2:W3].reshape(64, 64),
            w[W3:].reshape(48, 64))


class Encoding:
    """Level resolutions and the cubic box of the position encoding."""

    def __init__(self, lo, hi, log2_table=19, base_res=16.0, max_res=256.0):
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        ext = np.float32(max(float(np.max(hi - lo)), 1e-6))
        self.lo = (np.float32(0.5) * (lo + hi) - np.float32(0.5) * ext).astype(np.float32)
        self.inv_size = np.float32(1.0) / ext
        self.T = 1 << log2_table
        b = np.exp((np.log(max_res) - np.log(base_res)) / (LEVELS - 1))
        self.res = np.array([np.floor(base_res * b ** l + 1e-9) for l in range(LEVELS)], np.float64)
        self.dense = np.array([(r + 1) ** 3 <= self.T for r in self.res])

    def index(self, l, x, y, z):
        if self.dense[l]:
            r1 = np.uint64(int(self.res[l]) + 1)
            return (x.astype(np.uint64) + r1 * (y.astype(np.uint64) + r1 * z.astype(np.uint64))).astype(np.int64)
        h = x.astype(np.uint32) ^ (y.astype(np.uint32) * P1) ^ (z.astype(np.uint32) * P2)
        return (h & np.uint32(self.T - 1)).astype(np.int64)

    def corners(self, pos):
        """Per level: (indices n x 8, trilinear weights n x 8)."""
        p = np.asarray(pos, np.float32)
        q = np.clip((p - self.lo) * self.inv_size, 0.0, 1.0).astype(np.float64)
        out = []
        for l in range(LEVELS):
            r = self.res[l]
            g = q * r
            i = np.clip(np.floor(g), 0.0, r - 1.0)
            f = np.clip(g - i, 0.0, 1.0)
            i = i.astype(np.int64)
            idx = np.zeros((p.shape[0], 8), np.int64)
            wt = np.zeros((p.shape[0], 8))
            for k in range(8):
                d = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
                idx[:, k] = self.index(l, i[:, 0] + d[0], i[:, 1] + d[1], i[:, 2] + d[2])
                wt[:, k] = np.prod(np.where(d == 1, f, 1.0 - f), axis=1)
            out.append((idx, wt))
        return out

    def encode(self, table, pos, cs=None):
        """table: (LEVELS, T, 2) -> features (n, 32), level l at columns 2l, 2l + 1."""
        cs = cs if cs is not None else self.corners(pos)
        feats = np.zeros((len(pos), 2 * LEVELS))
        for l, (idx, wt) in enumerate(cs):
            feats[:, 2 * l:2 * l + 2] = np.einsum("nk,nkf->nf", wt, table[l][idx])
        return feats, cs


def inputs(feats, wo, nrm, rough):
    n = feats.shape[0]
    x = np.zeros((n, N_IN))
    x[:, :32] = feats
    x[:, 32:35] = wo
    x[:, 35:38] = nrm
    x[:, 38] = rough
    x[:, 39] = 1.0
    return x


def head(raw):
    """raw (n, >=41) -> pi (n,8), mu (n,8,3), kappa (n,8), c (n,), and intermediates."""
    o = raw[:, :40].reshape(-1, LOBES, 5)
    a, m, s = o[:, :, 0], o[:, :, 1:4], o[:, :, 4]
    e = np.exp(a - a.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    mnorm = np.maximum(np.linalg.norm(m, axis=2), 1e-12)
    mu = m / mnorm[:, :, None]
    kappa = np.exp(np.clip(s, LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    sig = 1.0 / (1.0 + np.exp(-raw[:, 40]))
    c = SEL_MIN + SEL_SPAN * sig
    return dict(pi=pi, mu=mu, kappa=kappa, c=c, mnorm=mnorm, s=s, sig=sig)


def vmf_norm(kappa):
    return kappa / (2.0 * np.pi * -np.expm1(-2.0 * kappa))


def vmf_pdf(mu, kappa, w):
    """Density of vMF(mu, kappa) on the unit sphere at unit directions w (broadcasting)."""
    return vmf_norm(kappa) * np.exp(kappa * (np.sum(mu * w, axis=-1) - 1.0))


def mix_pdf(pi, mu, kappa, w):
    """Mixture density; pi (n,8), mu (n,8,3), kappa (n,8), w (n,3)."""
    return np.sum(pi * vmf_pdf(mu, kappa, w[:, None, :]), axis=1)


def vmf_sample(mu, kappa, u1, u2):
    """W. Jakob (2012) stable vMF sampling around unit mu."""
    a = np.maximum(u1 + (1.0 - u1) * np.exp(-2.0 * kappa), 1e-300)
    w = np.clip(1.0 + np.log(a) / kappa, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    phi = 2.0 * np.pi * u2
    t = np.where(np.abs(mu[..., :1]) > 0.9, np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    t = np.cross(mu, t)
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    b = np.cross(mu, t)
    return (t * (s * np.cos(phi))[..., None] + b * (s * np.sin(phi))[..., None] + mu * w[..., None])


class Net:
    """Forward and backward of the network for weights w[14336] and table (LEVELS, T, 2)."""

    def __init__(self, enc: Encoding, weights, table, bf16_emulation=False):
        self.enc = enc
        self.bf = bf16_emulation
        w = np.asarray(weights, np.float64)
        t = np.asarray(table, np.float64).reshape(LEVELS, enc.T, 2)
        if bf16_emulation:
            w = bf16(w.astype(np.float32))
            t = bf16(t.astype(np.float32))
        self.W = split_weights(w)
        self.table = t.astype(np.float64)

    def _r(self, x):
        return bf16(x.astype(np.float32)).astype(np.float64) if self.bf else x

    def forward(self, pos, wo, nrm, rough):
        feats, cs = self.enc.encode(self.table, pos)
        x0 = self._r(inputs(feats, np.asarray(wo, np.float64), np.asarray(nrm, np.float64),
                            np.asarray(rough, np.float64)))
        W0_, W1_, W2_, W3_ = self.W
        h1 = self._r(np.maximum(x0 @ W0_.T, 0.0))
        h2 = self._r(np.maximum(h1 @ W1_.T, 0.0))
        h3 = self._r(np.maximum(h2 @ W2_.T, 0.0))
        raw = h3 @ W3_.T
        return dict(raw=raw, x0=x0, h1=h1, h2=h2, h3=h3, cs=cs)

    def loss_grad(self, s):
        """Mean loss, d/d weights (14336,) and d/d table (LEVELS, T, 2)."""
        n = len(s["weight"])
        f = self.forward(s["pos"], s["wo"], s["nrm"], s["rough"])
        hd = head(f["raw"])
        om = np.asarray(s["omega"], np.float64)
        w = np.asarray(s["weight"], np.float64)
        pb = np.asarray(s["pdf_bsdf"], np.float64)
        pi, mu, kappa, c = hd["pi"], hd["mu"], hd["kappa"], hd["c"]
        A = pi * vmf_norm(kappa)
        d = np.sum(mu * om[:, None, :], axis=2)
        u = A * np.exp(kappa * (d - 1.0))
        q = u.sum(axis=1)
        qp = np.maximum(c * q + (1.0 - c) * pb, 1e-20)
        loss = float(np.sum(-w * np.log(qp)) / n)
        G = -w / qp
        dq = G * c
        dt = G * (q - pb) * SEL_SPAN * hd["sig"] * (1.0 - hd["sig"])
        dO = np.zeros((n, N_OUT))
        da = dq[:, None] * (u - pi * q[:, None])
        dm = (dq[:, None] * u * kappa / hd["mnorm"])[:, :, None] * (om[:, None, :] - mu * d[:, :, None])
        inside = (hd["s"] > LOG_KAPPA_MIN) & (hd["s"] < LOG_KAPPA_MAX)
        coth = 1.0 / np.tanh(kappa)
        ds = np.where(inside, dq[:, None] * u * kappa * (1.0 / kappa - coth + d), 0.0)
        o = dO[:, :40].reshape(n, LOBES, 5)
        o[:, :, 0] = da
        o[:, :, 1:4] = dm
        o[:, :, 4] = ds
        dO[:, 40] = dt
        dO = self._r(dO)
        W0_, W1_, W2_, W3_ = self.W
        gW3 = dO.T @ f["h3"]
        dz3 = self._r((dO @ W3_) * (f["h3"] > 0))
        gW2 = dz3.T @ f["h2"]
        dz2 = self._r((dz3 @ W2_) * (f["h2"] > 0))
        gW1 = dz2.T @ f["h1"]
        dz1 = self._r((dz2 @ W1_) * (f["h1"] > 0))
        gW0 = dz1.T @ f["x0"]
        dx0 = dz1 @ W0_
        gt = np.zeros_like(self.table)
        for l, (idx, wt) in enumerate(f["cs"]):
            for k in range(8):
                np.add.at(gt[l], idx[:, k], wt[:, k:k + 1] * dx0[:, 2 * l:2 * l + 2])
        gw = np.concatenate([gW0.ravel(), gW1.ravel(), gW2.ravel(), gW3.ravel()]) / n
        return loss, gw, gt / n

    def loss(self, s):
        return self.loss_grad(s)[0]
