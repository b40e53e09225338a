"""Pins the oracle of the path-guiding network: gradients against central
finite differences (FP64), the vMF mixture's normalisation and sampler (chi
square), frequency-encoding values and the Adam step's closed form."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle


def init_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    dims = [(48, 64), (64, 64), (64, 64), (64, 64), (64, 32)]
    parts = []
    for fi, fo in dims:
        lim = scale * math.sqrt(6.0 / (fi + fo))
        parts.append(rng.uniform(-lim, lim, size=fi * fo))
        parts.append(rng.uniform(-0.1, 0.1, size=fo))
    return np.concatenate(parts).astype(np.float32)


def samples(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.random(n) + 0.1
    return (np.concatenate([pos, nrm], 1).astype(np.float32), d.astype(np.float32),
            w.astype(np.float32))


def test_param_count():
    assert oracle.mlp_param_count() == 17696
    assert init_params(0).size == 17696


def test_encoding_values():
    x = np.array([0.25, 0.5, 0.125, 0.0, 1.0, 0.0], np.float32)
    f = oracle.mlp_encode(x)
    assert abs(f[0] - math.sin(math.pi * 0.25)) < 1e-15 and abs(f[1] - math.cos(math.pi * 0.25)) < 1e-15
    assert abs(f[10] - math.sin(32 * math.pi * 0.25)) < 1e-12
    assert abs(f[36 + 4] - math.sin(math.pi * 1.0)) < 1e-12   # normal y, k = 0


def test_gradient_matches_finite_differences():
    # pure FP64 forward/backward (no bf16 rounding) against central differences
    params = init_params(1, scale=1.5)
    in6, d, w = samples(6, 2)
    loss, g = oracle.mlp_grad(params, in6, d, w, emulate_bf16=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(3072, 12, replace=False),            # W0
                          3072 + rng.choice(64, 4, replace=False),        # b0
                          rng.choice(np.arange(3136, 17696), 40, replace=False)])
    for i in idx:
        h = 1e-3 * max(1.0, abs(params[i]))
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        h_eff = (np.float64(p1[i]) - np.float64(p2[i]))
        l1, _ = oracle.mlp_grad(p1, in6, d, w, emulate_bf16=False)
        l2, _ = oracle.mlp_grad(p2, in6, d, w, emulate_bf16=False)
        fd = (l1 - l2) / h_eff
        assert abs(fd - g[i]) < 2e-3 * max(1e-3, abs(g[i])) + 2e-6, (i, fd, g[i])


def test_mixture_integrates_to_one_and_sampler_matches_pdf():
    params = init_params(4, scale=2.0)
    in6, _, _ = samples(3, 5)
    outs = oracle.mlp_forward(params, in6)
    nt, nphi = 400, 800
    th = (np.arange(nt) + 0.5) * math.pi / nt
    ph = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    for out in outs:
        # normalisation by quadrature
        total = 0.0
        for t in th[::1]:
            st = math.sin(t)
            dirs = np.stack([st * np.cos(ph), np.full(nphi, math.cos(t)), st * np.sin(ph)], 1).astype(np.float32)
            total += sum(oracle.vmf_pdf(out, dd) for dd in dirs[::8]) * 8 * st
        total *= (math.pi / nt) * (2 * math.pi / nphi)
        assert abs(total - 1.0) < 2e-2
        # sampler vs pdf: chi-square over equal-area (cos theta, phi) cells
        rng = np.random.default_rng(7)
        N = 40000
        u = rng.random((N, 3)).astype(np.float32)
        ds = np.array([oracle.vmf_sample(out, uu)[0] for uu in u])
        pd = np.array([oracle.vmf_sample(out, uu)[1] for uu in u[:200]])
        assert np.allclose(pd, [oracle.vmf_pdf(out, dd) for dd in ds[:200]], rtol=1e-9)
        nc, npc = 10, 20
        ci = np.clip(((ds[:, 1] + 1) / 2 * nc).astype(int), 0, nc - 1)
        pi = np.clip(((np.arctan2(ds[:, 2], ds[:, 0]) + math.pi) / (2 * math.pi) * npc).astype(int), 0, npc - 1)
        obs = np.bincount(ci * npc + pi, minlength=nc * npc)
        sub = 12
        cu = (np.arange(nc * sub) + 0.5) / (nc * sub) * 2 - 1
        pu = (np.arange(npc * sub) + 0.5) / (npc * sub) * 2 * math.pi - math.pi
        C, P = np.meshgrid(cu, pu, indexing="ij")
        s = np.sqrt(1 - C * C)
        dirs = np.stack([s * np.cos(P), C, s * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
        pg = np.array([oracle.vmf_pdf(out, dd) for dd in dirs]).reshape(nc, sub, npc, sub).sum((1, 3))
        exp = pg.reshape(-1) * (2 / (nc * sub)) * (2 * math.pi / (npc * sub)) * N
        big = exp > 5
        o = np.append(obs[big], obs[~big].sum())
        e = np.append(exp[big], exp[~big].sum())
        keep = e > 0
        o, e = o[keep], e[keep]
        e *= o.sum() / e.sum()
        chi2 = ((o - e) ** 2 / e).sum()
        assert stats.chi2.sf(chi2, len(e) - 1) > 1e-4


def test_adam_closed_form_on_quadratic():
    # minimise 0.5*||x - c||^2 with Adam; first step moves every coordinate by lr
    c = np.array([1.0, -2.0, 3.0, 0.5], np.float32)
    x = np.zeros(4, np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    oracle.adam_step(x, m, v, x - c, 1, lr=0.1)
    assert np.allclose(x, 0.1 * np.sign(c), atol=1e-6)
    for t in range(2, 3000):
        oracle.adam_step(x, m, v, x - c, t, lr=0.05)
    assert np.allclose(x, c, atol=1e-2)


def test_bf16_emulation_is_close_to_fp64():
    params = init_params(8)
    in6, d, w = samples(16, 9)
    a = oracle.mlp_forward(params, in6, emulate_bf16=True)
    b = oracle.mlp_forward(params, in6, emulate_bf16=False)
    assert np.max(np.abs(a - b)) < 0.05 * max(1.0, np.abs(b).max())
