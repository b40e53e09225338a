"""GPU guiding network against the FP64 oracle (with the same bf16 rounding
of weights, features, activations and gradients): forward outputs, mixture
pdf and sampler, gradients, and the Adam update.  Tolerances: the only
difference left is FP32 accumulation order on the matrix cores (and bf16
rounding flips it can cause), so outputs agree to ~1e-2 absolute."""
import math

import numpy as np
import pytest

import oracle
from kazen_proto_amd import native

pytestmark = pytest.mark.gpu


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
    return (np.concatenate([pos, nrm], 1).astype(np.float32), d.astype(np.float32), w.astype(np.float32))


@pytest.fixture(scope="module")
def gctx():
    c = native.Context(0)
    yield c
    c.close()


def test_forward_matches_oracle(gctx):
    params = init_params(1, scale=1.5)
    gctx.guiding_set_params(params)
    assert np.array_equal(gctx.guiding_get_params(), params)
    in6, _, _ = samples(1000, 2)
    gpu = gctx.guiding_eval(in6)
    ref = oracle.mlp_forward(params, in6, emulate_bf16=True)
    err = np.abs(gpu - ref)
    print("forward max abs err %.3e, 99.9th pct %.3e, |out| max %.2f" % (err.max(), np.quantile(err, 0.999),
                                                                       np.abs(ref).max()))
    assert np.quantile(err, 0.999) < 2e-2 * max(1.0, np.abs(ref).max())
    assert np.median(err) < 1e-3


def test_mixture_pdf_and_sampler_match_oracle(gctx):
    params = init_params(3, scale=2.0)
    gctx.guiding_set_params(params)
    in6, d, _ = samples(512, 4)
    pdf = gctx.guiding_pdf(in6, d)
    outs = gctx.guiding_eval(in6).astype(np.float64)
    ref = np.array([oracle.vmf_pdf(o, dd) for o, dd in zip(outs, d)])
    assert np.allclose(pdf, ref, rtol=2e-3, atol=1e-6)
    u = np.random.default_rng(5).random((512, 3)).astype(np.float32)
    dirs, spdf = gctx.guiding_sample(in6, u)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-5)
    close = 0
    for o, uu, dd, pp in zip(outs, u, dirs, spdf):
        rd, rp = oracle.vmf_sample(o, uu)
        if np.allclose(dd, rd, atol=2e-3):
            close += 1
            assert abs(pp - rp) <= 2e-3 * rp + 1e-6
    assert close >= 0.98 * len(u)


def test_gradient_matches_oracle(gctx):
    params = init_params(6, scale=1.5)
    gctx.guiding_set_params(params)
    in6, d, w = samples(3000, 7)
    g = gctx.guiding_train_step(in6, d, w, apply=False)
    _, ref = oracle.mlp_grad(params, in6, d, w, emulate_bf16=True)
    assert np.array_equal(gctx.guiding_get_params(), params)   # no update without apply
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    cos = float(np.dot(g, ref) / (np.linalg.norm(g) * np.linalg.norm(ref)))
    print("gradient relative L2 error %.3e, cosine %.6f" % (rel, cos))
    assert rel < 2e-2 and cos > 0.9998
    # per layer
    offs = [0, 3136, 7296, 11456, 15616, 17696]
    for a, b in zip(offs[:-1], offs[1:]):
        assert np.linalg.norm(g[a:b] - ref[a:b]) < 3e-2 * np.linalg.norm(ref[a:b]) + 1e-9


def test_adam_step_matches_oracle(gctx):
    params = init_params(8)
    gctx.guiding_set_params(params)
    in6, d, w = samples(2048, 9)
    g = gctx.guiding_train_step(in6, d, w, lr=1e-3, apply=True)
    p = params.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    oracle.adam_step(p, m, v, g, 1, lr=1e-3)
    assert np.allclose(gctx.guiding_get_params(), p, atol=1e-6)


def test_training_reduces_loss(gctx):
    """Directions drawn from a fixed lobe around +y must become more likely."""
    gctx.guiding_init(seed=11)
    rng = np.random.default_rng(12)
    n = 8192
    in6 = np.concatenate([rng.random((n, 3)), np.tile([0, 1, 0], (n, 1))], 1).astype(np.float32)
    d = rng.normal(size=(n, 3)) * 0.2 + np.array([0.3, 1.0, 0.0])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = np.ones(n, np.float32)
    before = np.mean(np.log(gctx.guiding_pdf(in6[:1024], d[:1024])))
    for _ in range(150):
        gctx.guiding_train_step(in6, d, w, lr=5e-3, apply=True)
    after = np.mean(np.log(gctx.guiding_pdf(in6[:1024], d[:1024])))
    print("mean log pdf of the training directions: %.3f -> %.3f" % (before, after))
    assert after > before + 1.5


def test_checkpoint_resumes_training_exactly(ctx):
    # save after some steps; further steps from the saved state (in another
    # context) equal the further steps of the original, bit for bit
    from kazen_proto_amd import native
    rng = np.random.default_rng(31)
    n = 4096
    in6 = np.concatenate([rng.random((n, 3)), rng.normal(size=(n, 3))], 1).astype(np.float32)
    in6[:, 3:] /= np.linalg.norm(in6[:, 3:], axis=1, keepdims=True)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    w = rng.random(n).astype(np.float32)
    ctx.guiding_init(7)
    for _ in range(3):
        ctx.guiding_train_step(in6, dirs, w, lr=1e-3)
    ck = ctx.guiding_save()
    assert len(ck) == 64 + 3 * 4 * ctx.GUIDE_PARAMS
    for _ in range(4):
        ctx.guiding_train_step(in6, dirs, w, lr=1e-3)
    a = ctx.guiding_get_params()
    other = native.Context(0)
    other.guiding_load(ck)
    for _ in range(4):
        other.guiding_train_step(in6, dirs, w, lr=1e-3)
    b = other.guiding_get_params()
    with pytest.raises(native.KpError):
        other.guiding_load(b"\0" * len(ck))
    other.close()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
