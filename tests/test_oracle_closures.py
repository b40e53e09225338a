"""Pins the oracle's closures against what the mathematics fixes: NDF
normalisation, the white-furnace tests of the reference
(tests/render_tests/microfacet_tests.cpp:8-21, src/shading/white_furnace.*),
chi-square agreement of every sampler with its pdf, reciprocity, energy
bounds, Fresnel closed forms, and the closure values themselves: smooth-limit
albedos (Schlick for the conductor, R + (1 - R)/eta^2 for the dielectric) and
rough albedos integrated over microfacet normals."""
import math
import zlib

import numpy as np
import pytest
from scipy import special, stats

import oracle

LAMBERT, CONDUCTOR, DIELECTRIC = 0, 1, 2
GGX, BECKMANN = 0, 1


def sphere_grid(nt, nphi, upper_only=False):
    th_max = math.pi / 2 if upper_only else math.pi
    th = (np.arange(nt) + 0.5) * th_max / nt
    ph = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (th_max / nt) * (2 * math.pi / nphi)).reshape(-1)
    return w.astype(np.float32), dw


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.7, 1.0])
def test_ndf_projected_area_is_one(dist, alpha):
    m, dw = sphere_grid(3000, 16, upper_only=True)
    D, _ = oracle.mf_eval_batch(dist, alpha, m)
    assert abs(np.sum(D * m[:, 1] * dw) - 1.0) < 2e-3


def _weak_furnace(dist, alpha, theta_o):
    wo = np.array([math.sin(theta_o), math.cos(theta_o), 0.0], np.float32)
    wi, dw = sphere_grid(1200, 2400)
    h = wo[None, :] + wi
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    D, _ = oracle.mf_eval_batch(dist, alpha, h.astype(np.float32))
    _, lam_o = oracle.mf_eval_batch(dist, alpha, wo[None, :])
    G1 = 1.0 / (1.0 + lam_o[0])
    return float(np.sum(G1 * D / (4.0 * abs(wo[1])) * dw))


@pytest.mark.parametrize("dist,alpha,theta_deg", [(BECKMANN, 0.9, 45), (BECKMANN, 1.0, 45),
                                                  (GGX, 1.0, 45), (GGX, 0.3, 60),
                                                  (BECKMANN, 0.3, 70)])
def test_weak_white_furnace(dist, alpha, theta_deg):
    # reference: Beckmann alpha 0.9 at 45 deg -> 1 +- 1%; alpha 1 at 45 deg
    # (both distributions) -> 1 +- 10%.  The exact Smith pair integrates to 1.
    assert abs(_weak_furnace(dist, alpha, math.radians(theta_deg)) - 1.0) < 5e-3


def test_beckmann_lambda_against_rational_fit():
    # microfacet.h:134-141 uses Walter et al.'s rational fit of the exact erf form.
    for a_rcp in np.linspace(0.05, 3.0, 60):
        a = 1.0 / a_rcp
        tan_t = a_rcp      # alpha = 1
        c = 1.0 / math.sqrt(1.0 + tan_t * tan_t)
        w = np.array([math.sqrt(1 - c * c), c, 0.0], np.float32)
        exact = oracle.mf_lambda(BECKMANN, 1.0, w)
        fit = (1 - 1.259 * a + 0.396 * a * a) / (3.535 * a + 2.181 * a * a) if a < 1.6 else 0.0
        assert abs(exact - fit) < 6e-3


def test_fresnel_closed_forms():
    eta = 1.5
    assert abs(oracle.fresnel_dielectric(1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-12
    assert abs(oracle.fresnel_dielectric(-1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-12
    assert oracle.fresnel_dielectric(1e-9, eta) > 0.999
    # total internal reflection from inside beyond the critical angle
    crit = math.asin(1 / eta)
    assert oracle.fresnel_dielectric(-math.cos(crit + 0.01), eta) == 1.0
    assert oracle.fresnel_dielectric(-math.cos(crit - 0.05), eta) < 1.0
    # Brewster angle: the p-polarised term vanishes
    tb = math.atan(eta)
    ci = math.cos(tb)
    ct = math.sqrt(1 - (1 - ci * ci) / eta ** 2)
    rs = (ci - eta * ct) / (ci + eta * ct)
    assert abs(oracle.fresnel_dielectric(ci, eta) - 0.5 * rs * rs) < 1e-12


def test_erfinv():
    x = np.linspace(-0.9999, 0.9999, 2001)
    y = np.array([oracle.erfinv(v) for v in x])
    assert np.max(np.abs(special.erf(y) - x)) < 1e-12
    assert np.max(np.abs(y - special.erfinv(x))) < 1e-9


MATERIALS = {
    "lambert": oracle.material(LAMBERT, albedo=(0.8, 0.5, 0.2)),
    "ggx_conductor": oracle.material(CONDUCTOR, GGX, F0=(0.9, 0.6, 0.3), alpha=0.3),
    "beckmann_conductor": oracle.material(CONDUCTOR, BECKMANN, F0=(1, 1, 1), alpha=0.5),
    "ggx_dielectric": oracle.material(DIELECTRIC, GGX, alpha=0.3, eta=1.5),
    "beckmann_dielectric": oracle.material(DIELECTRIC, BECKMANN, alpha=0.2, eta=1.5),
}


def _wo(theta_deg, inside=False):
    t = math.radians(theta_deg)
    w = np.array([math.sin(t) * 0.6, math.cos(t), math.sin(t) * 0.8], np.float32)
    return -w if inside else w


CHI2_CASES = [(name, th, False) for name in MATERIALS for th in (20, 65)] + \
             [("ggx_dielectric", 30, True), ("beckmann_dielectric", 40, True)]


@pytest.mark.parametrize("name,theta,inside", CHI2_CASES)
def test_sampler_matches_pdf_chi2(name, theta, inside):
    mat = MATERIALS[name]
    wo = _wo(theta, inside)
    rng = np.random.default_rng(zlib.crc32(("%s-%d-%d" % (name, theta, inside)).encode()))
    N = 300_000
    ok, wi, f, pdf = oracle.bsdf_sample_batch(mat, wo, rng.random((N, 3), dtype=np.float32))
    wi, pdf, f = wi[ok], pdf[ok], f[ok]
    # the sampler's pdf and f equal the evaluator's
    fe, pe = oracle.bsdf_eval_batch(mat, wo, wi)
    assert np.allclose(pe, pdf, rtol=2e-4, atol=1e-6)
    assert np.allclose(fe, f, rtol=2e-4, atol=1e-6)
    # histogram over (cos theta, phi) cells of equal solid angle
    # fine sub-cells: the transmission and glossy peaks sit near the poles,
    # where coarse quadrature of the pdf is visibly biased
    nc, nphi, sub = 24, 48, 24
    ci = np.clip(((wi[:, 1] + 1) * 0.5 * nc).astype(int), 0, nc - 1)
    pi = np.clip(((np.arctan2(wi[:, 2], wi[:, 0]) + math.pi) / (2 * math.pi) * nphi).astype(int), 0, nphi - 1)
    obs = np.bincount(ci * nphi + pi, minlength=nc * nphi).astype(np.float64)
    # expected: integrate the pdf over each cell
    cu = (np.arange(nc * sub) + 0.5) / (nc * sub) * 2 - 1
    pu = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2 * math.pi - math.pi
    C, P = np.meshgrid(cu, pu, indexing="ij")
    s = np.sqrt(np.maximum(0, 1 - C * C))
    dirs = np.stack([s * np.cos(P), C, s * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    _, pg = oracle.bsdf_eval_batch(mat, wo, dirs)
    cell_dw = (2.0 / (nc * sub)) * (2 * math.pi / (nphi * sub))
    pg = pg.reshape(nc, sub, nphi, sub).sum(axis=(1, 3)) * cell_dw
    exp = pg.reshape(-1) * N
    # the pdf integrates to the probability that a sample is valid
    assert abs(exp.sum() / N - ok.mean()) < 2e-3
    big = exp >= 8
    o = np.append(obs[big], obs[~big].sum())
    e = np.append(exp[big], exp[~big].sum())
    keep = e > 0
    chi2 = np.sum((o[keep] - e[keep]) ** 2 / e[keep])
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-4, (chi2, keep.sum(), p)


@pytest.mark.parametrize("name", ["lambert", "ggx_conductor", "beckmann_conductor",
                                  "ggx_dielectric", "beckmann_dielectric"])
def test_reflection_reciprocity(name):
    mat = MATERIALS[name]
    rng = np.random.default_rng(9)
    a = rng.normal(size=(500, 3))
    b = rng.normal(size=(500, 3))
    a[:, 1] = np.abs(a[:, 1])
    b[:, 1] = np.abs(b[:, 1])
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    for wo, wi in zip(a.astype(np.float32), b.astype(np.float32)):
        f1, _ = oracle.bsdf_eval(mat, wo, wi)
        f2, _ = oracle.bsdf_eval(mat, wi, wo)
        assert np.allclose(f1, f2, rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("name", list(MATERIALS))
@pytest.mark.parametrize("theta", [10, 50, 80])
def test_energy_bound(name, theta):
    mat = MATERIALS[name]
    wo = _wo(theta)
    rng = np.random.default_rng(theta)
    N = 200_000
    ok, wi, f, pdf = oracle.bsdf_sample_batch(mat, wo, rng.random((N, 3), dtype=np.float32))
    w = np.zeros((N, 3))
    w[ok] = f[ok] * np.abs(wi[ok, 1:2]) / pdf[ok, None]
    est = w.mean(0)
    err = w.std(0) / math.sqrt(N)
    assert np.all(est <= 1.0 + 4 * err + 1e-3), est
    if name == "lambert":
        assert np.allclose(est, [0.8, 0.5, 0.2], atol=5 * err.max() + 1e-3)
    if "dielectric" in name:
        # radiance entering the denser medium is scaled by 1/eta^2: the
        # throughput from outside is R + T/eta^2 with R + T = 1 up to masking
        # losses, so it lies in [1/eta^2, 1] (less a few percent of masking)
        assert 0.85 / 1.5 ** 2 < est[0] <= 1.0 + 4 * err[0] + 1e-3


# ---- closed forms that pin the closure values (not only their sampling) ----
# Smooth limit: at alpha = 1e-3 masking is negligible, so the directional
# albedo of a microfacet closure is its Fresnel term.  Rough: the albedo is
# an expectation over visible normals, written in the microfacet variable m
# (a change of variables away from the evaluator's wi-space formula).


def _albedo_mc(mat, wo, n=400_000, seed=1):
    rng = np.random.default_rng(seed)
    ok, wi, f, pdf = oracle.bsdf_sample_batch(mat, wo, rng.random((n, 3), dtype=np.float32))
    w = np.zeros((n, 3))
    w[ok] = f[ok] * np.abs(wi[ok, 1:2]) / pdf[ok, None]
    return w.mean(0)


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("theta", [20, 50, 75])
def test_smooth_conductor_albedo_is_schlick(dist, theta):
    # F = F0 + (1 - F0)(1 - cos)^5 per channel (fresnel.h:185-187); dropping
    # the Schlick term changes the 75-degree value by up to 52 %
    t = math.radians(theta)
    wo = np.array([math.sin(t), math.cos(t), 0.0], np.float32)
    F0 = np.array([0.9, 0.6, 0.3])
    mat = oracle.material(CONDUCTOR, dist, F0=tuple(F0), alpha=1e-3)
    want = F0 + (1.0 - F0) * (1.0 - math.cos(t)) ** 5
    assert np.allclose(_albedo_mc(mat, wo), want, rtol=5e-3)


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("theta", [20, 50, 75])
def test_smooth_dielectric_albedo_radiance_convention(dist, theta):
    # Radiance crossing into the denser medium is scaled by (n_view/n_light)^2:
    # from outside R + (1 - R)/eta^2, from inside R + (1 - R) eta^2 below the
    # critical angle (1 beyond it).  The reference's eval_refraction
    # (bsdfs.h:566-583) carries the adjoint factor instead, which would give
    # R + (1 - R) = 1 from either side.
    eta = 1.5
    t = math.radians(theta)
    wo = np.array([math.sin(t), math.cos(t), 0.0], np.float32)
    mat = oracle.material(DIELECTRIC, dist, alpha=1e-3, eta=eta)
    R = oracle.fresnel_dielectric(math.cos(t), eta)
    assert abs(_albedo_mc(mat, wo)[0] - (R + (1 - R) / eta ** 2)) < 3e-3
    Ri = oracle.fresnel_dielectric(-math.cos(t), eta)
    assert abs(_albedo_mc(mat, -wo)[0] - (Ri + (1 - Ri) * eta ** 2)) < 5e-3 * (Ri + (1 - Ri) * eta ** 2)


def _fresnel_np(c, eta):
    sin2t = (1 - c * c) / eta ** 2
    cost = np.sqrt(np.maximum(0.0, 1 - sin2t))
    rp = (eta * c - cost) / (eta * c + cost)
    rs = (c - eta * cost) / (c + eta * cost)
    return np.where(sin2t >= 1, 1.0, 0.5 * (rp * rp + rs * rs)), sin2t, cost


def _albedo_m_space(dist, alpha, wo, kind, eta=1.5):
    """Directional albedo as an integral over microfacet normals m (upper
    hemisphere grid): E = E_{m ~ VNDF(wo)}[F G2/G1] for reflection plus
    (n_view/n_light)^2 E[(1 - F) G2/G1] for refraction, using only the
    oracle's D, Lambda and the Fresnel closed form, which other tests pin."""
    m, dm = sphere_grid(1200, 2400, upper_only=True)
    m64 = m.astype(np.float64)
    D, _ = oracle.mf_eval_batch(dist, alpha, m)
    s = 1.0 if wo[1] > 0 else -1.0
    ms = m64 * s
    c = ms @ wo.astype(np.float64)
    _, lam_o = oracle.mf_eval_batch(dist, alpha, wo[None, :])
    vndf = np.maximum(c, 0) * D / abs(float(wo[1])) / (1 + lam_o[0])
    wr = 2 * c[:, None] * ms - wo
    _, lam_r = oracle.mf_eval_batch(dist, alpha, wr.astype(np.float32))
    okr = (c > 0) & (wr[:, 1] * wo[1] > 0)
    g_r = (1 + lam_o[0]) / (1 + lam_o[0] + lam_r)
    if kind == "conductor":   # F0 = 1: F = 1
        return float(np.sum((vndf * g_r * dm)[okr]))
    e = eta if wo[1] > 0 else 1 / eta
    F, sin2t, cost = _fresnel_np(c, e)
    wt = -wo / e + ms * (c / e - cost)[:, None]
    _, lam_t = oracle.mf_eval_batch(dist, alpha, wt.astype(np.float32))
    okt = (c > 0) & (sin2t < 1) & (wt[:, 1] * wo[1] < 0)
    g_t = (1 + lam_o[0]) / (1 + lam_o[0] + lam_t)
    Er = np.sum((vndf * F * g_r * dm)[okr])
    Et = np.sum((vndf * (1 - F) * g_t * dm)[okt])
    return float(Er + Et / e ** 2)


def _albedo_wi_space(mat, wo):
    w, dw = sphere_grid(2000, 4000)
    f, _ = oracle.bsdf_eval_batch(mat, wo, w)
    return float(np.sum(f[:, 0] * np.abs(w[:, 1]) * dw))


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("alpha,theta", [(0.2, 30), (0.5, 70), (0.9, 30)])
def test_rough_conductor_albedo_m_space(dist, alpha, theta):
    t = math.radians(theta)
    wo = np.array([math.sin(t) * 0.6, math.cos(t), math.sin(t) * 0.8], np.float32)
    mat = oracle.material(CONDUCTOR, dist, F0=(1, 1, 1), alpha=alpha)
    ref = _albedo_m_space(dist, alpha, wo, "conductor")
    assert abs(_albedo_wi_space(mat, wo) - ref) < 2e-3 * ref


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("alpha,theta,inside", [(0.1, 30, False), (0.3, 70, False), (0.6, 30, False),
                                                (0.3, 25, True), (0.6, 25, True)])
def test_rough_dielectric_albedo_m_space(dist, alpha, theta, inside):
    t = math.radians(theta)
    wo = np.array([math.sin(t) * 0.6, math.cos(t), math.sin(t) * 0.8], np.float32)
    if inside:
        wo = -wo
    mat = oracle.material(DIELECTRIC, dist, alpha=alpha, eta=1.5)
    ref = _albedo_m_space(dist, alpha, wo, "dielectric")
    assert abs(_albedo_wi_space(mat, wo) - ref) < 2e-3 * ref


# ---- anisotropic microfacets (microfacet.h:60-93, 186-291) ----------------
GLOSS = 3
ANISO = [(0.15, 0.6), (0.5, 0.1)]


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("ax,ay", ANISO)
def test_aniso_ndf_projected_area_is_one(dist, ax, ay):
    m, dw = sphere_grid(2000, 2000, upper_only=True)
    D, _ = oracle.mf_eval_batch(dist, ax, m, alpha_y=ay)
    assert abs(np.sum(D * m[:, 1] * dw) - 1.0) < 3e-3


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
def test_aniso_axes(dist):
    # stretched_roughness pairs m.x with xalpha and m.z with yalpha: a normal
    # tilted in the x-y plane sees alpha_x, one in the z-y plane alpha_y
    ax, ay = 0.2, 0.7
    for th in (0.1, 0.5, 1.0):
        mx = np.array([math.sin(th), math.cos(th), 0.0], np.float32)
        mz = np.array([0.0, math.cos(th), math.sin(th)], np.float32)
        # D carries 1/(ax ay): along x it is the isotropic D(ax) scaled by ax/ay
        # (float32 directions: 1 - cos^2 and sin^2 differ in the last bits)
        assert math.isclose(oracle.mf_D(dist, ax, mx, alpha_y=ay) * ay, oracle.mf_D(dist, ax, mx) * ax,
                            rel_tol=2e-4)
        assert math.isclose(oracle.mf_D(dist, ax, mz, alpha_y=ay) * ax, oracle.mf_D(dist, ay, mz) * ay,
                            rel_tol=2e-4)
        assert math.isclose(oracle.mf_lambda(dist, ax, mx, alpha_y=ay), oracle.mf_lambda(dist, ax, mx), rel_tol=2e-4,
                            abs_tol=1e-12)
        assert math.isclose(oracle.mf_lambda(dist, ax, mz, alpha_y=ay), oracle.mf_lambda(dist, ay, mz), rel_tol=2e-4,
                            abs_tol=1e-12)
        # and a transposed pair differs
        assert not math.isclose(oracle.mf_lambda(dist, ax, mx, alpha_y=ay), oracle.mf_lambda(dist, ay, mx),
                                rel_tol=1e-2)


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("ax,ay", ANISO)
@pytest.mark.parametrize("phi", [0.3, 1.9])
def test_aniso_weak_white_furnace(dist, ax, ay, phi):
    t = math.radians(55)
    wo = np.array([math.sin(t) * math.cos(phi), math.cos(t), math.sin(t) * math.sin(phi)], np.float32)
    wi, dw = sphere_grid(1500, 3000)
    h = wo[None, :] + wi
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    D, _ = oracle.mf_eval_batch(dist, ax, h.astype(np.float32), alpha_y=ay)
    _, lam_o = oracle.mf_eval_batch(dist, ax, wo[None, :], alpha_y=ay)
    assert abs(float(np.sum(D / (1 + lam_o[0]) / (4.0 * abs(wo[1])) * dw)) - 1.0) < 6e-3


ANISO_MATERIALS = {
    "ggx_aniso_conductor": oracle.material(CONDUCTOR, GGX, F0=(0.9, 0.6, 0.3), alpha=0.15, alpha_y=0.6),
    "beckmann_aniso_conductor": oracle.material(CONDUCTOR, BECKMANN, F0=(1, 1, 1), alpha=0.5, alpha_y=0.1),
    "ggx_aniso_dielectric": oracle.material(DIELECTRIC, GGX, alpha=0.4, eta=1.5, alpha_y=0.1),
    "ggx_gloss": oracle.material(GLOSS, GGX, alpha=0.3, eta=1.5),
    "beckmann_aniso_gloss": oracle.material(GLOSS, BECKMANN, alpha=0.1, eta=1.5, alpha_y=0.4),
}


def _chi2_against_pdf(sample, evalf, wo, seed, N=300_000):
    rng = np.random.default_rng(seed)
    ok, wi, f, pdf = sample(wo, rng.random((N, 3), dtype=np.float32))
    wi, pdf, f = wi[ok], pdf[ok], f[ok]
    fe, pe = evalf(wo, wi)
    assert np.allclose(pe, pdf, rtol=2e-4, atol=1e-6)
    assert np.allclose(fe, f, rtol=2e-4, atol=1e-6)
    nc, nphi, sub = 24, 48, 24
    ci = np.clip(((wi[:, 1] + 1) * 0.5 * nc).astype(int), 0, nc - 1)
    pi = np.clip(((np.arctan2(wi[:, 2], wi[:, 0]) + math.pi) / (2 * math.pi) * nphi).astype(int), 0, nphi - 1)
    obs = np.bincount(ci * nphi + pi, minlength=nc * nphi).astype(np.float64)
    cu = (np.arange(nc * sub) + 0.5) / (nc * sub) * 2 - 1
    pu = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2 * math.pi - math.pi
    C, P = np.meshgrid(cu, pu, indexing="ij")
    s = np.sqrt(np.maximum(0, 1 - C * C))
    dirs = np.stack([s * np.cos(P), C, s * np.sin(P)], -1).reshape(-1, 3).astype(np.float32)
    _, pg = evalf(wo, dirs)
    cell_dw = (2.0 / (nc * sub)) * (2 * math.pi / (nphi * sub))
    exp = pg.reshape(nc, sub, nphi, sub).sum(axis=(1, 3)).reshape(-1) * cell_dw * N
    assert abs(exp.sum() / N - ok.mean()) < 3e-3
    big = exp >= 8
    o = np.append(obs[big], obs[~big].sum())
    e = np.append(exp[big], exp[~big].sum())
    keep = e > 0
    chi2 = np.sum((o[keep] - e[keep]) ** 2 / e[keep])
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-4, (chi2, keep.sum(), p)


@pytest.mark.parametrize("name", list(ANISO_MATERIALS))
@pytest.mark.parametrize("theta,inside", [(25, False), (70, False), (35, True)])
def test_aniso_and_gloss_sampler_chi2(name, theta, inside):
    mat = ANISO_MATERIALS[name]
    if inside and "dielectric" not in name:
        pytest.skip("reflection-only closure")
    _chi2_against_pdf(lambda wo, u: oracle.bsdf_sample_batch(mat, wo, u),
                      lambda wo, wi: oracle.bsdf_eval_batch(mat, wo, wi), _wo(theta, inside),
                      zlib.crc32(("%s-%d-%d" % (name, theta, inside)).encode()))


@pytest.mark.parametrize("name", list(ANISO_MATERIALS))
def test_aniso_reflection_reciprocity(name):
    mat = ANISO_MATERIALS[name]
    rng = np.random.default_rng(19)
    a = rng.normal(size=(300, 3))
    b = rng.normal(size=(300, 3))
    a[:, 1] = np.abs(a[:, 1])
    b[:, 1] = np.abs(b[:, 1])
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    f1, _ = oracle.bsdf_eval_batch(mat, a.astype(np.float32), b.astype(np.float32))
    f2, _ = oracle.bsdf_eval_batch(mat, b.astype(np.float32), a.astype(np.float32))
    assert np.allclose(f1, f2, rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("theta", [20, 50, 75])
def test_smooth_gloss_albedo_is_dielectric_fresnel(dist, theta):
    # KpGloss (bsdfs.cpp:359-395): microfacet reflection with F(eta, m.wo)
    t = math.radians(theta)
    wo = np.array([math.sin(t), math.cos(t), 0.0], np.float32)
    mat = oracle.material(GLOSS, dist, alpha=1e-3, eta=1.5)
    want = oracle.fresnel_dielectric(math.cos(t), 1.5)
    assert abs(_albedo_mc(mat, wo)[0] - want) < 5e-3 * want + 1e-4


@pytest.mark.parametrize("dist", [GGX, BECKMANN])
@pytest.mark.parametrize("kind,ax,ay,theta", [("gloss", 0.3, 0.3, 40), ("gloss", 0.2, 0.6, 65),
                                               ("conductor", 0.15, 0.6, 50), ("conductor", 0.6, 0.2, 30)])
def test_rough_gloss_and_aniso_albedo_m_space(dist, kind, ax, ay, theta):
    t = math.radians(theta)
    wo = np.array([math.sin(t) * 0.6, math.cos(t), math.sin(t) * 0.8], np.float32)
    m, dm = sphere_grid(1500, 3000, upper_only=True)
    D, _ = oracle.mf_eval_batch(dist, ax, m, alpha_y=ay)
    c = m.astype(np.float64) @ wo.astype(np.float64)
    _, lam_o = oracle.mf_eval_batch(dist, ax, wo[None, :], alpha_y=ay)
    wr = 2 * c[:, None] * m - wo
    _, lam_r = oracle.mf_eval_batch(dist, ax, wr.astype(np.float32), alpha_y=ay)
    ok = (c > 0) & (wr[:, 1] > 0)
    F = _fresnel_np(c, 1.5)[0] if kind == "gloss" else 1.0
    ref = float(np.sum((np.maximum(c, 0) * D / wo[1] * F / (1 + lam_o[0] + lam_r) * dm)[ok]))
    mat = oracle.material(GLOSS if kind == "gloss" else CONDUCTOR, dist, F0=(1, 1, 1), alpha=ax, alpha_y=ay,
                          eta=1.5)
    assert abs(_albedo_wi_space(mat, wo) - ref) < 2e-3 * ref


# ---- composite closures (CompositeClosure, bsdf.h:132-196, bsdf.cpp:17-74) --
def _mix():
    return oracle.lobes([((0.5, 0.3, 0.2), oracle.material(LAMBERT, albedo=(0.9, 0.8, 0.7))),
                         ((0.2, 0.4, 0.6), oracle.material(CONDUCTOR, GGX, F0=(0.95, 0.9, 0.8), alpha=0.25)),
                         ((0.3, 0.3, 0.3), oracle.material(GLOSS, BECKMANN, alpha=0.4, alpha_y=0.15, eta=1.5))])


@pytest.mark.parametrize("beta", [None, (0.2, 0.5, 1.0)])
@pytest.mark.parametrize("theta", [30, 70])
def test_composite_sampler_chi2(beta, theta):
    mix = _mix()
    _chi2_against_pdf(lambda wo, u: oracle.composite_sample_batch(mix, wo, u, beta),
                      lambda wo, wi: oracle.composite_eval_batch(mix, wo, wi, beta), _wo(theta),
                      1000 + theta + (0 if beta is None else 1))


@pytest.mark.parametrize("beta", [None, (0.2, 0.5, 1.0)])
def test_composite_albedo_is_weighted_sum(beta):
    # linearity: E[f cos / pdf] of the mixture = sum_i w_i * albedo_i, whatever
    # the lobe-pick probabilities
    wo = _wo(40)
    mix = _mix()
    n = 400_000
    rng = np.random.default_rng(77)
    ok, wi, f, pdf = oracle.composite_sample_batch(mix, wo, rng.random((n, 3), dtype=np.float32), beta)
    w = np.zeros((n, 3))
    w[ok] = f[ok] * np.abs(wi[ok, 1:2]) / pdf[ok, None]
    est, err = w.mean(0), w.std(0) / math.sqrt(n)
    want = np.zeros(3)
    for lb in mix:
        want += np.array(lb.weight[:]) * _albedo_mc(lb.lobe, wo, n=300_000, seed=5)
    assert np.all(np.abs(est - want) < 5 * err + 2e-3), (est, want)


def test_composite_pick_follows_beta():
    # p_i = dot(w_i, beta) / sum: with beta = (1, 0, 0) a lobe weighted
    # (0, 1, 0) is never picked, so every sample's density is the other's
    lam = oracle.material(LAMBERT, albedo=(1, 1, 1))
    cond = oracle.material(CONDUCTOR, GGX, F0=(1, 1, 1), alpha=0.05)
    mix = oracle.lobes([((1, 0, 0), lam), ((0, 1, 0), cond)])
    wo = _wo(30)
    rng = np.random.default_rng(3)
    ok, wi, f, pdf = oracle.composite_sample_batch(mix, wo, rng.random((20000, 3), dtype=np.float32), (1, 0, 0))
    assert ok.all()
    assert np.allclose(pdf, wi[:, 1] / math.pi, rtol=1e-5)   # the Lambert lobe's cosine density


def test_single_lobe_composite_equals_material():
    for mat in list(MATERIALS.values()) + list(ANISO_MATERIALS.values()):
        one = oracle.lobes([((1, 1, 1), mat)])
        rng = np.random.default_rng(8)
        u = rng.random((5000, 3), dtype=np.float32)
        wo = _wo(45)
        a = oracle.composite_sample_batch(one, wo, u, (0.3, 0.2, 0.9))
        b = oracle.bsdf_sample_batch(mat, wo, u)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y))
