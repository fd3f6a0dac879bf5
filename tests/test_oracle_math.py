"""Pins the oracle's math layer against closed forms and independent computations.

Covers src/math_func.cpp (erf, erfinv), src/utils.cpp (Fresnel, Moller-Trumbore,
frames, warps), include/phase.h (HG), src/bsdf/* (diffuse, Phong, rough
dielectric) and include/pmf.h (discrete distributions).
"""
import math

import numpy as np
import pytest
from scipy import special, stats


def test_erf_within_abramowitz_stegun_bound(oracle):
    xs = np.linspace(-4, 4, 801)
    err = [abs(oracle.unit("erf", [x])[0] - math.erf(x)) for x in xs]
    assert max(err) < 1.5e-7   # A&S 7.1.26 maximum absolute error


def test_erfinv_inverts_erf(oracle):
    xs = np.linspace(-0.999, 0.999, 501)
    back = np.array([special.erf(oracle.unit("erfinv", [x])[0]) for x in xs])
    assert np.max(np.abs(back - xs)) < 5e-6   # Giles' single-precision fit


@pytest.mark.parametrize("eta", [1.33, 1.5, 2.4])
def test_fresnel_textbook(oracle, eta):
    assert abs(oracle.unit("fresnel", [1.0, eta])[0] - ((eta - 1) / (eta + 1)) ** 2) < 1e-12
    for ci in np.linspace(-1, 1, 41):
        if abs(ci) < 1e-3:
            continue
        F, ct = oracle.unit("fresnel", [ci, eta])
        n1, n2 = (1.0, eta) if ci > 0 else (eta, 1.0)
        c1 = abs(ci)
        s2 = (n1 / n2) ** 2 * (1 - c1 * c1)
        if s2 >= 1:
            assert F == 1.0 and ct == 0.0   # total internal reflection
            continue
        c2 = math.sqrt(1 - s2)
        rs = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)
        rp = (n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)
        assert abs(F - 0.5 * (rs * rs + rp * rp)) < 1e-9
        assert abs(abs(ct) - c2) < 1e-9 and np.sign(ct) == -np.sign(ci)


def test_frame_is_orthonormal_right_handed(oracle):
    rng = np.random.default_rng(0)
    for n in rng.normal(size=(200, 3)):
        f = oracle.unit("frame", n).reshape(3, 3)
        s, t, nn = f
        assert np.allclose(nn, n / np.linalg.norm(n), atol=1e-12)
        assert np.allclose(f @ f.T, np.eye(3), atol=1e-12)
        assert np.allclose(np.cross(s, t), nn, atol=1e-12)


def test_warps(oracle):
    rng = np.random.default_rng(1)
    u = rng.random((4000, 2))
    d = np.array([oracle.unit("cos_hemisphere", x) for x in u])
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12) and (d[:, 2] >= 0).all()
    assert stats.kstest(d[:, 2] ** 2, "uniform").pvalue > 1e-3   # cos^2 uniform under p = cos/pi
    s = np.array([oracle.unit("uniform_sphere", x) for x in u])
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-12)
    assert stats.kstest((s[:, 2] + 1) / 2, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.3, 0.8])
def test_hg_normalised_and_sampling_orientation(oracle, g):
    wi = np.array([0.0, 0.0, 1.0])
    mus = np.linspace(-1, 1, 4001)
    vals = np.array([oracle.unit("hg_eval", [g, 0, 0, 1, math.sqrt(max(0, 1 - m * m)), 0, m])[0] for m in mus])
    assert abs(np.trapezoid(2 * math.pi * vals, mus) - 1) < 1e-4
    # mean cosine of the evaluated lobe, taken against wi, is -g
    assert abs(np.trapezoid(2 * math.pi * vals * mus, mus) + g) < 1e-3
    # phase.h:38-51 builds the sampled direction in Frame(wi): the sampled mean
    # cosine against wi is +g, mirroring the evaluated lobe (reference behaviour).
    rng = np.random.default_rng(2)
    u = rng.random((20000, 2))
    smp = np.array([oracle.unit("hg_sample", [g, *wi, a, b]) for a, b in u])
    mu = smp @ wi
    assert abs(mu.mean() - g) < 4 * mu.std() / math.sqrt(len(mu)) + 1e-9


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.8])
def test_beckmann_normalisation_and_smith(oracle, alpha):
    th = np.linspace(0, math.pi / 2, 3001)
    D = np.array([oracle.unit("beckmann_eval", [alpha, math.sin(t), 0, math.cos(t)])[0] for t in th])
    assert abs(np.trapezoid(D * np.cos(th) * np.sin(th) * 2 * math.pi, th) - 1) < 2e-3
    # rational G1 fit vs the exact Beckmann Smith term, 0.35% claim (microfacet.h:131-134)
    for a in np.linspace(0.1, 1.55, 30):
        tan_t = 1.0 / (alpha * a)
        t = math.atan(tan_t)
        v = [math.sin(t), 0.0, math.cos(t)]
        g1 = oracle.unit("smith_g1", [alpha, *v, 0, 0, 1])[0]
        exact = 2.0 / (1.0 + math.erf(a) + math.exp(-a * a) / (a * math.sqrt(math.pi)))
        assert abs(g1 / exact - 1) < 4e-3


def test_visible_normal_pdf_integrates_to_one(oracle):
    alpha = 0.4
    wi = np.array([math.sin(0.7), 0.0, math.cos(0.7)])
    th = np.linspace(0, math.pi / 2, 121)
    ph = np.linspace(0, 2 * math.pi, 121)
    tot = np.zeros((len(th), len(ph)))
    for i, t in enumerate(th):
        for j, p in enumerate(ph):
            m = [math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)]
            tot[i, j] = oracle.unit("beckmann_pdf", [alpha, *wi, *m])[0] * math.sin(t)
    assert abs(np.trapezoid(np.trapezoid(tot, ph, axis=1), th) - 1) < 1e-2


def _bsdf(kind):
    if kind == "diffuse":
        return [0, 0.6, 0.5, 0.4, 0, 0, 0, 0]
    if kind == "phong":
        return [2, 0.3, 0.3, 0.3, 0.5, 0.5, 0.5, 20.0]
    return [3, 0.3, 1.5, 0, 0, 0, 0, 0]


@pytest.mark.parametrize("kind", ["diffuse", "phong", "dielectric"])
def test_bsdf_sample_weight_equals_eval_over_pdf(oracle, kind):
    rng = np.random.default_rng(3)
    params = _bsdf(kind)
    checked = 0
    for _ in range(400):
        wi = rng.normal(size=3)
        wi /= np.linalg.norm(wi)
        if kind != "dielectric":
            wi[2] = abs(wi[2])
        r = rng.random(3)
        out = oracle.unit("bsdf_sample", params + [*wi, *r], nout=8)
        w, wo, pdf = out[:3], out[3:6], out[6]
        if np.all(w == 0) or pdf <= 1e-6:
            continue
        ev = oracle.unit("bsdf_eval", params + [*wi, *wo])
        pd = oracle.unit("bsdf_pdf", params + [*wi, *wo])[0]
        assert abs(pd - pdf) <= 1e-6 * max(1.0, pdf)
        assert np.allclose(w, ev / pdf, rtol=1e-6, atol=1e-9)
        checked += 1
    assert checked > 100


def test_diffuse_closed_form(oracle):
    wi = [0, 0, 1]
    for th in np.linspace(0.05, 1.5, 10):
        wo = [math.sin(th), 0, math.cos(th)]
        ev = oracle.unit("bsdf_eval", _bsdf("diffuse") + wi + wo)
        assert np.allclose(ev, np.array([0.6, 0.5, 0.4]) * math.cos(th) / math.pi, rtol=1e-12)


def test_ray_triangle_against_linear_solve(oracle):
    rng = np.random.default_rng(4)
    for _ in range(200):
        v = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        A = np.stack([v[1] - v[0], v[2] - v[0], -d], axis=1)
        if abs(np.linalg.det(A)) < 1e-3:
            continue
        u, w, t = np.linalg.solve(A, o - v[0])
        got = oracle.unit("ray_triangle", [*v.ravel(), *o, *d])
        assert np.allclose(got, [u, w, t], rtol=1e-9, atol=1e-9)


def test_ray_triangle_tangent_vs_central_difference(oracle):
    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.normal(size=15)
        x[12:15] /= np.linalg.norm(x[12:15])
        g = rng.normal(size=15)
        out = oracle.unit("ray_triangle_ad", [*x, *g])
        h = 1e-6
        fp = oracle.unit("ray_triangle", x + h * g)
        fm = oracle.unit("ray_triangle", x - h * g)
        fd = (fp - fm) / (2 * h)
        assert np.allclose(out[3:6], fd, rtol=1e-5, atol=1e-5)


def test_distribution_sample_reuse(oracle):
    w = [1.0, 0.0, 3.0, 2.0]
    cdf = np.concatenate([[0], np.cumsum(w)]) / sum(w)
    for u in np.linspace(0.001, 0.999, 97):
        idx, pdf, ur = oracle.unit("distrib_sample", [len(w), *w, u])
        k = int(np.searchsorted(cdf, u, side="left") - 1)
        k = min(max(k, 0), len(w) - 1)
        while w[k] == 0:
            k += 1
        assert int(idx) == k
        assert abs(pdf - w[k] / sum(w)) < 1e-12
        assert abs(ur - (u - cdf[k]) / (cdf[k + 1] - cdf[k])) < 1e-12
