"""The oracle's world-mode BSDFs (oracle/oracle_bsdf.c) pinned by what the mathematics fixes:
normalisations of the Trowbridge-Reitz distribution and of its visible normals, the textbook
conductor Fresnel equations (real arithmetic, PBRT's form) against the reference's complex
arithmetic, the defining property of Gulbrandsen's parametrisation (reflectivity r at normal
incidence), reciprocity, the exact pdf mass of a reflection lobe at normal incidence, and the
principled mixture's limits (svm/surface/principled.rs:120-330, BsdfMixture mod.rs:1560-1593)."""
import numpy as np
import pytest
from scipy import integrate

import oracle as O

pytestmark = pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")


def _dir(theta, phi=0.0):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


N = np.array([0.0, 0.0, 1.0])


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 1.0])
def test_ggx_d_is_normalised(alpha):
    """int D(h) cos(theta_h) dw_h = 1 over the hemisphere"""
    f = lambda t: O.ggx_d(np.cos(t), alpha) * np.cos(t) * np.sin(t) * 2 * np.pi
    pts = [min(k * alpha, 1.5) for k in (0.25, 0.5, 1, 2, 4)]
    val, _ = integrate.quad(f, 0, np.pi / 2, points=pts, limit=400, epsabs=1e-12, epsrel=1e-10)
    assert abs(val - 1) < 1e-7, val


def _phi_integral(a, b):
    """int_0^2pi max(0, a cos(phi) + b) dphi, a >= 0"""
    if b >= a:
        return 2 * np.pi * b
    if b <= -a:
        return 0.0
    return 2 * (b * np.arccos(-b / a) + np.sqrt(a * a - b * b))


@pytest.mark.parametrize("alpha", [0.1, 0.4, 0.9])
@pytest.mark.parametrize("theta_o", [0.1, 0.6, 1.2, 1.5])
def test_visible_normals_are_normalised(alpha, theta_o):
    """G1(wo) int max(0, wo . h) D(h) dw_h = cos(theta_o): the identity that defines Smith's
    Lambda for a given D (so it pins Lambda against D)"""
    g1 = 1 / (1 + O.ggx_lambda(np.cos(theta_o), alpha))

    def f(t):
        a, b = np.sin(theta_o) * np.sin(t), np.cos(theta_o) * np.cos(t)
        return O.ggx_d(np.cos(t), alpha) * np.sin(t) * _phi_integral(a, b)

    pts = [min(k * alpha, 1.5) for k in (0.25, 0.5, 1, 2, 4)] + [np.pi / 2 - theta_o]
    val, _ = integrate.quad(f, 0, np.pi / 2, points=sorted(set(pts)), limit=400, epsabs=1e-12, epsrel=1e-10)
    assert abs(g1 * val - np.cos(theta_o)) < 1e-7 * max(1.0, np.cos(theta_o)), (g1 * val, np.cos(theta_o))


def _fr_conductor_textbook(c, n, k):
    """unpolarised conductor reflectance, real arithmetic (PBRT 3rd ed. FrConductor, eta_i = 1)"""
    c2 = c * c
    s2 = 1 - c2
    t0 = n * n - k * k - s2
    a2b2 = np.sqrt(t0 * t0 + 4 * n * n * k * k)
    t1 = a2b2 + c2
    a = np.sqrt(0.5 * (a2b2 + t0))
    t2 = 2 * c * a
    rs = (t1 - t2) / (t1 + t2)
    t3 = c2 * a2b2 + s2 * s2
    t4 = t2 * s2
    rp = rs * (t3 - t4) / (t3 + t4)
    return 0.5 * (rp + rs)


def test_fresnel_matches_textbook_conductor():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        c, n, k = rng.uniform(0, 1), rng.uniform(0.05, 4), rng.uniform(0, 6)
        want = _fr_conductor_textbook(min(c, 0.999), n, k)
        assert abs(O.fr_complex(c, n, k) - want) < 1e-12 * max(1, want), (c, n, k)
    # the clamp at 0.999 (and at 0 for a negative cosine)
    assert O.fr_complex(1.0, 0.2, 3.0) == O.fr_complex(0.999, 0.2, 3.0)
    assert O.fr_complex(-0.5, 0.2, 3.0) == O.fr_complex(0.0, 0.2, 3.0)
    assert abs(O.fr_complex(0.0, 0.2, 3.0) - 1) < 1e-12  # grazing incidence reflects everything


def test_conductor_parametrisation_keeps_the_reflectivity():
    """Gulbrandsen: whatever the edge tint g, the fitted (n, k) reflect r at normal incidence;
    g = 1 gives n = (1-r)/(1+r), k = 2 sqrt(r)/(1+r); g = 0 gives the dielectric k = 0"""
    rng = np.random.default_rng(4)
    for _ in range(500):
        r, g = rng.uniform(0, 0.99), rng.uniform(0, 1)
        n, k = O.conductor_nk(r, g)
        assert abs(_fr_conductor_textbook(1.0, n, k) - r) < 1e-12, (r, g)
    r = 0.64
    n, k = O.conductor_nk(r, 1.0)
    assert abs(n - 0.36 / 1.64) < 1e-14 and abs(k - 1.6 / 1.64) < 1e-14
    n, k = O.conductor_nk(r, 0.0)
    assert abs(n - 1.8 / 0.2) < 1e-12 and k == 0.0
    assert O.conductor_nk(1.5, 0.3) == O.conductor_nk(0.99, 0.3)  # reflectivity clamped to 0.99


def test_metal_retroreflection_closed_form():
    """wo = wi = n: h = n, D = 1/(pi alpha^2), G = 1, so f |cos| = F(0.999)/(4 pi alpha^2) and
    pdf = 1/(4 pi alpha^2), alpha = max(roughness^2, 1e-4)"""
    R = np.array([0.9, 0.6, 0.2])
    tint = (1.0, 0.7, 0.4)
    for rough in (0.5, 0.3, 0.0):
        alpha = float(max(np.float32(rough) * np.float32(rough), np.float32(1e-4)))
        e = O.bsdf(metallic=1.0, roughness=rough, specular_tint=tint)
        pdf, fc = O.bsdf_eval(e, R, N, N, N, N)
        assert abs(pdf - 1 / (4 * np.pi * alpha * alpha)) < 1e-12 * pdf
        for ch in range(3):
            n, k = O.conductor_nk(np.float32(R[ch]), np.float32(tint[ch]))
            want = _fr_conductor_textbook(0.999, n, k) / (4 * np.pi * alpha * alpha)
            assert abs(fc[ch] - want) < 1e-12 * want, (rough, ch)


@pytest.mark.parametrize("rough", [0.2, 0.5, 0.9])
def test_reflection_pdf_mass_at_normal_incidence(rough):
    """At wo = n the visible normals are D cos(theta_h), wi = reflect(n, h) has theta_i = 2
    theta_h, and wi leaves the hemisphere when theta_h > 45 degrees: the pdf integrates to
    P(tan^2 theta_h < 1) = 1 / (1 + alpha^2) over the upper hemisphere"""
    alpha = max(float(np.float32(rough) * np.float32(rough)), 1e-4)
    e = O.bsdf(metallic=1.0, roughness=rough)
    f = lambda t: O.bsdf_eval(e, (0.5, 0.5, 0.5), N, N, N, _dir(t, 0.3))[0] * np.sin(t) * 2 * np.pi
    pts = [min(k * alpha, 1.5) for k in (0.5, 1, 2, 4)]
    val, _ = integrate.quad(f, 0, np.pi / 2, points=pts, limit=400, epsabs=1e-11, epsrel=1e-9)
    assert abs(val - 1 / (1 + alpha * alpha)) < 1e-6, (val, 1 / (1 + alpha * alpha))


def test_reciprocity_and_height_correlated_masking():
    """f(wo, wi) = f(wi, wo); and the masking-shadowing term G = (f |cos_i| / F) / (pdf / G1(wo))
    lies strictly between the uncorrelated product G1(wo) G1(wi) and min(G1(wo), G1(wi))"""
    rng = np.random.default_rng(5)
    R = np.array([0.7, 0.5, 0.3])
    e = O.bsdf(metallic=1.0, roughness=0.6, specular_tint=(0.8, 0.9, 1.0))
    alpha = float(np.float32(0.6) * np.float32(0.6))
    checked = 0
    for _ in range(400):
        wo = _dir(rng.uniform(0.05, 1.45), rng.uniform(0, 2 * np.pi))
        wi = _dir(rng.uniform(0.05, 1.45), rng.uniform(0, 2 * np.pi))
        p1, f1 = O.bsdf_eval(e, R, N, N, wo, wi)
        p2, f2 = O.bsdf_eval(e, R, N, N, wi, wo)
        if p1 == 0:
            continue
        checked += 1
        np.testing.assert_allclose(f1 / wi[2], f2 / wo[2], rtol=1e-12)
        h = (wo + wi) / np.linalg.norm(wo + wi)
        n, k = O.conductor_nk(np.float32(R[0]), np.float32(0.8))
        F = O.fr_complex(abs(wi @ h), n, k)
        g1o = 1 / (1 + O.ggx_lambda(wo[2], alpha))
        g1i = 1 / (1 + O.ggx_lambda(wi[2], alpha))
        G = (f1[0] / F) / (p1 / g1o)
        assert g1o * g1i < G < min(g1o, g1i) + 1e-12, (G, g1o, g1i)
    assert checked > 300


def test_principled_mixture_limits():
    """metallic 0 (no specular, no transmission) is the diffuse BSDF bit for bit; metallic 1 the
    metal lobe alone; in between the lerp of both; BsdfMixture skips a side within 1e-4 of the
    end (the lerp then scales the other side by metallic)"""
    rng = np.random.default_rng(6)
    R = np.array([0.8, 0.4, 0.1])
    for _ in range(200):
        ns = _dir(rng.uniform(0, 0.3), rng.uniform(0, 6.3))
        wo = _dir(rng.uniform(0, 1.4), rng.uniform(0, 6.3))
        wi = _dir(rng.uniform(0, 1.4), rng.uniform(0, 6.3))
        pd, fd = O.bsdf_eval(None, R, ns, ns, wo, wi)
        p0, f0 = O.bsdf_eval(O.bsdf(metallic=0.0), R, ns, ns, wo, wi)
        assert p0 == pd and np.array_equal(f0, fd)
        pm, fm = O.bsdf_eval(O.bsdf(metallic=1.0, roughness=0.4), R, ns, ns, wo, wi)
        ph, fh = O.bsdf_eval(O.bsdf(metallic=0.5, roughness=0.4), R, ns, ns, wo, wi)
        np.testing.assert_allclose(ph, 0.5 * (pd + pm), rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(fh, 0.5 * (fd + fm), rtol=1e-14, atol=1e-300)
        m = float(np.float32(0.99995))
        pg, fg = O.bsdf_eval(O.bsdf(metallic=m, roughness=0.4), R, ns, ns, wo, wi)
        np.testing.assert_allclose(pg, m * pm, rtol=1e-14, atol=1e-300)
        m = float(np.float32(5e-5))
        pg, fg = O.bsdf_eval(O.bsdf(metallic=m, roughness=0.4), R, ns, ns, wo, wi)
        np.testing.assert_allclose(fg, (1 - m) * fd, rtol=1e-14, atol=1e-300)
    # emission-only surfaces reflect nothing
    p, f = O.bsdf_eval(O.bsdf(kind=O.BSDF_EMISSION), R, N, N, N, N)
    assert p == 0 and not f.any()


def test_leak_check_applies_to_every_lobe():
    """SurfaceClosure's check: a direction on the far side of the geometric normal gives zero"""
    ns = _dir(0.4)
    ng = N
    wo = _dir(1.3, np.pi)  # above ng, below ns's tangent plane on the far side
    wi = _dir(0.2)
    assert wo @ ng > 0 and wo @ ns < 0
    for e in (None, O.bsdf(metallic=1.0, roughness=0.5), O.bsdf(metallic=0.3, roughness=0.5)):
        p, f = O.bsdf_eval(e, (0.5, 0.5, 0.5), ns, ng, wo, wi)
        assert p == 0 and not f.any()
