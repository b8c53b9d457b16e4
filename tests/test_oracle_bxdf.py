"""The oracle's BxDF sampling (BxDF.h, Sampler.h, Utilities.h), pinned by closed forms on the CPU.

Pins: the shading frame of SampleSurface (GetLocalAxisFromNormal, Utilities.h:236-244) is a proper
orthonormal frame with w = n and the local <-> world round trip exact to rounding; cosine-weighted
hemisphere sampling (Sampler.h:79-95) has E[z] = 2/3, E[z^2] = 1/2 and z^2 uniform; the GGX visible
normals (Sampler.h:133-165) at normal incidence are the GGX NDF sampled by Walter et al. 2007, whose
tan^2(theta) / (alpha^2 + tan^2(theta)) is uniform with alpha = roughness^2, with a uniform azimuth,
and at any incidence face the viewer; the dielectric lobe (BxDF.h:112-141) obeys Snell's law,
reflects with the normal-incidence reflectance ((1 - eta) / (1 + eta))^2 and totally internally
reflects past the critical angle; the specular lobe's throughput for a white metal is at most 1
(and, as in the reference, which tests |cos|, a few of its reflections point below the horizon).
"""
import math

import numpy as np
import pytest

from oracle import refcpu

N = 400_000


def chi2_uniform(x, bins=20):
    """chi-square statistic of samples in [0, 1) against the uniform distribution."""
    h, _ = np.histogram(x, bins=bins, range=(0.0, 1.0))
    e = len(x) / bins
    return float(((h - e) ** 2 / e).sum())


CHI2_19_P1E4 = 50.0  # the 1e-4 upper quantile of chi-square with 19 degrees of freedom is ~49.4


def test_shading_frame_is_orthonormal_and_round_trips():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t = rng.normal(size=3)
        u, v, w = (x.astype(np.float64) for x in refcpu.local_axes(n.astype(np.float32), t.astype(np.float32)))
        B = np.stack([u, v, w])
        np.testing.assert_allclose(B @ B.T, np.eye(3), atol=2e-6)
        np.testing.assert_allclose(w, n, atol=1e-6)
        np.testing.assert_allclose(np.cross(u, v), w, atol=2e-6)
        d = rng.normal(size=3)
        np.testing.assert_allclose(B.T @ (B @ d), d, atol=1e-5)
    # a mistake the frame must not make: u along the tangent's component in the plane
    u, v, w = refcpu.local_axes(np.array([0, 0, 1], np.float32), np.array([1, 0, 0], np.float32))
    np.testing.assert_allclose(np.abs(u), [0, 1, 0], atol=1e-7)  # n x major


def test_lambert_cosine_hemisphere_moments():
    wi, thr, valid = refcpu.bxdf_sample(refcpu.BXDF_LAMBERT, (0, 0, 1), N, seed=11, base=(0.3, 0.5, 0.7))
    assert valid.all()
    np.testing.assert_allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-5)
    z = wi[:, 2].astype(np.float64)
    assert (z >= 0).all()
    assert abs(z.mean() - 2 / 3) < 3e-3
    assert abs((z * z).mean() - 0.5) < 3e-3
    assert chi2_uniform(z * z) < CHI2_19_P1E4
    np.testing.assert_array_equal(thr, np.tile(np.float32([0.3, 0.5, 0.7]), (N, 1)))


@pytest.mark.parametrize("roughness", [0.3, 0.7])
def test_vndf_at_normal_incidence_is_the_ggx_ndf(roughness):
    m, _, _ = refcpu.bxdf_sample(refcpu.BXDF_VNDF, (0, 0, 1), N, seed=5, roughness=roughness)
    m = m.astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-5)
    alpha = roughness * roughness
    tan2 = (m[:, 0] ** 2 + m[:, 1] ** 2) / np.maximum(m[:, 2] ** 2, 1e-30)
    xi = tan2 / (alpha * alpha + tan2)
    assert chi2_uniform(xi) < CHI2_19_P1E4
    phi = (np.arctan2(m[:, 1], m[:, 0]) / (2 * math.pi)) % 1.0
    assert chi2_uniform(phi) < CHI2_19_P1E4
    # the same statistic with alpha = roughness (a dropped square) is far from uniform
    xi_wrong = tan2 / (roughness * roughness + tan2)
    assert chi2_uniform(xi_wrong) > 10 * CHI2_19_P1E4


def test_vndf_faces_the_viewer_at_oblique_incidence():
    wo = np.float32([math.sin(1.2), 0.0, math.cos(1.2)])
    m, _, _ = refcpu.bxdf_sample(refcpu.BXDF_VNDF, wo, 100_000, seed=6, roughness=0.6)
    assert (m @ wo >= -1e-6).all() and (m[:, 2] >= 0).all()
    assert abs(m[:, 1].mean()) < 5e-3  # symmetric about the plane of incidence


def test_dielectric_normal_incidence_reflectance():
    wi, thr, valid = refcpu.bxdf_sample(refcpu.BXDF_DIELECTRIC, (0, 0, 1), N, seed=7, eta=1.5, base=(1, 1, 1))
    assert valid.all()
    refl = wi[:, 2] > 0
    assert abs(refl.mean() - ((1 - 1.5) / (1 + 1.5)) ** 2) < 3e-3
    np.testing.assert_allclose(wi[refl], np.tile([0, 0, 1], (refl.sum(), 1)), atol=1e-6)
    np.testing.assert_allclose(wi[~refl], np.tile([0, 0, -1], ((~refl).sum(), 1)), atol=1e-6)
    np.testing.assert_array_equal(thr, 1.0)


@pytest.mark.parametrize("sin_i,inside", [(0.6, False), (0.95, False), (0.5, True)])
def test_dielectric_refraction_obeys_snell(sin_i, inside):
    cos_i = math.sqrt(1 - sin_i * sin_i)
    wo = np.float32([sin_i, 0.0, -cos_i if inside else cos_i])
    wi, _, _ = refcpu.bxdf_sample(refcpu.BXDF_DIELECTRIC, wo, 50_000, seed=8, eta=1.5)
    wi = wi.astype(np.float64)
    refr = np.sign(wi[:, 2]) != np.sign(wo[2])
    assert refr.any() and (~refr).any()
    sin_t = np.hypot(wi[refr, 0], wi[refr, 1])
    ratio = 1.5 if inside else 1 / 1.5  # eta_i / eta_t
    np.testing.assert_allclose(sin_t, ratio * sin_i, atol=2e-6)
    assert (wi[refr, 0] < 0).all()  # continues across the normal
    # reflections are the mirror direction
    np.testing.assert_allclose(wi[~refr], np.tile([-wo[0], -wo[1], wo[2]], ((~refr).sum(), 1)), atol=1e-6)


def test_dielectric_total_internal_reflection_past_the_critical_angle():
    for sin_i in (0.7, 0.8, 0.99):  # critical sine 1 / 1.5 = 0.667 from inside
        wo = np.float32([sin_i, 0.0, -math.sqrt(1 - sin_i * sin_i)])
        wi, _, _ = refcpu.bxdf_sample(refcpu.BXDF_DIELECTRIC, wo, 20_000, seed=9, eta=1.5)
        np.testing.assert_allclose(wi, np.tile([-wo[0], -wo[1], wo[2]], (20_000, 1)), atol=1e-6)


def test_specular_white_metal_throughput_is_at_most_one():
    wo = np.float32([math.sin(0.7), 0.0, math.cos(0.7)])
    wi, thr, valid = refcpu.bxdf_sample(refcpu.BXDF_SPECULAR, wo, 100_000, seed=10, base=(1, 1, 1), roughness=0.4,
                                        metallic=1.0)
    assert valid.mean() > 0.9
    # the reference takes cosines as |z| (SaturatedCosTheta, Utilities.h:160-163), so a reflection
    # below the horizon still passes its > kEps checks; a small fraction at this roughness
    assert (np.abs(wi[valid, 2]) > 1e-6).all()
    assert (wi[valid, 2] < 0).mean() < 0.05
    np.testing.assert_allclose(np.linalg.norm(wi[valid], axis=1), 1.0, atol=1e-5)
    t = thr[valid]
    assert (t[:, 0] == t[:, 1]).all() and (t[:, 1] == t[:, 2]).all()
    assert (t <= 1.0 + 1e-5).all() and (t > 0).all()
