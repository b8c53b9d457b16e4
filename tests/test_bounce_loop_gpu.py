"""The device bounce loop on cases whose answer is fixed in closed form, and the dielectric lobe
against the oracle (MI355X).

* Nested IOR-1 shells around a pinhole camera inside an emissive sphere: every pixel is
  Le * prod(albedo) when the bounce limit lets the path reach the emitter, 0 when it does not, an
  invalid pixel when |T|^2 < 1e-6 first (CPURenderer.cpp:259-368); with Russian roulette the mean
  is the same product within 3.5 standard errors and each sample equals the oracle's.
* Glass (DieletricBxDF::Sample, BxDF.h:112-141, with FrDielectric and Refract, Utilities.h:211-220,
  301-324): meshes and analytic spheres of IOR 1.5, entering, exiting and total internal
  reflection, against the oracle per pixel within 3 sigma of the Monte Carlo error, with at least
  80% of the sample paths identical.
"""
import math

import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu
from tests.conftest import print_case
from tests.test_oracle_bounce_loop import LE, _product, shell_camera, shell_scene
from tests.test_render_gpu import Renderer

pytestmark = pytest.mark.gpu


def test_shells_closed_form_on_the_device(gpu_ctx):
    w, h = 48, 32
    cam = shell_camera(w, h)
    albedos = [(0.9, 0.6, 0.3), (0.5, 0.8, 0.7)]
    r = Renderer(gpu_ctx, shell_scene(albedos), cam, w, h)
    print_case("shells K=2")
    g, _, _ = r.render(5, bounce=3)
    expect = np.asarray(LE, np.float32) * _product(albedos)
    np.testing.assert_allclose(g[..., :3], np.broadcast_to(expect, (h, w, 3)), rtol=2e-6)
    assert (g[..., 3] == 1.0).all()
    g2, _, _ = r.render(5, bounce=2)
    assert (g2[..., :3] == 0.0).all() and (g2[..., 3] == 1.0).all()
    for a, early in ((0.02, True), (0.03, False)):
        print_case("shells early out", a)
        r = Renderer(gpu_ctx, shell_scene([(a, a, a), (a, a, a)]), cam, w, h)
        g, _, _ = r.render(0, bounce=6)
        if early:
            assert (g[..., 3] == -1.0).all() and (g[..., :3] == 0.0).all()
        else:
            assert (g[..., 3] == 1.0).all()
            np.testing.assert_allclose(g[..., :3], np.broadcast_to(np.asarray(LE, np.float32) * np.float32(a * a),
                                                                   (h, w, 3)), rtol=2e-6)


def test_shells_with_roulette_match_oracle_and_closed_form(gpu_ctx):
    albedos = [(0.9, 0.7, 0.5), (0.6, 0.9, 0.8), (0.8, 0.5, 0.9), (0.7, 0.8, 0.6), (0.5, 0.6, 0.9)]
    sc = shell_scene(albedos)
    w = h = 64
    cam = shell_camera(w, h)
    r = Renderer(gpu_ctx, sc, cam, w, h)
    o = refcpu.OracleScene(sc)
    frames = 32
    samples = []
    same = 0
    for f in range(frames):
        print_case("shells K=5 frame", f)
        g, _, _ = r.render(f, bounce=8)
        c, _ = o.render_frame(cam, w, h, frame_seed=f, max_bounce=8)
        same += int((np.abs(g[..., :3] - c[..., :3]).max(-1) <= 1e-5 * np.abs(c[..., :3]).max(-1) + 1e-7).sum())
        samples.append(g[..., :3].reshape(-1, 3).astype(np.float64))
    assert same >= 0.999 * frames * w * h, same / (frames * w * h)
    s = np.concatenate(samples)
    expect = np.asarray(LE, np.float64) * _product(albedos).astype(np.float64)
    se = s.std(0, ddof=1) / math.sqrt(s.shape[0])
    assert (np.abs(s.mean(0) - expect) < 3.5 * se).all(), (s.mean(0), expect, se)
    g, _, _ = r.render(0, bounce=5)
    assert (g[..., :3] == 0.0).all()


def glass_cornell(seed=0):
    """The Cornell room with its tall box and the icosphere made of IOR-1.5 glass, plus an analytic
    glass sphere on the floor (tinted), so rays enter and leave glass meshes and spheres and meet
    total internal reflection inside the box and the spheres."""
    sc = scenes.cornell_scene(seed)
    glass = sc.add_material(base_color=(1.0, 1.0, 1.0), eta=1.5, type=spk.MATERIAL_DIELECTRIC)
    tint = sc.add_material(base_color=(0.95, 0.85, 0.7), eta=1.5, type=spk.MATERIAL_DIELECTRIC)
    sc.instances[6]["material"] = glass   # tall box
    sc.instances[8]["material"] = glass   # icosphere
    sc.add_sphere((-0.45, -0.5, 0.25), 0.25, tint)
    return sc


@pytest.mark.parametrize("bounce", [5, 8])
def test_dielectric_matches_oracle(gpu_ctx, bounce):
    w = h = 48
    frames = 48
    sc = glass_cornell()
    cam = scenes.cornell_camera(w, h)
    r = Renderer(gpu_ctx, sc, cam, w, h)
    o = refcpu.OracleScene(sc)
    d = np.zeros((frames, h, w, 3))
    mg = np.zeros((h, w, 3))
    mc = np.zeros((h, w, 3))
    for k in range(frames):
        print_case("glass cornell frame", k, "bounce", bounce)
        gg, _, _ = r.render(k, bounce=bounce)
        cc, _ = o.render_frame(cam, w, h, frame_seed=k, max_bounce=bounce)
        d[k] = gg[..., :3] - cc[..., :3]
        mg += gg[..., :3] / frames
        mc += cc[..., :3] / frames
    se = d.std(0, ddof=1) / math.sqrt(frames)
    ok = np.abs(d.mean(0)) <= 3 * se + 1e-6
    same = (np.abs(d).max(-1) < 1e-3).mean()
    rel_mse = float(np.mean((mg - mc) ** 2 / (mc ** 2 + 1e-2)))
    # what the relMSE of the two means is expected to be from the divergent paths' noise alone
    floor = float(np.mean(d.var(0, ddof=1) / frames / (mc ** 2 + 1e-2)))
    print(f"glass: within 3 sigma {ok.mean():.4f}, identical paths {same:.4f}, relMSE(gpu mean vs oracle mean) "
          f"{rel_mse:.3g}, Monte Carlo floor {floor:.3g}")
    assert ok.mean() >= 0.99, ok.mean()
    assert same > 0.8, same
    assert rel_mse < 2.0 * floor + 1e-5
    # the glass is in view and carries light: the first-hit albedo of glass pixels is the glass colour
    g, _, _ = r.render(0, debug=spk.DEBUG_ALBEDO)
    c, _ = o.render_frame(cam, w, h, frame_seed=0, debug_mode=spk.DEBUG_ALBEDO)
    assert (np.abs(g[..., :3] - c[..., :3]).max(-1) < 1e-4).mean() > 0.99
    assert (np.abs(c[..., :3] - np.array([0.95, 0.85, 0.7], np.float32)).max(-1) == 0).sum() > 10


# ---- debug modes that rewrite the colour after the loop (CPURenderer.cpp:370-391) -------------
def test_post_loop_debug_modes_closed_form(gpu_ctx):
    """RayDepth is VisualizeInteger of the segment the path ended in, IndirectLighting and
    DirectionalLighting split the colour at segment 1, Debug is black: on the shells, where the
    segment is fixed, every pixel equals the closed form."""
    from tests.test_oracle_bounce_loop import PALETTE

    w, h = 16, 12
    cam = shell_camera(w, h)
    for k_shells, max_bounce, end in ((0, 4, 0), (2, 4, 2), (2, 2, 2), (5, 8, 5), (7, 7, 7)):
        print_case("ray depth shells", k_shells, max_bounce)
        r = Renderer(gpu_ctx, shell_scene([(1.0, 1.0, 1.0)] * k_shells), cam, w, h)
        g, _, _ = r.render(2, bounce=max_bounce, debug=spk.DEBUG_RAY_DEPTH)
        expect = PALETTE[end] if end < 7 else np.zeros(3, np.float32)
        np.testing.assert_array_equal(g[..., :3], np.broadcast_to(expect, (h, w, 3)))
    g, _, _ = r.render(2, bounce=0, debug=spk.DEBUG_RAY_DEPTH)  # the loop never runs: segment 0
    np.testing.assert_array_equal(g[..., :3], np.broadcast_to(PALETTE[0], (h, w, 3)))
    for k_shells in (1, 2):
        albedos = [(0.9, 0.8, 0.7)] * k_shells
        r = Renderer(gpu_ctx, shell_scene(albedos), cam, w, h)
        c, _, _ = r.render(0, bounce=6)
        ind, _, _ = r.render(0, bounce=6, debug=spk.DEBUG_INDIRECT_LIGHTING)
        dire, _, _ = r.render(0, bounce=6, debug=spk.DEBUG_DIRECTIONAL_LIGHTING)
        dbg, _, _ = r.render(0, bounce=6, debug=spk.DEBUG_DEBUG)
        assert (c[..., :3] > 0).all() and (dbg[..., :3] == 0).all()
        kept, dropped = (dire, ind) if k_shells == 1 else (ind, dire)
        np.testing.assert_array_equal(kept, c)
        assert (dropped[..., :3] == 0).all()


@pytest.mark.parametrize("mode", [spk.DEBUG_RAY_DEPTH, spk.DEBUG_INDIRECT_LIGHTING, spk.DEBUG_DIRECTIONAL_LIGHTING,
                                  spk.DEBUG_DEBUG])
def test_post_loop_debug_modes_match_oracle(gpu_ctx, mode):
    """The Cornell room with glass, 8 bounces and roulette: the same paths give the same RayDepth
    colour and the same indirect/direct split as the oracle on at least 97% of pixel samples, and the
    means agree within 3 sigma on 99% of pixels."""
    w = h = 40
    frames = 16
    sc = glass_cornell()
    cam = scenes.cornell_camera(w, h)
    r = Renderer(gpu_ctx, sc, cam, w, h)
    o = refcpu.OracleScene(sc)
    d = np.zeros((frames, h, w, 3))
    for k in range(frames):
        print_case("post-loop debug mode", mode, "frame", k)
        g, _, _ = r.render(k, bounce=8, debug=mode)
        c, _ = o.render_frame(cam, w, h, frame_seed=k, max_bounce=8, debug_mode=mode)
        d[k] = g[..., :3] - c[..., :3]
    same = (np.abs(d).max(-1) < 1e-5).mean()
    print(f"debug mode {mode}: identical samples {same:.4f}")
    assert same >= 0.97, same
    se = d.std(0, ddof=1) / math.sqrt(frames)
    assert (np.abs(d.mean(0)) <= 3 * se + 1e-6).mean() >= 0.99
    with pytest.raises(spk.SpkError) as e:
        r.render(0, bounce=4, debug=spk.DEBUG_RAY_DEPTH, flags=spk.FRAME_NEE)
    assert e.value.code == spk.SPK_ERR_UNSUPPORTED
