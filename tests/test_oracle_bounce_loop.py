"""Pins the oracle's multi-bounce SamplePixel loop and its random number generator (no GPU).

The loop (CPURenderer.cpp:259-368) is fixed here by cases whose expected value follows from the
reference's rules alone, not from the oracle's code:

* Nested transparent shells. The camera sits at the common centre of K analytic spheres of a
  dielectric with IOR 1 and albedo a_k, inside an emissive sphere of radiance Le. Every camera ray is
  radial, so it meets each shell at normal incidence, where the reflectance (1 - cos)^5 of
  FrDielectric (Utilities.h:301-324) is 0: the ray passes straight through and DieletricBxDF::Sample
  (BxDF.h:112-141) returns the albedo as throughput. A path therefore reaches the emitter after
  exactly K + 1 rays with throughput prod(a_k), and the pixel's expected value is Le * prod(a_k)
  whatever the roulette does (CPURenderer.cpp:350-364), 0 if the bounce limit (:259) stops it
  first, and an invalid pixel (:343-347) if |T|^2 drops below 1e-6 (Types.h:54) on the way. A
  wrong throughput order (:330), emitter weighting (:286-291), early out, bounce count or roulette
  weight changes the answer.
* Russian roulette is unbiased on a scene with Lambertian inter-reflection (the Cornell room): the
  mean with roulette equals the mean of the same estimator with the roulette's outcome ignored
  (an oracle-only switch; both consume the same random numbers).
* xoshiro128+: the state transition equals the published generator's, checked through its published
  jump polynomial (Blackman and Vigna, xoshiro128plus.c: jump() is equivalent to 2^64 calls of
  next()) against the 2^64-th power of the oracle's one-step transition over GF(2).
"""
import math

import numpy as np
import pytest

import sparkle_amd as spk
from sparkle_amd import scenes
from oracle import refcpu

LE = (2.0, 3.0, 4.0)


def shell_scene(albedos, le=LE, emitter_radius=20.0):
    """Camera-centred nested shells: sphere k (radius k + 1) is an IOR-1 dielectric of albedo
    albedos[k]; an emissive sphere encloses them all. A small triangle outside the emitter keeps
    the BVH non-empty; no ray reaches it."""
    sc = spk.Scene()
    far = sc.add_mesh(np.array([[100, 0, 0], [101, 0, 0], [100, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    grey = sc.add_material(base_color=(0.5, 0.5, 0.5), metallic=0.0)
    sc.add_instance(far, grey)
    for k, a in enumerate(albedos):
        m = sc.add_material(base_color=tuple(a), type=spk.MATERIAL_DIELECTRIC, eta=1.0)
        sc.add_sphere((0.0, 0.0, 0.0), float(k + 1), m)
    em = sc.add_material(base_color=(1, 1, 1), emissive=le, metallic=0.0)
    sc.add_sphere((0.0, 0.0, 0.0), emitter_radius, em)
    return sc


def shell_camera(w, h):
    # pinhole at the centre, 100 degree field of view: every ray is radial
    return spk.camera_desc(position=(0.0, 0.0, 0.0), front=(0.3, 1.0, 0.2), vertical_fov=math.radians(100.0),
                           focus_distance=1.0, aperture_radius=0.0, aspect=w / h)


def _product(albedos):
    T = np.ones(3, np.float32)
    for a in albedos:
        T = T * np.asarray(a, np.float32)
    return T


def test_shells_without_roulette_give_the_throughput_product_on_every_pixel():
    albedos = [(0.9, 0.6, 0.3), (0.5, 0.8, 0.7)]
    sc = shell_scene(albedos)
    o = refcpu.OracleScene(sc)
    w, h = 24, 16
    cam = shell_camera(w, h)
    expect = np.asarray(LE, np.float32) * _product(albedos)
    # K = 2 shells need 3 rays: the limit of 3 reaches the emitter (bounce 2 < 3 is never rouletted)
    rgba, nrm = o.render_frame(cam, w, h, frame_seed=5, max_bounce=3)
    np.testing.assert_allclose(rgba[..., :3], np.broadcast_to(expect, (h, w, 3)), rtol=2e-6)
    assert (rgba[..., 3] == 1.0).all()
    # the first-hit normal is the innermost shell's outward normal = the ray direction
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=-1), 1.0, rtol=1e-5)
    # a limit of 2 stops every path on the outer shell: nothing is added, the pixel stays valid
    rgba2, _ = o.render_frame(cam, w, h, frame_seed=5, max_bounce=2)
    assert (rgba2[..., :3] == 0.0).all() and (rgba2[..., 3] == 1.0).all()
    # the shells alone, with the emitter's radiance above the clamp: per channel min(Le * T, 6)
    big = shell_scene(albedos, le=(40.0, 1.0, 9.0))
    rgba3, _ = refcpu.OracleScene(big).render_frame(cam, w, h, frame_seed=1, max_bounce=3)
    expect3 = np.minimum(np.asarray((40.0, 1.0, 9.0), np.float32) * _product(albedos), 6.0)
    np.testing.assert_allclose(rgba3[0, 0, :3], expect3, rtol=2e-6)


@pytest.mark.parametrize("a,early_out", [(0.02, True), (0.03, False)])
def test_early_out_on_squared_throughput_norm(a, early_out):
    """Two grey shells: T = a^2 per channel, |T|^2 = 3 a^4. At a = 0.02 that is 4.8e-7 < 1e-6 and the
    path ends invalid before the emitter; at a = 0.03 it is 2.4e-6 and the emitter is reached. (A
    test on |T| instead of |T|^2 would keep both.)"""
    sc = shell_scene([(a, a, a), (a, a, a)])
    w, h = 8, 8
    rgba, _ = refcpu.OracleScene(sc).render_frame(shell_camera(w, h), w, h, frame_seed=0, max_bounce=6)
    if early_out:
        assert (rgba[..., 3] == -1.0).all() and (rgba[..., :3] == 0.0).all()
    else:
        assert (rgba[..., 3] == 1.0).all()
        np.testing.assert_allclose(rgba[..., :3], np.broadcast_to(np.asarray(LE, np.float32) * np.float32(a * a),
                                                                  (h, w, 3)), rtol=2e-6)


def test_shells_with_roulette_are_unbiased_and_counted():
    """Five shells: the path reaches the emitter at bounce 5, after roulette at bounces 3 and 4
    (survival max(T) clamped to [0.05, 1], survivors divided by it). Per sample the result is 0 or
    Le * prod(a) / (p3 p4); its mean over 64 frames of 64x64 pixels is Le * prod(a) within 3.5 standard
    errors. Without the roulette every pixel is Le * prod(a); with a limit of 5 rays it is 0."""
    albedos = [(0.9, 0.7, 0.5), (0.6, 0.9, 0.8), (0.8, 0.5, 0.9), (0.7, 0.8, 0.6), (0.5, 0.6, 0.9)]
    sc = shell_scene(albedos)
    o = refcpu.OracleScene(sc)
    w = h = 64
    cam = shell_camera(w, h)
    expect = np.asarray(LE, np.float64) * _product(albedos).astype(np.float64)
    off, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=8, roulette=False)
    np.testing.assert_allclose(off[..., :3], np.broadcast_to(expect, (h, w, 3)), rtol=5e-6)
    short, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=5)
    assert (short[..., :3] == 0.0).all()
    frames = 64
    samples = np.concatenate([o.render_frame(cam, w, h, frame_seed=f, max_bounce=8)[0][..., :3].reshape(-1, 3)
                              for f in range(frames)]).astype(np.float64)
    survived = (samples[:, 0] > 0).mean()
    # both roulette draws survive with probability p3 * p4 = max(T after 4 shells) * max(T after 5 / p3)
    t4 = _product(albedos[:4]).astype(np.float64)
    p3 = min(max(t4.max(), 0.05), 1.0)
    t5 = t4 / p3 * np.asarray(albedos[4])
    p4 = min(max(t5.max(), 0.05), 1.0)
    n = samples.shape[0]
    assert abs(survived - p3 * p4) < 4 * math.sqrt(p3 * p4 * (1 - p3 * p4) / n)
    mean = samples.mean(0)
    se = samples.std(0, ddof=1) / math.sqrt(n)
    assert (np.abs(mean - expect) < 3.5 * se).all(), (mean, expect, se)


def test_roulette_is_unbiased_on_interreflection():
    """Cornell room, 10 bounces: paired per-sample difference between the estimator with roulette and
    the same paths with the roulette's outcome ignored has mean 0 within 3.5 standard errors on the
    image mean and on 99% of 8x8 pixel blocks, and roulette ended paths (it is active)."""
    sc = scenes.cornell_scene(0)
    o = refcpu.OracleScene(sc)
    w = h = 32
    cam = scenes.cornell_camera(w, h)
    frames = 96
    d = np.zeros((frames, h, w), np.float64)
    fewer_rays = 0
    for f in range(frames):
        on, _ = o.render_frame(cam, w, h, frame_seed=f, max_bounce=10, output_limit=1e30)
        rays_on = o.last_render_rays()
        off, _ = o.render_frame(cam, w, h, frame_seed=f, max_bounce=10, output_limit=1e30, roulette=False)
        rays_off = o.last_render_rays()
        fewer_rays += rays_off - rays_on
        d[f] = (on[..., :3] - off[..., :3]).astype(np.float64).mean(-1)
    assert fewer_rays > 0.02 * frames * w * h  # roulette ended paths early
    x = d.reshape(-1)
    z = x.mean() / (x.std(ddof=1) / math.sqrt(x.size))
    assert abs(z) < 3.5, z
    blocks = d.reshape(frames, h // 8, 8, w // 8, 8).mean((2, 4)).reshape(frames, -1)
    zb = blocks.mean(0) / (blocks.std(0, ddof=1) / math.sqrt(frames) + 1e-12)
    assert (np.abs(zb) < 3.5).mean() >= 0.95


# ---- xoshiro128+ transition against the published jump polynomial --------------------------
JUMP_2_64 = (0x8764000B, 0xF542D2D3, 0x6FA035C3, 0x77F2DB5B)  # xoshiro128plus.c, jump()


def _to_int(st):
    return int(st[0]) | (int(st[1]) << 32) | (int(st[2]) << 64) | (int(st[3]) << 96)


def _to_state(v):
    return np.array([(v >> (32 * k)) & 0xFFFFFFFF for k in range(4)], dtype=np.uint32)


def _apply(cols, v):
    r = 0
    b = 0
    while v:
        if v & 1:
            r ^= cols[b]
        v >>= 1
        b += 1
    return r


def test_xoshiro_transition_matches_published_jump_polynomial():
    # one step of the oracle's generator as a 128x128 matrix over GF(2) (columns = images of unit states)
    step = [_to_int(refcpu.xoshiro_advance(_to_state(1 << b), 1)) for b in range(128)]
    cols = step
    for _ in range(64):  # M^(2^64) by repeated squaring
        cols = [_apply(cols, c) for c in cols]
    for seed_state in ([0x12345678, 0x9ABCDEF0, 0x0FEDCBA9, 0x87654321], [1, 0, 0, 0]):
        st = np.array(seed_state, dtype=np.uint32)
        acc = 0
        cur = st
        for word in JUMP_2_64:
            for b in range(32):
                if (word >> b) & 1:
                    acc ^= _to_int(cur)
                cur = refcpu.xoshiro_advance(cur, 1)
        assert acc == _apply(cols, _to_int(st))
    # the jump differs from fewer or more steps: M^(2^64) is not M^(2^63) or M^(2^65) here
    half = step
    for _ in range(63):
        half = [_apply(half, c) for c in half]
    assert [_apply(half, c) for c in half] == cols and half != cols


def test_xoshiro_plus_output_is_first_plus_last_word():
    """The '+' scrambler of xoshiro128+ returns s[0] + s[3] (mod 2^32) of the state before the step,
    and a seeded stream is the raw-state stream from the four low SplitMix64 words of the seed."""
    st = np.array([0xDEADBEEF, 0x01234567, 0xCAFEBABE, 0xFFFFFFFF], dtype=np.uint32)
    cur = st
    for _ in range(100):
        nxt, out = refcpu.xoshiro_advance(cur, 1, outputs=True)
        assert int(out[0]) == (int(cur[0]) + int(cur[3])) & 0xFFFFFFFF
        cur = nxt
    seeded = refcpu.xoshiro_stream(42, 16)
    x = 42
    words = []
    for _ in range(4):
        x = (x + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        words.append((z ^ (z >> 31)) & 0xFFFFFFFF)
    _, raw = refcpu.xoshiro_advance(np.array(words, np.uint32), 16, outputs=True)
    np.testing.assert_array_equal(seeded, raw)


# ---- debug modes that rewrite the colour after the loop (CPURenderer.cpp:370-391) -------------
PALETTE = np.array([[137, 49, 239], [242, 202, 25], [255, 0, 189], [0, 87, 233], [135, 233, 17], [225, 24, 69],
                    [255, 255, 255]], np.float32) / np.float32(255.0)  # VisualizeInteger (Utilities.h:93-112)


@pytest.mark.parametrize("k_shells,max_bounce,end", [(0, 4, 0), (1, 4, 1), (2, 4, 2), (3, 8, 3), (2, 2, 2),
                                                     (5, 8, 5), (6, 7, 6), (7, 7, 7), (7, 9, 7)])
def test_ray_depth_is_the_segment_the_path_ended_in(k_shells, max_bounce, end):
    """K shells put the emitter at segment K; a limit below K + 1 ends the loop at max_bounce. RayDepth
    shows VisualizeInteger of that segment, black from 7 on. Roulette is on from segment 3: with
    albedo 1 it never ends a path (survival max(T) = 1)."""
    sc = shell_scene([(1.0, 1.0, 1.0)] * k_shells)
    w, h = 6, 4
    rgba, _ = refcpu.OracleScene(sc).render_frame(shell_camera(w, h), w, h, frame_seed=2, max_bounce=max_bounce,
                                                  debug_mode=spk.DEBUG_RAY_DEPTH)
    expect = PALETTE[end] if end < 7 else np.zeros(3, np.float32)
    np.testing.assert_array_equal(rgba[..., :3], np.broadcast_to(expect, (h, w, 3)))
    assert (rgba[..., 3] == 1.0).all()


@pytest.mark.parametrize("k_shells", [0, 1, 2, 3])
def test_indirect_and_directional_lighting_split_at_segment_one(k_shells):
    """IndirectLighting keeps paths that ended after segment 1, DirectionalLighting those that ended
    at segment 0 or 1; together they add up to the colour. Debug mode 1 is black everywhere."""
    albedos = [(0.9, 0.8, 0.7)] * k_shells
    sc = shell_scene(albedos)
    o = refcpu.OracleScene(sc)
    w, h = 6, 4
    cam = shell_camera(w, h)
    color, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=6)
    indirect, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=6, debug_mode=spk.DEBUG_INDIRECT_LIGHTING)
    direct, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=6, debug_mode=spk.DEBUG_DIRECTIONAL_LIGHTING)
    expect = np.broadcast_to(np.asarray(LE, np.float32) * _product(albedos), (h, w, 3))
    np.testing.assert_allclose(color[..., :3], expect, rtol=2e-6)
    if k_shells <= 1:
        assert (indirect[..., :3] == 0).all()
        np.testing.assert_array_equal(direct[..., :3], color[..., :3])
    else:
        assert (direct[..., :3] == 0).all()
        np.testing.assert_array_equal(indirect[..., :3], color[..., :3])
    dbg, _ = o.render_frame(cam, w, h, frame_seed=0, max_bounce=6, debug_mode=spk.DEBUG_DEBUG)
    assert (dbg[..., :3] == 0).all() and (dbg[..., 3] == 1).all()
