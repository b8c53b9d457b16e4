"""Pins the FP64 cache oracle (no GPU): hash indexing, trilinear interpolation, gradients by
central differences, and Adam's first steps in closed form."""
import numpy as np

from oracle import refcpu


def test_hash_index_dense_and_hashed_levels():
    cfg = refcpu.cache_cfg(log2_table=12, base_resolution=4, per_level_scale=2.0)
    res = refcpu.level_resolutions(cfg)
    assert res[:3] == [4, 8, 16]
    # level 0 is dense: (4+1)^3 <= 4096 -> x + 5 (y + 5 z)
    assert refcpu.hashgrid_index(cfg, 0, 1, 2, 3) == 1 + 5 * (2 + 5 * 3)
    # level 3 (res 32) is hashed with the Instant NGP primes
    h = (5 * 1) ^ ((7 * 2654435761) & 0xFFFFFFFF) ^ ((9 * 805459861) & 0xFFFFFFFF)
    assert refcpu.hashgrid_index(cfg, 3, 5, 7, 9) == (h & 0xFFF)


def test_trilinear_weights_sum_to_one_and_reproduce_linear_fields():
    cfg = refcpu.cache_cfg(log2_table=14, base_resolution=4, per_level_scale=1.5)
    T = 1 << cfg.log2_table
    table = np.full(cfg.levels * T * 2, 0.75, dtype=np.float32)
    rng = np.random.Generator(np.random.PCG64(0))
    pts = rng.random((200, 3)).astype(np.float32)
    f = refcpu.hashgrid_encode(cfg, table, pts)
    np.testing.assert_allclose(f, 0.75, rtol=1e-12)
    # a dense level holding f(x,y,z) = x at its grid points interpolates x exactly
    res0 = refcpu.level_resolutions(cfg)[0]
    table = np.zeros(cfg.levels * T * 2, dtype=np.float32)
    for z in range(res0 + 1):
        for y in range(res0 + 1):
            for x in range(res0 + 1):
                table[refcpu.hashgrid_index(cfg, 0, x, y, z) * 2] = x / res0
    f = refcpu.hashgrid_encode(cfg, table, pts)
    np.testing.assert_allclose(f[:, 0], pts[:, 0].astype(np.float64), atol=1e-6)


def test_mlp_gradients_match_central_differences():
    cfg = refcpu.cache_cfg()
    rng = np.random.Generator(np.random.PCG64(1))
    nw = 32 * 64 + 3 * 64 * 64 + 64 * 3
    w = rng.normal(0, 0.2, nw)
    x = rng.random((5, 32))
    t = rng.random((5, 3))
    loss, out, dw, dx = refcpu.mlp_loss_grad(cfg, w, x, t)
    assert loss > 0 and out.shape == (5, 3)

    def loss_at(wv, xv):
        # the relative-L2 denominator is held fixed at the base point (stop-gradient)
        _, o, _, _ = refcpu.mlp_loss_grad(cfg, wv, xv, t)
        den = out ** 2 + 0.01
        return float(np.sum((o - t) ** 2 / den) / (5 * 3))

    h = 1e-6
    for k in rng.choice(nw, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (loss_at(wp, x) - loss_at(wm, x)) / (2 * h)
        assert abs(fd - dw[k]) <= 1e-6 + 1e-4 * abs(dw[k])
    for (i, j) in [(0, 0), (2, 17), (4, 31)]:
        xp, xm = x.copy(), x.copy()
        xp[i, j] += h
        xm[i, j] -= h
        fd = (loss_at(w, xp) - loss_at(w, xm)) / (2 * h)
        assert abs(fd - dx[i, j]) <= 1e-6 + 1e-4 * abs(dx[i, j])


def test_adam_first_steps_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.01, 0.0])
    m, v = np.zeros(3), np.zeros(3)
    refcpu.adam(p, g, m, v, 0.1, 0.9, 0.99, 1e-8, 1)
    # after one step m_hat = g, v_hat = g^2: the update is lr * g / (|g| + eps)
    np.testing.assert_allclose(p, [1.0 - 0.1 * 0.3 / (0.3 + 1e-8), -2.0 + 0.1 * 0.01 / (0.01 + 1e-8), 0.5])
    refcpu.adam(p, g, m, v, 0.1, 0.9, 0.99, 1e-8, 2)
    mh = (0.9 * 0.1 * g + 0.1 * g) / (1 - 0.81)
    vh = (0.99 * 0.01 * g * g + 0.01 * g * g) / (1 - 0.99 ** 2)
    exp = np.array([1.0 - 0.1 * 0.3 / (0.3 + 1e-8), -2.0 + 0.1 * 0.01 / (0.01 + 1e-8), 0.5]) - 0.1 * mh / (
        np.sqrt(vh) + 1e-8)
    np.testing.assert_allclose(p, exp)


def test_table_backward_is_the_adjoint_of_encode():
    cfg = refcpu.cache_cfg(log2_table=10, base_resolution=2, per_level_scale=1.5)
    T = 1 << cfg.log2_table
    rng = np.random.Generator(np.random.PCG64(2))
    pts = rng.random((6, 3)).astype(np.float32)
    table = rng.normal(size=cfg.levels * T * 2).astype(np.float32)
    d = rng.normal(size=(6, 32))
    g = refcpu.table_backward(cfg, pts, d)
    # <encode(table), d> is linear in the table: its gradient is g
    e = rng.normal(size=table.size).astype(np.float32)
    lhs = np.sum(refcpu.hashgrid_encode(cfg, e, pts) * d)
    assert abs(lhs - float(np.dot(g, e.astype(np.float64)))) < 1e-6 * (1 + abs(lhs))
