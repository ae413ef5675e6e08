"""CPU: the surrogate network's oracle (oracle/mlp_oracle.py) pinned by
central finite differences and by the bfloat16 rounding rule."""

import numpy as np

from conftest import mlp_oracle as mo


def _random_net(rng, d, H, scale=0.3):
    D = (d + 63) // 64 * 64
    p = {
        "W1": np.zeros((H, D)), "b1": rng.normal(0, 0.2, H), "W2": rng.normal(0, scale, (H, H)) / 4,
        "b2": rng.normal(0, 0.2, H), "w3": rng.normal(0, scale, H), "b3": np.zeros(64),
        "v": np.zeros(D), "q": np.zeros(D),
    }
    p["W1"][:, :d] = rng.normal(0, scale, (H, d))
    p["b3"][0] = 0.7
    p["v"][:d] = rng.normal(0, 0.3, d)
    p["q"][:d] = rng.normal(0, 0.1, d)
    return np.concatenate([p[k].reshape(-1) for k in mo.BLOCKS]), D


def test_bf16_rounding_is_nearest_even():
    one = np.float32(1.0)
    vals = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -2.5,
                     3.0e-39], dtype=np.float32)
    r = mo.bf16(vals)
    assert r[0] == one
    assert r[1] == one                        # tie -> even (1.0)
    assert r[2] == np.float32(1 + 2.0 ** -6)  # tie -> even (mantissa ...10)
    assert r[3] == np.float32(1 + 2.0 ** -7)  # above the tie -> up
    assert r[4] == np.float32(-2.5)           # exactly representable
    assert r[5] > 0                           # subnormal survives
    # every output has a zero low half
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, H, n = 10, 64, 23
    flat, D = _random_net(rng, d, H)
    X = np.exp(rng.normal(-2.5, 0.6, (n, d)))
    xm, xs = np.log(X).mean(0), np.log(X).std(0)
    logl = rng.normal(-100, 3, n)
    ym, ys = logl.mean(), logl.std()
    loss, g = mo.param_grad(flat, H, D, X, logl, xm, xs, ym, ys)
    offs = np.cumsum([0, H * D, H, H * H, H, H, 64, D, D])
    h = 1e-6
    valid = [[r * D + c for r in range(H) for c in range(d)], range(H), range(H * H), range(H),
             range(H), [0], range(d), range(d)]
    picks = [offs[b] + int(j) for b, idxs in enumerate(valid)
             for j in rng.choice(list(idxs), size=min(4, len(idxs)), replace=False)]
    for i in picks:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp, _ = mo.param_grad(fp, H, D, X, logl, xm, xs, ym, ys)
        lm, _ = mo.param_grad(fm, H, D, X, logl, xm, xs, ym, ys)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # padding columns of W1 / v / q have zero gradient
    assert np.all(g[offs[0]:offs[1]].reshape(H, D)[:, d:] == 0.0)
    assert np.all(g[offs[6] + d:offs[7]] == 0.0) and np.all(g[offs[7] + d:offs[8]] == 0.0)
    assert np.all(g[offs[5] + 1:offs[6]] == 0.0)


def test_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    d, H, n = 12, 64, 5
    flat, D = _random_net(rng, d, H)
    X = np.exp(rng.normal(-2.0, 0.5, (n, d)))
    xm, xs = np.log(X).mean(0), np.log(X).std(0) + 0.1
    ym, ys = -500.0, 7.0
    ll, g = mo.predict(flat, H, D, X, xm, xs, ym, ys)
    for r in range(n):
        for k in range(d):
            h = 1e-6 * X[r, k]
            Xp, Xm = X.copy(), X.copy()
            Xp[r, k] += h
            Xm[r, k] -= h
            fd = (mo.predict(flat, H, D, Xp[r], xm, xs, ym, ys)[0][0]
                  - mo.predict(flat, H, D, Xm[r], xm, xs, ym, ys)[0][0]) / (2 * h)
            assert abs(fd - g[r, k]) <= 2e-5 * max(1.0, abs(g[r, k])), (r, k, fd, g[r, k])


def test_linear_and_quadratic_skip_alone():
    """W1 = W2 = w3 = 0: logL~ = y_sd (b3 + v.z + q.z^2) + y_mean in closed form."""
    rng = np.random.default_rng(2)
    d, H = 7, 64
    flat, D = _random_net(rng, d, H)
    blocks = mo.unflatten(flat, H, D)
    blocks["W1"][:] = 0
    blocks["W2"][:] = 0
    blocks["w3"][:] = 0
    flat = np.concatenate([blocks[k].reshape(-1) for k in mo.BLOCKS])
    X = np.exp(rng.normal(-1.0, 0.4, (3, d)))
    xm, xs = np.zeros(d), np.ones(d)
    ll, g = mo.predict(flat, H, D, X, xm, xs, 10.0, 2.0)
    z = np.log(X)
    v, q = blocks["v"][:d], blocks["q"][:d]
    y = blocks["b3"][0] + z @ v + (z * z) @ q
    assert np.allclose(ll, 2.0 * y + 10.0, rtol=1e-12, atol=1e-9)
    assert np.allclose(g, 2.0 * (v + 2 * q * z) / X, rtol=1e-12)


def test_bf16_mode_is_close_to_exact():
    rng = np.random.default_rng(3)
    d, H, n = 30, 128, 40
    flat, D = _random_net(rng, d, H)
    X = np.exp(rng.normal(-2.5, 0.6, (n, d)))
    xm, xs = np.log(X).mean(0), np.log(X).std(0)
    logl = rng.normal(-100, 3, n)
    l1, g1 = mo.param_grad(flat, H, D, X, logl, xm, xs, logl.mean(), logl.std())
    l2, g2 = mo.param_grad(flat, H, D, X, logl, xm, xs, logl.mean(), logl.std(), mode="bf16")
    assert abs(l1 - l2) < 2e-2 * l1
    assert np.linalg.norm(g1 - g2) < 2e-2 * np.linalg.norm(g1)


def test_gradient_matching_loss_matches_finite_differences():
    rng = np.random.default_rng(4)
    d, H, n = 9, 64, 11
    flat, D = _random_net(rng, d, H)
    X = np.exp(rng.normal(-2.5, 0.6, (n, d)))
    xm, xs = np.log(X).mean(0), np.log(X).std(0)
    logl = rng.normal(-100, 3, n)
    dlogl = rng.normal(0, 20, (n, d))
    ym, ys = logl.mean(), logl.std()
    kw = dict(dlogl=dlogl, grad_weight=3.0)
    loss, g = mo.param_grad(flat, H, D, X, logl, xm, xs, ym, ys, **kw)
    l0, _ = mo.param_grad(flat, H, D, X, logl, xm, xs, ym, ys)
    assert loss > l0                       # the matching term adds a positive amount
    offs = np.cumsum([0, H * D, H, H * H, H, H, 64, D, D])
    valid = [[r * D + c for r in range(H) for c in range(d)], range(H), range(H * H), range(H),
             range(H), [0], range(d), range(d)]
    picks = [offs[b] + int(j) for b, idxs in enumerate(valid)
             for j in rng.choice(list(idxs), size=min(5, len(idxs)), replace=False)]
    h = 1e-6
    for i in picks:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp, _ = mo.param_grad(fp, H, D, X, logl, xm, xs, ym, ys, **kw)
        lm, _ = mo.param_grad(fm, H, D, X, logl, xm, xs, ym, ys, **kw)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # with weight 0 the two paths agree
    l1, g1 = mo.param_grad(flat, H, D, X, logl, xm, xs, ym, ys, dlogl=dlogl, grad_weight=0.0)
    assert l1 == l0
