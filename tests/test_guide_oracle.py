"""The learned guide's CPU oracle, pinned by mathematics rather than by its
own formulas:
- its gradients (through the loss on the lobes alone and on the one-sample
  MIS density with the selection probability, the four layers and the
  hash grid) equal central finite differences of its FP64 loss;
- each vMF lobe, and so the mixture, integrates to one over the sphere
  (a fine spherical quadrature), and sampling by Jakob's inversion draws
  directions in proportion to the density (binned counts within 5 sigma);
- the hash grid's dense levels are a bijection of their corners, the
  resolutions run 16 .. 2048.
"""
import numpy as np
import pytest

from oracle import guide_ref as R


def _guide(seed, grid_scale=0.05, max_res=2048, learned=True):
    rng = np.random.default_rng(seed)
    grid = (rng.standard_normal((R.LEVELS, R.TABLE, R.FEATURES)) * grid_scale).astype(np.float64)
    mlp = rng.uniform(-0.3, 0.3, (R.LAYERS, R.WIDTH, R.WIDTH))
    return R.Guide(grid, mlp, lower=[-2, -2, -1], size=[4, 4, 4], max_res=max_res, cap=0.9, learned=learned)


def _unit(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _samples(rng, n):
    S = np.zeros((n, 16))
    S[:, 0:3] = rng.uniform([-2, -2, -1], [2, 2, 3], (n, 3))
    S[:, 3:6] = _unit(rng, n)
    S[:, 6:9] = _unit(rng, n)
    S[:, 9] = rng.uniform(380, 780, n)
    S[:, 10:13] = _unit(rng, n)
    S[:, 13] = rng.uniform(0.1, 2.0, n)
    # half with the BSDF's density (the loss on the MIS density, the
    # selection trained too), half without (the loss on the lobes alone)
    S[:, 14] = np.where(np.arange(n) % 2 == 0, rng.uniform(0.05, 0.6, n), -1.0)
    return S


def _fib(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def test_resolutions_and_dense_levels():
    res, dense = R.resolutions(2048)
    assert res[0] == 16 and res[-1] == 2048 and res == sorted(res)
    assert R.resolutions()[0][-1] == 128
    for level in range(R.LEVELS):
        side = res[level] + 1
        if dense >> level & 1:
            assert side ** 3 <= R.TABLE
        else:
            assert side ** 3 > R.TABLE
    # a dense level maps its corners one to one
    g = _guide(1)
    g.res, g.dense = res, dense
    level = 0
    n = res[level]
    pts = np.stack(np.meshgrid(*[np.arange(n + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = np.minimum(pts, n - 1) / n
    idx, _ = R._corners(g, level, p.astype(np.float64), np.float64)
    assert len(np.unique(idx)) == (n + 1) ** 3


def test_loss_gradient_equals_finite_differences_of_the_outputs():
    rng = np.random.default_rng(3)
    O = rng.standard_normal((6, R.WIDTH))
    w_in = _unit(rng, 6)
    wgt = rng.uniform(0.5, 2, 6)
    bsdf = np.array([0.3, -1, 0.05, -1, 1.2, 0.0])
    loss, G = R.loss_gradient(O, w_in, wgt, bsdf, cap=0.9)
    h = 1e-6
    for b in range(6):
        for c in list(range(R.OUTPUTS)) + [R.OUTPUTS + 3]:
            Op, Om = O.copy(), O.copy()
            Op[b, c] += h
            Om[b, c] -= h
            fd = (R.loss_gradient(Op, w_in, wgt, bsdf, cap=0.9)[0][b] - R.loss_gradient(Om, w_in, wgt, bsdf, cap=0.9)[0][b]) / (2 * h)
            assert abs(fd - G[b, c]) <= 1e-6 + 1e-5 * abs(fd), (b, c, fd, G[b, c])


@pytest.mark.parametrize("learned", [True, False])
def test_network_gradients_equal_finite_differences(learned):
    g = _guide(5, learned=learned)
    rng = np.random.default_rng(7)
    S = _samples(rng, 24)
    loss, grid, dW = R.gradients(g, S, bf=False)
    assert abs(loss - R.loss_of(g, S)) < 1e-9 * max(1, abs(loss))
    h = 1e-6
    # weights of every layer
    for layer in range(R.LAYERS):
        for _ in range(6):
            i, o = rng.integers(0, R.WIDTH, 2)
            if layer == R.LAYERS - 1:
                o = rng.integers(0, R.OUTPUTS)
            keep = g.mlp[layer, i, o]
            g.mlp[layer, i, o] = keep + h
            lp = R.loss_of(g, S)
            g.mlp[layer, i, o] = keep - h
            lm = R.loss_of(g, S)
            g.mlp[layer, i, o] = keep
            fd = (lp - lm) / (2 * h)
            assert abs(fd - dW[layer, i, o]) <= 1e-6 + 1e-4 * abs(fd), (layer, i, o, fd, dW[layer, i, o])
    # grid entries the samples touch, coarse and fine levels
    keys = sorted(grid)
    for key in [keys[j] for j in rng.choice(len(keys), 12, replace=False)]:
        level, entry = key
        for f in range(2):
            keep = g.grid[level, entry, f]
            g.grid[level, entry, f] = keep + h
            lp = R.loss_of(g, S)
            g.grid[level, entry, f] = keep - h
            lm = R.loss_of(g, S)
            g.grid[level, entry, f] = keep
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grid[key][f]) <= 1e-6 + 1e-4 * abs(fd), (key, f, fd, grid[key][f])


@pytest.mark.parametrize("log_sharp", [-3.0, 0.0, 2.0, 6.2])
def test_lobes_integrate_to_one(log_sharp):
    dirs = _fib(400000)
    rng = np.random.default_rng(11)
    O = np.zeros((1, R.WIDTH))
    O[0, 0:5 * R.LOBES:5] = rng.standard_normal(R.LOBES)
    for k in range(R.LOBES):
        O[0, 5 * k + 1:5 * k + 4] = rng.standard_normal(3)
    O[0, 4:5 * R.LOBES:5] = log_sharp
    share, axis, sharp = R.decode(O)
    assert abs(share.sum() - 1) < 1e-12
    q = R.mixture_pdf(share, axis, sharp, dirs[None])[0]
    assert abs(q.mean() * 4 * np.pi - 1) < 2e-3


def test_sampling_follows_the_density():
    rng = np.random.default_rng(13)
    O = np.zeros((1, R.WIDTH))
    O[0, 0:5 * R.LOBES:5] = rng.standard_normal(R.LOBES)
    for k in range(R.LOBES):
        O[0, 5 * k + 1:5 * k + 4] = rng.standard_normal(3)
        O[0, 5 * k + 4] = rng.uniform(0, 4)
    share, axis, sharp = (a[0] for a in R.decode(O))
    n = 200000
    w = R.mixture_sample(share, axis, sharp, rng.random(n), rng.random(n), rng.random(n))
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-9)
    # bins: the nearest of 48 directions; expected shares by quadrature
    centres = _fib(48)
    quad = _fib(400000)
    qv = R.mixture_pdf(share[None], axis[None], sharp[None], quad[None])[0]
    expected = np.bincount(np.argmax(quad @ centres.T, axis=1), weights=qv, minlength=48)
    expected /= expected.sum()
    counts = np.bincount(np.argmax(w @ centres.T, axis=1), minlength=48)
    sigma = np.sqrt(n * expected * (1 - expected)) + 1
    assert np.all(np.abs(counts - n * expected) < 5 * sigma)
