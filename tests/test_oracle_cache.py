"""FP64 radiance-cache oracle, pinned by finite differences, the published
Instant-NGP level layout, hand-computed hash indices and the Adam closed form."""
import numpy as np
import pytest

MASK19 = (1 << 19) - 1


@pytest.fixture(scope="module")
def problem(oracle):
    rng = np.random.default_rng(0)
    mlp = rng.uniform(-0.3, 0.3, oracle.CACHE_MLP_PARAMS).astype(np.float32)
    grid = rng.uniform(-0.1, 0.1, oracle.CACHE_GRID_PARAMS).astype(np.float32)
    q = np.concatenate([rng.uniform(0, 1, (6, 3)), rng.normal(size=(6, 3))], 1).astype(np.float32)
    t = rng.uniform(0, 1, (6, 3)).astype(np.float32)
    return mlp, grid, q, t


def test_levels_match_instant_ngp(oracle):
    res, mask = oracle.cache_levels()
    b = np.exp((np.log(2048) - np.log(16)) / 15)
    assert list(res) == [int(np.floor(16 * b ** l + 1e-9)) for l in range(16)]
    assert res[0] == 16 and res[-1] == 2048
    assert mask == 0b11111  # (N+1)^3 <= 2^19 for levels 0..4 only
    assert (res[4] + 1) ** 3 <= 2 ** 19 < (res[5] + 1) ** 3


def test_backward_matches_finite_differences(oracle, problem):
    mlp, grid, q, t = problem

    def loss(m, g):
        return oracle.cache_loss_grad(m, g, q, t, [0, 0, 0], 1.0, round_fp16=False, plain_l2=True,
                                      want_grid=False)[0]

    _, _, dm, dg = oracle.cache_loss_grad(mlp, grid, q, t, [0, 0, 0], 1.0, round_fp16=False, plain_l2=True)
    h = 1e-3
    for idx in [7, 2000, 3500, 9000, 15000, 19500, 20470]:
        a, b = mlp.copy(), mlp.copy()
        a[idx] += h
        b[idx] -= h
        fd = (loss(a, grid) - loss(b, grid)) / (2 * h)
        assert abs(fd - dm[idx]) <= 1e-4 * max(1.0, abs(fd)), (idx, fd, dm[idx])
    nz = np.nonzero(dg)[0]
    assert nz.size > 0
    for idx in nz[:: max(1, nz.size // 6)]:
        a, b = grid.copy(), grid.copy()
        a[idx] += h
        b[idx] -= h
        fd = (loss(mlp, a) - loss(mlp, b)) / (2 * h)
        assert abs(fd - dg[idx]) <= 1e-4 * max(1.0, abs(fd)), (idx, fd, dg[idx])


def test_glossy_inputs_reach_the_network(oracle):
    """Outgoing direction, roughness and the glossy flag feed input rows
    40..44: finite differences agree there, and Lambert queries (direction 0,
    roughness 0) leave those rows without gradient."""
    rng = np.random.default_rng(5)
    mlp = rng.uniform(-0.3, 0.3, oracle.CACHE_MLP_PARAMS).astype(np.float32)
    grid = rng.uniform(-0.1, 0.1, oracle.CACHE_GRID_PARAMS).astype(np.float32)
    n = 6
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.normal(size=(n, 3)), d, rng.uniform(0.5, 1, (n, 1))],
                       1).astype(np.float32)
    t = rng.uniform(0, 1, (n, 3)).astype(np.float32)

    def loss(m, qq):
        return oracle.cache_loss_grad(m, grid, qq, t, [0, 0, 0], 1.0, round_fp16=False, plain_l2=True,
                                      want_grid=False)[0]

    _, _, dm, _ = oracle.cache_loss_grad(mlp, grid, q, t, [0, 0, 0], 1.0, round_fp16=False, plain_l2=True,
                                         want_grid=False)
    rows = dm[40 * 64:45 * 64]
    assert np.all(np.abs(rows) > 0) or np.count_nonzero(rows) > 200
    h = 1e-3
    for idx in [40 * 64 + 3, 43 * 64 + 17, 44 * 64 + 60]:
        a, b = mlp.copy(), mlp.copy()
        a[idx] += h
        b[idx] -= h
        fd = (loss(a, q) - loss(b, q)) / (2 * h)
        assert abs(fd - dm[idx]) <= 1e-4 * max(1.0, abs(fd)), (idx, fd, dm[idx])
    lam = q.copy()
    lam[:, 6:] = 0
    _, _, dm0, _ = oracle.cache_loss_grad(mlp, grid, lam, t, [0, 0, 0], 1.0, round_fp16=False, plain_l2=True,
                                          want_grid=False)
    assert not dm0[40 * 64:48 * 64].any()
    assert dm0[32 * 64:40 * 64].any()


def test_relative_l2_gradient_holds_denominator(oracle, problem):
    """Relative L2 differentiates (p - y)^2 / (p0^2 + 0.01) with p0 frozen."""
    mlp, grid, q, t = problem
    _, p0, dm, _ = oracle.cache_loss_grad(mlp, grid, q, t, [0, 0, 0], 1.0, round_fp16=False, want_grid=False)
    den = p0.astype(np.float64) ** 2 + 0.01

    def surrogate(m):
        p = oracle.cache_loss_grad(m, grid, q, t, [0, 0, 0], 1.0, round_fp16=False, want_grid=False)[1]
        return float(np.sum((p.astype(np.float64) - t) ** 2 / den))

    h = 1e-3
    for idx in [11, 5000, 19460, 20001]:
        a, b = mlp.copy(), mlp.copy()
        a[idx] += h
        b[idx] -= h
        fd = (surrogate(a) - surrogate(b)) / (2 * h)
        assert abs(fd - dm[idx]) <= 2e-3 * max(1.0, abs(fd)), (idx, fd, dm[idx])


def test_hash_indices_hand_computed(oracle):
    """A single query at a point of level 5 (N = 80, hashed): the grid
    gradient lands on the 8 corners given by (x ^ y*2654435761 ^ z*805459861) mod 2^19."""
    mlp = np.zeros(oracle.CACHE_MLP_PARAMS, np.float32)
    mlp[10 * 64 + 0] = 1.0     # feature 10 = level 5, feature 0 -> hidden 0
    for l in range(1, 5):
        mlp[3072 + (l - 1) * 4096 + 0] = 1.0   # hidden 0 -> hidden 0
    mlp[19456 + 0] = 1.0       # hidden 0 -> output 0
    grid = np.full(oracle.CACHE_GRID_PARAMS, 0.01, np.float32)
    x = np.array([[0.5 + 0.3 / 80, 0.25 + 0.6 / 80, 0.75 + 0.2 / 80, 0, 0, 1]], np.float32)
    _, _, _, dg = oracle.cache_loss_grad(mlp, grid, x, np.zeros((1, 3), np.float32), [0, 0, 0], 1.0,
                                         round_fp16=False, plain_l2=True)
    level = 5
    base = level * (1 << 19) * 2
    got = set((np.nonzero(dg[base:base + (1 << 20)])[0] // 2).tolist())
    ix, iy, iz = 40, 20, 60
    want = set()
    for c in range(8):
        cx, cy, cz = ix + (c & 1), iy + ((c >> 1) & 1), iz + (c >> 2)
        want.add((cx ^ ((cy * 2654435761) & 0xFFFFFFFF) ^ ((cz * 805459861) & 0xFFFFFFFF)) & MASK19)
    assert got == want
    # a dense level: index x + y (N+1) + z (N+1)^2
    res, _ = oracle.cache_levels()
    mlp2 = np.zeros_like(mlp)
    mlp2[0 * 64 + 0] = 1.0     # level 0, feature 0
    for l in range(1, 5):
        mlp2[3072 + (l - 1) * 4096 + 0] = 1.0
    mlp2[19456] = 1.0
    _, _, _, dg2 = oracle.cache_loss_grad(mlp2, grid, x, np.zeros((1, 3), np.float32), [0, 0, 0], 1.0,
                                          round_fp16=False, plain_l2=True)
    got0 = set((np.nonzero(dg2[: (1 << 20)])[0] // 2).tolist())
    n = res[0]
    sx, sy, sz = np.floor(np.array(x[0, :3], np.float64) * n).astype(int)
    want0 = {(sx + (c & 1)) + (sy + ((c >> 1) & 1)) * (n + 1) + (sz + (c >> 2)) * (n + 1) ** 2 for c in range(8)}
    assert got0 == want0


def test_adam_first_step_is_signed_learning_rate(oracle):
    g = np.array([3.0, -0.25, 1e-3, -7.0], np.float64)
    p = np.array([1.0, 1.0, 1.0, 1.0])
    m = np.zeros(4)
    v = np.zeros(4)
    oracle.adam_step(p, m, v, g, 1e-2, 1)
    np.testing.assert_allclose(p, 1.0 - 1e-2 * np.sign(g), rtol=0, atol=1e-12)
    # second step with the same gradient: the bias-corrected ratio is still 1
    oracle.adam_step(p, m, v, g, 1e-2, 2)
    np.testing.assert_allclose(p, 1.0 - 2e-2 * np.sign(g), rtol=0, atol=1e-12)
