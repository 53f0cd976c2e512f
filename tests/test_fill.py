"""Depression filling (an extension: the reference has none; its README
claims depression handling, README.md:35-37).

The oracle's Priority-Flood+epsilon (a heap-based flood from rings 0 and 1)
is pinned by an independent computation of the same surface: the unique
fixed point of F = max(h, nextafter(min over the D8 neighbours of F, +inf))
on interior cells, with rings 0 and 1 held at h, reached here by plain
Jacobi iteration from +inf (a monotone operator, so the iteration decreases
to that fixed point).  Properties checked besides: F >= h, rings unchanged,
idempotence, and every interior cell of F gets a receiver.
"""
import numpy as np
import pytest

from conftest import cone_terrain


def jacobi_fill(h, W, H):
    h = np.asarray(h, dtype=np.float64).reshape(H, W)
    F = h.copy()
    inner = np.zeros((H, W), dtype=bool)
    inner[2:H - 2, 2:W - 2] = True
    F[inner] = np.inf
    for _ in range(W * H + 1):
        P = np.pad(F, 1, constant_values=np.inf)
        m = np.full((H, W), np.inf)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dy or dx:
                    m = np.minimum(m, P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W])
        G = np.where(inner, np.fmax(h, np.nextafter(m, np.inf)), h)
        if np.array_equal(G, F):
            return F.ravel()
        F = G
    raise AssertionError("no fixed point")


def pit_terrain(W, H, seed):
    """Random terrain with a wide basin, a flat plateau and a deep pit."""
    rng = np.random.default_rng(seed)
    h = rng.random((H, W))
    y, x = np.mgrid[0:H, 0:W]
    h += 2.0 * np.exp(-((x - W / 3) ** 2 + (y - H / 2) ** 2) / (W * H / 20))
    h[H // 4:H // 2, W // 2:W // 2 + W // 5] = 1.5
    h[H // 2, 2 * W // 3] = -3.0
    h[:2, :] = h[-2:, :] = 0
    h[:, :2] = h[:, -2:] = 0
    return h.ravel()


CASES = [(5, 5, 0), (17, 13, 1), (13, 17, 2), (64, 48, 3), (40, 40, 4)]


@pytest.mark.parametrize("W,H,seed", CASES)
def test_priority_flood_equals_fixed_point(oracle, W, H, seed):
    for h in (oracle.init_terrain(W, H, seed), pit_terrain(W, H, seed)):
        np.testing.assert_array_equal(oracle.fill_depressions(h, W, H), jacobi_fill(h, W, H))


@pytest.mark.parametrize("W,H,seed", CASES)
def test_filled_surface_properties(oracle, W, H, seed):
    h = pit_terrain(W, H, seed)
    F = oracle.fill_depressions(h, W, H)
    assert np.all(F >= h)
    ring = np.ones((H, W), dtype=bool)
    ring[2:H - 2, 2:W - 2] = False
    np.testing.assert_array_equal(F.reshape(H, W)[ring], h.reshape(H, W)[ring])
    np.testing.assert_array_equal(oracle.fill_depressions(F, W, H), F)
    rec = oracle.receivers(F, W, H, oracle.preset(1)).reshape(H, W)
    assert np.all(rec[2:H - 2, 2:W - 2] >= 0), "an interior cell of the filled surface is a pit"
    # cells with a descending path to the rings keep their elevation
    if W >= 40:
        assert np.count_nonzero(F == h) > 0.3 * W * H


def test_fill_leaves_drained_terrain_alone(oracle):
    """Conical basins draining to their centres, with the centres cut down
    to the base level: nothing to fill."""
    W, H = 60, 50
    h = cone_terrain(W, H, 60, 0)
    np.testing.assert_array_equal(oracle.fill_depressions(h, W, H), jacobi_fill(h, W, H))


def test_fill_without_epsilon_leaves_pits(oracle):
    """Without the one-ulp rise a filled depression is a flat whose cells
    have no receiver: the receiver check of the properties test catches a
    fill that drops the epsilon."""
    W, H = 40, 40
    h = pit_terrain(W, H, 4).reshape(H, W)
    F = h.copy()
    inner = np.zeros((H, W), dtype=bool)
    inner[2:H - 2, 2:W - 2] = True
    F[inner] = np.inf
    for _ in range(W * H):
        P = np.pad(F, 1, constant_values=np.inf)
        m = np.full((H, W), np.inf)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dy or dx:
                    m = np.minimum(m, P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W])
        G = np.where(inner, np.fmax(h, m), h)
        if np.array_equal(G, F):
            break
        F = G
    rec = oracle.receivers(F.ravel(), W, H, oracle.preset(1)).reshape(H, W)
    assert np.any(rec[2:H - 2, 2:W - 2] < 0)
