"""CPU check of the premise behind the list walk's cell skip (k_fill_neighbors):
a coordinate x binned to cell coordinate i = floor((x + s) / s) in float32
(neighbors.h:110-111) lies within [(i - 1) s - d, i s + d] with d = s / 1024,
so a cell whose widened extent is at least s (1 + 2^-10) away from a particle
holds no candidate the float32 test d^2 < s^2 accepts.  Checked on random
float32 coordinates over the benchmark boxes, bin spacings of the
configurations, and on the cell faces themselves."""
import numpy as np


def _bin(x, s):
    return np.floor((x + s) / s).astype(np.float32)


def test_binned_coordinates_lie_in_their_widened_cell():
    rng = np.random.default_rng(0)
    for s in (np.float32(0.1), np.float32(0.108), np.float32(0.12), np.float32(0.11), np.float32(0.105)):
        x = rng.uniform(0.0, 26.0, size=2_000_000).astype(np.float32)
        # also every cell face and its float neighbours
        faces = (np.arange(0, 260, dtype=np.float32) * s).astype(np.float32)
        x = np.concatenate([x, faces, np.nextafter(faces, np.float32(-1)), np.nextafter(faces, np.float32(1e9))])
        x = x[x >= 0]
        i = _bin(x, s)
        d = np.float32(s * np.float32(2.0 ** -10))
        lo = ((i - np.float32(1.0)) * s - d).astype(np.float32)
        hi = (i * s + d).astype(np.float32)
        assert np.all(x >= lo) and np.all(x <= hi), s


def test_out_of_reach_cells_hold_no_accepted_candidate():
    """Brute force in float32: pairs whose cell test says out of reach never
    pass the list's acceptance test."""
    rng = np.random.default_rng(1)
    s = np.float32(0.12)
    ss = np.float32(s * s)
    d = np.float32(s * np.float32(2.0 ** -10))
    p = rng.uniform(0.5, 2.0, size=(400_000, 3)).astype(np.float32)
    q = (p + rng.uniform(-2.5 * s, 2.5 * s, size=p.shape).astype(np.float32)).astype(np.float32)
    c = _bin(q, s)
    lo = ((c - np.float32(1.0)) * s - d).astype(np.float32)
    hi = (c * s + d).astype(np.float32)
    gap = np.where(p < lo, lo - p, np.where(p > hi, p - hi, np.float32(0.0))).astype(np.float32)
    far = ((gap[:, 0] * gap[:, 0] + gap[:, 1] * gap[:, 1]) + gap[:, 2] * gap[:, 2]) >= ss * np.float32(1.0 + 2.0 ** -10)
    dd = (p - q).astype(np.float32)
    d2 = ((dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]).astype(np.float32)
    assert far.sum() > 100_000
    assert not np.any(far & (d2 < ss))
