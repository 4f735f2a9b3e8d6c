"""Legalization-shaped auction instances: sites on a W x H grid (`per`
objects per tile), bidders with targets clustered around the die centre,
each bidder's domain every site within Manhattan radius `radius` of its
target at base cost = the distance (the ring costs of
legalization/mod.rs:1128-1154); sigma <= 0 draws the targets uniformly."""
import numpy as np


def grid_instance(seed: int, side: int, per: int, bidders: int, radius: int, sigma: float):
    rng = np.random.default_rng(seed)
    if sigma > 0:
        tx = np.clip(np.rint(side / 2 + rng.normal(0, sigma, bidders)), 0, side - 1).astype(np.int64)
        ty = np.clip(np.rint(side / 2 + rng.normal(0, sigma, bidders)), 0, side - 1).astype(np.int64)
    else:  # uniform targets
        tx = rng.integers(0, side, bidders)
        ty = rng.integers(0, side, bidders)
    offs = [(dx, dy) for dx in range(-radius, radius + 1) for dy in range(-radius, radius + 1)
            if abs(dx) + abs(dy) <= radius]
    dx = np.array([o[0] for o in offs]); dy = np.array([o[1] for o in offs])
    x = tx[:, None] + dx[None, :]; y = ty[:, None] + dy[None, :]
    ok = (x >= 0) & (x < side) & (y >= 0) & (y < side)
    dist = (np.abs(dx) + np.abs(dy))[None, :].repeat(bidders, 0)
    tile = y * side + x
    counts = ok.sum(1) * per
    start = np.zeros(bidders + 1, np.uint32)
    start[1:] = np.cumsum(counts)
    obj = np.concatenate([(tile[ok][:, None] * per + np.arange(per)[None, :]).reshape(-1)]).astype(np.uint32)
    cost = np.repeat(dist[ok], per).astype(np.uint64)
    # per bidder, rows are in ring-offset order, per sites each
    return start, obj, cost, side * side * per, 2 * (side - 1)
