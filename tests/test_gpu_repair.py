"""legalization::project end to end (crates/texo-pnr/src/legalization/mod.rs:
322-389): the relaxed auction on the GPU (texo_gp_auction, ring domains of a
small radius as the reference's candidate search gives it) leaves the
bidders of a congested region unmatched; the transactional repair
(texo_gp_repair) then places every unit over its whole candidate set, as the
reference's ChoiceCursor walks it.  The product repair equals the oracle's
on the GPU auction's proposal, and the result is complete and legal."""
import numpy as np
import pytest

from texo_amd import auction as A
from texo_amd import repair as R
from tests.test_repair import legal, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,bidders", [(3, 300), (4, 480)])
def test_gpu_auction_then_repair_places_every_unit(seed, bidders):
    side, radius = 24, 2
    rng = np.random.default_rng(seed)
    tx = np.clip(np.rint(side / 2 + rng.normal(0, 3.0, bidders)), 0, side - 1).astype(int)
    ty = np.clip(np.rint(side / 2 + rng.normal(0, 3.0, bidders)), 0, side - 1).astype(int)
    sites = [(x, y) for y in range(side) for x in range(side)]
    domains = []
    for b in range(bidders):
        d = [(y * side + x, abs(x - tx[b]) + abs(y - ty[b])) for (x, y) in sites
             if abs(x - tx[b]) + abs(y - ty[b]) <= radius]
        domains.append(d)
    start, obj, cost = A.domains_csr(domains)
    assignment, stats = A.auction(start, obj, cost, len(sites), 2 * (side - 1))
    unmatched = int((assignment < 0).sum())
    print(f"seed {seed}: {bidders} bidders, {unmatched} unmatched after the GPU auction", flush=True)
    assert unmatched > 0
    # repair candidates: every site, candidate j = site j
    units = [R.RepairUnit([R.Candidate([s]) for s in range(len(sites))], target=(int(tx[b]), int(ty[b])),
                          preferred=None if assignment[b] < 0 else int(assignment[b]), first_cell=b)
             for b in range(bidders)]
    p = R.RepairProblem(side, side, sites, units)
    got, want = run(p, "texo_gp"), run(p, "oracle")
    assert got == want
    status, placed, rstats, _ = got
    assert status == 0 and legal(p, placed)
    assert rstats["initial_pending"] == unmatched
    # matched units keep their auction sites
    assert all(placed[b] == assignment[b] for b in range(bidders) if assignment[b] >= 0) or \
        rstats["units_evicted"] > 0
