"""The legalization auction's oracle (CPU): assign_implicitly over explicit
domains, pinned to the reference's own auction tests
(legalization/auction.rs:255-507): the known assignments, the dummy's
maximum-cardinality guarantee, the malformed-input errors, and the exhaustive
comparison with a brute-force optimum on every small graph."""
import ctypes as C

import numpy as np
import pytest

from texo_amd import abi

U32, U64, I64 = C.c_uint32, C.c_uint64, C.c_int64


def declare(lib):
    P = C.POINTER
    lib.oracle_auction.argtypes = [U64, U64, U64, P(U32), P(U32), P(U64), P(I64), P(U64)]
    lib.oracle_auction.restype = C.c_int


@pytest.fixture(scope="module")
def lib(oracle_lib):
    declare(oracle_lib)
    return oracle_lib


def csr(domains):
    start = np.zeros(len(domains) + 1, np.uint32)
    start[1:] = np.cumsum([len(d) for d in domains])
    obj = np.array([o for d in domains for o, _ in d] or [0], np.uint32)
    cost = np.array([c for d in domains for _, c in d] or [0], np.uint64)
    return start, obj, cost


def auction(lib, domains, objects, max_base):
    start, obj, cost = csr(domains)
    out = np.zeros(max(len(domains), 1), np.int64)
    stats = np.zeros(3, np.uint64)
    rc = lib.oracle_auction(len(domains), objects, max_base, abi.ptr(start, U32), abi.ptr(obj, U32),
                            abi.ptr(cost, U64), abi.ptr(out, I64), abi.ptr(stats, U64))
    return rc, [None if a < 0 else int(a) for a in out[:len(domains)]], stats


def quality(domains, assignment):
    unmatched = sum(a is None for a in assignment)
    cost = sum(dict(domains[b])[a] for b, a in enumerate(assignment) if a is not None)
    return unmatched, cost


def brute_force(domains, objects):
    """assignment_quality's optimum (auction.rs:303-338): fewest unmatched, then least cost."""
    best = [(len(domains) + 1, 1 << 62)]
    used = [False] * objects

    def visit(b, unmatched, cost):
        if (unmatched, cost) >= best[0]:
            return
        if b == len(domains):
            best[0] = (unmatched, cost)
            return
        visit(b + 1, unmatched + 1, cost)
        for o, c in domains[b]:
            if not used[o]:
                used[o] = True
                visit(b + 1, unmatched, cost + c)
                used[o] = False

    visit(0, 0, 0)
    return best[0]


def check_against_brute_force(lib, domains, objects, max_base):
    rc, a, _ = auction(lib, domains, objects, max_base)
    assert rc == 0
    got = [o for o in a if o is not None]
    assert len(got) == len(set(got))
    actual, optimum = quality(domains, a), brute_force(domains, objects)
    assert actual[0] == optimum[0], (domains, a)
    assert actual[1] <= optimum[1] + len(domains), (domains, a)


def test_flexible_bidder_moves_instead_of_spilling(lib):
    # flexible_bidder_moves_instead_of_spilling_constrained_bidder (auction.rs:354-364)
    rc, a, st = auction(lib, [[(0, 0), (1, 1), (2, 100)], [(0, 0), (2, 100)]], 3, 100)
    assert rc == 0 and a == [1, 0] and st[2] == 0


def test_restricted_graph_uses_the_real_solution(lib):
    # restricted_graph_uses_real_solution_instead_of_false_no_bel (auction.rs:366-376)
    rc, a, st = auction(lib, [[(0, 0), (1, 1)], [(0, 0)]], 2, 1)
    assert rc == 0 and a == [1, 0] and st[2] == 0


def test_private_dummy_gap_keeps_a_perfect_matching(lib):
    # private_dummy_gap_does_not_hide_a_perfect_matching (auction.rs:378-388)
    rc, a, st = auction(lib, [[(1, 0), (2, 0)], [(0, 0), (1, 0)], [(0, 0)]], 3, 0)
    assert rc == 0 and a == [2, 1, 0] and st[2] == 0


def test_candidate_after_sixty_four_remains_visible(lib):
    # candidate_after_legacy_sixty_four_limit_remains_visible (auction.rs:390-407)
    domains = [[(o, 0)] for o in range(65)] + [[(o, 0) for o in range(66)]]
    r1, a1, s1 = auction(lib, domains, 66, 0)
    r2, a2, s2 = auction(lib, domains, 66, 0)
    assert r1 == r2 == 0 and a1 == a2 and list(s1) == list(s2)
    assert s1[2] == 0 and a1[65] == 65


def test_private_dummy_reports_the_deficiency(lib):
    # private_dummy_reports_maximum_cardinality_deficiency (auction.rs:409-418)
    rc, a, st = auction(lib, [[(0, 0)], [(0, 0)]], 1, 0)
    assert rc == 0 and sum(o is not None for o in a) == 1 and st[2] == 1


def test_price_range_exhaustion_is_reported_before_bidding(lib):
    # native_price_range_exhaustion_is_reported_before_bidding (auction.rs:437-445)
    rc, _, _ = auction(lib, [[]], 0, 2 ** 64 - 1)
    assert abi.STATUS[rc] == "AuctionArithmeticOverflow"


def test_malformed_domains_are_reported(lib):
    # the reference's closure-contract errors (auction.rs:174-199): an object
    # out of range, a base cost above the bound, an object twice in a domain
    assert abi.STATUS[auction(lib, [[(3, 0)]], 3, 0)[0]] == "ObjectOutOfRange"
    assert abi.STATUS[auction(lib, [[(0, 2)]], 1, 1)[0]] == "BaseCostOutOfRange"
    assert abi.STATUS[auction(lib, [[(0, 0), (0, 1)]], 1, 1)[0]] == "DuplicateBestObject"


def test_empty_auction(lib):
    rc, a, st = auction(lib, [], 0, 0)
    assert rc == 0 and a == [] and list(st) == [0, 0, 0]


def test_exhaustive_small_graphs_against_brute_force(lib):
    # implicit_auction_has_exact_cardinality_and_bounded_cost_on_small_graphs
    # (auction.rs:447-506): every absent / cost-0 / cost-1 edge pattern
    # through 3x3, then every 4 x 0..4 domain mask with (bidder + object) & 1 costs
    for bidders in range(1, 4):
        for objects in range(0, 4):
            for code in range(3 ** (bidders * objects)):
                domains = []
                for _ in range(bidders):
                    d = []
                    for o in range(objects):
                        e = code % 3
                        code //= 3
                        if e:
                            d.append((o, e - 1))
                    domains.append(d)
                check_against_brute_force(lib, domains, objects, 1)
    for objects in range(0, 5):
        for mask in range(1 << (4 * objects)):
            domains = [[(o, (b + o) & 1) for o in range(objects) if mask >> (b * objects + o) & 1]
                       for b in range(4)]
            check_against_brute_force(lib, domains, objects, 1)
