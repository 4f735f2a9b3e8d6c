"""The transactional legalization repair (repair_relaxed,
crates/texo-pnr/src/legalization/repair.rs:60-155) over explicit candidate
footprints: the oracle restatement pinned by the reference's own repair tests
(repair.rs:1319-1929) and by brute force on tiny instances, and the product
(texo_gp_repair, a host search in libtexo_gp.so) equal to the oracle on the
same inputs.  The repair needs no device, so these run on the CPU.

Device and design of each reference test are restated as explicit
footprints: a BEL is an object at its point, a constraint group's
assignments are the unit's candidates, a shared resource becomes the
(resource, value) pairs of the candidates that use it, and a BEL pin on a
wire carrying a net becomes a (wire, net) pair.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle
from texo_amd import abi
from texo_amd import repair as R

OK = 0


def olib():
    lib = oracle.load()
    if not getattr(lib, "_repair_hooks", False):
        P = C.POINTER
        lib.oracle_repair_cursor.argtypes = [P(R.texo_gp_repair_problem), C.c_uint64, P(C.c_uint64), C.c_int]
        lib.oracle_repair_cursor.restype = C.c_int
        lib.oracle_repair_pin_branches.argtypes = [P(R.texo_gp_repair_problem), C.c_uint64, P(C.c_uint64),
                                                   P(C.c_uint64), C.c_uint64, C.c_uint64, C.c_uint64,
                                                   P(C.c_uint64), C.c_int]
        lib.oracle_repair_pin_branches.restype = C.c_int
        lib._repair_hooks = True
    return lib


def run(problem, prefix="oracle"):
    """(status, placed or None, stats or None, error unit)."""
    try:
        placed, stats = R.repair(problem, olib() if prefix == "oracle" else None, prefix)
        return OK, list(placed), stats, -1
    except R.RepairError as e:
        return e.status, None, None, e.unit


def line_bels(n):
    """add_logic_bels (repair.rs tests): BEL i at (i, 0)."""
    return [(i, 0) for i in range(n)]


def single(objects, **kw):
    return R.RepairUnit([R.Candidate([o]) for o in objects], **kw)


# ---- the reference's tests --------------------------------------------------

def test_choice_cursor_is_preferred_first_then_point_stable_and_exhaustive():
    # repair.rs:1319-1345: 3x3 device, BELs at (2,1) (0,1) (1,2) (1,1) (1,0),
    # target (1,1), preferred choice 0
    p = R.RepairProblem(3, 3, [(2, 1), (0, 1), (1, 2), (1, 1), (1, 0)],
                        [single(range(5), target=(1, 1), preferred=0)])
    st = p.struct()
    out = (C.c_uint64 * 8)()
    n = olib().oracle_repair_cursor(C.byref(st), 0, out, 8)
    assert list(out[:n]) == [0, 3, 1, 4, 2]


def test_twelve_thousand_unit_augmenting_chain_uses_explicit_frames():
    # repair.rs:1347-1408: unit 0 wants BEL 0 only, unit u > 0 prefers BEL
    # u - 1 and may take BEL u; unit 0 starts unmatched
    n = 12_000
    units = [single([0], first_cell=0)] + [single([u - 1, u], preferred=0, first_cell=u) for u in range(1, n)]
    p = R.RepairProblem(1, 1, [(0, 0)] * n, units)
    status, placed, stats, _ = run(p)
    assert status == OK
    assert placed == [0] + [1] * (n - 1)  # unit u on BEL u
    assert stats["direct_fast_fallbacks"] == 1
    assert stats["augmenting_roots"] == 1
    assert stats["augmenting_fallbacks"] == 0
    assert stats["states_visited"] >= n


def test_multi_blocker_candidate_falls_back_to_exact_iterative_frames():
    # repair.rs:1421-1477: a two-cell bundle that needs BELs 0 and 1, held by
    # two single-cell units that can move to BELs 2 and 3
    units = [R.RepairUnit([R.Candidate([0, 1])], cells=2, first_cell=0),
             single([0, 2], preferred=0, first_cell=2),
             single([1, 3], preferred=0, first_cell=3)]
    p = R.RepairProblem(4, 1, line_bels(4), units)
    status, placed, stats, _ = run(p)
    assert status == OK
    assert placed == [0, 1, 1]  # bundle on BELs 0, 1; first on 2; second on 3
    assert stats["direct_fast_fallbacks"] == 1
    assert stats["augmenting_fallbacks"] == 1
    assert stats["states_visited"] >= 3


def test_partial_direct_fast_path_rolls_back_before_augmenting():
    # repair.rs:1479-1540
    units = [single([3], first_cell=0), single([0, 1], first_cell=1), single([0, 2], preferred=0, first_cell=2),
             single([1], preferred=0, fixed=True, first_cell=3)]
    p = R.RepairProblem(1, 1, [(0, 0)] * 4, units)
    status, placed, stats, _ = run(p)
    assert status == OK
    assert placed == [0, 0, 1, 0]  # BELs 3, 0, 2, 1
    assert stats["direct_fast_moves"] == 1
    assert stats["direct_fast_fallbacks"] == 1
    assert stats["augmenting_roots"] == 2
    assert stats["augmenting_fallbacks"] == 0


def test_far_spill_displaces_a_clean_bel_owner_instead_of_returning_no_bel():
    # repair.rs:1542-1572: 65x1 device; `wanted` may take BEL 0 or 1 (1 is
    # fixed), `flexible` sits on 0 and may spill to BEL 64
    units = [single([0, 1], first_cell=0), single([0, 64], preferred=0, first_cell=1),
             single([1], preferred=0, fixed=True, first_cell=2)]
    p = R.RepairProblem(65, 1, line_bels(65), units)
    status, placed, _, _ = run(p)
    assert status == OK
    assert placed == [0, 1, 0]  # wanted on BEL 0, flexible on 64, fixed on 1


def test_overlapping_bundle_rows_are_repaired_atomically():
    # repair.rs:1574-1613: two two-cell units over the rows {0,1} {1,2} {2,3}
    rows = [[0, 1], [1, 2], [2, 3]]
    units = [R.RepairUnit([R.Candidate(r) for r in rows], cells=2, target=(0, 0), preferred=0, first_cell=0),
             R.RepairUnit([R.Candidate(r) for r in rows], cells=2, target=(1, 0), preferred=1, first_cell=2)]
    p = R.RepairProblem(4, 1, line_bels(4), units)
    status, placed, _, _ = run(p)
    assert status == OK
    used = [b for u, c in enumerate(placed) for b in rows[c]]
    assert sorted(used) == [0, 1, 2, 3]


def test_incompatible_shared_values_move_to_different_resources():
    # repair.rs:1615-1645: BELs 0, 1 share resource 0 and BELs 2, 3 resource
    # 1; `zero` needs value 0 (BEL 0 or 2), `one` value 1 (BEL 1 or 3)
    zero = R.RepairUnit([R.Candidate([0], shared=[(0, 0)]), R.Candidate([2], shared=[(1, 0)])], preferred=0,
                        first_cell=0)
    one = R.RepairUnit([R.Candidate([1], shared=[(0, 1)]), R.Candidate([3], shared=[(1, 1)])], preferred=0,
                       first_cell=1)
    p = R.RepairProblem(4, 1, line_bels(4), [zero, one])
    status, placed, stats, _ = run(p)
    assert status == OK
    bel = [[0, 2][placed[0]], [1, 3][placed[1]]]
    assert bel[0] // 2 != bel[1] // 2
    assert stats["initial_pending"] == 1


def test_initial_pending_keeps_a_fixed_net_class_and_fills_remaining_capacity():
    # repair.rs:1647-1720: one wire of capacity 2 carries the nets of four
    # units: the fixed net class (units 0 and 1), class b (unit 2), class c
    # (unit 3); class c is set aside.  With no other site for unit 3 the
    # repair then reports it as the unit that could not be placed.
    units = [R.RepairUnit([R.Candidate([i], wires=[(0, net)])], preferred=0, fixed=(i == 0), first_cell=i)
             for i, net in enumerate([0, 0, 1, 2])]
    p = R.RepairProblem(1, 1, [(0, 0)] * 4, units, wire_capacity=[2])
    status, _, _, unit = run(p)
    assert status == abi.STATUS_CODE["NoBel"]
    assert unit == 3


def test_unavoidable_shared_value_conflict_exhausts_finite_states():
    # repair.rs:1722-1753: every BEL maps to resource 0, the two units need
    # values 0 and 1
    zero = R.RepairUnit([R.Candidate([0], shared=[(0, 0)]), R.Candidate([1], shared=[(0, 0)])], preferred=0,
                        first_cell=0)
    one = R.RepairUnit([R.Candidate([2], shared=[(0, 1)]), R.Candidate([3], shared=[(0, 1)])], preferred=0,
                       first_cell=1)
    p = R.RepairProblem(4, 1, line_bels(4), [zero, one])
    status, _, _, _ = run(p)
    assert status == abi.STATUS_CODE["NoBel"]


def pin_branches(p, installed, unit, choice, mandatory):
    st = p.struct()
    iu = (C.c_uint64 * max(1, len(installed)))(*[u for u, _ in installed])
    ic = (C.c_uint64 * max(1, len(installed)))(*[c for _, c in installed])
    out = (C.c_uint64 * 16)()
    mask = sum(1 << m for m in mandatory)
    n = olib().oracle_repair_pin_branches(C.byref(st), len(installed), iu, ic, unit, choice, mask, out, 16)
    return [{u for u in range(64) if out[k] >> u & 1} for k in range(n)]


def test_pin_class_remains_until_its_last_owner_is_blocked():
    # repair.rs:1755-1849: wire 0 ("shared", capacity 1) carries net a of
    # the source and sink BELs; the candidate's net b needs it: with the
    # source already a blocker, the sink (the class's last owner) joins it
    wires = [1, 1, 1]  # shared, alternate-a, alternate-b
    NET_A, NET_B = 0, 1
    units = [R.RepairUnit([R.Candidate([0], wires=[(0, NET_A)]), R.Candidate([3], wires=[(1, NET_A)])],
                          preferred=0, first_cell=0),
             R.RepairUnit([R.Candidate([1], wires=[(0, NET_A)]), R.Candidate([4], wires=[(2, NET_A)])],
                          preferred=0, first_cell=1),
             R.RepairUnit([R.Candidate([2], wires=[(0, NET_B)])], first_cell=2)]
    p = R.RepairProblem(3, 1, [(0, 0), (0, 0), (0, 0), (1, 0), (2, 0)], units, wire_capacity=wires)
    assert pin_branches(p, [(0, 0), (1, 0)], 2, 0, {0}) == [{0, 1}]
    # the whole repair moves both net-a owners to their alternates
    status, placed, _, _ = run(p)
    assert status == OK and placed == [1, 1, 0]


def test_candidate_distinct_nets_share_the_candidate_wire_capacity():
    # repair.rs:1851-1928: a two-cell candidate brings nets c and d onto a
    # wire of capacity 2 that carries the incumbent's net a
    units = [R.RepairUnit([R.Candidate([0], wires=[(0, 0)])], preferred=0, first_cell=0),
             R.RepairUnit([R.Candidate([1, 2], wires=[(0, 1), (0, 2)])], cells=2, first_cell=1)]
    p = R.RepairProblem(1, 1, [(0, 0)] * 3, units, wire_capacity=[2])
    assert pin_branches(p, [(0, 0)], 1, 0, set()) == [{0}]


def test_projection_errors():
    # repair.rs:70-86 and initial_pending_units' fixed-unit conflicts
    two = R.RepairProblem(2, 1, line_bels(2), [single([0], preferred=0, fixed=True),
                                                single([0], preferred=0, fixed=True, first_cell=1)])
    assert run(two)[0] == abi.STATUS_CODE["InvalidPlacement"]
    unmatched_fixed = R.RepairProblem(2, 1, line_bels(2), [single([0], fixed=True)])
    assert run(unmatched_fixed)[0] == abi.STATUS_CODE["InvalidPlacement"]
    out_of_range = R.RepairProblem(2, 1, line_bels(2), [single([0], preferred=3)])
    assert run(out_of_range)[0] == abi.STATUS_CODE["InvalidPlacement"]
    for prefix in ("oracle", "texo_gp"):
        for prob in (two, unmatched_fixed, out_of_range):
            assert run(prob, prefix)[0] == abi.STATUS_CODE["InvalidPlacement"]


# ---- brute force and product equality ---------------------------------------

def random_problem(rng, n_units, n_objects, width=4, height=2, wires=2, resources=2, p_shared=0.3, p_wire=0.3,
                   max_cells=2, p_fixed=0.1):
    xy = [(int(rng.integers(width)), int(rng.integers(height))) for _ in range(n_objects)]
    caps = [int(rng.integers(1, 3)) for _ in range(wires)]
    units = []
    used_fixed = set()
    for u in range(n_units):
        cells = int(rng.integers(1, max_cells + 1))
        cands = []
        for _ in range(int(rng.integers(1, 4))):
            objs = [int(o) for o in rng.choice(n_objects, size=cells, replace=False)]
            shared = [(int(rng.integers(resources)), int(rng.integers(2)))] if rng.random() < p_shared else []
            w = [(int(rng.integers(wires)), int(rng.integers(3)))] if rng.random() < p_wire else []
            cands.append(R.Candidate(objs, shared, w))
        pref = int(rng.integers(len(cands))) if rng.random() < 0.8 else None
        fixed = pref is not None and rng.random() < p_fixed and not (set(cands[pref].objects) & used_fixed)
        if fixed:
            used_fixed |= set(cands[pref].objects)
        units.append(R.RepairUnit(cands, target=(int(rng.integers(width)), int(rng.integers(height))),
                                  preferred=pref, fixed=fixed, cells=cells, first_cell=2 * u))
    return R.RepairProblem(width, height, xy, units, wire_capacity=caps)


def legal(p, placed):
    """Every unit placed on one of its candidates, objects used once, one
    value per shared resource, wire capacities kept."""
    objs, values, nets = set(), {}, {}
    for unit, c in zip(p.units, placed):
        if c < 0 or c >= len(unit.candidates):
            return False
        cand = unit.candidates[c]
        for o in cand.objects:
            if o in objs:
                return False
            objs.add(o)
        for r, v in set(cand.shared):
            values.setdefault(r, set()).add(v)
        for w, n in set(cand.wires):
            nets.setdefault(w, set()).add(n)
    return all(len(v) <= 1 for v in values.values()) and all(len(n) <= p.wire_capacity[w] for w, n in nets.items())


def test_repair_result_is_legal_and_complete_iff_brute_force_finds_one():
    rng = np.random.default_rng(7)
    checked = feasible = 0
    for _ in range(400):
        p = random_problem(rng, n_units=int(rng.integers(2, 6)), n_objects=int(rng.integers(4, 8)))
        status, placed, _, _ = run(p)
        if status == abi.STATUS_CODE["InvalidPlacement"]:
            continue
        exists = any(legal(p, list(a)) for a in itertools.product(*[range(len(u.candidates)) for u in p.units])
                     if all(not u.fixed or a[i] == u.preferred for i, u in enumerate(p.units)))
        checked += 1
        feasible += exists
        if status == OK:
            assert legal(p, placed)
            assert all(not u.fixed or placed[i] == u.preferred for i, u in enumerate(p.units))
        else:
            assert status == abi.STATUS_CODE["NoBel"]
            assert not exists
    assert checked > 300 and feasible > 100


def test_product_repair_equals_the_oracle():
    rng = np.random.default_rng(11)
    for k in range(300):
        p = random_problem(rng, n_units=int(rng.integers(2, 24)), n_objects=int(rng.integers(8, 30)), width=6,
                           height=3, wires=3, resources=3, max_cells=3)
        a, b = run(p, "oracle"), run(p, "texo_gp")
        assert a == b, (k, a, b)


def test_product_repair_equals_the_oracle_on_a_long_chain():
    n = 3000
    units = [single([0], first_cell=0)] + [single([u - 1, u], preferred=0, first_cell=u) for u in range(1, n)]
    p = R.RepairProblem(1, 1, [(0, 0)] * n, units)
    assert run(p, "oracle") == run(p, "texo_gp")
