"""GP input builders and provenance (CPU): the oracle restatements pinned to
the reference's known answers (routing_demand.rs:315-457,
texo-flow lib.rs:6625-6688, checkpoint.rs:1218-1250), and the host-side
product functions (channel names, timing weights, placement model) against
them bit for bit."""
import numpy as np
import pytest

from texo_amd import builders, synth


@pytest.fixture(scope="module")
def olib(oracle_lib):
    return oracle_lib


def test_channel_orientation_names(olib):
    # channel_orientation (routing_demand.rs:185-198)
    cases = {"R0C0/H02E0001": 1, "R1C2/V06N0003": 2, "H01": 1, "R0C0/JA": 0, "R0C0/HX2": 0, "R0C0/V0": 0,
             "a/b/V12": 2, "": 0, "R3C3/X02": 0}
    for name, code in cases.items():
        assert olib.oracle_channel_orientation(name.encode()) == code, name
        assert builders.channel_orientation(name) == code, name


def capacity_case():
    # blocked_only_incident_pip_removes_channel_capacity (routing_demand.rs:358-394):
    # wires H(0,0) cap 2, H(1,0) cap 3 (its only pip blocked), V(2,1) cap 4,
    # and three local endpoint wires
    names = ["R0C0/H02E0001", "R0C1/H02E0001", "R1C2/V06N0003", "R0C0/JA", "R0C1/JB", "R1C2/JC"]
    xy = [(0, 0), (1, 0), (2, 1), (0, 0), (1, 0), (2, 1)]
    cap = [2, 3, 4, 1, 1, 1]
    orient = [builders.channel_orientation(n) for n in names]
    return 3, 2, orient, xy, cap, [0, 1, 2], [3, 4, 5], [0, 1, 0]


def test_blocked_pip_removes_channel_capacity(olib):
    W, H, orient, xy, cap, pf, pt, pb = capacity_case()
    h, v = builders.routing_capacity_map(W, H, orient, xy, cap, pf, pt, pb, lib=olib, prefix="oracle")
    assert h[0, 0] == 2 and h[0, 1] == 0 and v[1, 2] == 4
    assert h.sum() == 2 and v.sum() == 4


def test_capacity_saturates_at_u32(olib):
    n = 70000  # 70000 * 65535 > 2^32
    h, v = builders.routing_capacity_map(1, 1, [1] * n, [(0, 0)] * n, [65535] * n, list(range(n)),
                                         list(range(n)), None, lib=olib, prefix="oracle")
    assert h[0, 0] == 0xFFFFFFFF and v[0, 0] == 0


def test_directional_demand_conserves_bbox_hpwl(olib):
    # directional_demand_conserves_bbox_hpwl (routing_demand.rs:396-416)
    h, v, inc, exc = builders.routing_demand_map(4, 3, [0, 2], [(0, 0), (3, 2)], [0], lib=olib, prefix="oracle")
    assert inc == 1 and exc == 0
    assert abs(h.sum() - 3.0) < 1e-9 and abs(v.sum() - 2.0) < 1e-9
    assert abs(h[1, 1] - 0.25) < 1e-9 and abs(v[1, 1] - 1.0 / 6.0) < 1e-9


def test_clock_driven_nets_are_excluded(olib):
    # clock_driven_nets_are_excluded (routing_demand.rs:418-428)
    h, v, inc, exc = builders.routing_demand_map(4, 3, [0, 2], [(0, 0), (3, 2)], [1], lib=olib, prefix="oracle")
    assert inc == 0 and exc == 1 and h.sum() == 0.0 and v.sum() == 0.0


def test_bin_aggregation_conserves_capacity_and_demand(olib):
    # bin_aggregation_conserves_capacity_and_demand (routing_demand.rs:430-456):
    # 2x2 bins over the 4x3 grid keep the totals; the lower-right bin holds
    # the vertical wire's capacity 3
    names = ["R1C1/H02E0001", "R2C3/V02N0001", "R1C1/JH", "R2C3/JV"]
    orient = [builders.channel_orientation(n) for n in names]
    ch, cv = builders.routing_capacity_map(4, 3, orient, [(1, 1), (3, 2), (1, 1), (3, 2)], [2, 3, 1, 1], [0, 1],
                                           [2, 3], None, lib=olib, prefix="oracle")
    h, v, _, _ = builders.routing_demand_map(4, 3, [0, 2], [(0, 0), (3, 2)], [0], lib=olib, prefix="oracle")
    bins = [(x, y, ch[y:y + 2, x:x + 2].sum(), cv[y:y + 2, x:x + 2].sum(), h[y:y + 2, x:x + 2].sum(),
             v[y:y + 2, x:x + 2].sum()) for y in (0, 2) for x in (0, 2)]
    assert len(bins) == 4
    assert sum(b[2] for b in bins) == 2 and sum(b[3] for b in bins) == 3
    assert abs(sum(b[4] for b in bins) - 3.0) < 1e-9 and abs(sum(b[5] for b in bins) - 2.0) < 1e-9
    assert next(b for b in bins if b[:2] == (2, 2))[3] == 3


def test_unplaced_sinks_and_drivers(olib):
    # placement.point(...) is None: a sink without a placement is skipped, a
    # net whose driver has none is skipped entirely (routing_demand.rs:289-301)
    U = 0xFFFFFFFF
    h, v, inc, _ = builders.routing_demand_map(4, 3, [0, 3, 5], [(0, 0), (U, U), (2, 1), (U, U), (3, 2)], [0, 0],
                                               lib=olib, prefix="oracle")
    assert inc == 1 and abs(h.sum() - 2.0) < 1e-12 and abs(v.sum() - 1.0) < 1e-12


def weight(olib, path, worst, k):
    m = (1 << 64) - 1
    return olib.oracle_criticality_weight(path & m, path >> 64, worst & m, worst >> 64, k)


def test_criticality_weight_known_answers(olib):
    # ecp5_placement_weight_is_quartic_by_default_without_delay_attenuation (texo-flow lib.rs:6631-6646)
    for path, worst, expect in ((0, 4000, 1), (1000, 4000, 1), (2000, 4000, 2), (3000, 4000, 4), (4000, 4000, 11),
                                (5000, 4000, 11), (1, 0, 1), ((1 << 128) - 1, (1 << 128) - 1, 11)):
        assert weight(olib, path, worst, 4) == expect, (path, worst)
    # placement_weight_exponent_controls_the_ecp5_power (:6677-6688)
    assert [weight(olib, 2000, 4000, k) for k in (1, 2, 4)] == [6, 4, 2]
    assert weight(olib, 2000, 4000, 0) == weight(olib, 2000, 4000, 1)
    assert weight(olib, 2000, 4000, 2 ** 32 - 1) == 1


def test_timing_weights_are_domain_normalised(olib):
    # ecp5_weights_remain_domain_normalized_after_timing_closes (texo-flow lib.rs:6648-6675)
    w = builders.timing_placement_weights([(2, 20, 1, 2), (3, 30, 7, 7)], 4, lib=olib, prefix="oracle")
    assert w == {(2, 20): 2, (3, 30): 11}
    # the maximum per (net, sink) wins
    w = builders.timing_placement_weights([(2, 20, 1, 2), (2, 20, 7, 7), (2, 20, 0, 7)], 4, lib=olib,
                                          prefix="oracle")
    assert w == {(2, 20): 11}


def test_product_timing_weights_match_oracle_bit_for_bit(olib):
    rng = np.random.default_rng(9)
    edges = []
    for i in range(4000):
        worst = int(rng.integers(0, 10 ** 6)) if i % 7 else int(rng.integers(0, 2 ** 62)) << 60
        path = (int(rng.integers(0, 2 ** 62)) * (2 * max(worst, 1)) >> 62) if worst else int(rng.integers(0, 100))
        edges.append((int(rng.integers(0, 300)), int(rng.integers(0, 50)), path, worst))
    for k in (0, 1, 2, 4, 6, 13, 100):
        assert builders.timing_placement_weights(edges, k) == \
            builders.timing_placement_weights(edges, k, lib=olib, prefix="oracle")
    # and the test-data generator's weight formula
    for net, sink, path, worst in edges[:500]:
        if worst < 2 ** 64:
            assert synth.criticality_weight(path, worst, 4) == weight(olib, path, worst, 4)


def test_placement_model_record():
    # checkpoints_the_effective_timing_driven_placement_model (checkpoint.rs:1218-1250)
    rec = builders.placement_model("timing_routability_electrostatic", 4)
    assert {k: rec[k] for k in list(rec) if k != "global_placement"} == {
        "initial_algorithm": "ecp5_timing_routability_electrostatic_v1",
        "timing_weight_model": "ecp5_1_plus_10_criticality_power_v1", "criticality_exponent": 4,
        "routability_model": "directional_rudy_area_adjustment_v1", "initial_predicted_detail": False}
    rec = builders.placement_model("routability_electrostatic", 4)
    assert rec["initial_algorithm"] == "routability_electrostatic_v1" and rec["timing_weight_model"] is None
    assert rec["criticality_exponent"] is None and rec["routability_model"] == "directional_rudy_area_adjustment_v1"
    assert builders.placement_model("imported", 4) == {
        "initial_algorithm": "imported_v1", "timing_weight_model": None, "criticality_exponent": None,
        "routability_model": None, "initial_predicted_detail": False}
    # the engine record names the run's precision and options
    from texo_amd import abi
    gp = builders.placement_model("routability_electrostatic", 4,
                                  abi.options(dct_precision=abi.DCT_FP16_SCALED, fixed_iterations=200,
                                              gamma_base=4.0))["global_placement"]
    assert gp["dct_precision"] == "fp16_scaled" and gp["fixed_iterations"] == 200 and gp["gamma_base"] == 4.0
    assert gp["engine"] == "texo_amd_mi355x_electrostatic_v1" and gp["max_iterations"] == 3000
