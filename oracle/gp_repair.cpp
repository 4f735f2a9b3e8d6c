// gp_repair.cpp — CPU restatement of texo-pnr's transactional legalization
// repair.  TEST INFRASTRUCTURE ONLY: the parity oracle for texo_gp_repair.
//
// Follows crates/texo-pnr/src/legalization/repair.rs (repair_relaxed and its
// searches, :60-1266) over explicit candidate footprints: the reference
// derives each candidate's BELs, shared-resource values and pin wires from
// the device graph (assignment_footprint, :162-192); here they are inputs.
// The legality rules are those of prepared_resource_lists_are_legal
// (legalization/mod.rs:1080-1128).  The containers mirror the reference's
// (ordered maps and sets where it iterates them in order).

#include <algorithm>
#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gp_oracle.h"

namespace {

using Unit = uint64_t;
using Choice = uint64_t;
constexpr Choice UNPLACED = UINT64_MAX;  // UNPLACED_CHOICE of the state key

struct ProjectionError {
  std::string reason;
};
struct NoBel {
  Unit unit;
};

struct Footprint {
  std::vector<uint32_t> objects;
  std::vector<std::pair<uint64_t, uint64_t>> pin_wires;  // (wire, net), sorted, unique
  std::vector<std::pair<uint64_t, uint64_t>> shared;     // (resource, value), sorted, unique
};

struct Instance {
  const texo_gp_repair_problem *p;
  uint64_t n_units;
  uint64_t choices(Unit u) const { return p->cand_start[u + 1] - p->cand_start[u]; }
  uint64_t cand(Unit u, Choice c) const { return p->cand_start[u] + c; }
  Footprint footprint(Unit u, Choice c) const {
    // assignment_footprint (repair.rs:162-192): sorted, deduplicated lists
    const uint64_t k = cand(u, c);
    Footprint f;
    for (uint64_t i = p->cand_object_start[k]; i < p->cand_object_start[k + 1]; ++i) f.objects.push_back(p->cand_object[i]);
    for (uint64_t i = p->cand_wire_start[k]; i < p->cand_wire_start[k + 1]; ++i)
      f.pin_wires.push_back({p->cand_wire[2 * i], p->cand_wire[2 * i + 1]});
    for (uint64_t i = p->cand_shared_start[k]; i < p->cand_shared_start[k + 1]; ++i)
      f.shared.push_back({p->cand_shared[2 * i], p->cand_shared[2 * i + 1]});
    std::sort(f.pin_wires.begin(), f.pin_wires.end());
    f.pin_wires.erase(std::unique(f.pin_wires.begin(), f.pin_wires.end()), f.pin_wires.end());
    std::sort(f.shared.begin(), f.shared.end());
    f.shared.erase(std::unique(f.shared.begin(), f.shared.end()), f.shared.end());
    return f;
  }
  // (choices, Reverse(cells), first cell): next_unplaced (repair.rs:573-583)
  bool before(Unit a, Unit b) const {
    if (choices(a) != choices(b)) return choices(a) < choices(b);
    if (p->unit_cells[a] != p->unit_cells[b]) return p->unit_cells[a] > p->unit_cells[b];
    return p->unit_first_cell[a] < p->unit_first_cell[b];
  }
  // (choices, Reverse(cells), unit): the survivor of a shared BEL (:356-364)
  bool survives_before(Unit a, Unit b) const {
    if (choices(a) != choices(b)) return choices(a) < choices(b);
    if (p->unit_cells[a] != p->unit_cells[b]) return p->unit_cells[a] > p->unit_cells[b];
    return a < b;
  }
};

// PlacementResourceUsage: multiplicity of each net on a wire and each value
// of a shared resource.
struct Usage {
  std::map<uint64_t, std::map<uint64_t, uint64_t>> pin_wires, shared;
};

bool lists_are_legal(const Instance &in, const Usage &usage, const Footprint &f) {
  // prepared_resource_lists_are_legal (legalization/mod.rs:1080-1128)
  for (size_t s = 0; s < f.pin_wires.size();) {
    const uint64_t wire = f.pin_wires[s].first;
    size_t e = s + 1;
    while (e < f.pin_wires.size() && f.pin_wires[e].first == wire) ++e;
    const auto it = usage.pin_wires.find(wire);
    size_t fresh = 0;
    for (size_t k = s; k < e; ++k)
      if (it == usage.pin_wires.end() || !it->second.count(f.pin_wires[k].second)) fresh += 1;
    const size_t have = it == usage.pin_wires.end() ? 0 : it->second.size();
    if (have + fresh > in.p->wire_capacity[wire]) return false;
    s = e;
  }
  for (size_t s = 0; s < f.shared.size();) {
    const uint64_t res = f.shared[s].first;
    size_t e = s + 1;
    while (e < f.shared.size() && f.shared[e].first == res) ++e;
    const auto it = usage.shared.find(res);
    size_t fresh = 0;
    for (size_t k = s; k < e; ++k)
      if (it == usage.shared.end() || !it->second.count(f.shared[k].second)) fresh += 1;
    const size_t have = it == usage.shared.end() ? 0 : it->second.size();
    if (have + fresh > 1) return false;
    s = e;
  }
  return true;
}

void update_usage(Usage &usage, const Footprint &f, bool add) {
  auto apply = [&](std::map<uint64_t, std::map<uint64_t, uint64_t>> &m, const std::pair<uint64_t, uint64_t> &kv) {
    if (add) {
      m[kv.first][kv.second] += 1;
      return;
    }
    auto &inner = m[kv.first];
    if (--inner[kv.second] == 0) inner.erase(kv.second);
    if (inner.empty()) m.erase(kv.first);
  };
  for (const auto &kv : f.pin_wires) apply(usage.pin_wires, kv);
  for (const auto &kv : f.shared) apply(usage.shared, kv);
}

// PlacementOwners (repair.rs:194-283)
struct Owners {
  std::vector<int64_t> objects;
  std::map<uint64_t, std::map<uint64_t, std::set<Unit>>> pin_wires, shared;
  void insert(Unit u, const Footprint &f) {
    for (uint32_t o : f.objects) objects[o] = (int64_t)u;
    for (const auto &kv : f.pin_wires) pin_wires[kv.first][kv.second].insert(u);
    for (const auto &kv : f.shared) shared[kv.first][kv.second].insert(u);
  }
  void remove(Unit u, const Footprint &f) {
    for (uint32_t o : f.objects) objects[o] = -1;
    auto drop = [&](std::map<uint64_t, std::map<uint64_t, std::set<Unit>>> &m, const std::pair<uint64_t, uint64_t> &kv) {
      auto &inner = m[kv.first];
      auto &set = inner[kv.second];
      set.erase(u);
      if (set.empty()) inner.erase(kv.second);
      if (inner.empty()) m.erase(kv.first);
    };
    for (const auto &kv : f.pin_wires) drop(pin_wires, kv);
    for (const auto &kv : f.shared) drop(shared, kv);
  }
};

// initial_pending_units (repair.rs:322-425)
std::set<Unit> initial_pending_units(const Instance &in) {
  const auto *p = in.p;
  std::vector<std::vector<Unit>> object_owners(p->n_objects);
  std::map<uint64_t, std::map<uint64_t, std::set<Unit>>> pin_wires, shared;
  std::set<Unit> pending;
  for (Unit u = 0; u < in.n_units; ++u) {
    if (p->preferred[u] < 0) {
      pending.insert(u);
      continue;
    }
    const Footprint f = in.footprint(u, (Choice)p->preferred[u]);
    for (uint32_t o : f.objects) object_owners[o].push_back(u);
    for (const auto &kv : f.pin_wires) pin_wires[kv.first][kv.second].insert(u);
    for (const auto &kv : f.shared) shared[kv.first][kv.second].insert(u);
  }
  for (const auto &owners : object_owners) {
    if (owners.size() <= 1) continue;
    std::vector<Unit> fixed;
    for (Unit u : owners)
      if (p->unit_fixed[u]) fixed.push_back(u);
    if (fixed.size() > 1) throw ProjectionError{"multiple fixed units occupy one BEL during repair"};
    Unit survivor;
    if (!fixed.empty()) {
      survivor = fixed[0];
    } else {
      survivor = owners[0];
      for (Unit u : owners)
        if (in.survives_before(u, survivor)) survivor = u;
    }
    for (Unit u : owners)
      if (u != survivor) pending.insert(u);
  }
  for (const auto &res : shared) {
    const auto &values = res.second;
    if (values.size() <= 1) continue;
    std::vector<uint64_t> fixed_values;
    for (const auto &v : values)
      for (Unit u : v.second)
        if (p->unit_fixed[u]) {
          fixed_values.push_back(v.first);
          break;
        }
    if (fixed_values.size() > 1) throw ProjectionError{"fixed units require incompatible shared-resource values"};
    uint64_t survivor;
    if (!fixed_values.empty()) {
      survivor = fixed_values[0];
    } else {
      // min by (Reverse(owner count), value)
      survivor = values.begin()->first;
      size_t best = values.begin()->second.size();
      for (const auto &v : values)
        if (v.second.size() > best) {
          best = v.second.size();
          survivor = v.first;
        }
    }
    for (const auto &v : values)
      if (v.first != survivor) pending.insert(v.second.begin(), v.second.end());
  }
  for (const auto &wire : pin_wires) {
    const uint64_t capacity = p->wire_capacity[wire.first];
    const auto &nets = wire.second;
    if (nets.size() <= capacity) continue;
    struct Ranked {
      bool fixed;
      size_t owners;
      uint64_t net;
      const std::set<Unit> *set;
    };
    std::vector<Ranked> ranked;
    for (const auto &n : nets) {
      bool fixed = false;
      for (Unit u : n.second) fixed = fixed || p->unit_fixed[u];
      ranked.push_back({fixed, n.second.size(), n.first, &n.second});
    }
    // (Reverse(fixed), Reverse(owners), net)
    std::sort(ranked.begin(), ranked.end(), [](const Ranked &a, const Ranked &b) {
      if (a.fixed != b.fixed) return a.fixed;
      if (a.owners != b.owners) return a.owners > b.owners;
      return a.net < b.net;
    });
    size_t fixed_nets = 0;
    for (const auto &r : ranked) fixed_nets += r.fixed;
    if (fixed_nets > capacity) throw ProjectionError{"fixed units exceed a placement pin-wire capacity"};
    for (size_t k = capacity; k < ranked.size(); ++k) pending.insert(ranked[k].set->begin(), ranked[k].set->end());
  }
  return pending;
}

struct JournalEntry {
  bool installed;
  Unit unit;
  Choice choice;
};

// RepairState (repair.rs:432-635)
struct State {
  const Instance &in;
  std::vector<int64_t> choice_by_unit;  // -1: unplaced
  size_t unplaced_count;
  std::map<Unit, Choice> deviations;
  Usage usage;
  Owners owners;
  std::vector<JournalEntry> journal;

  explicit State(const Instance &i) : in(i), choice_by_unit(i.n_units, -1), unplaced_count(i.n_units) {
    owners.objects.assign(i.p->n_objects, -1);
    for (Unit u = 0; u < i.n_units; ++u)
      if (i.p->preferred[u] >= 0) deviations[u] = UNPLACED;
  }
  void record_choice(Unit u, int64_t c) {
    if (c == in.p->preferred[u])
      deviations.erase(u);
    else
      deviations[u] = c < 0 ? UNPLACED : (Choice)c;
  }
  bool can_install(Unit u, Choice c) const {
    const Footprint f = in.footprint(u, c);
    for (uint32_t o : f.objects)
      if (owners.objects[o] >= 0) return false;
    return lists_are_legal(in, usage, f);
  }
  void raw_install(Unit u, Choice c) {
    const Footprint f = in.footprint(u, c);
    owners.insert(u, f);
    update_usage(usage, f, true);
    choice_by_unit[u] = (int64_t)c;
    unplaced_count -= 1;
    record_choice(u, (int64_t)c);
  }
  void raw_remove(Unit u, Choice c) {
    const Footprint f = in.footprint(u, c);
    update_usage(usage, f, false);
    owners.remove(u, f);
    choice_by_unit[u] = -1;
    unplaced_count += 1;
    record_choice(u, -1);
  }
  void remove_for_branch(Unit u) {
    const Choice c = (Choice)choice_by_unit[u];
    raw_remove(u, c);
    journal.push_back({false, u, c});
  }
  void install_for_branch(Unit u, Choice c) {
    raw_install(u, c);
    journal.push_back({true, u, c});
  }
  void rollback(size_t checkpoint) {
    while (journal.size() > checkpoint) {
      const JournalEntry e = journal.back();
      journal.pop_back();
      if (e.installed)
        raw_remove(e.unit, e.choice);
      else
        raw_install(e.unit, e.choice);
    }
  }
  int64_t next_unplaced() const {
    int64_t best = -1;
    for (Unit u = 0; u < in.n_units; ++u)
      if (choice_by_unit[u] < 0 && (best < 0 || in.before(u, (Unit)best))) best = (int64_t)u;
    return best;
  }
  std::vector<std::pair<Unit, Choice>> key() const { return {deviations.begin(), deviations.end()}; }
};

// ChoiceCursor (repair.rs:649-793): the preferred choice, then the unit's
// choices on Manhattan rings around the target: radius 0, 1, ...; on a ring
// x ascending, at each x the point below the target (or on its row) before
// the point above; at one point in choice order.
struct Cursor {
  std::vector<Choice> order;
  size_t next = 0;
  Cursor(const Instance &in, Unit u) {
    const auto *p = in.p;
    const int64_t pref = p->preferred[u];
    if (pref >= 0) order.push_back((Choice)pref);
    const int64_t tx = p->unit_target_xy[2 * u], ty = p->unit_target_xy[2 * u + 1];
    struct Key {
      uint64_t radius, x, above, choice;
    };
    std::vector<Key> keys;
    for (Choice c = 0; c < in.choices(u); ++c) {
      if ((int64_t)c == pref) continue;
      const uint32_t o = p->cand_object[p->cand_object_start[in.cand(u, c)]];
      const int64_t x = p->object_xy[2 * (size_t)o], y = p->object_xy[2 * (size_t)o + 1];
      const uint64_t r = (uint64_t)(std::llabs(x - tx) + std::llabs(y - ty));
      keys.push_back({r, (uint64_t)x, (uint64_t)(y > ty), c});
    }
    std::sort(keys.begin(), keys.end(), [](const Key &a, const Key &b) {
      if (a.radius != b.radius) return a.radius < b.radius;
      if (a.x != b.x) return a.x < b.x;
      if (a.above != b.above) return a.above < b.above;
      return a.choice < b.choice;
    });
    for (const Key &k : keys) order.push_back(k.choice);
  }
  bool take(Choice &c) {
    if (next >= order.size()) return false;
    c = order[next++];
    return true;
  }
};

// enumerate_combinations (repair.rs:1244-1266)
std::vector<std::vector<size_t>> enumerate_combinations(size_t items, size_t choose) {
  if (choose > items) return {};
  if (choose == 0) return {std::vector<size_t>{}};
  std::vector<std::vector<size_t>> result;
  std::vector<size_t> cur(choose);
  for (size_t i = 0; i < choose; ++i) cur[i] = i;
  for (;;) {
    result.push_back(cur);
    int64_t pivot = -1;
    for (int64_t i = (int64_t)choose - 1; i >= 0; --i)
      if (cur[i] < items - choose + (size_t)i) {
        pivot = i;
        break;
      }
    if (pivot < 0) break;
    cur[pivot] += 1;
    for (size_t i = (size_t)pivot + 1; i < choose; ++i) cur[i] = cur[i - 1] + 1;
  }
  return result;
}

struct Stats {
  uint64_t initial_pending = 0, direct_fast_choices = 0, direct_fast_moves = 0, direct_fast_fallbacks = 0,
           augmenting_roots = 0, augmenting_fallbacks = 0, states_visited = 0, failed_states = 0, max_unplaced = 0,
           max_blocker_branches = 0, units_evicted = 0, choices_examined = 0;
};

using Key = std::vector<std::pair<Unit, Choice>>;

struct Search {
  State state;
  std::set<Key> visiting, failed;
  Stats stats;

  explicit Search(const Instance &in) : state(in) {}

  void record_search_frame() {
    stats.states_visited += 1;
    stats.max_unplaced = std::max<uint64_t>(stats.max_unplaced, state.unplaced_count);
  }

  // mandatory_blockers (repair.rs:1119-1167); false = the reference's None
  bool mandatory_blockers(Unit u, Choice c, const Footprint &f, std::set<Unit> &blockers) const {
    const auto *p = state.in.p;
    blockers.clear();
    for (uint32_t o : f.objects) {
      const int64_t owner = state.owners.objects[o];
      if (owner >= 0) {
        if (p->unit_fixed[owner]) return false;
        blockers.insert((Unit)owner);
      }
    }
    (void)u;
    (void)c;
    for (size_t s = 0; s < f.shared.size();) {
      const uint64_t res = f.shared[s].first;
      size_t e = s + 1;
      while (e < f.shared.size() && f.shared[e].first == res) ++e;
      if (e - s > 1) return false;  // the candidate itself needs two values
      const uint64_t value = f.shared[s].second;
      const auto it = state.owners.shared.find(res);
      if (it != state.owners.shared.end())
        for (const auto &kv : it->second) {
          if (kv.first == value) continue;
          for (Unit owner : kv.second) {
            if (p->unit_fixed[owner]) return false;
            blockers.insert(owner);
          }
        }
      s = e;
    }
    return true;
  }

  // pin_blocker_branches (repair.rs:1169-1241)
  std::vector<std::set<Unit>> pin_blocker_branches(const Footprint &f, const std::set<Unit> &mandatory) const {
    const auto *p = state.in.p;
    std::map<uint64_t, std::set<uint64_t>> candidate_by_wire;
    for (const auto &kv : f.pin_wires) candidate_by_wire[kv.first].insert(kv.second);
    std::vector<std::set<Unit>> branches{mandatory};
    for (const auto &w : candidate_by_wire) {
      const uint64_t wire = w.first;
      const std::set<uint64_t> &candidate_nets = w.second;
      const uint64_t capacity = p->wire_capacity[wire];
      if (candidate_nets.size() > capacity) return {};
      std::vector<std::set<Unit>> next;
      for (const auto &blockers : branches) {
        std::vector<std::pair<uint64_t, std::vector<Unit>>> effective;
        const auto ex = state.owners.pin_wires.find(wire);
        if (ex != state.owners.pin_wires.end())
          for (const auto &kv : ex->second) {
            std::vector<Unit> remaining;
            for (Unit o : kv.second)
              if (!blockers.count(o)) remaining.push_back(o);
            if (!remaining.empty()) effective.push_back({kv.first, remaining});
          }
        std::set<uint64_t> distinct(candidate_nets.begin(), candidate_nets.end());
        for (const auto &e : effective) distinct.insert(e.first);
        const size_t required = distinct.size() > capacity ? distinct.size() - capacity : 0;
        if (required == 0) {
          next.push_back(blockers);
          continue;
        }
        std::vector<std::pair<uint64_t, std::vector<Unit>>> removable;
        for (const auto &e : effective) {
          if (candidate_nets.count(e.first)) continue;
          bool any_fixed = false;
          for (Unit o : e.second) any_fixed = any_fixed || p->unit_fixed[o];
          if (!any_fixed) removable.push_back(e);
        }
        if (removable.size() < required) continue;
        for (const auto &combo : enumerate_combinations(removable.size(), required)) {
          std::set<Unit> branch = blockers;
          for (size_t i : combo) branch.insert(removable[i].second.begin(), removable[i].second.end());
          next.push_back(branch);
        }
      }
      std::sort(next.begin(), next.end());
      next.erase(std::unique(next.begin(), next.end()), next.end());
      branches = std::move(next);
      if (branches.empty()) break;
    }
    return branches;
  }

  // try_direct_empty_fast_path (repair.rs:873-904)
  bool direct_fast_path() {
    const size_t checkpoint = state.journal.size();
    for (;;) {
      const int64_t u = state.next_unplaced();
      if (u < 0) break;
      Cursor cursor(state.in, (Unit)u);
      Choice c;
      bool installed = false;
      while (cursor.take(c)) {
        stats.direct_fast_choices += 1;
        stats.choices_examined += 1;
        if (state.can_install((Unit)u, c)) {
          state.install_for_branch((Unit)u, c);
          stats.direct_fast_moves += 1;
          installed = true;
          break;
        }
      }
      if (installed) continue;
      state.rollback(checkpoint);
      stats.direct_fast_fallbacks += 1;
      return false;
    }
    state.journal.resize(checkpoint);
    return true;
  }

  // augment_one (repair.rs:931-1017)
  bool augment_one(Unit root) {
    const size_t root_checkpoint = state.journal.size();
    std::set<Unit> visited_units{root};
    std::set<std::pair<Unit, Choice>> visited_assignments;
    record_search_frame();
    struct Frame {
      Unit unit;
      Cursor choices;
      int64_t resume;  // -1: none
    };
    std::vector<Frame> stack;
    stack.push_back({root, Cursor(state.in, root), -1});
    std::set<Unit> mandatory;
    while (!stack.empty()) {
      Frame frame = std::move(stack.back());
      stack.pop_back();
      if (frame.resume >= 0) {
        state.rollback((size_t)frame.resume);
        frame.resume = -1;
      }
      Choice c;
      if (!frame.choices.take(c)) continue;
      if (!visited_assignments.insert({frame.unit, c}).second) {
        stack.push_back(std::move(frame));
        continue;
      }
      stats.choices_examined += 1;
      const Footprint f = state.in.footprint(frame.unit, c);
      if (!mandatory_blockers(frame.unit, c, f, mandatory)) {
        stack.push_back(std::move(frame));
        continue;
      }
      const auto branches = pin_blocker_branches(f, mandatory);
      stats.max_blocker_branches = std::max<uint64_t>(stats.max_blocker_branches, branches.size());
      if (branches.size() != 1 || branches[0].size() > 1) {
        stack.push_back(std::move(frame));
        continue;
      }
      const bool has_blocker = !branches[0].empty();
      const Unit blocker = has_blocker ? *branches[0].begin() : 0;
      if (has_blocker && visited_units.count(blocker)) {
        stack.push_back(std::move(frame));
        continue;
      }
      const size_t checkpoint = state.journal.size();
      if (has_blocker) {
        state.remove_for_branch(blocker);
        stats.units_evicted += 1;
      }
      if (!state.can_install(frame.unit, c)) {
        state.rollback(checkpoint);
        stack.push_back(std::move(frame));
        continue;
      }
      state.install_for_branch(frame.unit, c);
      if (!has_blocker) return true;
      visited_units.insert(blocker);
      record_search_frame();
      frame.resume = (int64_t)checkpoint;
      stack.push_back(std::move(frame));
      stack.push_back({blocker, Cursor(state.in, blocker), -1});
    }
    state.rollback(root_checkpoint);
    return false;
  }

  // try_single_blocker_fast_path (repair.rs:913-929)
  bool single_blocker_fast_path() {
    const size_t checkpoint = state.journal.size();
    for (;;) {
      const int64_t root = state.next_unplaced();
      if (root < 0) break;
      stats.augmenting_roots += 1;
      if (augment_one((Unit)root)) continue;
      state.rollback(checkpoint);
      stats.augmenting_fallbacks += 1;
      return false;
    }
    state.journal.resize(checkpoint);
    return true;
  }

  struct ExactFrame {
    Unit unit;
    Key key;
    Cursor choices;
    int64_t current_choice = -1;
    std::vector<std::set<Unit>> branches;
    size_t next_branch = 0;
    int64_t resume = -1;
  };

  // begin_exact_frame (repair.rs:1019-1036); false = None
  bool begin_exact_frame(std::vector<ExactFrame> &out) {
    const int64_t u = state.next_unplaced();
    if (u < 0) return false;
    Key key = state.key();
    if (failed.count(key) || !visiting.insert(key).second) return false;
    record_search_frame();
    ExactFrame frame{(Unit)u, std::move(key), Cursor(state.in, (Unit)u), -1, {}, 0, -1};
    out.push_back(std::move(frame));
    return true;
  }

  // next_exact_attempt (repair.rs:1038-1066)
  bool next_exact_attempt(ExactFrame &frame, Choice &choice, std::set<Unit> &blockers) {
    std::set<Unit> mandatory;
    for (;;) {
      if (frame.next_branch < frame.branches.size()) {
        blockers = frame.branches[frame.next_branch++];
        choice = (Choice)frame.current_choice;
        return true;
      }
      Choice c;
      if (!frame.choices.take(c)) return false;
      stats.choices_examined += 1;
      const Footprint f = state.in.footprint(frame.unit, c);
      if (!mandatory_blockers(frame.unit, c, f, mandatory)) continue;
      auto branches = pin_blocker_branches(f, mandatory);
      stats.max_blocker_branches = std::max<uint64_t>(stats.max_blocker_branches, branches.size());
      frame.current_choice = (int64_t)c;
      frame.branches = std::move(branches);
      frame.next_branch = 0;
    }
  }

  // solve_exact_iterative (repair.rs:1068-1117)
  bool solve_exact() {
    std::vector<ExactFrame> stack;
    if (!begin_exact_frame(stack)) return state.next_unplaced() < 0;
    while (!stack.empty()) {
      ExactFrame frame = std::move(stack.back());
      stack.pop_back();
      if (frame.resume >= 0) {
        state.rollback((size_t)frame.resume);
        frame.resume = -1;
      }
      Choice c;
      std::set<Unit> blockers;
      if (!next_exact_attempt(frame, c, blockers)) {
        visiting.erase(frame.key);
        failed.insert(frame.key);
        if (stack.empty()) return false;
        continue;
      }
      const size_t checkpoint = state.journal.size();
      for (Unit b : blockers)
        if (state.choice_by_unit[b] >= 0) {
          state.remove_for_branch(b);
          stats.units_evicted += 1;
        }
      if (state.can_install(frame.unit, c)) {
        state.install_for_branch(frame.unit, c);
        if (state.next_unplaced() < 0) {
          visiting.clear();
          return true;
        }
        std::vector<ExactFrame> child;
        if (begin_exact_frame(child)) {
          frame.resume = (int64_t)checkpoint;
          stack.push_back(std::move(frame));
          stack.push_back(std::move(child.back()));
          continue;
        }
      }
      state.rollback(checkpoint);
      stack.push_back(std::move(frame));
    }
    return false;
  }

  bool solve() {
    if (state.next_unplaced() < 0) return true;
    if (direct_fast_path()) return true;
    if (single_blocker_fast_path()) return true;
    return solve_exact();
  }
};

bool valid(const texo_gp_repair_problem *p) {
  if (!p || p->width == 0 || p->height == 0) return false;
  if (p->n_units && (!p->unit_cells || !p->unit_first_cell || !p->unit_target_xy || !p->unit_fixed ||
                     !p->preferred || !p->cand_start))
    return false;
  if (p->n_units == 0) return true;
  if (p->cand_start[0] != 0) return false;
  const uint64_t n_cand = p->cand_start[p->n_units];
  for (uint64_t u = 0; u < p->n_units; ++u) {
    if (p->cand_start[u + 1] < p->cand_start[u] || p->unit_cells[u] == 0) return false;
    if (p->unit_target_xy[2 * u] >= p->width || p->unit_target_xy[2 * u + 1] >= p->height) return false;
  }
  if (n_cand && (!p->cand_object_start || !p->cand_shared_start || !p->cand_wire_start)) return false;
  for (uint64_t c = 0; c < n_cand; ++c)
    if (p->cand_object_start[c + 1] < p->cand_object_start[c] || p->cand_shared_start[c + 1] < p->cand_shared_start[c] ||
        p->cand_wire_start[c + 1] < p->cand_wire_start[c])
      return false;
  for (uint64_t u = 0; u < p->n_units; ++u)
    for (uint64_t c = p->cand_start[u]; c < p->cand_start[u + 1]; ++c) {
      if (p->cand_object_start[c + 1] - p->cand_object_start[c] != p->unit_cells[u]) return false;
      for (uint64_t i = p->cand_object_start[c]; i < p->cand_object_start[c + 1]; ++i) {
        const uint32_t o = p->cand_object[i];
        if (o >= p->n_objects || p->object_xy[2 * (size_t)o] >= p->width || p->object_xy[2 * (size_t)o + 1] >= p->height)
          return false;
        for (uint64_t k = p->cand_object_start[c]; k < i; ++k)
          if (p->cand_object[k] == o) return false;
      }
      for (uint64_t i = p->cand_wire_start[c]; i < p->cand_wire_start[c + 1]; ++i)
        if (p->cand_wire[2 * i] >= p->n_wires) return false;
    }
  return true;
}

}  // namespace

extern "C" int oracle_repair(const texo_gp_repair_problem *p, int64_t *placed, uint64_t *stats_out,
                             int64_t *error_unit) {
  if (error_unit) *error_unit = -1;
  if (!valid(p) || (!placed && p->n_units)) return TEXO_GP_ERR_INVALID_ARGUMENT;
  Instance in{p, p->n_units};
  try {
    // repair_relaxed (repair.rs:60-155)
    for (Unit u = 0; u < in.n_units; ++u) {
      if (p->preferred[u] >= 0 && (uint64_t)p->preferred[u] >= in.choices(u))
        throw ProjectionError{"transactional repair received an out-of-range assignment"};
      if (p->unit_fixed[u] && p->preferred[u] < 0)
        throw ProjectionError{"transactional repair received an unmatched fixed unit"};
    }
    const std::set<Unit> pending = initial_pending_units(in);
    Search search(in);
    search.stats.initial_pending = pending.size();
    search.stats.max_unplaced = pending.size();
    for (int pass = 0; pass < 2; ++pass)  // fixed units first, then the movable ones
      for (Unit u = 0; u < in.n_units; ++u) {
        if ((p->unit_fixed[u] != 0) != (pass == 0)) continue;
        if (p->preferred[u] < 0 || pending.count(u)) continue;
        if (!search.state.can_install(u, (Choice)p->preferred[u]))
          throw ProjectionError{"relaxed-conflict removal did not leave a legal repair base"};
        search.state.raw_install(u, (Choice)p->preferred[u]);
      }
    const int64_t first_pending = search.state.next_unplaced();
    const bool ok = search.solve();
    search.stats.failed_states = search.failed.size();
    if (stats_out) {
      const Stats &s = search.stats;
      const uint64_t v[12] = {s.initial_pending,      s.direct_fast_choices, s.direct_fast_moves,
                              s.direct_fast_fallbacks, s.augmenting_roots,   s.augmenting_fallbacks,
                              s.states_visited,        s.failed_states,      s.max_unplaced,
                              s.max_blocker_branches,  s.units_evicted,      s.choices_examined};
      std::copy(v, v + 12, stats_out);
    }
    if (!ok) throw NoBel{first_pending < 0 ? 0 : (Unit)first_pending};
    for (Unit u = 0; u < in.n_units; ++u) placed[u] = search.state.choice_by_unit[u];
    return 0;
  } catch (const ProjectionError &) {
    return TEXO_GP_ERR_REPAIR_PROJECTION;
  } catch (const NoBel &e) {
    if (error_unit) *error_unit = (int64_t)e.unit;
    return TEXO_GP_ERR_REPAIR_NO_BEL;
  }
}

// pin_blocker_branches on a state with the given units installed, for the
// reference's tests of the branch enumeration (repair.rs:1755-1929): the
// branches of (unit, choice) with `mandatory` blockers, written as unit
// bitmasks (units < 64) into out[max_out]; returns the branch count.
extern "C" int oracle_repair_pin_branches(const texo_gp_repair_problem *p, uint64_t n_installed,
                                          const uint64_t *installed_unit, const uint64_t *installed_choice,
                                          uint64_t unit, uint64_t choice, uint64_t mandatory_mask, uint64_t *out,
                                          int max_out) {
  if (!valid(p)) return -1;
  Instance in{p, p->n_units};
  Search search(in);
  for (uint64_t i = 0; i < n_installed; ++i) search.state.raw_install(installed_unit[i], installed_choice[i]);
  std::set<Unit> mandatory;
  for (Unit u = 0; u < 64 && u < in.n_units; ++u)
    if (mandatory_mask >> u & 1u) mandatory.insert(u);
  const auto branches = search.pin_blocker_branches(in.footprint(unit, choice), mandatory);
  int k = 0;
  for (const auto &b : branches) {
    if (k >= max_out) break;
    uint64_t m = 0;
    for (Unit u : b) m |= uint64_t{1} << u;
    out[k++] = m;
  }
  return (int)branches.size();
}

// ChoiceCursor order of `unit` (repair.rs:1319-1345 tests it): writes up to
// max_out choices, returns their number.
extern "C" int oracle_repair_cursor(const texo_gp_repair_problem *p, uint64_t unit, uint64_t *out, int max_out) {
  if (!valid(p) || unit >= p->n_units) return -1;
  Instance in{p, p->n_units};
  Cursor cursor(in, unit);
  int k = 0;
  Choice c;
  while (k < max_out && cursor.take(c)) out[k++] = c;
  return k;
}
