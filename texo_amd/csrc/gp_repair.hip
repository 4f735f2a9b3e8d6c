// gp_repair.hip — the transactional legalization repair on the host
// (repair_relaxed, crates/texo-pnr/src/legalization/repair.rs:60-155).
//
// After the relaxed auction, legalization::project (legalization/mod.rs:
// 377-386) repairs the conflicting part of the proposal with a serial search:
// greedy placement on free candidates, single-blocker alternating paths, and
// a complete transactional search with memoised failed states.  Its control
// flow is inherently sequential, so it runs on the host here, over explicit
// candidate footprints (the reference derives them from the device graph,
// assignment_footprint, repair.rs:162-192).  Data layout: every candidate's
// sorted footprint is flattened once; ownership is kept per object, per
// wire (its nets and their owners) and per shared resource (its values and
// their owners); the unplaced units sit in an ordered set keyed by the
// reference's next_unplaced order; each unit's candidate order (preferred,
// then Manhattan rings around the target, ChoiceCursor :649-793) is computed
// once and cached; the transactional search keys states by their sorted
// deviations from the proposal and hashes them.

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <set>
#include <tuple>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "gp_host.h"

namespace texo {
namespace {

using Pair = std::pair<uint64_t, uint64_t>;
constexpr uint64_t NONE = UINT64_MAX;

struct RepairProjection {
  const char *reason;
};
struct RepairNoBel {
  uint64_t unit;
};

// Flattened, sorted and deduplicated footprints of every candidate.
struct Footprints {
  std::vector<uint64_t> obj_start, wire_start, shared_start;
  std::vector<uint32_t> obj;
  std::vector<Pair> wire, shared;
  explicit Footprints(const texo_gp_repair_problem &p) {
    const uint64_t n = p.n_units ? p.cand_start[p.n_units] : 0;
    obj_start.assign(1, 0);
    wire_start.assign(1, 0);
    shared_start.assign(1, 0);
    for (uint64_t c = 0; c < n; ++c) {
      for (uint64_t i = p.cand_object_start[c]; i < p.cand_object_start[c + 1]; ++i) obj.push_back(p.cand_object[i]);
      obj_start.push_back(obj.size());
      const size_t w0 = wire.size();
      for (uint64_t i = p.cand_wire_start[c]; i < p.cand_wire_start[c + 1]; ++i)
        wire.push_back({p.cand_wire[2 * i], p.cand_wire[2 * i + 1]});
      std::sort(wire.begin() + w0, wire.end());
      wire.erase(std::unique(wire.begin() + w0, wire.end()), wire.end());
      wire_start.push_back(wire.size());
      const size_t s0 = shared.size();
      for (uint64_t i = p.cand_shared_start[c]; i < p.cand_shared_start[c + 1]; ++i)
        shared.push_back({p.cand_shared[2 * i], p.cand_shared[2 * i + 1]});
      std::sort(shared.begin() + s0, shared.end());
      shared.erase(std::unique(shared.begin() + s0, shared.end()), shared.end());
      shared_start.push_back(shared.size());
    }
  }
};

// Owners of one value of a wire (a net) or of a shared resource (a value),
// kept sorted by unit.
struct Holders {
  uint64_t key;
  std::vector<uint64_t> units;
};
using HolderMap = std::unordered_map<uint64_t, std::vector<Holders>>;  // sorted by key inside

std::vector<Holders> *find(HolderMap &m, uint64_t id) {
  auto it = m.find(id);
  return it == m.end() ? nullptr : &it->second;
}
const std::vector<Holders> *find(const HolderMap &m, uint64_t id) {
  auto it = m.find(id);
  return it == m.end() ? nullptr : &it->second;
}
void hold(HolderMap &m, uint64_t id, uint64_t key, uint64_t unit) {
  auto &v = m[id];
  auto it = std::lower_bound(v.begin(), v.end(), key, [](const Holders &h, uint64_t k) { return h.key < k; });
  if (it == v.end() || it->key != key) it = v.insert(it, Holders{key, {}});
  it->units.insert(std::lower_bound(it->units.begin(), it->units.end(), unit), unit);
}
void release(HolderMap &m, uint64_t id, uint64_t key, uint64_t unit) {
  auto mi = m.find(id);
  auto &v = mi->second;
  auto it = std::lower_bound(v.begin(), v.end(), key, [](const Holders &h, uint64_t k) { return h.key < k; });
  it->units.erase(std::lower_bound(it->units.begin(), it->units.end(), unit));
  if (it->units.empty()) v.erase(it);
  if (v.empty()) m.erase(mi);
}
bool holds(const std::vector<Holders> *v, uint64_t key) {
  if (!v) return false;
  auto it = std::lower_bound(v->begin(), v->end(), key, [](const Holders &h, uint64_t k) { return h.key < k; });
  return it != v->end() && it->key == key;
}

struct KeyHash {
  size_t operator()(const std::vector<Pair> &k) const {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ k.size();
    for (const Pair &p : k) {
      h ^= p.first + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
      h ^= p.second + 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    }
    return (size_t)h;
  }
};

struct KeyHashPair {
  size_t operator()(const Pair &p) const { return (size_t)(p.first * 0x9E3779B97F4A7C15ull ^ (p.second + 0x632BE59BD9B4E019ull)); }
};

class Repair {
 public:
  explicit Repair(const texo_gp_repair_problem &p) : p_(p), fp_(p) {
    choice_.assign(p.n_units, NONE);
    object_owner_.assign(p.n_objects, NONE);
    order_.resize(p.n_units);
    for (uint64_t u = 0; u < p.n_units; ++u) {
      unplaced_.insert(rank(u));
      if (p.preferred[u] >= 0) deviation_.insert({u, NONE});
    }
  }

  uint64_t stats[12] = {};  // the reference's RepairStats, in its field order

  std::set<uint64_t> initial_pending() const;
  void place_proposal(const std::set<uint64_t> &pending);
  bool solve();
  uint64_t first_unplaced() const { return unplaced_.empty() ? NONE : std::get<3>(*unplaced_.begin()); }
  uint64_t choice(uint64_t u) const { return choice_[u]; }
  size_t failed_states() const { return failed_.size(); }

 private:
  using Rank = std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>;  // choices, -cells, first cell, unit
  struct Frame {
    uint64_t unit;
    size_t next = 0;  // position in the unit's candidate order
    size_t resume = SIZE_MAX;
  };
  struct ExactFrame {
    uint64_t unit;
    std::vector<Pair> key;
    size_t next = 0;
    uint64_t current = NONE;
    std::vector<std::vector<uint64_t>> branches;
    size_t next_branch = 0;
    size_t resume = SIZE_MAX;
  };
  struct Entry {
    bool installed;
    uint64_t unit, choice;
  };

  const texo_gp_repair_problem &p_;
  Footprints fp_;
  std::vector<uint64_t> choice_, object_owner_;
  HolderMap wires_, shared_;
  std::set<Rank> unplaced_;
  std::set<Pair> deviation_;  // (unit, choice or NONE) off the proposal
  std::vector<Entry> journal_;
  std::vector<std::vector<uint64_t>> order_;  // cached candidate orders
  std::unordered_set<std::vector<Pair>, KeyHash> visiting_, failed_;

  uint64_t choices(uint64_t u) const { return p_.cand_start[u + 1] - p_.cand_start[u]; }
  uint64_t cand(uint64_t u, uint64_t c) const { return p_.cand_start[u] + c; }
  Rank rank(uint64_t u) const { return Rank{choices(u), UINT64_MAX - p_.unit_cells[u], p_.unit_first_cell[u], u}; }
  bool fixed(uint64_t u) const { return p_.unit_fixed[u] != 0; }

  const std::vector<uint64_t> &order(uint64_t u);
  bool legal(uint64_t u, uint64_t c) const;
  bool can_install(uint64_t u, uint64_t c) const;
  void set_choice(uint64_t u, uint64_t c);
  void raw_install(uint64_t u, uint64_t c);
  void raw_remove(uint64_t u);
  void evict(uint64_t u) {
    journal_.push_back({false, u, choice_[u]});
    raw_remove(u);
  }
  void install(uint64_t u, uint64_t c) {
    raw_install(u, c);
    journal_.push_back({true, u, c});
  }
  void rollback(size_t checkpoint);
  void frame_visited() {
    stats[6] += 1;
    stats[8] = std::max<uint64_t>(stats[8], unplaced_.size());
  }
  bool blockers(uint64_t u, uint64_t c, std::vector<uint64_t> &out) const;
  std::vector<std::vector<uint64_t>> branches(uint64_t u, uint64_t c, const std::vector<uint64_t> &mandatory) const;
  bool direct();
  bool augment(uint64_t root);
  bool single_blocker();
  bool begin_exact(std::vector<ExactFrame> &stack);
  bool next_attempt(ExactFrame &f, uint64_t &c, std::vector<uint64_t> &out);
  bool exact();
};

// Preferred first, then the unit's candidates by (Manhattan radius from the
// target, x, below-or-on-row before above, candidate index).
const std::vector<uint64_t> &Repair::order(uint64_t u) {
  std::vector<uint64_t> &o = order_[u];
  if (!o.empty() || choices(u) == 0) return o;
  const int64_t pref = p_.preferred[u];
  const int64_t tx = p_.unit_target_xy[2 * u], ty = p_.unit_target_xy[2 * u + 1];
  std::vector<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>> keys;
  keys.reserve(choices(u));
  for (uint64_t c = 0; c < choices(u); ++c) {
    if ((int64_t)c == pref) continue;
    const uint32_t first = fp_.obj[fp_.obj_start[cand(u, c)]];
    const int64_t x = p_.object_xy[2 * (size_t)first], y = p_.object_xy[2 * (size_t)first + 1];
    keys.emplace_back((uint64_t)(std::llabs(x - tx) + std::llabs(y - ty)), (uint64_t)x, (uint64_t)(y > ty), c);
  }
  std::sort(keys.begin(), keys.end());
  if (pref >= 0) o.push_back((uint64_t)pref);
  for (const auto &k : keys) o.push_back(std::get<3>(k));
  return o;
}

// At most `capacity` distinct nets per wire and one value per shared
// resource after adding candidate c (legalization/mod.rs:1080-1128).
bool Repair::legal(uint64_t u, uint64_t c) const {
  const uint64_t k = cand(u, c);
  for (uint64_t s = fp_.wire_start[k], e; s < fp_.wire_start[k + 1]; s = e) {
    const uint64_t w = fp_.wire[s].first;
    const auto *held = find(wires_, w);
    uint64_t fresh = 0;
    for (e = s; e < fp_.wire_start[k + 1] && fp_.wire[e].first == w; ++e) fresh += !holds(held, fp_.wire[e].second);
    if ((held ? held->size() : 0) + fresh > p_.wire_capacity[w]) return false;
  }
  for (uint64_t s = fp_.shared_start[k], e; s < fp_.shared_start[k + 1]; s = e) {
    const uint64_t r = fp_.shared[s].first;
    const auto *held = find(shared_, r);
    uint64_t fresh = 0;
    for (e = s; e < fp_.shared_start[k + 1] && fp_.shared[e].first == r; ++e) fresh += !holds(held, fp_.shared[e].second);
    if ((held ? held->size() : 0) + fresh > 1) return false;
  }
  return true;
}

bool Repair::can_install(uint64_t u, uint64_t c) const {
  const uint64_t k = cand(u, c);
  for (uint64_t i = fp_.obj_start[k]; i < fp_.obj_start[k + 1]; ++i)
    if (object_owner_[fp_.obj[i]] != NONE) return false;
  return legal(u, c);
}

void Repair::set_choice(uint64_t u, uint64_t c) {
  // record_choice (repair.rs:478-485): the key holds the units off the proposal
  const uint64_t pref = p_.preferred[u] < 0 ? NONE : (uint64_t)p_.preferred[u];
  const uint64_t old = choice_[u];
  if (old != pref) deviation_.erase({u, old});
  choice_[u] = c;
  if (c != pref) deviation_.insert({u, c});
}

void Repair::raw_install(uint64_t u, uint64_t c) {
  const uint64_t k = cand(u, c);
  for (uint64_t i = fp_.obj_start[k]; i < fp_.obj_start[k + 1]; ++i) object_owner_[fp_.obj[i]] = u;
  for (uint64_t i = fp_.wire_start[k]; i < fp_.wire_start[k + 1]; ++i) hold(wires_, fp_.wire[i].first, fp_.wire[i].second, u);
  for (uint64_t i = fp_.shared_start[k]; i < fp_.shared_start[k + 1]; ++i)
    hold(shared_, fp_.shared[i].first, fp_.shared[i].second, u);
  unplaced_.erase(rank(u));
  set_choice(u, c);
}

void Repair::raw_remove(uint64_t u) {
  const uint64_t k = cand(u, choice_[u]);
  for (uint64_t i = fp_.obj_start[k]; i < fp_.obj_start[k + 1]; ++i) object_owner_[fp_.obj[i]] = NONE;
  for (uint64_t i = fp_.wire_start[k]; i < fp_.wire_start[k + 1]; ++i)
    release(wires_, fp_.wire[i].first, fp_.wire[i].second, u);
  for (uint64_t i = fp_.shared_start[k]; i < fp_.shared_start[k + 1]; ++i)
    release(shared_, fp_.shared[i].first, fp_.shared[i].second, u);
  unplaced_.insert(rank(u));
  set_choice(u, NONE);
}

void Repair::rollback(size_t checkpoint) {
  while (journal_.size() > checkpoint) {
    const Entry e = journal_.back();
    journal_.pop_back();
    if (e.installed)
      raw_remove(e.unit);
    else
      raw_install(e.unit, e.choice);
  }
}

// initial_pending_units (repair.rs:322-425): the proposal's conflicts keep a
// deterministic legal subset.
std::set<uint64_t> Repair::initial_pending() const {
  std::set<uint64_t> pending;
  std::vector<std::vector<uint64_t>> on_object(p_.n_objects);
  HolderMap wires, shared;
  for (uint64_t u = 0; u < p_.n_units; ++u) {
    if (p_.preferred[u] < 0) {
      pending.insert(u);
      continue;
    }
    const uint64_t k = cand(u, (uint64_t)p_.preferred[u]);
    for (uint64_t i = fp_.obj_start[k]; i < fp_.obj_start[k + 1]; ++i) on_object[fp_.obj[i]].push_back(u);
    for (uint64_t i = fp_.wire_start[k]; i < fp_.wire_start[k + 1]; ++i) hold(wires, fp_.wire[i].first, fp_.wire[i].second, u);
    for (uint64_t i = fp_.shared_start[k]; i < fp_.shared_start[k + 1]; ++i)
      hold(shared, fp_.shared[i].first, fp_.shared[i].second, u);
  }
  for (const auto &owners : on_object) {
    if (owners.size() < 2) continue;
    uint64_t survivor = NONE, n_fixed = 0;
    for (uint64_t u : owners)
      if (fixed(u)) {
        n_fixed += 1;
        if (survivor == NONE) survivor = u;
      }
    if (n_fixed > 1) throw RepairProjection{"multiple fixed units occupy one BEL during repair"};
    if (survivor == NONE)
      for (uint64_t u : owners) {
        // (choices, Reverse(cells), unit)
        if (survivor == NONE || std::make_tuple(choices(u), UINT64_MAX - p_.unit_cells[u], u) <
                                    std::make_tuple(choices(survivor), UINT64_MAX - p_.unit_cells[survivor], survivor))
          survivor = u;
      }
    for (uint64_t u : owners)
      if (u != survivor) pending.insert(u);
  }
  for (const auto &res : shared) {
    const auto &values = res.second;
    if (values.size() < 2) continue;
    uint64_t keep = NONE, n_fixed = 0;
    for (const Holders &h : values)
      for (uint64_t u : h.units)
        if (fixed(u)) {
          n_fixed += 1;
          keep = h.key;
          break;
        }
    if (n_fixed > 1) throw RepairProjection{"fixed units require incompatible shared-resource values"};
    if (keep == NONE) {
      // most owners, then the smallest value
      size_t most = 0;
      for (const Holders &h : values)
        if (h.units.size() > most) {
          most = h.units.size();
          keep = h.key;
        }
    }
    for (const Holders &h : values)
      if (h.key != keep) pending.insert(h.units.begin(), h.units.end());
  }
  for (const auto &w : wires) {
    const uint64_t capacity = p_.wire_capacity[w.first];
    if (w.second.size() <= capacity) continue;
    std::vector<std::tuple<int, uint64_t, uint64_t, const Holders *>> ranked;  // (!fixed, -owners, net)
    uint64_t fixed_nets = 0;
    for (const Holders &h : w.second) {
      bool any = false;
      for (uint64_t u : h.units) any = any || fixed(u);
      fixed_nets += any;
      ranked.emplace_back(any ? 0 : 1, UINT64_MAX - h.units.size(), h.key, &h);
    }
    if (fixed_nets > capacity) throw RepairProjection{"fixed units exceed a placement pin-wire capacity"};
    std::sort(ranked.begin(), ranked.end());
    for (size_t i = capacity; i < ranked.size(); ++i) {
      const Holders *h = std::get<3>(ranked[i]);
      pending.insert(h->units.begin(), h->units.end());
    }
  }
  return pending;
}

void Repair::place_proposal(const std::set<uint64_t> &pending) {
  for (int pass = 0; pass < 2; ++pass)  // fixed units first
    for (uint64_t u = 0; u < p_.n_units; ++u) {
      if (fixed(u) != (pass == 0) || p_.preferred[u] < 0 || pending.count(u)) continue;
      if (!can_install(u, (uint64_t)p_.preferred[u]))
        throw RepairProjection{"relaxed-conflict removal did not leave a legal repair base"};
      raw_install(u, (uint64_t)p_.preferred[u]);
    }
}

// mandatory_blockers (repair.rs:1119-1167): the owners of the candidate's
// objects and of conflicting values of its shared resources; false when one
// of them is fixed or the candidate needs two values of one resource.
bool Repair::blockers(uint64_t u, uint64_t c, std::vector<uint64_t> &out) const {
  out.clear();
  const uint64_t k = cand(u, c);
  for (uint64_t i = fp_.obj_start[k]; i < fp_.obj_start[k + 1]; ++i) {
    const uint64_t owner = object_owner_[fp_.obj[i]];
    if (owner == NONE) continue;
    if (fixed(owner)) return false;
    out.push_back(owner);
  }
  for (uint64_t s = fp_.shared_start[k], e; s < fp_.shared_start[k + 1]; s = e) {
    const uint64_t r = fp_.shared[s].first;
    for (e = s; e < fp_.shared_start[k + 1] && fp_.shared[e].first == r; ++e) {
    }
    if (e - s > 1) return false;
    if (const auto *held = find(shared_, r))
      for (const Holders &h : *held) {
        if (h.key == fp_.shared[s].second) continue;
        for (uint64_t owner : h.units) {
          if (fixed(owner)) return false;
          out.push_back(owner);
        }
      }
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return true;
}

// pin_blocker_branches (repair.rs:1169-1241): every way of evicting whole
// net classes from the candidate's over-full wires.
std::vector<std::vector<uint64_t>> Repair::branches(uint64_t u, uint64_t c, const std::vector<uint64_t> &mandatory) const {
  const uint64_t k = cand(u, c);
  std::vector<std::vector<uint64_t>> out{mandatory};
  for (uint64_t s = fp_.wire_start[k], e; s < fp_.wire_start[k + 1] && !out.empty(); s = e) {
    const uint64_t w = fp_.wire[s].first;
    std::vector<uint64_t> nets;
    for (e = s; e < fp_.wire_start[k + 1] && fp_.wire[e].first == w; ++e) nets.push_back(fp_.wire[e].second);
    const uint64_t capacity = p_.wire_capacity[w];
    if (nets.size() > capacity) return {};
    const auto *held = find(wires_, w);
    std::vector<std::vector<uint64_t>> next;
    for (const auto &taken : out) {
      // the wire's nets whose owners are not all evicted already
      std::vector<std::pair<uint64_t, std::vector<uint64_t>>> live;
      if (held)
        for (const Holders &h : *held) {
          std::vector<uint64_t> rest;
          for (uint64_t o : h.units)
            if (!std::binary_search(taken.begin(), taken.end(), o)) rest.push_back(o);
          if (!rest.empty()) live.push_back({h.key, std::move(rest)});
        }
      size_t distinct = nets.size();
      for (const auto &l : live) distinct += !std::binary_search(nets.begin(), nets.end(), l.first);
      const size_t need = distinct > capacity ? distinct - capacity : 0;
      if (need == 0) {
        next.push_back(taken);
        continue;
      }
      std::vector<const std::vector<uint64_t> *> removable;
      for (const auto &l : live) {
        if (std::binary_search(nets.begin(), nets.end(), l.first)) continue;
        bool any = false;
        for (uint64_t o : l.second) any = any || fixed(o);
        if (!any) removable.push_back(&l.second);
      }
      if (removable.size() < need) continue;
      // combinations of `need` of them in lexicographic order (:1244-1266)
      std::vector<size_t> pick(need);
      for (size_t i = 0; i < need; ++i) pick[i] = i;
      for (;;) {
        std::vector<uint64_t> b = taken;
        for (size_t i : pick) b.insert(b.end(), removable[i]->begin(), removable[i]->end());
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        next.push_back(std::move(b));
        int64_t pivot = -1;
        for (int64_t i = (int64_t)need - 1; i >= 0; --i)
          if (pick[i] < removable.size() - need + (size_t)i) {
            pivot = i;
            break;
          }
        if (pivot < 0) break;
        pick[pivot] += 1;
        for (size_t i = (size_t)pivot + 1; i < need; ++i) pick[i] = pick[i - 1] + 1;
      }
    }
    std::sort(next.begin(), next.end());
    next.erase(std::unique(next.begin(), next.end()), next.end());
    out = std::move(next);
  }
  return out;
}

// try_direct_empty_fast_path (repair.rs:873-904)
bool Repair::direct() {
  const size_t checkpoint = journal_.size();
  while (!unplaced_.empty()) {
    const uint64_t u = first_unplaced();
    bool placed = false;
    for (uint64_t c : order(u)) {
      stats[1] += 1;
      stats[11] += 1;
      if (can_install(u, c)) {
        install(u, c);
        stats[2] += 1;
        placed = true;
        break;
      }
    }
    if (placed) continue;
    rollback(checkpoint);
    stats[3] += 1;
    return false;
  }
  journal_.resize(checkpoint);
  return true;
}

// augment_one (repair.rs:931-1017): an alternating path in which every
// candidate displaces at most one unit.
bool Repair::augment(uint64_t root) {
  const size_t root_checkpoint = journal_.size();
  std::unordered_set<uint64_t> seen_units{root};
  std::unordered_set<Pair, KeyHashPair> seen_pairs;
  frame_visited();
  std::vector<Frame> stack{Frame{root}};
  std::vector<uint64_t> mandatory;
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    if (f.resume != SIZE_MAX) {
      rollback(f.resume);
      f.resume = SIZE_MAX;
    }
    const std::vector<uint64_t> &ord = order(f.unit);
    if (f.next >= ord.size()) continue;
    const uint64_t c = ord[f.next++];
    if (!seen_pairs.insert({f.unit, c}).second) {
      stack.push_back(f);
      continue;
    }
    stats[11] += 1;
    if (!blockers(f.unit, c, mandatory)) {
      stack.push_back(f);
      continue;
    }
    const auto br = branches(f.unit, c, mandatory);
    stats[9] = std::max<uint64_t>(stats[9], br.size());
    if (br.size() != 1 || br[0].size() > 1) {
      stack.push_back(f);
      continue;
    }
    const uint64_t blocker = br[0].empty() ? NONE : br[0][0];
    if (blocker != NONE && seen_units.count(blocker)) {
      stack.push_back(f);
      continue;
    }
    const size_t checkpoint = journal_.size();
    if (blocker != NONE) {
      evict(blocker);
      stats[10] += 1;
    }
    if (!can_install(f.unit, c)) {
      rollback(checkpoint);
      stack.push_back(f);
      continue;
    }
    install(f.unit, c);
    if (blocker == NONE) return true;
    seen_units.insert(blocker);
    frame_visited();
    f.resume = checkpoint;
    stack.push_back(f);
    stack.push_back(Frame{blocker});
  }
  rollback(root_checkpoint);
  return false;
}

// try_single_blocker_fast_path (repair.rs:913-929)
bool Repair::single_blocker() {
  const size_t checkpoint = journal_.size();
  while (!unplaced_.empty()) {
    stats[4] += 1;
    if (augment(first_unplaced())) continue;
    rollback(checkpoint);
    stats[5] += 1;
    return false;
  }
  journal_.resize(checkpoint);
  return true;
}

bool Repair::begin_exact(std::vector<ExactFrame> &stack) {
  if (unplaced_.empty()) return false;
  std::vector<Pair> key(deviation_.begin(), deviation_.end());
  if (failed_.count(key) || !visiting_.insert(key).second) return false;
  frame_visited();
  ExactFrame f;
  f.unit = first_unplaced();
  f.key = std::move(key);
  stack.push_back(std::move(f));
  return true;
}

bool Repair::next_attempt(ExactFrame &f, uint64_t &c, std::vector<uint64_t> &out) {
  std::vector<uint64_t> mandatory;
  for (;;) {
    if (f.next_branch < f.branches.size()) {
      out = f.branches[f.next_branch++];
      c = f.current;
      return true;
    }
    const std::vector<uint64_t> &ord = order(f.unit);
    if (f.next >= ord.size()) return false;
    const uint64_t cc = ord[f.next++];
    stats[11] += 1;
    if (!blockers(f.unit, cc, mandatory)) continue;
    auto br = branches(f.unit, cc, mandatory);
    stats[9] = std::max<uint64_t>(stats[9], br.size());
    f.current = cc;
    f.branches = std::move(br);
    f.next_branch = 0;
  }
}

// solve_exact_iterative (repair.rs:1068-1117)
bool Repair::exact() {
  std::vector<ExactFrame> stack;
  if (!begin_exact(stack)) return unplaced_.empty();
  std::vector<uint64_t> evict_list;
  while (!stack.empty()) {
    ExactFrame f = std::move(stack.back());
    stack.pop_back();
    if (f.resume != SIZE_MAX) {
      rollback(f.resume);
      f.resume = SIZE_MAX;
    }
    uint64_t c;
    if (!next_attempt(f, c, evict_list)) {
      visiting_.erase(f.key);
      failed_.insert(std::move(f.key));
      if (stack.empty()) return false;
      continue;
    }
    const size_t checkpoint = journal_.size();
    for (uint64_t b : evict_list)
      if (choice_[b] != NONE) {
        evict(b);
        stats[10] += 1;
      }
    if (can_install(f.unit, c)) {
      install(f.unit, c);
      if (unplaced_.empty()) {
        visiting_.clear();
        return true;
      }
      if (begin_exact(stack)) {
        ExactFrame child = std::move(stack.back());
        stack.pop_back();
        f.resume = checkpoint;
        stack.push_back(std::move(f));
        stack.push_back(std::move(child));
        continue;
      }
    }
    rollback(checkpoint);
    stack.push_back(std::move(f));
  }
  return false;
}

bool Repair::solve() {
  if (unplaced_.empty()) return true;
  if (direct()) return true;
  if (single_blocker()) return true;
  return exact();
}

bool well_formed(const texo_gp_repair_problem *p) {
  if (!p || p->width == 0 || p->height == 0) return false;
  if (p->n_units == 0) return true;
  if (!p->unit_cells || !p->unit_first_cell || !p->unit_target_xy || !p->unit_fixed || !p->preferred ||
      !p->cand_start || p->cand_start[0] != 0)
    return false;
  const uint64_t n = p->cand_start[p->n_units];
  if (n && (!p->cand_object_start || !p->cand_shared_start || !p->cand_wire_start || !p->object_xy)) return false;
  for (uint64_t u = 0; u < p->n_units; ++u) {
    if (p->cand_start[u + 1] < p->cand_start[u] || p->unit_cells[u] == 0) return false;
    if (p->unit_target_xy[2 * u] >= p->width || p->unit_target_xy[2 * u + 1] >= p->height) return false;
    for (uint64_t c = p->cand_start[u]; c < p->cand_start[u + 1]; ++c) {
      if (p->cand_object_start[c + 1] < p->cand_object_start[c] || p->cand_shared_start[c + 1] < p->cand_shared_start[c] ||
          p->cand_wire_start[c + 1] < p->cand_wire_start[c])
        return false;
      if (p->cand_object_start[c + 1] - p->cand_object_start[c] != p->unit_cells[u]) return false;
      for (uint64_t i = p->cand_object_start[c]; i < p->cand_object_start[c + 1]; ++i) {
        const uint32_t o = p->cand_object[i];
        if (o >= p->n_objects || p->object_xy[2 * (size_t)o] >= p->width || p->object_xy[2 * (size_t)o + 1] >= p->height)
          return false;
        for (uint64_t j = p->cand_object_start[c]; j < i; ++j)
          if (p->cand_object[j] == o) return false;
      }
      for (uint64_t i = p->cand_wire_start[c]; i < p->cand_wire_start[c + 1]; ++i)
        if (p->cand_wire[2 * i] >= p->n_wires) return false;
    }
  }
  return true;
}

}  // namespace
}  // namespace texo

extern "C" int texo_gp_repair(const texo_gp_repair_problem *problem, int64_t *placed, uint64_t *stats,
                              int64_t *error_unit) {
  using namespace texo;
  if (error_unit) *error_unit = -1;
  try {
    if (!well_formed(problem) || (problem->n_units && !placed)) {
      g_last_error = "texo_gp_repair: malformed problem arrays";
      return TEXO_GP_ERR_INVALID_ARGUMENT;
    }
    const texo_gp_repair_problem &p = *problem;
    for (uint64_t u = 0; u < p.n_units; ++u) {
      if (p.preferred[u] >= 0 && (uint64_t)p.preferred[u] >= p.cand_start[u + 1] - p.cand_start[u])
        throw RepairProjection{"transactional repair received an out-of-range assignment"};
      if (p.unit_fixed[u] && p.preferred[u] < 0) throw RepairProjection{"transactional repair received an unmatched fixed unit"};
    }
    Repair r(p);
    const std::set<uint64_t> pending = r.initial_pending();
    r.stats[0] = pending.size();
    r.stats[8] = pending.size();
    r.place_proposal(pending);
    const uint64_t first = r.first_unplaced();
    const bool ok = r.solve();
    r.stats[7] = r.failed_states();
    if (stats) std::copy(r.stats, r.stats + 12, stats);
    if (!ok) throw RepairNoBel{first == NONE ? 0 : first};
    for (uint64_t u = 0; u < p.n_units; ++u) placed[u] = (int64_t)r.choice(u);
    return 0;
  } catch (const RepairProjection &e) {
    g_last_error = std::string("texo_gp_repair: ") + e.reason;
    return TEXO_GP_ERR_REPAIR_PROJECTION;
  } catch (const RepairNoBel &e) {
    if (error_unit) *error_unit = (int64_t)e.unit;
    g_last_error = "texo_gp_repair: no complete legal assignment";
    return TEXO_GP_ERR_REPAIR_NO_BEL;
  } catch (const std::bad_alloc &) {
    g_last_error = "texo_gp_repair: host memory exhausted";
    return TEXO_GP_ERR_OUT_OF_MEMORY;
  } catch (...) {
    g_last_error = "texo_gp_repair: unexpected failure";
    return TEXO_GP_ERR_DEVICE;
  }
}
