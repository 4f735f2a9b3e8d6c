// gp_auction.hip — the legalization auction (assign_implicitly,
// crates/texo-pnr/src/legalization/auction.rs:123-253) as a parallel
// (Jacobi) auction on the GPU, over explicit domains.
//
// The reference bids one bidder at a time from a FIFO queue.  Here every
// unassigned bidder bids in the same round at the same prices: it finds its
// two cheapest objects (reduced cost = base cost + price, ties by object),
// its private dummy included at the reference's cost n * max + n + 1, and
// bids second - best + 1 above its best object's price, as the reference's
// bidder does.  Each object takes its highest bid, ties to the lowest bidder
// index; the displaced owner and the losers bid again in the next round; a
// bidder whose cheapest choice is its dummy stays unmatched.  The winner is
// decided by atomics on integer keys, so the assignment does not depend on
// the order threads run in: it is the same on every run.  It carries the
// reference's guarantee (epsilon-complementary slackness with epsilon = 1
// and the dummy gap: exact maximum cardinality, base cost within n of the
// minimum of that cardinality) but, where several assignments meet it, need
// not be the sequential queue's.
//
// A round is four grid launches over the active list (bid, tie-break,
// resolve, reset) and one swap of the list; the host enqueues rounds in
// batches and reads the active count back between batches.  Once at most
// AUCTION_TAIL_BIDDERS bidders are active the remaining ones bid one at a
// time on the host, by the reference's rule, from the device's prices and
// owners: the invariants the guarantee rests on (every assigned object owned,
// unowned objects at price zero, each assignment within 1 of its bidder's
// best) hold at every round boundary, so the guarantee carries over.

#include "gp_host.h"

#include <algorithm>
#include <cstring>
#include <deque>
#include <vector>

namespace texo {
namespace {

constexpr uint32_t NO_WINNER = 0xffffffffu;
constexpr int AUCTION_ROUNDS_PER_SYNC = 8;
// From this few active bidders on, a round's launches cost more than its
// bids: the rest of the auction (typically a price war among the last
// bidders of a congested region, which needs many small price steps) bids
// one bidder at a time on the host.
constexpr uint32_t AUCTION_TAIL_BIDDERS = 512;

struct AuctionState {
  const uint32_t *start, *object;
  const uint64_t *cost;
  uint64_t *price;          // [objects]
  int32_t *owner;           // [objects] bidder or -1
  unsigned long long *best_bid;  // [objects] highest bid this round (0: none)
  uint32_t *winner;         // [objects] lowest bidder with the highest bid
  int64_t *assignment;      // [bidders] object or -1
  int64_t *target;          // [bidders] object bid on this round, -1: dummy chosen
  uint64_t *bid;            // [bidders] bid value this round
  uint32_t *list[2];        // active bidders, current / next by round parity
  uint32_t *count;          // [2] list lengths
  unsigned long long *stats;  // [2] bids won, candidates examined
  uint64_t dummy, max_base;
};

__device__ __forceinline__ bool before(uint64_t ra, uint64_t oa, uint64_t rb, uint64_t ob) {
  return ra < rb || (ra == rb && oa < ob);
}

__global__ void __launch_bounds__(BLOCK) k_auction_bid(AuctionState s, int parity) {
  const uint32_t n = s.count[parity];
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  unsigned long long examined = 0;
  if (i < n) {
    const uint32_t b = s.list[parity][i];
    // the two best real objects at this round's prices
    uint64_t r0 = ~0ull, o0 = ~0ull, c0 = 0, r1 = ~0ull, o1 = ~0ull;
    for (uint32_t k = s.start[b]; k < s.start[b + 1]; ++k) {
      const uint64_t o = s.object[k], c = s.cost[k];
      const uint64_t r = c + s.price[o];
      if (before(r, o, r0, o0)) {
        r1 = r0;
        o1 = o0;
        r0 = r;
        o0 = o;
        c0 = c;
      } else if (before(r, o, r1, o1)) {
        r1 = r;
        o1 = o;
      }
    }
    examined = s.start[b + 1] - s.start[b];
    // the private dummy (object index above every real one: a real object
    // wins an equal-cost tie)
    const uint64_t od = ~0ull - 1;
    if (before(s.dummy, od, r0, o0)) {
      s.target[b] = -1;  // the dummy: unmatched for good
      s.assignment[b] = -1;
    } else {
      if (before(s.dummy, od, r1, o1)) r1 = s.dummy;
      const uint64_t v = r1 - c0 + 1;  // price + (second - best) + 1
      s.target[b] = (int64_t)o0;
      s.bid[b] = v;
      atomicMax(&s.best_bid[o0], (unsigned long long)v);
    }
  }
  examined = lane_reduce<64>(examined, OpAdd{});
  if ((threadIdx.x & 63) == 0 && examined) atomicAdd(&s.stats[1], examined);
}

__global__ void __launch_bounds__(BLOCK) k_auction_tiebreak(AuctionState s, int parity) {
  const uint32_t n = s.count[parity];
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = s.list[parity][i];
  const int64_t o = s.target[b];
  if (o >= 0 && s.bid[b] == s.best_bid[o]) atomicMin(&s.winner[o], b);
}

__global__ void __launch_bounds__(BLOCK) k_auction_resolve(AuctionState s, int parity) {
  const uint32_t n = s.count[parity];
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  unsigned long long won = 0;
  if (i < n) {
    const uint32_t b = s.list[parity][i];
    const int64_t o = s.target[b];
    uint32_t *next = s.list[parity ^ 1];
    if (o >= 0) {
      if (s.winner[o] == b) {
        const int32_t prev = s.owner[o];
        s.owner[o] = (int32_t)b;
        s.price[o] = s.bid[b];
        s.assignment[b] = o;
        won = 1;
        if (prev >= 0) {
          s.assignment[prev] = -1;
          next[atomicAdd(&s.count[parity ^ 1], 1u)] = (uint32_t)prev;
        }
      } else {
        next[atomicAdd(&s.count[parity ^ 1], 1u)] = b;
      }
    }
  }
  won = lane_reduce<64>(won, OpAdd{});
  if ((threadIdx.x & 63) == 0 && won) atomicAdd(&s.stats[0], won);
}

__global__ void __launch_bounds__(BLOCK) k_auction_reset(AuctionState s, int parity) {
  const uint32_t n = s.count[parity];
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const int64_t o = s.target[s.list[parity][i]];
  if (o >= 0) {
    s.best_bid[o] = 0;
    s.winner[o] = NO_WINNER;
  }
}

// The round's list is spent: its count becomes the next round's empty list.
__global__ void k_auction_swap(AuctionState s, int parity) {
  if (threadIdx.x == 0) s.count[parity] = 0u;
}

__global__ void __launch_bounds__(BLOCK) k_auction_init(AuctionState s, uint64_t bidders, uint64_t objects) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < objects) {
    s.price[i] = 0;
    s.owner[i] = -1;
    s.best_bid[i] = 0;
    s.winner[i] = NO_WINNER;
  }
  if (i < bidders) {
    s.assignment[i] = -1;
    s.target[i] = -1;
    s.list[0][i] = (uint32_t)i;
  }
}

// The reference's bidding loop (auction.rs:169-248) from a state the
// parallel rounds reached: every assigned bidder owns its object, every
// object with a price has an owner, the queue holds the unassigned bidders
// that have not chosen their dummy.  The same bid, tie and dummy rules.
void sequential_tail(const std::vector<uint32_t> &queue_init, const uint32_t *start, const uint32_t *object,
                     const uint64_t *cost, uint64_t dummy, std::vector<uint64_t> &price, std::vector<int32_t> &owner,
                     int64_t *assignment, uint64_t &bids, uint64_t &examined) {
  std::deque<uint32_t> queue(queue_init.begin(), queue_init.end());
  const uint64_t od = ~0ull - 1;  // the dummy's object index: above every real one
  auto before = [](uint64_t ra, uint64_t oa, uint64_t rb, uint64_t ob) { return ra < rb || (ra == rb && oa < ob); };
  while (!queue.empty()) {
    const uint32_t b = queue.front();
    queue.pop_front();
    uint64_t r0 = ~0ull, o0 = ~0ull, c0 = 0, r1 = ~0ull, o1 = ~0ull;
    for (uint32_t k = start[b]; k < start[b + 1]; ++k) {
      const uint64_t o = object[k], c = cost[k], r = c + price[o];
      if (before(r, o, r0, o0)) {
        r1 = r0;
        o1 = o0;
        r0 = r;
        o0 = o;
        c0 = c;
      } else if (before(r, o, r1, o1)) {
        r1 = r;
        o1 = o;
      }
    }
    examined += start[b + 1] - start[b];
    if (before(dummy, od, r0, o0)) {
      assignment[b] = -1;
      continue;
    }
    if (before(dummy, od, r1, o1)) r1 = dummy;
    price[o0] = r1 - c0 + 1;
    const int32_t prev = owner[o0];
    owner[o0] = (int32_t)b;
    if (prev >= 0) {
      assignment[prev] = -1;
      queue.push_back((uint32_t)prev);
    }
    assignment[b] = (int64_t)o0;
    bids += 1;
  }
}

}  // namespace
}  // namespace texo

extern "C" int texo_gp_auction(int32_t device, uint64_t n_bidders, uint64_t n_objects, uint64_t max_base_cost,
                               const uint32_t *domain_start, const uint32_t *domain_object,
                               const uint64_t *domain_cost, int64_t *assignment, uint64_t *stats) {
  using namespace texo;
  if (stats) std::memset(stats, 0, 4 * sizeof(uint64_t));
  if (n_bidders == 0) return 0;
  if (!domain_start || !assignment || n_bidders >= 0x7fffffffull || n_objects >= 0x7fffffffull)
    return TEXO_GP_ERR_INVALID_ARGUMENT;
  const uint64_t nnz = domain_start[n_bidders];
  if (nnz && (!domain_object || !domain_cost)) return TEXO_GP_ERR_INVALID_ARGUMENT;
  // the domains, as the oracle validates them
  {
    std::vector<uint32_t> seen;
    for (uint64_t b = 0; b < n_bidders; ++b) {
      if (domain_start[b + 1] < domain_start[b]) return TEXO_GP_ERR_INVALID_ARGUMENT;
      seen.assign(domain_object + domain_start[b], domain_object + domain_start[b + 1]);
      for (uint32_t k = domain_start[b]; k < domain_start[b + 1]; ++k) {
        if (domain_object[k] >= n_objects) return TEXO_GP_ERR_AUCTION_OBJECT_OUT_OF_RANGE;
        if (domain_cost[k] > max_base_cost) return TEXO_GP_ERR_AUCTION_BASE_COST_OUT_OF_RANGE;
      }
      std::sort(seen.begin(), seen.end());
      if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return TEXO_GP_ERR_AUCTION_DUPLICATE_BEST_OBJECT;
    }
  }
  // the dummy's cost and the price bound (auction.rs:145-167)
  uint64_t dummy = 0, bound = 0;
  if (__builtin_mul_overflow(n_bidders, max_base_cost, &dummy) || __builtin_add_overflow(dummy, n_bidders, &dummy) ||
      __builtin_add_overflow(dummy, (uint64_t)1, &dummy) || __builtin_add_overflow(dummy, (uint64_t)1, &bound) ||
      __builtin_add_overflow(bound, max_base_cost, &bound))
    return TEXO_GP_ERR_AUCTION_OVERFLOW;
  hipStream_t s = nullptr;
  char *blk = nullptr;
  auto cleanup = [&] {
    if (s) (void)hipStreamSynchronize(s);
    if (blk) (void)hipFree(blk);
    if (s) (void)hipStreamDestroy(s);
  };
  try {
    HIP(hipSetDevice(device));
    HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t n = n_bidders, m = std::max<uint64_t>(n_objects, 1), z = std::max<uint64_t>(nnz, 1);
    size_t at = 0;
    auto take = [&](size_t bytes) {
      const size_t o = at;
      at = (at + bytes + 255) & ~size_t(255);
      return o;
    };
    const size_t o_start = take(4 * (n + 1)), o_obj = take(4 * z), o_cost = take(8 * z), o_price = take(8 * m),
                 o_owner = take(4 * m), o_best = take(8 * m), o_win = take(4 * m), o_asg = take(8 * n),
                 o_tgt = take(8 * n), o_bid = take(8 * n), o_l0 = take(4 * n), o_l1 = take(4 * n), o_st = take(16),
                 o_cnt = take(8);
    HIP(hipMalloc(&blk, at));
    const uint32_t counts0[2] = {(uint32_t)n, 0u};
    HIP(hipMemcpyAsync(blk + o_cnt, counts0, sizeof counts0, hipMemcpyHostToDevice, s));
    HIP(hipMemcpyAsync(blk + o_start, domain_start, 4 * (n + 1), hipMemcpyHostToDevice, s));
    if (nnz) {
      HIP(hipMemcpyAsync(blk + o_obj, domain_object, 4 * nnz, hipMemcpyHostToDevice, s));
      HIP(hipMemcpyAsync(blk + o_cost, domain_cost, 8 * nnz, hipMemcpyHostToDevice, s));
    }
    HIP(hipMemsetAsync(blk + o_st, 0, 16, s));
    AuctionState st{};
    st.start = reinterpret_cast<const uint32_t *>(blk + o_start);
    st.object = reinterpret_cast<const uint32_t *>(blk + o_obj);
    st.cost = reinterpret_cast<const uint64_t *>(blk + o_cost);
    st.price = reinterpret_cast<uint64_t *>(blk + o_price);
    st.owner = reinterpret_cast<int32_t *>(blk + o_owner);
    st.best_bid = reinterpret_cast<unsigned long long *>(blk + o_best);
    st.winner = reinterpret_cast<uint32_t *>(blk + o_win);
    st.assignment = reinterpret_cast<int64_t *>(blk + o_asg);
    st.target = reinterpret_cast<int64_t *>(blk + o_tgt);
    st.bid = reinterpret_cast<uint64_t *>(blk + o_bid);
    st.list[0] = reinterpret_cast<uint32_t *>(blk + o_l0);
    st.list[1] = reinterpret_cast<uint32_t *>(blk + o_l1);
    st.count = reinterpret_cast<uint32_t *>(blk + o_cnt);
    st.stats = reinterpret_cast<unsigned long long *>(blk + o_st);
    st.dummy = dummy;
    st.max_base = max_base_cost;
    const unsigned init_grid = (unsigned)((std::max(n, m) + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_auction_init, dim3(init_grid), dim3(BLOCK), 0, s, st, (uint64_t)n, (uint64_t)n_objects);
    HIP(hipGetLastError());
    const unsigned grid = (unsigned)((n + BLOCK - 1) / BLOCK);
    uint64_t rounds = 0;
    int parity = 0;
    // every round raises a price by at least 1 or retires a bidder, and no
    // price exceeds the bound: the loop ends; the cap only guards the host
    const uint64_t cap = 1ull << 40;
    uint32_t active = 0;
    for (;;) {
      for (int r = 0; r < AUCTION_ROUNDS_PER_SYNC; ++r, parity ^= 1) {
        hipLaunchKernelGGL(k_auction_bid, dim3(grid), dim3(BLOCK), 0, s, st, parity);
        hipLaunchKernelGGL(k_auction_tiebreak, dim3(grid), dim3(BLOCK), 0, s, st, parity);
        hipLaunchKernelGGL(k_auction_resolve, dim3(grid), dim3(BLOCK), 0, s, st, parity);
        hipLaunchKernelGGL(k_auction_reset, dim3(grid), dim3(BLOCK), 0, s, st, parity);
        hipLaunchKernelGGL(k_auction_swap, dim3(1), dim3(64), 0, s, st, parity);
      }
      HIP(hipGetLastError());
      HIP(hipMemcpyAsync(&active, st.count + parity, sizeof active, hipMemcpyDeviceToHost, s));
      HIP(hipStreamSynchronize(s));
      rounds += AUCTION_ROUNDS_PER_SYNC;
      if (active <= AUCTION_TAIL_BIDDERS) break;
      if (rounds >= cap) fail(TEXO_GP_ERR_DEVICE);
    }
    unsigned long long dev_stats[2] = {0, 0};
    HIP(hipMemcpyAsync(assignment, st.assignment, 8 * n, hipMemcpyDeviceToHost, s));
    HIP(hipMemcpyAsync(dev_stats, st.stats, 16, hipMemcpyDeviceToHost, s));
    uint64_t tail_bids = 0, tail_examined = 0;
    if (active > 0) {
      // the few bidders still active bid one at a time on the host, from the
      // device's prices and owners, in bidder order
      std::vector<uint64_t> price(m);
      std::vector<int32_t> owner(m);
      std::vector<uint32_t> queue_init(active);
      HIP(hipMemcpyAsync(price.data(), st.price, 8 * m, hipMemcpyDeviceToHost, s));
      HIP(hipMemcpyAsync(owner.data(), st.owner, 4 * m, hipMemcpyDeviceToHost, s));
      HIP(hipMemcpyAsync(queue_init.data(), st.list[parity], 4 * (size_t)active, hipMemcpyDeviceToHost, s));
      HIP(hipStreamSynchronize(s));
      std::sort(queue_init.begin(), queue_init.end());
      sequential_tail(queue_init, domain_start, domain_object, domain_cost, dummy, price, owner, assignment,
                      tail_bids, tail_examined);
    }
    HIP(hipStreamSynchronize(s));
    if (stats) {
      stats[0] = dev_stats[0] + tail_bids;
      stats[1] = dev_stats[1] + tail_examined;
      uint64_t unmatched = 0;
      for (size_t b = 0; b < n; ++b) unmatched += assignment[b] < 0 ? 1 : 0;
      stats[2] = unmatched;
      stats[3] = rounds;
    }
  } catch (...) {
    cleanup();
    return current_failure().status;
  }
  cleanup();
  return 0;
}
