// CPU test of the persistent batched kernel's work plan
// (gpu_fft_amd/csrc/gfft_batch_sched.h): the plan invariants on random
// batches, and a simulation of the kernel with threads as workgroups, only R
// of which may be resident at once (the others wait for a slot, as
// workgroups wait for a free CU).  Each simulated workgroup runs the way the
// kernel does: it takes the next local slot of its domain (round robin over
// the domains, or every workgroup in domain 0, or a random domain); a team
// member, per level: claims each of its own shares in turn (compare-and-
// swap), runs it, marks it done, then waits until every share of the level
// is done, stealing unclaimed shares after a short spin; then every workgroup
// takes solo
// transforms from the shared queue, then helps the teams of its domain whose
// first member has arrived (claims free shares of each one's current level);
// the last workgroup to leave runs the
// orphaned teams (no member arrived) and resets the counters.  Checked:
// every tile runs exactly once and only after its transform's previous level
// is complete, every team tile run by a member runs in the team's domain, no
// wait hits its bound, the counters are zero after each of two launches in a
// row, for R down to 1.  Planted mistake: with stealing switched off and
// fewer resident workgroups than slots, a team waits for a member that
// cannot start; the bounded wait ends the simulation (failure reported)
// instead of hanging.  Planted mistake: helpers that also help teams of which
// no member has arrived run tiles the orphan rule runs again.
// The global-level plans (team factor 0, WPlan) get the same treatment: their
// invariants on random batches, and a simulation of the kernel's wide body
// (per level: claim the slot's list of shares with one compare-and-swap, run
// it, add the count to the domain's done word; the domain's first workgroup
// sums the done words, marks its domain's write-back, waits for every working
// domain's and raises the domain's go flag, which the others watch; after a
// delay, waiting workgroups claim and run lists of absent workgroups) under the
// same residencies and spreads; checked: every tile once, every tile of level
// g after every tile of level g - 1 has finished, counters zero after two
// launches; planted mistake: no stealing with absent workgroups ends at the
// bound.
// Built and run by tests/test_batch_sched_host.py with g++ -pthread.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "gfft_batch_sched.h"

static int failures = 0;
#define CHECK(c)                                               \
  do {                                                         \
    if (!(c)) {                                                \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                              \
    }                                                          \
  } while (0)

static const gfft::PLevelSpec kRoundTrip[6] = {{0, 0}, {1, 0}, {2, 0}, {2, 1}, {1, 1}, {0, 1}};

struct Batch {
  std::vector<std::vector<uint32_t>> tiles;  // [transform][level]
  std::vector<int> cls;
};

static Batch random_batch(std::mt19937& rng, int count) {
  Batch b;
  b.tiles.resize(count);
  for (int t = 0; t < count; ++t) {
    b.cls.push_back((int)(rng() % 8));
    // heavy-tailed sizes: a few large transforms, many small ones
    const uint32_t base = rng() % 5 == 0 ? 40 + rng() % 200 : 1 + rng() % 12;
    const bool no_x = rng() % 7 == 0, no_y = rng() % 9 == 0;  // length-1 axes
    for (int l = 0; l < 6; ++l) {
      const int axis = kRoundTrip[l].axis;
      uint32_t n = base + rng() % 5;
      if ((axis == 2 && no_x) || (axis == 1 && no_y)) n = 0;
      b.tiles[t].push_back(n);
    }
  }
  return b;
}

static gfft::PPlan plan_of(const Batch& b, uint32_t domains, uint32_t per_dom, uint32_t q = 2) {
  return gfft::build_plan(
      (uint32_t)b.tiles.size(), 6, kRoundTrip, [&](uint32_t t, int l) { return b.tiles[t][l]; },
      [&](uint32_t t, int) { return b.cls[t]; },
      [&](uint32_t t) {
        double s = 0;
        for (uint32_t n : b.tiles[t]) s += n;
        return s;
      },
      domains, per_dom, 2.0, 2, q);
}

class Slots {  // residency: a workgroup needs one of R places to run, until it leaves
 public:
  explicit Slots(int n) : free_(n) {}
  void take() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return free_ > 0; });
    --free_;
  }
  void give() {
    std::lock_guard<std::mutex> lk(mu_);
    ++free_;
    cv_.notify_one();
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  int free_;
};

struct SimResult {
  bool failed = false;
  bool order_ok = true;
  bool exactly_once = true;
  bool domain_ok = true;
  bool reset = true;
};

using Ctr = std::vector<std::atomic<uint32_t>>;

// HalfOutside: the first half of the workgroups round robin over the domains,
// the rest on an XCD the plan does not use (no slot, no domain to help)
enum class Spread { RoundRobin, AllInZero, Random, HalfOutside };

static SimResult simulate(const gfft::PPlan& p, Ctr& ctr, int resident, bool steal,
                          std::chrono::milliseconds wait_limit, unsigned seed, Spread spread,
                          bool help_checks_arrival = true, bool release_parts = true) {
  const size_t nseg = p.segs.size();
  std::vector<std::vector<std::atomic<int>>> seen(nseg);
  std::vector<std::atomic<uint32_t>> finished(nseg);  // test-only: tiles actually run per level
  for (size_t s = 0; s < nseg; ++s) {
    seen[s] = std::vector<std::atomic<int>>(p.segs[s].ntiles);
    for (auto& v : seen[s]) v.store(0);
    finished[s].store(0);
  }
  std::atomic<bool> order_bad{false}, domain_bad{false};
  Slots places(resident);
  const int groups = (int)(p.domains * p.per_dom);
  auto failed = [&] { return ctr[gfft::kCtrFail].load(std::memory_order_relaxed) != 0; };
  auto words = [&](uint32_t s) { return &ctr[gfft::kCtrShares + p.segs[s].claim0]; };
  auto part_dom = [&](const gfft::PTeam& tm, uint32_t i) { return p.parts[tm.part0 + i].dom; };
  auto run_tiles = [&](uint32_t s, uint32_t k, uint32_t m, std::mt19937& rng) {
    const gfft::PSeg& sg = p.segs[s];
    const gfft::PTeam& tm = p.teams[sg.xf];
    if (s > tm.seg0 && finished[s - 1].load() != p.segs[s - 1].ntiles) order_bad = true;
    for (uint32_t j = m; j < sg.ntiles; j += k) seen[s][j].fetch_add(1);
    finished[s].fetch_add(gfft::share_tiles(sg.ntiles, k, m));  // k: shares of the level, m: the share
    if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 40));
  };
  // What the kernel's memory ordering needs before a team share of level s
  // runs on domain d: the level before ran on d only, or (span teams) every
  // other part that ran shares of it has released (written its L2 back).
  auto prev_visible = [&](uint32_t s, uint32_t d) {
    const gfft::PTeam& tm = p.teams[p.segs[s].xf];
    if (s == tm.seg0) return true;
    const uint32_t n = tm.k * tm.q;
    std::atomic<uint32_t>* w = words(s - 1);
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t v = w[j].load(std::memory_order_acquire);
      if (v < gfft::kShareDone) return false;
      const uint32_t q = v - gfft::kShareDone;
      if (part_dom(tm, q) != d && (tm.nparts < 2 || w[n + q].load(std::memory_order_acquire) != gfft::kRelDone))
        return false;
    }
    return true;
  };
  // level s of team tm complete for a workgroup of part `mine`: every share
  // done and, for a span team, every part that ran shares released (its own
  // part released by it when it is the first to find that missing)
  auto level_complete = [&](const gfft::PTeam& tm, uint32_t s, uint32_t mine) {
    const uint32_t n = tm.k * tm.q;
    std::atomic<uint32_t>* w = words(s);
    uint32_t mask = 0;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t v = w[j].load(std::memory_order_acquire);
      if (v < gfft::kShareDone) return false;
      mask |= 1u << (v - gfft::kShareDone);
    }
    if (tm.nparts < 2 || !release_parts) return true;
    std::atomic<uint32_t>* rel = w + n;
    if ((mask >> mine) & 1u) {
      uint32_t z = gfft::kRelNone;
      if (rel[mine].compare_exchange_strong(z, gfft::kRelClaimed))  // the write-back, then the flag
        rel[mine].store(gfft::kRelDone, std::memory_order_release);
    }
    for (uint32_t q = 0; q < tm.nparts; ++q)
      if (((mask >> q) & 1u) && rel[q].load(std::memory_order_acquire) != gfft::kRelDone) return false;
    return true;
  };
  // a team share: run on domain d by a workgroup of part `part`
  auto run_share = [&](uint32_t s, uint32_t j, uint32_t d, uint32_t part, std::mt19937& rng) {
    const gfft::PTeam& tm = p.teams[p.segs[s].xf];
    if (part_dom(tm, part) != d) domain_bad = true;  // a part's hand-over needs its own L2
    if (!prev_visible(s, d)) order_bad = true;
    run_tiles(s, tm.k * tm.q, j, rng);
    words(s)[j].store(gfft::kShareDone + part, std::memory_order_release);
  };
  auto claim = [&](uint32_t s, uint32_t j) {
    uint32_t z = gfft::kShareFree;
    return words(s)[j].compare_exchange_strong(z, gfft::kShareClaimed);
  };
  auto run_whole = [&](uint32_t t, std::mt19937& rng) {
    const gfft::PTeam& tm = p.teams[t];
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg && !failed(); ++s) run_tiles(s, 1, 0, rng);
  };
  auto run_member = [&](const gfft::PEntry& e, uint32_t d, std::mt19937& rng) {
    const gfft::PTeam& tm = p.teams[e.team];
    const uint32_t n = tm.k * tm.q;
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg && !failed(); ++s) {
      for (uint32_t i = 0; i < tm.q; ++i)
        if (claim(s, e.member + i * tm.k)) run_share(s, e.member + i * tm.k, d, e.part, rng);
      const auto t0 = std::chrono::steady_clock::now();
      uint32_t spins = 0, next = (e.member + 1) % n;
      while (!level_complete(tm, s, e.part)) {
        if (failed()) break;
        if (std::chrono::steady_clock::now() - t0 > wait_limit) {
          ctr[gfft::kCtrFail].fetch_or(gfft::kFailTimeout);
          break;
        }
        if (steal && ++spins > 20) {  // look at one other share per spin
          if (claim(s, next)) run_share(s, next, d, e.part, rng);
          next = (next + 1) % n;
        }
        std::this_thread::yield();
      }
    }
  };
  // a helper: free shares of the current level of each team with a part in
  // domain d whose first member there has arrived, until every such team is
  // complete
  auto help = [&](uint32_t d, std::mt19937& rng) {
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      if (failed()) return;
      bool unfinished = false, ran = false;
      for (uint32_t i = p.dom_off[d]; i < p.dom_off[d + 1] && !ran; ++i) {
        const uint32_t t = p.dom_teams[i];
        const gfft::PTeam& tm = p.teams[t];
        uint32_t part = 0;
        while (part < tm.nparts && part_dom(tm, part) != d) ++part;
        if (part == tm.nparts) domain_bad = true;  // listed under a domain it has no part in
        if (help_checks_arrival &&
            ctr[gfft::kCtrArrive + d].load() <= p.parts[tm.part0 + std::min(part, tm.nparts - 1)].slot0)
          continue;
        const uint32_t n = tm.k * tm.q;
        uint32_t s = tm.seg0;
        while (s < tm.seg0 + tm.nseg && level_complete(tm, s, part)) ++s;
        if (s == tm.seg0 + tm.nseg) continue;  // complete
        unfinished = true;
        for (uint32_t j = 0; j < n && !ran; ++j)
          if (claim(s, j)) {
            run_share(s, j, d, part, rng);
            ran = true;
          }
      }
      if (!unfinished) return;
      if (ran) {
        t0 = std::chrono::steady_clock::now();
      } else if (std::chrono::steady_clock::now() - t0 > wait_limit) {
        ctr[gfft::kCtrFail].fetch_or(gfft::kFailTimeout);
        return;
      }
      std::this_thread::yield();
    }
  };
  auto orphaned = [&](const gfft::PTeam& tm) {
    for (uint32_t i = 0; i < tm.nparts; ++i) {
      const gfft::PPart& pt = p.parts[tm.part0 + i];
      if (ctr[gfft::kCtrArrive + pt.dom].load() > pt.slot0) return false;
    }
    return true;
  };
  auto wg = [&](int w) {
    places.take();
    std::mt19937 rng(seed * 131 + w);
    const uint32_t d = spread == Spread::RoundRobin ? (uint32_t)w % p.domains
                       : spread == Spread::AllInZero ? 0u
                       : spread == Spread::HalfOutside
                           ? (w < groups / 2 ? (uint32_t)w % p.domains : p.domains)
                           : (uint32_t)(rng() % p.domains);
    if (d < p.domains) {  // a workgroup on an XCD outside the plan takes no slot
      const uint32_t li = ctr[gfft::kCtrArrive + d].fetch_add(1);
      if (li < p.per_dom) {
        const gfft::PEntry& e = p.local[(size_t)d * p.per_dom + li];
        if (e.team != gfft::kNoTeam) run_member(e, d, rng);
      }
    }
    while (!failed()) {
      const uint32_t j = ctr[gfft::kCtrSolo].fetch_add(1);
      if (j >= p.solo.size()) break;
      run_whole(p.solo[j], rng);
    }
    if (d < p.domains) help(d, rng);
    const uint32_t prev = ctr[gfft::kCtrExited].fetch_add(1, std::memory_order_acq_rel);
    if (prev == (uint32_t)groups - 1) {  // last one out
      for (uint32_t t : p.team_ids)
        if (!failed() && orphaned(p.teams[t])) run_whole(t, rng);
      ctr[gfft::kCtrStatus].fetch_or(ctr[gfft::kCtrFail].load());
      for (size_t i = 0; i < ctr.size(); ++i)
        if (i != gfft::kCtrStatus) ctr[i].store(0);
    }
    places.give();
  };
  std::vector<std::thread> th;
  for (int w = 0; w < groups; ++w) th.emplace_back(wg, w);
  for (auto& t : th) t.join();
  SimResult r;
  r.failed = ctr[gfft::kCtrStatus].exchange(0) != 0;  // the host reads and clears
  r.order_ok = !order_bad.load();
  r.domain_ok = !domain_bad.load();
  for (size_t s = 0; s < nseg; ++s)
    for (auto& v : seen[s]) r.exactly_once = r.exactly_once && v.load() == 1;
  for (auto& c : ctr) r.reset = r.reset && c.load() == 0;
  return r;
}

static void plan_invariants_random() {
  std::mt19937 rng(5);
  for (int trial = 0; trial < 300; ++trial) {
    const Batch b = random_batch(rng, 1 + (int)(rng() % 80));
    const uint32_t domains = 1 + rng() % 8, per_dom = 1 + rng() % 80;
    const gfft::PPlan p = plan_of(b, domains, per_dom, 1 + rng() % 4);
    CHECK(gfft::check_plan(p) == 0);
    CHECK(p.local.size() == (size_t)domains * per_dom);
    size_t tiles = 0, planned = 0;
    for (const auto& t : b.tiles)
      for (uint32_t n : t) tiles += n;
    for (const auto& s : p.segs) planned += s.ntiles;
    CHECK(tiles == planned);
  }
  // the bench's shape of batch: the large transform's share of members
  // exceeds a domain's slots, so its team spans two domains
  Batch b;
  b.tiles = {{500, 500, 500, 500, 500, 500}};
  b.cls = {6};
  for (int i = 0; i < 50; ++i) {
    b.tiles.push_back({8, 9, 9, 9, 9, 8});
    b.cls.push_back(0);
  }
  const gfft::PPlan p = plan_of(b, 8, 8);
  CHECK(gfft::check_plan(p) == 0 && p.teams[0].nparts == 2 && p.teams[0].k > 8 && p.teams[0].k <= 16);
  CHECK(p.parts[p.teams[0].part0].dom != p.parts[p.teams[0].part0 + 1].dom);
  CHECK(p.solo.size() == 50 && p.team_ids.size() == 1);
  // spanning switched off: one domain, at most its slots
  const gfft::PPlan one = gfft::build_plan(
      (uint32_t)b.tiles.size(), 6, kRoundTrip, [&](uint32_t t, int l) { return b.tiles[t][l]; },
      [&](uint32_t t, int) { return b.cls[t]; },
      [&](uint32_t t) {
        double x = 0;
        for (uint32_t n : b.tiles[t]) x += n;
        return x;
      },
      8, 8, 2.0, 2, 2, 1);
  CHECK(gfft::check_plan(one) == 0 && one.teams[0].nparts == 1 && one.teams[0].k == 8);
}

static void simulated_kernel_completes() {
  std::mt19937 rng(9);
  const int resident_counts[] = {1, 3, 8, 1000};
  const Spread spreads[] = {Spread::RoundRobin, Spread::AllInZero, Spread::Random, Spread::HalfOutside};
  int spans = 0;
  for (int trial = 0; trial < 6; ++trial) {
    const Batch b = random_batch(rng, 6 + (int)(rng() % 20));
    const gfft::PPlan p = plan_of(b, 4, 6, 1 + trial % 3);
    CHECK(gfft::check_plan(p) == 0);
    for (uint32_t t : p.team_ids) spans += p.teams[t].nparts > 1;
    for (Spread sp : spreads)
      for (int r : resident_counts) {
        Ctr ctr(gfft::counter_words(p));
        for (auto& c : ctr) c.store(0);
        for (int rep = 0; rep < 2; ++rep) {  // the second launch relies on the first one's reset
          const SimResult s =
              simulate(p, ctr, r, true, std::chrono::milliseconds(20000), trial * 7 + r + rep, sp);
          CHECK(!s.failed);
          CHECK(s.order_ok);
          CHECK(s.exactly_once);
          CHECK(s.domain_ok);
          CHECK(s.reset);
        }
      }
  }
  CHECK(spans > 0);  // span teams were among the simulated plans
}

// One transform whose team spans two domains of four slots, every workgroup
// resident and spread round robin, so both parts run shares of every level:
// without the parts' releases a level's shares start before the other part's
// stores are written back.
static void span_team_levels_wait_for_every_part() {
  Batch b;
  b.tiles = {{64, 64, 64, 64, 64, 64}};
  b.cls = {6};
  const gfft::PPlan p = plan_of(b, 2, 4);
  CHECK(gfft::check_plan(p) == 0 && p.teams[0].nparts == 2 && p.teams[0].k == 8);
  Ctr ctr(gfft::counter_words(p));
  for (auto& c : ctr) c.store(0);
  for (int rep = 0; rep < 2; ++rep) {
    const SimResult good = simulate(p, ctr, 1000, true, std::chrono::milliseconds(20000), 7 + rep,
                                    Spread::RoundRobin);
    CHECK(!good.failed && good.exactly_once && good.order_ok && good.domain_ok && good.reset);
  }
  const SimResult bad = simulate(p, ctr, 1000, true, std::chrono::milliseconds(20000), 9, Spread::RoundRobin,
                                 true, false);
  CHECK(!bad.order_ok);
}

static void no_stealing_with_absent_members_ends_by_the_bound() {
  Batch b;
  b.tiles = {{64, 64, 64, 64, 64, 64}};
  b.cls = {7};
  const gfft::PPlan p = plan_of(b, 1, 8);
  CHECK(p.teams[0].k >= 2);
  Ctr ctr(gfft::counter_words(p));
  for (auto& c : ctr) c.store(0);
  const SimResult s = simulate(p, ctr, 1, false, std::chrono::milliseconds(200), 3, Spread::RoundRobin);
  CHECK(s.failed);
  CHECK(s.reset);
  const SimResult t = simulate(p, ctr, 1, true, std::chrono::milliseconds(20000), 4, Spread::RoundRobin);
  CHECK(!t.failed && t.exactly_once && t.order_ok && t.domain_ok && t.reset);
}


static gfft::WPlan wide_of(const Batch& b, uint32_t slots, uint32_t q = 2, double solo_factor = 0.0) {
  return gfft::build_wide_plan(
      (uint32_t)b.tiles.size(), 6, kRoundTrip, [&](uint32_t t, int l) { return b.tiles[t][l]; },
      [&](uint32_t t, int) { return b.cls[t]; }, [&](uint32_t t, int l) { return 1.0 + (t * 7 + l) % 3; }, slots, q,
      solo_factor);
}

static void wide_plan_invariants_random() {
  std::mt19937 rng(11);
  for (int trial = 0; trial < 300; ++trial) {
    const Batch b = random_batch(rng, 1 + (int)(rng() % 80));
    const uint32_t slots = 1 + rng() % 300;
    const double solo_factor = trial % 3 == 0 ? 0.0 : (trial % 3 == 1 ? 0.5 : 4.0);
    const gfft::WPlan p = wide_of(b, slots, 1 + rng() % 4, solo_factor);
    CHECK(gfft::check_wide_plan(p) == 0);
    if (solo_factor == 0.0) CHECK(p.solo.empty());
    size_t tiles = 0, planned = 0;
    for (const auto& t : b.tiles)
      for (uint32_t n : t) tiles += n;
    for (const auto& sg : p.segs) planned += sg.ntiles;
    CHECK(tiles == planned);
    CHECK(gfft::wide_counter_words(p) ==
          gfft::kCtrShares + (size_t)p.nlevels * p.slots + (size_t)p.nlevels * gfft::kLevelWords);
  }
  // one large transform among many small ones: its levels are cut over many slots
  Batch b;
  b.tiles = {{500, 500, 500, 500, 500, 500}};
  b.cls = {6};
  for (int i = 0; i < 50; ++i) {
    b.tiles.push_back({8, 9, 9, 9, 9, 8});
    b.cls.push_back(0);
  }
  const gfft::WPlan p = wide_of(b, 64);
  CHECK(gfft::check_wide_plan(p) == 0 && p.nlevels == 6);
  uint32_t big = 0;
  for (uint32_t j = p.lvl_off[0]; j < p.lvl_off[1]; ++j) big += p.segs[p.shares[j].seg].xf == 0 ? 1u : 0u;
  CHECK(big >= 32);
}

static SimResult simulate_wide(const gfft::WPlan& p, uint32_t domains, uint32_t per_dom, Ctr& ctr, int resident,
                               bool steal, std::chrono::milliseconds wait_limit, unsigned seed, Spread spread) {
  const size_t nseg = p.segs.size();
  std::vector<std::vector<std::atomic<int>>> seen(nseg);
  std::vector<std::atomic<uint32_t>> finished(p.nlevels);  // test-only: tiles run per global level
  std::vector<uint32_t> level_tiles(p.nlevels, 0);
  for (size_t i = 0; i < nseg; ++i) {
    seen[i] = std::vector<std::atomic<int>>(p.segs[i].ntiles);
    for (auto& v : seen[i]) v.store(0);
    if (p.seg_level[i] != gfft::kNoLevel) level_tiles[p.seg_level[i]] += p.segs[i].ntiles;
  }
  for (auto& f : finished) f.store(0);
  std::atomic<bool> order_bad{false};
  Slots places(resident);
  const int groups = (int)(domains * per_dom);
  const size_t claim0 = gfft::kCtrShares, lvl0 = gfft::wide_level_base(p);
  // test-only: which domains wrote their L2 back per level, and which ran shares
  std::vector<std::vector<std::atomic<int>>> wb(p.nlevels), worked(p.nlevels);
  for (uint32_t g = 0; g < p.nlevels; ++g) {
    wb[g] = std::vector<std::atomic<int>>(domains);
    worked[g] = std::vector<std::atomic<int>>(domains);
    for (uint32_t k = 0; k < domains; ++k) {
      wb[g][k].store(0);
      worked[g][k].store(0);
    }
  }
  std::atomic<bool> unreleased{false};
  auto failed = [&] { return ctr[gfft::kCtrFail].load(std::memory_order_relaxed) != 0; };
  // the slot's list of level g, by a workgroup of domain d; returns its share count
  auto run_list = [&](uint32_t g, uint32_t s, uint32_t d, std::mt19937& rng) {
    const uint32_t b0 = p.own_off[g * p.slots + s], b1 = p.own_off[g * p.slots + s + 1];
    if (b1 > b0) {
      if (g > 0 && finished[g - 1].load() != level_tiles[g - 1]) order_bad = true;
      if (g > 0)  // the previous level's data must be in memory: every domain that wrote it wrote back
        for (uint32_t k = 0; k < domains; ++k)
          if (worked[g - 1][k].load() && !wb[g - 1][k].load()) unreleased = true;
      worked[g][d].store(1);
    }
    for (uint32_t i = b0; i < b1; ++i) {
      const gfft::WShare& w = p.shares[p.own[i]];
      for (uint32_t k = 0; k < w.count; ++k) seen[w.seg][w.first + k * w.step].fetch_add(1);
      if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 40));
      finished[g].fetch_add(w.count);
    }
    return b1 - b0;
  };
  auto claim = [&](uint32_t g, uint32_t s) {
    uint32_t z = gfft::kShareFree;
    return ctr[claim0 + (size_t)g * p.slots + s].compare_exchange_strong(z, gfft::kShareClaimed);
  };
  auto run_whole = [&](uint32_t t, std::mt19937& rng) {  // a solo transform, level after level, by one workgroup
    for (uint32_t i = p.xseg0[t]; i < p.xseg0[t] + p.xnseg[t]; ++i) {
      for (uint32_t k = 0; k < p.segs[i].ntiles; ++k) seen[i][k].fetch_add(1);
      if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 20));
    }
  };
  auto wg = [&](int wgi) {
    places.take();
    std::mt19937 rng(seed * 131 + wgi);
    const uint32_t d = spread == Spread::RoundRobin ? (uint32_t)wgi % domains
                       : spread == Spread::AllInZero ? 0u
                                                     : (uint32_t)(rng() % domains);
    const uint32_t li = ctr[gfft::kCtrArrive + d].fetch_add(1);
    const uint32_t slot = li < per_dom ? d * per_dom + li : p.slots;
    const bool leader = li == 0;  // the domain's first workgroup leads its hand-overs
    bool solo_left = !p.solo.empty();
    auto take_solo = [&] {  // one ticket of the solo queue; false once it is empty
      const uint32_t j = ctr[gfft::kCtrSolo].fetch_add(1);
      if (j >= p.solo.size()) return false;
      run_whole(p.solo[j], rng);
      return true;
    };
    for (uint32_t g = 0; g < p.nlevels && !failed(); ++g) {
      const size_t L = lvl0 + (size_t)g * gfft::kLevelWords;
      std::atomic<uint32_t>& done = ctr[L + gfft::kLevelDone + d * gfft::kFlagStride];
      uint32_t ran = 0;
      if (slot < p.slots && claim(g, slot)) ran = run_list(g, slot, d, rng);
      if (ran) done.fetch_add(ran, std::memory_order_relaxed);  // stores in the domain's L2
      const uint32_t n = p.lvl_off[g + 1] - p.lvl_off[g];
      int phase = 0;
      auto level_done = [&] {
        if (!leader) return ctr[L + gfft::kLevelGo + d * gfft::kFlagStride].load(std::memory_order_acquire) != 0;
        if (phase == 0) {
          uint32_t sum = 0;
          for (uint32_t k = 0; k <= gfft::kMaxDomains; ++k) sum += ctr[L + gfft::kLevelDone + k * gfft::kFlagStride].load();
          if (sum >= n) phase = 1;
        }
        if (phase == 1) {  // this domain's L2 written back, if it did work
          if (ctr[L + gfft::kLevelDone + d * gfft::kFlagStride].load() != 0) {
            wb[g][d].store(1);  // test-only: the write-back happened
            ctr[L + gfft::kLevelReleased + d * gfft::kFlagStride].store(1, std::memory_order_release);
          }
          phase = 2;
        }
        for (uint32_t k = 0; k < domains; ++k)
          if (ctr[L + gfft::kLevelDone + k * gfft::kFlagStride].load() != 0 &&
              ctr[L + gfft::kLevelReleased + k * gfft::kFlagStride].load(std::memory_order_acquire) == 0)
            return false;
        ctr[L + gfft::kLevelGo + d * gfft::kFlagStride].store(1, std::memory_order_release);
        return true;
      };
      const auto t0 = std::chrono::steady_clock::now();
      auto steal_at = t0 + std::chrono::microseconds(50);
      uint32_t next = slot < p.slots ? (slot + 1) % p.slots : 0;
      while (!level_done()) {
        if (failed()) break;
        const auto now = std::chrono::steady_clock::now();
        if (now - t0 > wait_limit) {
          ctr[gfft::kCtrFail].fetch_or(gfft::kFailTimeout);
          break;
        }
        if (steal && phase == 0 && now >= steal_at) {  // lists of absent workgroups
          steal_at = now + std::chrono::microseconds(50);
          for (uint32_t k = 0; k < p.slots; ++k) {
            const uint32_t s = (next + k) % p.slots;
            if (p.own_off[g * p.slots + s + 1] > p.own_off[g * p.slots + s] &&
                ctr[claim0 + (size_t)g * p.slots + s].load() == gfft::kShareFree && claim(g, s)) {
              done.fetch_add(run_list(g, s, d, rng), std::memory_order_relaxed);
              next = (s + 1) % p.slots;
              break;
            }
          }
        } else if (!leader && solo_left) {  // a solo transform while the level hands over
          solo_left = take_solo();
        }
        std::this_thread::yield();
      }
    }
    while (!failed() && solo_left) solo_left = take_solo();
    const uint32_t prev = ctr[gfft::kCtrExited].fetch_add(1, std::memory_order_acq_rel);
    if (prev == (uint32_t)groups - 1) {  // last one out
      ctr[gfft::kCtrStatus].fetch_or(ctr[gfft::kCtrFail].load());
      for (size_t i = 0; i < ctr.size(); ++i)
        if (i != gfft::kCtrStatus) ctr[i].store(0);
    }
    places.give();
  };
  std::vector<std::thread> th;
  for (int w = 0; w < groups; ++w) th.emplace_back(wg, w);
  for (auto& t : th) t.join();
  SimResult r;
  r.failed = ctr[gfft::kCtrStatus].exchange(0) != 0;
  r.order_ok = !order_bad.load() && !unreleased.load();
  for (size_t i = 0; i < nseg; ++i)
    for (auto& v : seen[i]) r.exactly_once = r.exactly_once && v.load() == 1;
  for (auto& c : ctr) r.reset = r.reset && c.load() == 0;
  return r;
}

static void simulated_wide_kernel_completes() {
  std::mt19937 rng(13);
  const int resident_counts[] = {1, 3, 8, 1000};
  const Spread spreads[] = {Spread::RoundRobin, Spread::AllInZero, Spread::Random};
  for (int trial = 0; trial < 6; ++trial) {
    const Batch b = random_batch(rng, 6 + (int)(rng() % 20));
    const uint32_t domains = 4, per_dom = 6;
    const gfft::WPlan p = wide_of(b, domains * per_dom, 1 + trial % 3, trial % 2 ? 2.0 : 0.0);
    if (trial % 2) CHECK(!p.solo.empty());
    CHECK(gfft::check_wide_plan(p) == 0);
    for (Spread sp : spreads)
      for (int r : resident_counts) {
        Ctr ctr(gfft::wide_counter_words(p));
        for (auto& c : ctr) c.store(0);
        for (int rep = 0; rep < 2; ++rep) {
          const SimResult s = simulate_wide(p, domains, per_dom, ctr, r, true, std::chrono::milliseconds(20000),
                                            trial * 7 + r + rep, sp);
          CHECK(!s.failed);
          CHECK(s.order_ok);
          CHECK(s.exactly_once);
          CHECK(s.reset);
        }
      }
  }
}

static void wide_no_stealing_with_absent_workgroups_ends_by_the_bound() {
  Batch b;
  b.tiles = {{64, 64, 64, 64, 64, 64}};
  b.cls = {7};
  const gfft::WPlan p = wide_of(b, 8);
  Ctr ctr(gfft::wide_counter_words(p));
  for (auto& c : ctr) c.store(0);
  const SimResult s = simulate_wide(p, 1, 8, ctr, 1, false, std::chrono::milliseconds(200), 3, Spread::RoundRobin);
  CHECK(s.failed);
  CHECK(s.reset);
  const SimResult t = simulate_wide(p, 1, 8, ctr, 1, true, std::chrono::milliseconds(20000), 4, Spread::RoundRobin);
  CHECK(!t.failed && t.exactly_once && t.order_ok && t.reset);
}

// Two teams of four in one domain of eight slots; only the first four
// workgroups land there (the rest on an XCD outside the plan), so the second
// team is orphaned while the first team's members, done, help.  Helpers that
// skip the arrival check run the orphan's tiles, and the last workgroup out
// runs them again.
static void helpers_leave_orphaned_teams_alone() {
  Batch b;
  b.tiles = {{40, 40, 40, 40, 40, 40}, {40, 40, 40, 40, 40, 40}};
  b.cls = {5, 5};
  for (int i = 0; i < 4; ++i) {
    b.tiles.push_back({1, 1, 1, 1, 1, 1});
    b.cls.push_back(0);
  }
  const gfft::PPlan p = plan_of(b, 1, 8);
  CHECK(gfft::check_plan(p) == 0);
  CHECK(p.team_ids.size() == 2 && p.teams[0].k == 4 && p.teams[1].k == 4);
  CHECK(p.teams[0].slot0 + p.teams[1].slot0 == 4);  // one team takes slots 0-3, the other 4-7
  Ctr ctr(gfft::counter_words(p));
  for (auto& c : ctr) c.store(0);
  const SimResult good =
      simulate(p, ctr, 1000, true, std::chrono::milliseconds(20000), 5, Spread::HalfOutside, true);
  CHECK(!good.failed && good.exactly_once && good.order_ok && good.domain_ok && good.reset);
  const SimResult bad =
      simulate(p, ctr, 1000, true, std::chrono::milliseconds(20000), 6, Spread::HalfOutside, false);
  CHECK(!bad.exactly_once);
}

int main() {
  plan_invariants_random();
  simulated_kernel_completes();
  no_stealing_with_absent_members_ends_by_the_bound();
  helpers_leave_orphaned_teams_alone();
  span_team_levels_wait_for_every_part();
  wide_plan_invariants_random();
  simulated_wide_kernel_completes();
  wide_no_stealing_with_absent_workgroups_ends_by_the_bound();
  if (failures) {
    std::printf("%d failure(s)\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
