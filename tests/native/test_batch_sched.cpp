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
// transforms from the shared queue; the last workgroup to leave runs the
// orphaned teams (no member arrived) and resets the counters.  Checked:
// every tile runs exactly once and only after its transform's previous level
// is complete, every team tile run by a member runs in the team's domain, no
// wait hits its bound, the counters are zero after each of two launches in a
// row, for R down to 1.  Planted mistake: with stealing switched off and
// fewer resident workgroups than slots, a team waits for a member that
// cannot start; the bounded wait ends the simulation (failure reported)
// instead of hanging.
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

enum class Spread { RoundRobin, AllInZero, Random };

static SimResult simulate(const gfft::PPlan& p, Ctr& ctr, int resident, bool steal,
                          std::chrono::milliseconds wait_limit, unsigned seed, Spread spread) {
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
  auto run_tiles = [&](uint32_t s, uint32_t k, uint32_t m, std::mt19937& rng) {
    const gfft::PSeg& sg = p.segs[s];
    const gfft::PTeam& tm = p.teams[sg.xf];
    if (s > tm.seg0 && finished[s - 1].load() != p.segs[s - 1].ntiles) order_bad = true;
    for (uint32_t j = m; j < sg.ntiles; j += k) seen[s][j].fetch_add(1);
    finished[s].fetch_add(gfft::share_tiles(sg.ntiles, k, m));  // k: shares of the level, m: the share
    if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 40));
  };
  auto run_whole = [&](uint32_t t, std::mt19937& rng) {
    const gfft::PTeam& tm = p.teams[t];
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg && !failed(); ++s) run_tiles(s, 1, 0, rng);
  };
  auto run_member = [&](const gfft::PEntry& e, uint32_t d, std::mt19937& rng) {
    const gfft::PTeam& tm = p.teams[e.team];
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg && !failed(); ++s) {
      const gfft::PSeg& sg = p.segs[s];
      std::atomic<uint32_t>* share = &ctr[gfft::kCtrShares + sg.claim0];
      auto claim = [&](uint32_t m) {
        uint32_t z = gfft::kShareFree;
        return share[m].compare_exchange_strong(z, gfft::kShareClaimed);
      };
      const uint32_t n = tm.k * tm.q;
      auto run_share = [&](uint32_t j) {
        if (d != tm.dom) domain_bad = true;  // the kernel's light hand-over needs one L2
        run_tiles(s, n, j, rng);
        share[j].store(gfft::kShareDone, std::memory_order_release);
      };
      auto level_done = [&] {
        for (uint32_t j = 0; j < n; ++j)
          if (share[j].load(std::memory_order_acquire) != gfft::kShareDone) return false;
        return true;
      };
      for (uint32_t i = 0; i < tm.q; ++i)
        if (claim(e.member + i * tm.k)) run_share(e.member + i * tm.k);
      const auto t0 = std::chrono::steady_clock::now();
      uint32_t spins = 0, next = (e.member + 1) % n;
      while (!level_done()) {
        if (failed()) break;
        if (std::chrono::steady_clock::now() - t0 > wait_limit) {
          ctr[gfft::kCtrFail].fetch_or(gfft::kFailTimeout);
          break;
        }
        if (steal && ++spins > 20) {  // look at one other share per spin
          if (claim(next)) run_share(next);
          next = (next + 1) % n;
        }
        std::this_thread::yield();
      }
    }
  };
  auto wg = [&](int w) {
    places.take();
    std::mt19937 rng(seed * 131 + w);
    const uint32_t d = spread == Spread::RoundRobin ? (uint32_t)w % p.domains
                       : spread == Spread::AllInZero ? 0u
                                                     : (uint32_t)(rng() % p.domains);
    const uint32_t li = ctr[gfft::kCtrArrive + d].fetch_add(1);
    if (li < p.per_dom) {
      const gfft::PEntry& e = p.local[(size_t)d * p.per_dom + li];
      if (e.team != gfft::kNoTeam) run_member(e, d, rng);
    }
    while (!failed()) {
      const uint32_t j = ctr[gfft::kCtrSolo].fetch_add(1);
      if (j >= p.solo.size()) break;
      run_whole(p.solo[j], rng);
    }
    const uint32_t prev = ctr[gfft::kCtrExited].fetch_add(1, std::memory_order_acq_rel);
    if (prev == (uint32_t)groups - 1) {  // last one out
      for (uint32_t t : p.team_ids)
        if (!failed() && ctr[gfft::kCtrArrive + p.teams[t].dom].load() <= p.teams[t].slot0) run_whole(t, rng);
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
  // the bench's shape of batch: the large transform gets a team in one domain
  Batch b;
  b.tiles = {{500, 500, 500, 500, 500, 500}};
  b.cls = {6};
  for (int i = 0; i < 50; ++i) {
    b.tiles.push_back({8, 9, 9, 9, 9, 8});
    b.cls.push_back(0);
  }
  const gfft::PPlan p = plan_of(b, 8, 8);
  CHECK(gfft::check_plan(p) == 0 && p.teams[0].k >= 2 && p.teams[0].k <= 8);
  CHECK(p.solo.size() == 50 && p.team_ids.size() == 1);
}

static void simulated_kernel_completes() {
  std::mt19937 rng(9);
  const int resident_counts[] = {1, 3, 8, 1000};
  const Spread spreads[] = {Spread::RoundRobin, Spread::AllInZero, Spread::Random};
  for (int trial = 0; trial < 6; ++trial) {
    const Batch b = random_batch(rng, 6 + (int)(rng() % 20));
    const gfft::PPlan p = plan_of(b, 4, 6, 1 + trial % 3);
    CHECK(gfft::check_plan(p) == 0);
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

int main() {
  plan_invariants_random();
  simulated_kernel_completes();
  no_stealing_with_absent_members_ends_by_the_bound();
  if (failures) {
    std::printf("%d failure(s)\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
