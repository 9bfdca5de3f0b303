// Work plan of the persistent batched kernel (host side, no device calls,
// so it is unit-tested and simulated on the CPU).
//
// One launch runs every axis pass ("level") of every transform of a batch
// mode (R2C: z, y, x; C2R: x, y, z; round trip: the six in that order).
// A level's tiles are independent; every tile of a level needs every tile of
// the transform's previous level (an axis pass reads the whole output of the
// pass before it).  Transforms are independent of each other.
//
// Domains.  The device's workgroups run in D cache domains (the XCDs of an
// MI355X, each with its own L2).  A workgroup learns its domain at run time
// (the hardware XCC id) and takes the next *local slot* of that domain (an
// arrival counter per domain); the plan gives each domain P local slots.
//
// Teams.  A transform whose work exceeds about two slots' shares gets a
// *team* of k >= 2 members, all in ONE domain: member m is local slot
// slot0 + m of the team's domain.  Each level is cut into n = q k *shares*
// (share s: the tiles s, s+n, s+2n, ..., so neighbouring shares work on
// neighbouring tiles); member m owns the shares m, m+k, ..., m+(q-1)k and
// runs them in that order.  Because every workgroup that touches a team's
// array shares one L2, the members' level hand-over needs only the L2-level
// ordering (stores complete to L2, L1 invalidated), not a write-back of the
// L2 to memory.
// Teams are placed on domains longest first, onto the least-loaded domain
// with room.
//
// Solo transforms (the rest) form one queue, longest first.  Every
// workgroup, once its team work (if any) is over, takes solo transforms from
// the queue (one atomic ticket each) and runs each whole, level after level,
// with nothing shared outside the workgroup.
//
// Team synchronisation, per level: every share has a state word (0 =
// unclaimed, 1 = claimed, 2 = done).  A member claims each of its shares in
// turn (one compare-and-swap on the share's own word, so no two workgroups
// contend on one address in the common case), runs it and marks it done,
// then waits until every share of the level is done.  While it waits it
// *steals*: claims any unclaimed share of the level (a slower member's later
// share) and runs it, which evens out the members' finishing times.
//
// Progress argument.  A wait ends once every share of the level has run.
// A share is claimed only by a running workgroup, which then runs it to the
// end; an unclaimed share can be claimed by any waiting member.  So if some
// member of the team is running, every share of the level gets run (by its
// owner or by a waiting member that steals it), for any number of resident
// workgroups; members that start late find their shares taken or done and
// walk through the levels.  A team of which no member arrives (its domain
// received fewer workgroups than slot0 + 1, or none) is *orphaned*: the last
// workgroup to leave finds it from the arrival counters and runs it whole.
// Solo transforms wait for nobody.  Hence the launch completes, with every
// tile run exactly once, whatever the residency and whatever the spread of
// workgroups over the domains.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

#if defined(__HIPCC__)
#define GFFT_HD __host__ __device__
#else
#define GFFT_HD
#endif

namespace gfft {

constexpr uint32_t kMaxDomains = 8;  // at most this many L2 domains (XCDs)
constexpr uint32_t kNoTeam = 0xffffffffu;

struct PSeg {             // one level of one transform
  uint32_t xf;            // transform (position in the batch's descriptor array)
  int32_t axis;           // 0 = z rows, 1 = y columns, 2 = x columns
  int32_t inv;            // 1 = the C2R direction
  int32_t cls;            // size class of the pass (selects the tile code)
  uint32_t ntiles;        // tiles of the level
  uint32_t claim0;        // first claim word of its shares (teams), in the counter array
};

struct PTeam {            // one transform
  uint32_t seg0, nseg;    // its non-empty levels, in order
  uint32_t k;             // members (1 = solo)
  uint32_t q;             // shares per member (teams)
  uint32_t dom, slot0;    // teams: domain, first local slot (members slot0 .. slot0 + k - 1)
};

struct PEntry {           // what a local slot runs: member `member` of team `team` (kNoTeam: none)
  uint32_t team, member;
};

struct PPlan {
  std::vector<PSeg> segs;
  std::vector<PTeam> teams;     // one per transform
  uint32_t domains = 1, per_dom = 1;
  std::vector<PEntry> local;    // domains * per_dom local slots, domain-major
  std::vector<uint32_t> solo;   // solo transforms, longest first
  std::vector<uint32_t> team_ids;  // transforms with k >= 2
  uint32_t claim_words = 0;     // claim words of all team shares
};

struct PLevelSpec {
  int32_t axis, inv;
};

// Tiles of share s of a level of `ntiles` tiles cut into n shares.
GFFT_HD inline uint32_t share_tiles(uint32_t ntiles, uint32_t n, uint32_t s) {
  return s < ntiles ? (ntiles - s + n - 1) / n : 0u;
}

// Counter words of one launch (device memory, zero between launches): the
// count of workgroups that left, the launch's failure bits (kFail*), a
// sticky status word, the solo queue's ticket, one arrival counter per
// domain, then the share state words of every team level.  The last
// workgroup to leave ORs the failure bits into the status word (which only
// the host clears) and zeroes every other word.
constexpr uint32_t kCtrExited = 0, kCtrFail = 1, kCtrStatus = 2, kCtrSolo = 3, kCtrArrive = 4,
                   kCtrShares = kCtrArrive + kMaxDomains;
constexpr uint32_t kFailTimeout = 1, kFailAborted = 2;
constexpr uint32_t kShareFree = 0, kShareClaimed = 1, kShareDone = 2;
inline size_t counter_words(const PPlan& p) { return kCtrShares + p.claim_words; }

// tiles(t, l), cls(t, l): callables; work(t): the transform's relative
// cost; domains D and local slots per domain P.  A transform gets a team
// when its work exceeds team_factor slot shares (share = total work / (D P)),
// of about work / share members (at most P); min_team_tiles: a member owns
// at least this many tiles of the transform's smallest level; each member
// owns q = shares_per_member shares of a level (fewer where a level has
// too few tiles).
template <typename TilesFn, typename ClsFn, typename WorkFn>
PPlan build_plan(uint32_t ntransforms, int nlevels, const PLevelSpec* levels, TilesFn tiles, ClsFn cls,
                 WorkFn work, uint32_t domains, uint32_t per_dom, double team_factor = 2.0,
                 uint32_t min_team_tiles = 2, uint32_t shares_per_member = 2) {
  PPlan p;
  domains = std::min(std::max(domains, 1u), kMaxDomains);
  per_dom = std::max(per_dom, 1u);
  p.domains = domains;
  p.per_dom = per_dom;
  std::vector<double> w(ntransforms);
  double total = 0;
  for (uint32_t t = 0; t < ntransforms; ++t) total += (w[t] = work(t));
  const double share = total > 0 ? total / ((double)domains * per_dom) : 1.0;
  std::vector<uint32_t> order(ntransforms);
  for (uint32_t t = 0; t < ntransforms; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return w[a] > w[b]; });
  // team sizes (about one slot share of work per member), then domains:
  // longest first onto the least-loaded domain that has room
  std::vector<uint32_t> k(ntransforms, 1), q(ntransforms, 1), dom(ntransforms, 0), slot0(ntransforms, 0);
  std::vector<uint32_t> used(domains, 0);
  std::vector<double> load(domains, 0.0);
  for (uint32_t t : order) {
    uint32_t min_tiles = 0xffffffffu;
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (n) min_tiles = std::min(min_tiles, n);
    }
    if (min_tiles == 0xffffffffu) min_tiles = 0;
    uint32_t kk = w[t] > team_factor * share ? (uint32_t)(w[t] / share + 0.5) : 1u;
    kk = std::min(kk, std::max(1u, min_tiles / std::max(1u, min_team_tiles)));
    kk = std::min(kk, per_dom);
    if (kk < 2) continue;
    uint32_t best = domains;
    for (uint32_t d = 0; d < domains; ++d)
      if (used[d] + kk <= per_dom && (best == domains || load[d] < load[best])) best = d;
    if (best == domains) {  // no domain has room for kk: the roomiest one, with a smaller team
      best = 0;
      for (uint32_t d = 1; d < domains; ++d)
        if (used[d] < used[best]) best = d;
      kk = per_dom - used[best];
      if (kk < 2) continue;
    }
    k[t] = kk;
    q[t] = std::max(1u, std::min(std::max(1u, shares_per_member), min_tiles / kk));
    dom[t] = best;
    slot0[t] = used[best];
    used[best] += kk;
    load[best] += w[t];
  }
  // levels and teams
  p.teams.resize(ntransforms);
  for (uint32_t t = 0; t < ntransforms; ++t) {
    PTeam& tm = p.teams[t];
    tm.seg0 = (uint32_t)p.segs.size();
    tm.k = k[t];
    tm.q = k[t] >= 2 ? q[t] : 1u;
    tm.dom = k[t] >= 2 ? dom[t] : 0u;
    tm.slot0 = k[t] >= 2 ? slot0[t] : 0u;
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (!n) continue;
      PSeg s;
      s.xf = t;
      s.axis = levels[l].axis;
      s.inv = levels[l].inv;
      s.cls = cls(t, l);
      s.ntiles = n;
      s.claim0 = tm.k >= 2 ? p.claim_words : 0u;
      if (tm.k >= 2) p.claim_words += tm.k * tm.q;
      p.segs.push_back(s);
    }
    tm.nseg = (uint32_t)p.segs.size() - tm.seg0;
  }
  // local slots of the members; the solo queue, longest first
  p.local.assign((size_t)domains * per_dom, PEntry{kNoTeam, 0});
  for (uint32_t t : order) {
    if (k[t] >= 2) {
      p.team_ids.push_back(t);
      for (uint32_t m = 0; m < k[t]; ++m) p.local[(size_t)dom[t] * per_dom + slot0[t] + m] = PEntry{t, m};
    } else if (p.teams[t].nseg > 0) {
      p.solo.push_back(t);
    }
  }
  return p;
}

// The plan's invariants: every team's members occupy consecutive local
// slots of its domain, each local slot holds at most one member; every
// transform with levels is a team or in the solo queue exactly once; claim
// words of teams are disjoint; shares cover each level's tiles exactly
// once.  Returns 0, or a positive code.
inline int check_plan(const PPlan& p) {
  if (p.domains < 1 || p.domains > kMaxDomains || p.local.size() != (size_t)p.domains * p.per_dom) return 1;
  std::vector<int> seen(p.teams.size(), 0);
  for (size_t i = 0; i < p.local.size(); ++i) {
    const PEntry& e = p.local[i];
    if (e.team == kNoTeam) continue;
    if (e.team >= p.teams.size()) return 2;
    const PTeam& tm = p.teams[e.team];
    if (tm.k < 2 || e.member >= tm.k) return 2;
    if (i != (size_t)tm.dom * p.per_dom + tm.slot0 + e.member) return 3;
    seen[e.team]++;
  }
  for (uint32_t t : p.solo) {
    if (t >= p.teams.size() || p.teams[t].k != 1) return 4;
    seen[t] += 1000;
  }
  for (uint32_t t : p.team_ids)
    if (t >= p.teams.size() || p.teams[t].k < 2) return 4;
  for (size_t t = 0; t < p.teams.size(); ++t) {
    const PTeam& tm = p.teams[t];
    if (tm.nseg == 0) continue;
    if (tm.k >= 2 ? seen[t] != (int)tm.k : seen[t] != 1000) return 5;
  }
  std::vector<int> claim(p.claim_words, 0);
  for (const PTeam& tm : p.teams) {
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg; ++s) {
      const PSeg& sg = p.segs[s];
      if (&p.teams[sg.xf] != &tm) return 6;
      if (tm.q < 1 || (tm.k < 2 && tm.q != 1)) return 7;
      const uint32_t n = tm.k * tm.q;
      uint32_t sum = 0;
      for (uint32_t j = 0; j < n; ++j) sum += share_tiles(sg.ntiles, n, j);
      if (sum != sg.ntiles) return 7;
      if (tm.k >= 2)
        for (uint32_t j = 0; j < n; ++j) {
          if (sg.claim0 + j >= p.claim_words) return 8;
          if (claim[sg.claim0 + j]++) return 9;
        }
    }
  }
  return 0;
}

}  // namespace gfft
