// Work plan of the persistent batched kernel (host side, no device calls,
// so it is unit-tested and simulated on the CPU).
//
// One launch runs every axis pass ("level") of every transform of a batch
// mode (R2C: z, y, x; C2R: x, y, z; round trip: the six in that order).
// A level's tiles are independent; every tile of a level needs every tile of
// the transform's previous level (an axis pass reads the whole output of the
// pass before it).  Transforms are independent of each other.
//
// Plan.  The launch has G workgroups ("slots", G = the workgroups the device
// holds at once).  A transform whose work exceeds about two slots' shares
// gets a *team* of k >= 2 member slots; member m owns the tiles m, m+k,
// m+2k, ... of every level (its *share*), so neighbouring members work on
// neighbouring tiles.  Smaller transforms are *solo*: packed, longest first,
// onto the remaining slots, each run whole by one workgroup, level after
// level, with nothing shared outside the workgroup.
//
// Team synchronisation, per level: every share has a state word (0 =
// unclaimed, 1 = claimed, 2 = done).  A member claims its share (one
// compare-and-swap on the share's own word, so no two workgroups contend on
// one address in the common case), runs it, marks it done, and waits until
// every share of the level is done.  While it waits it may *steal*: claim
// any unclaimed share of the level and run it.
//
// Progress argument.  A wait ends once every share of the level has run.
// A share is claimed only by a running workgroup, which then runs it to the
// end; an unclaimed share can be claimed by any waiting member.  So if some
// member of the team is running, every share of the level gets run (by its
// owner or by a waiting member that steals it), for any number of resident
// workgroups; members that start late find their shares taken or done and
// walk through the levels.  Solo slots wait for nobody.  Hence the launch
// completes whatever the residency.
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
};

struct PSlot {            // one workgroup of the launch
  uint32_t first, count;  // entries [first, first + count) of the entry list
};

struct PEntry {           // what a slot runs: member `member` of team `team`
  uint32_t team, member;
};

struct PPlan {
  std::vector<PSeg> segs;
  std::vector<PTeam> teams;
  std::vector<PSlot> slots;
  std::vector<PEntry> entries;
  uint32_t claim_words = 0;  // claim words of all team shares
};

struct PLevelSpec {
  int32_t axis, inv;
};

// Tiles of member m's share of a level of `ntiles` tiles in a team of k.
GFFT_HD inline uint32_t share_tiles(uint32_t ntiles, uint32_t k, uint32_t m) {
  return m < ntiles ? (ntiles - m + k - 1) / k : 0u;
}

// Counter words of one launch (device memory, zero between launches): the
// count of workgroups that left, the launch's failure bits (kFail*), a
// sticky status word, then the share state words of every team level.  The
// last workgroup to leave ORs the failure bits into the status word (which
// only the host clears) and zeroes every other word.
constexpr uint32_t kCtrExited = 0, kCtrFail = 1, kCtrStatus = 2, kCtrShares = 4;
constexpr uint32_t kFailTimeout = 1, kFailAborted = 2;
constexpr uint32_t kShareFree = 0, kShareClaimed = 1, kShareDone = 2;
inline size_t counter_words(const PPlan& p) { return kCtrShares + p.claim_words; }

// tiles(t, l), cls(t, l): callables; work(t): the transform's relative
// cost; slots: G.  A transform gets a team when its work exceeds team_factor
// slot shares (share = total work / G), of about work / share members;
// min_team_tiles: a member owns at least this many tiles of the transform's
// smallest level.
template <typename TilesFn, typename ClsFn, typename WorkFn>
PPlan build_plan(uint32_t ntransforms, int nlevels, const PLevelSpec* levels, TilesFn tiles, ClsFn cls,
                 WorkFn work, uint32_t slots, double team_factor = 2.0, uint32_t min_team_tiles = 2) {
  PPlan p;
  if (slots < 1) slots = 1;
  std::vector<double> w(ntransforms);
  double total = 0;
  for (uint32_t t = 0; t < ntransforms; ++t) total += (w[t] = work(t));
  const double share = total > 0 ? total / slots : 1.0;
  // team sizes: about one slot share of work per member
  std::vector<uint32_t> order(ntransforms);
  for (uint32_t t = 0; t < ntransforms; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return w[a] > w[b]; });
  std::vector<uint32_t> k(ntransforms, 1);
  uint32_t members = 0, nsolo = 0;
  for (uint32_t t : order) {
    uint32_t min_tiles = 0xffffffffu;
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (n) min_tiles = std::min(min_tiles, n);
    }
    if (min_tiles == 0xffffffffu) min_tiles = 0;
    uint32_t kk = w[t] > team_factor * share ? (uint32_t)(w[t] / share + 0.5) : 1u;
    kk = std::min(kk, std::max(1u, min_tiles / std::max(1u, min_team_tiles)));
    k[t] = kk;
    if (kk >= 2)
      members += kk;
    else
      ++nsolo;
  }
  // keep slots for the solo transforms: scale the teams down if needed
  auto keep = [&] { return nsolo ? std::max(1u, std::min(nsolo, slots / 4)) : 0u; };
  while (members + keep() > slots && members > 0) {
    uint32_t big = order[0];
    for (uint32_t t : order)
      if (k[t] > k[big]) big = t;
    if (k[big] < 2) break;
    const uint32_t drop = k[big] == 2 ? 2u : 1u;
    k[big] = k[big] == 2 ? 1u : k[big] - 1;
    members -= drop;
    if (k[big] == 1) ++nsolo;
  }
  // levels and teams
  p.teams.resize(ntransforms);
  for (uint32_t t = 0; t < ntransforms; ++t) {
    PTeam& tm = p.teams[t];
    tm.seg0 = (uint32_t)p.segs.size();
    tm.k = k[t];
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
      if (tm.k >= 2) p.claim_words += tm.k;
      p.segs.push_back(s);
    }
    tm.nseg = (uint32_t)p.segs.size() - tm.seg0;
  }
  // slots: team members first (largest teams first), then the solo
  // transforms packed longest first onto the least loaded remaining slot
  std::vector<std::vector<PEntry>> sl;
  for (uint32_t t : order)
    if (k[t] >= 2)
      for (uint32_t m = 0; m < k[t]; ++m) sl.push_back({PEntry{t, m}});
  const uint32_t nteam_slots = (uint32_t)sl.size();
  const uint32_t free_slots = slots > nteam_slots ? slots - nteam_slots : 0u;
  const uint32_t nsolo_slots = std::min(free_slots, std::max(1u, nsolo));
  if (nsolo) {
    std::vector<double> load(nsolo_slots, 0.0);
    std::vector<std::vector<PEntry>> ss(nsolo_slots);
    for (uint32_t t : order) {
      if (k[t] >= 2) continue;
      const size_t j = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
      ss[j].push_back(PEntry{t, 0});
      load[j] += w[t];
    }
    for (auto& v : ss)
      if (!v.empty()) sl.push_back(v);
  }
  for (auto& v : sl) {
    PSlot s;
    s.first = (uint32_t)p.entries.size();
    s.count = (uint32_t)v.size();
    p.entries.insert(p.entries.end(), v.begin(), v.end());
    p.slots.push_back(s);
  }
  return p;
}

// The plan's invariants: every transform's levels appear in one team with
// contiguous levels; every member of every team and every solo transform
// appears in exactly one slot; claim words of teams are disjoint; shares
// cover each level's tiles exactly once.  Returns 0, or a positive code.
inline int check_plan(const PPlan& p) {
  std::vector<std::vector<int>> seen(p.teams.size());
  for (size_t t = 0; t < p.teams.size(); ++t) seen[t].assign(p.teams[t].k, 0);
  for (const PSlot& s : p.slots)
    for (uint32_t i = 0; i < s.count; ++i) {
      const PEntry& e = p.entries[s.first + i];
      if (e.team >= p.teams.size() || e.member >= p.teams[e.team].k) return 1;
      seen[e.team][e.member]++;
      if (p.teams[e.team].k >= 2 && s.count != 1) return 2;  // a member slot runs only that member
    }
  for (size_t t = 0; t < p.teams.size(); ++t)
    for (int c : seen[t])
      if (c != 1 && p.teams[t].nseg > 0) return 3;
  std::vector<int> claim(p.claim_words, 0);
  for (const PTeam& tm : p.teams) {
    for (uint32_t s = tm.seg0; s < tm.seg0 + tm.nseg; ++s) {
      const PSeg& sg = p.segs[s];
      if (&p.teams[sg.xf] != &tm) return 4;
      uint32_t sum = 0;
      for (uint32_t m = 0; m < tm.k; ++m) sum += share_tiles(sg.ntiles, tm.k, m);
      if (sum != sg.ntiles) return 5;
      if (tm.k >= 2)
        for (uint32_t m = 0; m < tm.k; ++m) {
          if (sg.claim0 + m >= p.claim_words) return 6;
          if (claim[sg.claim0 + m]++) return 7;
        }
    }
  }
  return 0;
}

}  // namespace gfft
