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
// Span teams.  A transform whose share of members exceeds a domain's P slots
// (the largest transforms of a batch) gets a team split into n = ceil(k / P)
// *parts* of near-equal size, each on its own domain (the least-loaded with
// room).  Its level hand-over crosses L2s: a share's done word records the
// part of the workgroup that ran it (kShareDone + part); once every share of
// a level is done, for each part that ran shares one workgroup of that part
// (the first to claim the part's release word) writes its XCD's L2 back
// (agent-scope release) and marks the part released; the team's workgroups
// go on to the next level (agent-scope acquire) once every such part has
// released.  A part ran shares only if one of its workgroups arrived, and
// every arrived member walks every level, so each needed release happens.
//
// Solo transforms (the rest) form one queue, longest first.  Every
// workgroup, once its team work (if any) is over, takes solo transforms from
// the queue (one atomic ticket each) and runs each whole, level after level,
// with nothing shared outside the workgroup.
//
// Helpers.  A workgroup that finds the solo queue empty helps the teams of
// its own domain (same L2 as the team's part there, so the hand-over holds
// for it as for a member of that part): for each team with a part in the
// domain whose first member has arrived, it finds the current level (the
// first not complete: not all shares done, or, for a span team, a part that
// ran shares not yet released; it releases its own part when that is what
// is missing), claims a free share of it (the same compare-and-swap as a
// stealing member), runs it and marks it done; it leaves when every such
// team is complete.  Teams of which no member has arrived are left alone, so
// the orphan rule below still finds them untouched.
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
// walk through the levels.  A team of which no member arrives (none of its
// parts' domains received more workgroups than the part's first slot) is
// *orphaned*: the last workgroup to leave finds it from the arrival counters
// and runs it whole.
// Solo transforms wait for nobody, and helpers wait only while a team they
// may help is incomplete, which ends as above.  Hence the launch completes,
// with every tile run exactly once, whatever the residency and whatever the
// spread of workgroups over the domains.
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
  uint32_t dom, slot0;    // teams: part 0's domain and first local slot
  uint32_t nparts;        // teams: domains spanned (1: members slot0 .. slot0 + k - 1 of dom)
  uint32_t part0;         // teams: its parts are parts[part0 .. part0 + nparts)
};

struct PPart {            // the members of a team on one domain
  uint32_t dom, slot0;    // domain, first local slot
  uint32_t k, m0;         // members m0 .. m0 + k - 1 on local slots slot0 .. slot0 + k - 1
};

struct PEntry {           // what a local slot runs: member `member` (of part `part`) of team `team`
  uint32_t team, member, part;  // (team kNoTeam: none)
};

struct PPlan {
  std::vector<PSeg> segs;
  std::vector<PTeam> teams;     // one per transform
  std::vector<PPart> parts;     // of every team (one for a team in one domain)
  uint32_t domains = 1, per_dom = 1;
  std::vector<PEntry> local;    // domains * per_dom local slots, domain-major
  std::vector<uint32_t> solo;   // solo transforms, longest first
  std::vector<uint32_t> team_ids;  // transforms with k >= 2
  uint32_t claim_words = 0;     // claim words of all team shares
  // the teams of domain d, longest first: dom_teams[dom_off[d] .. dom_off[d + 1])
  std::vector<uint32_t> dom_off;
  std::vector<uint32_t> dom_teams;
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
// domain, then the share state words of every team level (a span team's
// level: its k q share words, then one release word per part).  The last
// workgroup to leave ORs the failure bits into the status word (which only
// the host clears) and zeroes every other word.
constexpr uint32_t kCtrExited = 0, kCtrFail = 1, kCtrStatus = 2, kCtrSolo = 3, kCtrArrive = 4,
                   kCtrShares = kCtrArrive + kMaxDomains;
constexpr uint32_t kFailTimeout = 1, kFailAborted = 2;
constexpr uint32_t kShareFree = 0, kShareClaimed = 1, kShareDone = 2;  // done: kShareDone + part
constexpr uint32_t kRelNone = 0, kRelClaimed = 1, kRelDone = 2;        // a span team part's release word
// claim words of one level of a team
GFFT_HD inline uint32_t level_words(uint32_t k, uint32_t q, uint32_t nparts) {
  return k * q + (nparts > 1 ? nparts : 0u);
}
inline size_t counter_words(const PPlan& p) { return kCtrShares + p.claim_words; }

// tiles(t, l), cls(t, l): callables; work(t): the transform's relative
// cost; domains D and local slots per domain P.  A transform gets a team
// when its work exceeds team_factor slot shares (share = total work / (D P)),
// of about work / share members (at most P, or max_parts P split over that
// many domains); min_team_tiles: a member owns at least this many tiles of
// the transform's smallest level; each member owns q = shares_per_member
// shares of a level (fewer where a level has too few tiles).
template <typename TilesFn, typename ClsFn, typename WorkFn>
PPlan build_plan(uint32_t ntransforms, int nlevels, const PLevelSpec* levels, TilesFn tiles, ClsFn cls,
                 WorkFn work, uint32_t domains, uint32_t per_dom, double team_factor = 2.0,
                 uint32_t min_team_tiles = 2, uint32_t shares_per_member = 2, uint32_t max_parts = 2) {
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
  // longest first onto the least-loaded domain that has room; a team of more
  // than P members splits into near-equal parts on as many domains
  std::vector<uint32_t> k(ntransforms, 1), q(ntransforms, 1);
  std::vector<std::vector<PPart>> tp(ntransforms);
  std::vector<uint32_t> used(domains, 0);
  std::vector<double> load(domains, 0.0);
  const uint32_t span = std::max(1u, std::min(max_parts, domains));
  for (uint32_t t : order) {
    uint32_t min_tiles = 0xffffffffu;
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (n) min_tiles = std::min(min_tiles, n);
    }
    if (min_tiles == 0xffffffffu) min_tiles = 0;
    uint32_t kk = w[t] > team_factor * share ? (uint32_t)(w[t] / share + 0.5) : 1u;
    kk = std::min(kk, std::max(1u, min_tiles / std::max(1u, min_team_tiles)));
    kk = std::min(kk, per_dom * span);
    if (kk < 2) continue;
    std::vector<PPart> parts;
    const uint32_t np = (kk + per_dom - 1) / per_dom;
    if (np > 1) {
      std::vector<char> taken(domains, 0);
      uint32_t m0 = 0;
      for (uint32_t i = 0; i < np; ++i) {
        const uint32_t kp = kk / np + (i < kk % np ? 1u : 0u);
        uint32_t best = domains;
        for (uint32_t d = 0; d < domains; ++d)
          if (!taken[d] && used[d] + kp <= per_dom && (best == domains || load[d] < load[best])) best = d;
        if (best == domains) break;
        taken[best] = 1;
        parts.push_back(PPart{best, used[best], kp, m0});
        m0 += kp;
      }
      if (parts.size() != np) {  // no room for the parts: one domain, as below
        parts.clear();
        kk = per_dom;
      }
    }
    if (parts.empty()) {
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
      parts.push_back(PPart{best, used[best], kk, 0u});
    }
    k[t] = kk;
    q[t] = std::max(1u, std::min(std::max(1u, shares_per_member), min_tiles / kk));
    for (const PPart& pt : parts) {
      used[pt.dom] += pt.k;
      load[pt.dom] += w[t] * pt.k / kk;
    }
    tp[t] = parts;
  }
  // levels and teams
  p.teams.resize(ntransforms);
  for (uint32_t t = 0; t < ntransforms; ++t) {
    PTeam& tm = p.teams[t];
    const bool team = k[t] >= 2;
    tm.seg0 = (uint32_t)p.segs.size();
    tm.k = k[t];
    tm.q = team ? q[t] : 1u;
    tm.dom = team ? tp[t][0].dom : 0u;
    tm.slot0 = team ? tp[t][0].slot0 : 0u;
    tm.nparts = team ? (uint32_t)tp[t].size() : 0u;
    tm.part0 = (uint32_t)p.parts.size();
    if (team) p.parts.insert(p.parts.end(), tp[t].begin(), tp[t].end());
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (!n) continue;
      PSeg s;
      s.xf = t;
      s.axis = levels[l].axis;
      s.inv = levels[l].inv;
      s.cls = cls(t, l);
      s.ntiles = n;
      s.claim0 = team ? p.claim_words : 0u;
      if (team) p.claim_words += level_words(tm.k, tm.q, tm.nparts);
      p.segs.push_back(s);
    }
    tm.nseg = (uint32_t)p.segs.size() - tm.seg0;
  }
  // local slots of the members; the solo queue, longest first
  p.local.assign((size_t)domains * per_dom, PEntry{kNoTeam, 0, 0});
  for (uint32_t t : order) {
    if (k[t] >= 2) {
      p.team_ids.push_back(t);
      for (uint32_t i = 0; i < (uint32_t)tp[t].size(); ++i) {
        const PPart& pt = tp[t][i];
        for (uint32_t m = 0; m < pt.k; ++m) p.local[(size_t)pt.dom * per_dom + pt.slot0 + m] = PEntry{t, pt.m0 + m, i};
      }
    } else if (p.teams[t].nseg > 0) {
      p.solo.push_back(t);
    }
  }
  // each domain's teams for its helpers (a span team under each of its
  // parts' domains), longest first (team_ids' order)
  p.dom_off.assign(domains + 1, 0u);
  for (uint32_t t : p.team_ids)
    for (const PPart& pt : tp[t]) p.dom_off[pt.dom + 1]++;
  for (uint32_t d = 0; d < domains; ++d) p.dom_off[d + 1] += p.dom_off[d];
  p.dom_teams.resize(p.dom_off.back());
  std::vector<uint32_t> fill(p.dom_off.begin(), p.dom_off.end() - 1);
  for (uint32_t t : p.team_ids)
    for (const PPart& pt : tp[t]) p.dom_teams[fill[pt.dom]++] = t;
  return p;
}

// The plan's invariants: every team's members occupy consecutive local
// slots of its domain, each local slot holds at most one member; every
// transform with levels is a team or in the solo queue exactly once; claim
// words of teams are disjoint; shares cover each level's tiles exactly
// once; each team is listed once, under its own domain, for the helpers.
// Returns 0, or a positive code.
inline int check_plan(const PPlan& p) {
  if (p.domains < 1 || p.domains > kMaxDomains || p.local.size() != (size_t)p.domains * p.per_dom) return 1;
  // parts: members m0 .. m0 + k - 1 in order, distinct domains, part 0 = (dom, slot0)
  for (const PTeam& tm : p.teams) {
    if (tm.k < 2) continue;
    if (tm.nparts < 1 || tm.nparts > p.domains || (size_t)tm.part0 + tm.nparts > p.parts.size()) return 12;
    uint32_t m0 = 0, doms = 0;
    for (uint32_t i = 0; i < tm.nparts; ++i) {
      const PPart& pt = p.parts[tm.part0 + i];
      if (pt.dom >= p.domains || pt.k < 1 || pt.m0 != m0 || pt.slot0 + pt.k > p.per_dom || (doms >> pt.dom) & 1u)
        return 12;
      doms |= 1u << pt.dom;
      m0 += pt.k;
    }
    if (m0 != tm.k || p.parts[tm.part0].dom != tm.dom || p.parts[tm.part0].slot0 != tm.slot0) return 12;
  }
  std::vector<int> seen(p.teams.size(), 0);
  for (size_t i = 0; i < p.local.size(); ++i) {
    const PEntry& e = p.local[i];
    if (e.team == kNoTeam) continue;
    if (e.team >= p.teams.size()) return 2;
    const PTeam& tm = p.teams[e.team];
    if (tm.k < 2 || e.member >= tm.k || e.part >= tm.nparts) return 2;
    const PPart& pt = p.parts[tm.part0 + e.part];
    if (e.member < pt.m0 || e.member >= pt.m0 + pt.k) return 3;
    if (i != (size_t)pt.dom * p.per_dom + pt.slot0 + (e.member - pt.m0)) return 3;
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
        for (uint32_t j = 0; j < level_words(tm.k, tm.q, tm.nparts); ++j) {
          if (sg.claim0 + j >= p.claim_words) return 8;
          if (claim[sg.claim0 + j]++) return 9;
        }
    }
  }
  // the helpers' lists: every team once under each of its parts' domains
  if (p.dom_off.size() != (size_t)p.domains + 1 || p.dom_off[0] != 0 || p.dom_off.back() != p.dom_teams.size())
    return 10;
  size_t want = 0;
  for (uint32_t t : p.team_ids) want += p.teams[t].nparts;
  if (p.dom_teams.size() != want) return 10;
  std::vector<uint32_t> listed(p.teams.size(), 0);  // bit d: listed under domain d
  for (uint32_t d = 0; d < p.domains; ++d) {
    if (p.dom_off[d] > p.dom_off[d + 1]) return 10;
    for (uint32_t i = p.dom_off[d]; i < p.dom_off[d + 1]; ++i) {
      const uint32_t t = p.dom_teams[i];
      if (t >= p.teams.size() || p.teams[t].k < 2 || (listed[t] >> d) & 1u) return 11;
      bool has = false;
      for (uint32_t j = 0; j < p.teams[t].nparts; ++j) has = has || p.parts[p.teams[t].part0 + j].dom == d;
      if (!has) return 11;
      listed[t] |= 1u << d;
    }
  }
  return 0;
}


// ---------------------------------------------------------------------------
// Global-level plan (team factor 0).  A team confined to one domain moves its
// transform through that domain's share of the memory system only (one XCD of
// eight: a 192^3 FP32 round trip took 0.58 ms that way against 0.09 ms over
// the whole device).  Here every workgroup takes part in every transform:
// level g of the launch holds the g-th non-empty level of every transform,
// each level's tiles are cut into shares of about equal cost (share j of a
// transform level of m shares: tiles j, j + m, ..., so neighbouring shares
// work on neighbouring tiles), and the host assigns the shares of a level to
// the slots longest first onto the least-loaded slot.  A workgroup claims its
// slot's list of the level with one compare-and-swap on the list's word, runs
// it, and publishes the number of shares it ran (see the hand-over below); it
// then waits for the level, and after a wait of steal_ticks claims and runs
// any list still unclaimed (a slot whose workgroup has not arrived).  The
// progress argument of the team plans holds unchanged: a list is claimed only
// by a running workgroup, and any waiting workgroup may claim an unclaimed
// one, so every level completes if one workgroup runs.
// ---------------------------------------------------------------------------
struct WShare {           // tiles first, first + step, ... (count of them) of one transform level
  uint32_t seg;           // the transform level (index into WPlan::segs)
  uint32_t first, step, count;
};

constexpr uint32_t kNoLevel = 0xffffffffu;

struct WPlan {
  std::vector<PSeg> segs;         // every transform's non-empty levels, in order per transform
  std::vector<uint32_t> seg_level;  // global level of each seg (kNoLevel: a solo transform's)
  std::vector<uint32_t> xseg0, xnseg;  // per transform: its segs
  std::vector<uint32_t> solo;     // transforms run whole by one workgroup, longest first
  uint32_t nlevels = 0;           // global levels of the launch
  uint32_t slots = 1;             // workgroups the shares are assigned to
  std::vector<WShare> shares;     // level-major
  std::vector<uint32_t> lvl_off;  // nlevels + 1: level g's shares are [lvl_off[g], lvl_off[g + 1])
  std::vector<uint32_t> own;      // share indices per (level, slot), in the order the slot runs them
  std::vector<uint32_t> own_off;  // nlevels * slots + 1: (g, s) owns own[own_off[g*slots+s] .. own_off[g*slots+s+1])
};

// Counter words of a global-level launch: the common header, per level one
// claim word per slot (its list of shares), then per level a block of
// kLevelWords words, each group on cache lines of its own (kFlagStride words
// apart):
//   done[d], d <= kMaxDomains: shares completed by the workgroups of domain d
//                (the last for workgroups without a domain);
//   released[d]: domain d's L2 is written back for the level;
//   go[d]:       domain d's workgroups may start the next level.
// Level hand-over: a workgroup's stores complete (they are in its domain's
// L2), then it adds its count to done[its domain] (no write-back).  The
// domain's first workgroup (its leader) sums the done words; when they reach
// the level's share count it writes its domain's L2 back once (an agent-scope
// release) if its domain did any of the level's work and sets released[d],
// waits until every domain that did work has done the same, and sets go[d],
// which the domain's other workgroups watch.  Each XCD's L2 is written back
// once per level, and no cache line is polled by more than one domain.
constexpr uint32_t kFlagStride = 32;
constexpr uint32_t kLevelDone = 0, kLevelReleased = (kMaxDomains + 1) * kFlagStride,
                   kLevelGo = kLevelReleased + kMaxDomains * kFlagStride,
                   kLevelWords = kLevelGo + kMaxDomains * kFlagStride;
inline size_t wide_level_base(const WPlan& p) { return kCtrShares + (size_t)p.nlevels * p.slots; }
inline size_t wide_counter_words(const WPlan& p) { return wide_level_base(p) + (size_t)p.nlevels * kLevelWords; }

// tiles(t, l), cls(t, l) as for build_plan; cost(t, l): relative cost of one
// tile of level l of transform t; every transform's k-th non-empty level goes
// to global level k; a level is cut into about slots * shares_per_slot shares.
// A transform whose work (sum of its levels' tile costs) is below
// solo_factor slot shares of the total (share = total / slots) goes to the
// solo queue instead: one workgroup runs it whole, taken from the queue while
// workgroups wait at level hand-overs (and after the last level).
template <typename TilesFn, typename ClsFn, typename CostFn>
WPlan build_wide_plan(uint32_t ntransforms, int nlevels, const PLevelSpec* levels, TilesFn tiles, ClsFn cls,
                      CostFn cost, uint32_t slots, uint32_t shares_per_slot = 2, double solo_factor = 0.0) {
  WPlan p;
  p.slots = std::max(slots, 1u);
  std::vector<double> work(ntransforms, 0.0);
  double all_work = 0;
  for (uint32_t t = 0; t < ntransforms; ++t) {
    for (int l = 0; l < nlevels; ++l) work[t] += (double)tiles(t, l) * std::max((double)cost(t, l), 1e-9);
    all_work += work[t];
  }
  std::vector<uint8_t> is_solo(ntransforms, 0);
  for (uint32_t t = 0; t < ntransforms; ++t)
    is_solo[t] = work[t] > 0 && work[t] < solo_factor * all_work / p.slots ? 1 : 0;
  std::vector<double> seg_cost;  // cost of one tile of each seg
  p.xseg0.assign(ntransforms, 0);
  p.xnseg.assign(ntransforms, 0);
  for (uint32_t t = 0; t < ntransforms; ++t) {
    uint32_t k = 0;
    p.xseg0[t] = (uint32_t)p.segs.size();
    for (int l = 0; l < nlevels; ++l) {
      const uint32_t n = tiles(t, l);
      if (!n) continue;
      PSeg s;
      s.xf = t;
      s.axis = levels[l].axis;
      s.inv = levels[l].inv;
      s.cls = cls(t, l);
      s.ntiles = n;
      s.claim0 = 0;  // unused in global-level plans
      p.segs.push_back(s);
      p.seg_level.push_back(is_solo[t] ? kNoLevel : k);
      seg_cost.push_back(std::max((double)cost(t, l), 1e-9));
      if (!is_solo[t]) p.nlevels = std::max(p.nlevels, k + 1);
      ++k;
    }
    p.xnseg[t] = (uint32_t)p.segs.size() - p.xseg0[t];
  }
  for (uint32_t t = 0; t < ntransforms; ++t)
    if (is_solo[t]) p.solo.push_back(t);
  std::stable_sort(p.solo.begin(), p.solo.end(), [&](uint32_t a, uint32_t b) { return work[a] > work[b]; });
  p.lvl_off.push_back(0);
  std::vector<std::vector<uint32_t>> per_slot(p.slots);
  for (uint32_t g = 0; g < p.nlevels; ++g) {
    double total = 0;
    for (size_t i = 0; i < p.segs.size(); ++i)
      if (p.seg_level[i] == g) total += seg_cost[i] * p.segs[i].ntiles;
    const double target = total / ((double)p.slots * std::max(1u, shares_per_slot));
    const uint32_t base = (uint32_t)p.shares.size();
    std::vector<double> share_cost;
    for (size_t i = 0; i < p.segs.size(); ++i) {
      if (p.seg_level[i] != g) continue;
      const uint32_t n = p.segs[i].ntiles;
      const double c = seg_cost[i] * n;
      uint32_t m = target > 0 ? (uint32_t)std::min<double>(n, std::max(1.0, std::floor(c / target + 0.5))) : 1u;
      m = std::max(1u, std::min(m, n));
      for (uint32_t j = 0; j < m; ++j) {
        const uint32_t cnt = share_tiles(n, m, j);
        p.shares.push_back(WShare{(uint32_t)i, j, m, cnt});
        share_cost.push_back(seg_cost[i] * cnt);
      }
    }
    // longest first onto the least-loaded slot (ties: the lowest slot)
    std::vector<uint32_t> order(share_cost.size());
    for (uint32_t j = 0; j < order.size(); ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return share_cost[a] > share_cost[b]; });
    std::vector<double> load(p.slots, 0.0);
    for (auto& v : per_slot) v.clear();
    // slots are a few hundred: a linear scan for the least-loaded one
    for (uint32_t j : order) {
      uint32_t best = 0;
      for (uint32_t s = 1; s < p.slots; ++s)
        if (load[s] < load[best]) best = s;
      load[best] += share_cost[j];
      per_slot[best].push_back(base + j);
    }
    for (uint32_t s = 0; s < p.slots; ++s) {
      p.own_off.push_back((uint32_t)p.own.size());
      for (uint32_t j : per_slot[s]) p.own.push_back(j);
    }
    p.lvl_off.push_back((uint32_t)p.shares.size());
  }
  p.own_off.push_back((uint32_t)p.own.size());
  return p;
}

// Invariants of a global-level plan: every tile of every seg in exactly one
// share, shares of level g only over segs of level g, a transform's levels on
// consecutive global levels in order, every share owned by exactly one slot of
// its level.  Returns 0, or a positive code.
inline int check_wide_plan(const WPlan& p) {
  if (p.slots < 1 || p.lvl_off.size() != (size_t)p.nlevels + 1 || p.seg_level.size() != p.segs.size() ||
      p.own_off.size() != (size_t)p.nlevels * p.slots + 1 || p.lvl_off.back() != p.shares.size())
    return 1;
  std::vector<uint8_t> solo(p.xseg0.size(), 0);
  for (uint32_t t : p.solo) {
    if (t >= solo.size() || solo[t]++) return 2;
  }
  for (size_t t = 0; t < p.xseg0.size(); ++t)
    for (uint32_t i = p.xseg0[t]; i < p.xseg0[t] + p.xnseg[t]; ++i)
      if (i >= p.segs.size() || p.segs[i].xf != t || (p.seg_level[i] == kNoLevel) != (solo[t] != 0)) return 2;
  for (size_t i = 0; i < p.segs.size(); ++i) {
    if (p.seg_level[i] == kNoLevel) continue;
    if (p.seg_level[i] >= p.nlevels) return 2;
    if (i > 0 && p.segs[i - 1].xf == p.segs[i].xf && p.seg_level[i] != p.seg_level[i - 1] + 1) return 2;
    if ((i == 0 || p.segs[i - 1].xf != p.segs[i].xf) && p.seg_level[i] != 0) return 2;
  }
  std::vector<std::vector<uint8_t>> seen(p.segs.size());
  for (size_t i = 0; i < p.segs.size(); ++i) seen[i].assign(p.segs[i].ntiles, 0);
  for (uint32_t g = 0; g < p.nlevels; ++g)
    for (uint32_t j = p.lvl_off[g]; j < p.lvl_off[g + 1]; ++j) {
      const WShare& w = p.shares[j];
      if (w.seg >= p.segs.size() || p.seg_level[w.seg] != g || w.step < 1) return 3;
      if (w.count != share_tiles(p.segs[w.seg].ntiles, w.step, w.first)) return 4;
      for (uint32_t k = 0; k < w.count; ++k) {
        const uint32_t tile = w.first + k * w.step;
        if (tile >= p.segs[w.seg].ntiles || seen[w.seg][tile]++) return 4;
      }
    }
  for (size_t i = 0; i < p.segs.size(); ++i)
    for (uint8_t v : seen[i])
      if (v != (p.seg_level[i] == kNoLevel ? 0 : 1)) return 5;
  std::vector<int> owned(p.shares.size(), 0);
  for (uint32_t g = 0; g < p.nlevels; ++g)
    for (uint32_t s = 0; s < p.slots; ++s)
      for (uint32_t i = p.own_off[g * p.slots + s]; i < p.own_off[g * p.slots + s + 1]; ++i) {
        const uint32_t j = p.own[i];
        if (j < p.lvl_off[g] || j >= p.lvl_off[g + 1] || owned[j]++) return 6;
      }
  for (int v : owned)
    if (v != 1) return 6;
  return 0;
}

}  // namespace gfft
