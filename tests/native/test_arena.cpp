// CPU unit test of the serving runtime's host bookkeeping (gfft_arena.h):
// first-fit arena with coalescing, strict-FIFO admission under a budget.
// Built and run by tests/test_serving_host.py with g++.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "gfft_arena.h"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
      ++failures;                                                  \
    }                                                              \
  } while (0)

static void arena_basics() {
  gfft::Arena a(1000);  // rounds down to 768 (three 256-byte blocks)
  CHECK(a.size() == 768);
  const size_t x = a.alloc(1);
  const size_t y = a.alloc(300);  // two blocks
  CHECK(x == 0 && y == 256);
  CHECK(a.used() == 768);
  CHECK(a.alloc(1) == gfft::Arena::kNone);
  a.release(x, 1);
  CHECK(a.largest_free() == 256);
  CHECK(a.alloc(257) == gfft::Arena::kNone);  // fragmentation: only one block free
  a.release(y, 300);
  CHECK(a.free_blocks() == 1 && a.largest_free() == 768 && a.used() == 0);
}

static void arena_coalesces_both_sides() {
  gfft::Arena a(256 * 5);
  size_t b[5];
  for (int i = 0; i < 5; ++i) b[i] = a.alloc(256);
  a.release(b[1], 256);
  a.release(b[3], 256);
  CHECK(a.free_blocks() == 2);
  a.release(b[2], 256);  // joins 1..3
  CHECK(a.free_blocks() == 1 && a.largest_free() == 768);
  CHECK(a.alloc(768) == 256);
}

// Random alloc/free against a byte map: no overlap, everything back at the end.
static void arena_random() {
  std::mt19937_64 rng(7);
  const size_t units = 4096;
  gfft::Arena a(units * 256);
  std::vector<int> owner(units, -1);
  struct B { size_t off, n; };
  std::vector<B> live;
  for (int step = 0; step < 20000; ++step) {
    if (live.empty() || rng() % 3 != 0) {
      const size_t n = 1 + rng() % (256 * 40);
      const size_t off = a.alloc(n);
      if (off == gfft::Arena::kNone) continue;
      CHECK(off % 256 == 0);
      for (size_t u = off / 256; u < (off + gfft::Arena::round(n)) / 256; ++u) {
        CHECK(u < units && owner[u] == -1);
        if (u < units) owner[u] = step;
      }
      live.push_back({off, n});
    } else {
      const size_t k = rng() % live.size();
      for (size_t u = live[k].off / 256; u < (live[k].off + gfft::Arena::round(live[k].n)) / 256; ++u) owner[u] = -1;
      a.release(live[k].off, live[k].n);
      live[k] = live.back();
      live.pop_back();
    }
  }
  for (const B& b : live) a.release(b.off, b.n);
  CHECK(a.used() == 0 && a.free_blocks() == 1 && a.largest_free() == units * 256);
}

static void admission_fifo() {
  gfft::Admission<int> q(100);
  CHECK(!q.fits(101));  // larger than the whole budget: never admissible
  q.push(1);
  q.push(2);
  q.push(3);
  // sizes: 1 -> 60, 2 -> 50, 3 -> 10; head-of-line: 3 must wait behind 2
  CHECK(q.fits(60));
  q.pop_admitted(60);
  CHECK(q.head() == 2 && !q.fits(50));
  CHECK(q.fits(10));  // 3 would fit, but the queue is strict FIFO: the caller stops at the head
  q.finished(60);
  CHECK(q.fits(50));
  q.pop_admitted(50);
  q.pop_admitted(10);
  CHECK(q.empty() && q.in_flight() == 60 && q.peak() == 60);
  q.finished(50);
  q.finished(10);
  CHECK(q.in_flight() == 0);
}

// Two FIFOs share one budget: the small queue's head is admitted while the
// large queue's head still waits; each queue stays in its own arrival order.
static void admission_two_queues() {
  gfft::Admission<int> q(100);
  q.push(1, 0);
  q.push(2, 0);
  q.push(3, 1);
  q.push(4, 1);
  CHECK(q.waiting() == 4);
  q.pop_admitted(70, 0);           // 1 -> 70
  CHECK(q.head(0) == 2 && !q.fits(50));
  CHECK(q.head(1) == 3 && q.fits(20));
  q.pop_admitted(20, 1);           // 3 runs beside 1 while 2 waits
  CHECK(q.head(1) == 4 && q.in_flight() == 90 && !q.empty(0));
  q.finished(70);
  q.pop_admitted(50, 0);
  q.pop_admitted(5, 1);
  CHECK(q.empty() && q.empty(0) && q.empty(1) && q.peak() == 90 && q.in_flight() == 75);
}

// A host-staged request whose rounded size exceeds the arena can never get a
// block, even when its unrounded size is within the budget.
static void arena_can_hold() {
  gfft::Arena a(1000);  // 768 usable
  CHECK(a.can_hold(768) && a.can_hold(700) && !a.can_hold(769) && !a.can_hold(1000));
  gfft::Arena b(0);
  CHECK(!b.can_hold(1));
}

static void stream_choice() {
  const size_t depth[4] = {1, 0, 2, 0}, bytes[4] = {50, 90, 10, 5};
  // no reservation, no depth limit: least bytes
  CHECK(gfft::pick_stream(depth, bytes, 4, false, false, 0) == 3);
  // the last stream is reserved: a large request takes the least loaded of 0..2
  CHECK(gfft::pick_stream(depth, bytes, 4, true, false, 0) == 2);
  CHECK(gfft::pick_stream(depth, bytes, 4, true, true, 0) == 3);
  // depth 2: stream 2 is full
  CHECK(gfft::pick_stream(depth, bytes, 4, true, false, 2) == 0);
  // depth 1: only streams 1 and 3 have room
  CHECK(gfft::pick_stream(depth, bytes, 4, true, false, 1) == 1);
  const size_t full[2] = {2, 2}, b2[2] = {1, 1};
  CHECK(gfft::pick_stream(full, b2, 2, false, false, 2) == -1);  // stays in its FIFO
  const size_t tie[3] = {0, 0, 0}, bt[3] = {7, 7, 7};
  CHECK(gfft::pick_stream(tie, bt, 3, false, false, 2) == 0);
}

int main() {
  admission_two_queues();
  arena_can_hold();
  stream_choice();
  arena_basics();
  arena_coalesces_both_sides();
  arena_random();
  admission_fifo();
  if (failures) {
    std::printf("%d failure(s)\n", failures);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
