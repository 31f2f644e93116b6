// CPU model of the work-sharing protocol of the persistent engines (ws_queue.h), restated with
// C++ atomics on many host threads, plus an exhaustive check of the range arithmetic both GPU
// engines share (ws_split.h, compiled here unchanged).
//
// The reference's substrate (TaskBook.cuh:13-109 task append, GCQueue.cuh:30-107 tagged CAS
// queue, CBarrier.cuh:9-46 idle barrier) has no capacity check; the protocol modelled here
// reserves slots with a clamped fetch-add, publishes each record with a release store of its
// own flag, claims with a CAS on the head, and ends when no unit is left, no published task is
// unclaimed and nobody holds work, with every wait bounded.
//
// A "unit" u owns W(u) children (heavily skewed); a worker walks a range 64 children at a
// time and, when more than split + 64 are left, shares the tail as tasks, exactly as
// ws_ldescend / dyn_descend do. Every child must be visited exactly once, every record read
// only after its flag, every worker must leave, and a record whose flag never comes must end
// in the error flag, not a hang. The program stops at its first failure (exit code 1).
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "ws_split.h"

using namespace amj;

#define CHECK(c, ...)                                          \
  do {                                                         \
    if (!(c)) {                                                \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                       \
      std::fprintf(stderr, "\n");                              \
      std::exit(1);                                            \
    }                                                          \
  } while (0)

static uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// ---- 1. range arithmetic ----------------------------------------------------------------
static void check_split_math() {
  uint64_t cases = 0;
  for (uint64_t split = 64; split <= 512; split += 64)
    for (uint64_t c = 0; c <= 640; c += 64)
      for (uint64_t c1 = c; c1 <= c + 4 * 1024 + 70; ++c1) {
        if (!ws_should_split(c, c1, split)) {
          CHECK(c1 - c <= split + 64, "no split with %llu left", (unsigned long long)(c1 - c));
          continue;
        }
        const uint64_t k = ws_pieces(c, c1, split);
        CHECK(k >= 1, "no pieces");
        for (uint64_t got = 0; got <= k; ++got) {
          // kept [c, keep_end) and the published pieces must tile [c, c1) in order
          const uint64_t keep_end = ws_keep_end(c, c1, split, k, got);
          CHECK(keep_end >= c + split && keep_end <= c1, "keep_end out of range");
          CHECK(got == 0 || keep_end < c1, "a published piece is empty");
          uint64_t at = keep_end;
          for (uint64_t j = 0; j < got; ++j) {
            uint64_t p0 = 0, p1 = 0;
            ws_piece(c, c1, split, k, got, j, &p0, &p1);
            CHECK(p0 == at, "piece %llu starts at %llu, expected %llu", (unsigned long long)j,
                  (unsigned long long)p0, (unsigned long long)at);
            CHECK(p1 > p0 && p1 - p0 <= split && p1 <= c1, "bad piece [%llu,%llu)",
                  (unsigned long long)p0, (unsigned long long)p1);
            CHECK(p0 % 64 == c % 64, "piece not aligned with the walk");
            at = p1;
          }
          CHECK(at == (got ? c1 : keep_end), "pieces do not end at c1");
          if (got == 0) CHECK(keep_end == c1, "nothing published but the range shrank");
          ++cases;
        }
      }
  for (uint64_t cap = 1; cap <= 64; ++cap)
    for (uint64_t tail = 0; tail <= 80; ++tail)
      for (uint64_t want = 1; want <= 20; ++want) {
        const uint64_t n = ws_granted(tail, want, cap);
        CHECK(n <= want, "granted more than asked");
        CHECK(tail + n <= cap || n == 0, "granted past the capacity");
        CHECK(tail >= cap || n == std::min(want, cap - tail), "granted less than room");
      }
  std::printf("split arithmetic: %llu cases ok\n", (unsigned long long)cases);
}

// ---- 2. the protocol ------------------------------------------------------------------------
struct Ctl {
  std::atomic<uint64_t> next_unit{0}, q_tail{0}, q_head{0}, pushed{0}, progress{0};
  std::atomic<uint32_t> busy{0}, flags{0};
};

struct Task {
  uint64_t unit, c0, c1;  // the record: which unit's children [c0, c1)
};

struct Model {
  uint64_t units, split, cap, stall_ns;
  uint64_t max_claim = 64;  // WsArgs::max_claim (kMaxClaim when 0 there)
  int workers;
  bool withhold_one;  // a worker reserves a slot and never publishes it (the stuck case)
  std::vector<uint64_t> width;          // children per unit
  std::vector<uint64_t> first;          // offset of a unit's children in `visits`
  std::vector<std::atomic<uint32_t>> visits;
  std::vector<Task> tasks;
  std::vector<std::atomic<uint32_t>> ready;
  Ctl ctl;
  std::atomic<uint64_t> withheld{0};

  Model(uint64_t u, uint64_t sp, uint64_t cp, int w, uint64_t seed, bool withhold)
      : units(u), split(sp), cap(cp), stall_ns(2'000'000'000ull), workers(w),
        withhold_one(withhold), width(u), first(u + 1), tasks(cp), ready(cp) {
    uint64_t total = 0;
    for (uint64_t i = 0; i < u; ++i) {
      const uint64_t r = mix(seed ^ i);
      // skew: most units small, a few hubs with tens of thousands of children
      width[i] = (r % 100 < 3) ? 20000 + r % 40000 : r % 200;
      first[i] = total;
      total += width[i];
    }
    first[u] = total;
    visits = std::vector<std::atomic<uint32_t>>(total);
  }

  static uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }

  // reserve_tasks: one fetch-add, clamped
  uint64_t reserve(uint64_t k, uint64_t* got, bool* full) {
    const uint64_t tl = ctl.q_tail.fetch_add(k, std::memory_order_relaxed);
    *got = ws_granted(tl, k, cap);
    *full = tl + k >= cap;
    return *got ? tl : 0;
  }
  uint64_t queue_end() const {
    const uint64_t tl = ctl.q_tail.load(std::memory_order_relaxed);
    return tl < cap ? tl : cap;
  }

  // ws_ldescend over one unit's children [c0, c1)
  void walk(uint64_t unit, uint64_t c0, uint64_t c1, bool& qfull) {
    uint64_t end = c1;
    for (uint64_t c = c0; c < end; c += 64) {
      if (!qfull && ws_should_split(c, end, split)) {
        const uint64_t k = ws_pieces(c, end, split);
        uint64_t got = 0;
        bool full = false;
        const uint64_t slot = reserve(k, &got, &full);
        if (full) qfull = true;
        if (got) {
          for (uint64_t j = 0; j < got; ++j) {
            uint64_t p0 = 0, p1 = 0;
            ws_piece(c, end, split, k, got, j, &p0, &p1);
            CHECK(slot + j < cap, "slot %llu past capacity", (unsigned long long)(slot + j));
            tasks[slot + j] = Task{unit, p0, p1};
            if (withhold_one && withheld.exchange(1) == 0) continue;  // never published
            ready[slot + j].store(1, std::memory_order_release);
          }
          ctl.pushed.fetch_add(got, std::memory_order_relaxed);
          end = ws_keep_end(c, end, split, k, got);
        }
      }
      const uint64_t stop = std::min(c + 64, end);
      for (uint64_t g = c; g < stop; ++g) visits[first[unit] + g].fetch_add(1);
    }
  }

  // try_pop: -1 none, -2 the record never became ready
  int64_t try_pop() {
    uint64_t h = ctl.q_head.load(std::memory_order_relaxed);
    const uint64_t tl = queue_end();
    if (h < tl && ctl.q_head.compare_exchange_strong(h, h + 1, std::memory_order_relaxed)) {
      const uint64_t t0 = now_ns();
      while (ready[h].load(std::memory_order_relaxed) == 0) {
        if (now_ns() - t0 > stall_ns) {
          ctl.flags.fetch_or(1, std::memory_order_relaxed);
          return -2;
        }
        std::this_thread::yield();
      }
      (void)ready[h].load(std::memory_order_acquire);
      return (int64_t)h;
    }
    return -1;
  }

  // ws_serve
  void serve() {
    bool qfull = false;
    uint64_t seen = ~0ull, since = 0;
    for (;;) {
      uint64_t u = ~0ull, nu = 0;
      const uint64_t seen_u = ctl.next_unit.load(std::memory_order_relaxed);
      if (seen_u < units) {
        // guided claims: half the remaining units per worker and at least 4 with 64 or more
        // units per worker, else a quarter and at least 1; at most max_claim
        const bool many = units >= 64 * (uint64_t)workers;
        const uint64_t least = (many && max_claim >= 4) ? 4 : 1;
        uint64_t k = (units - seen_u) / ((many ? 2 : 4) * (uint64_t)workers);
        k = k < least ? least : (k > max_claim ? max_claim : k);
        u = ctl.next_unit.fetch_add(k, std::memory_order_relaxed);
        nu = k;
      }
      if (u < units) {
        const uint64_t ue = std::min(u + nu, units);
        // busy and progress serve only the waits for tasks
        if (split) ctl.busy.fetch_add(1, std::memory_order_relaxed);
        for (uint64_t v = u; v < ue; ++v) walk(v, 0, width[v], qfull);
        if (split) {
          ctl.progress.fetch_add(1, std::memory_order_relaxed);
          ctl.busy.fetch_sub(1, std::memory_order_relaxed);
        }
        continue;
      }
      if (split == 0) break;
      const int64_t slot = try_pop();
      if (slot >= 0) {
        ctl.busy.fetch_add(1, std::memory_order_relaxed);
        const Task t = tasks[(uint64_t)slot];
        CHECK(t.unit < units && t.c0 < t.c1 && t.c1 <= width[t.unit], "bad task record");
        walk(t.unit, t.c0, t.c1, qfull);
        ctl.progress.fetch_add(1, std::memory_order_relaxed);
        ctl.busy.fetch_sub(1, std::memory_order_relaxed);
        continue;
      }
      if (slot == -2) break;
      if (ctl.busy.load(std::memory_order_relaxed) == 0 &&
          ctl.q_head.load(std::memory_order_relaxed) >= queue_end())
        break;
      const uint64_t prog = ctl.progress.load(std::memory_order_relaxed);
      const uint64_t now = now_ns();
      if (prog != seen) {
        seen = prog;
        since = now;
      } else if (now - since > stall_ns) {
        ctl.flags.fetch_or(1, std::memory_order_relaxed);
        break;
      }
      std::this_thread::yield();
    }
  }

  // runs the workers; returns the error flag; every worker must have left
  uint32_t run() {
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w) th.emplace_back([this] { serve(); });
    for (auto& t : th) t.join();
    return ctl.flags.load();
  }
};

static void run_case(uint64_t units, uint64_t split, uint64_t cap, int workers, uint64_t seed,
                     uint64_t max_claim = 64) {
  Model m(units, split, cap, workers, seed, false);
  m.max_claim = max_claim;
  const uint32_t flags = m.run();
  CHECK(flags == 0, "error flag raised (units %llu split %llu cap %llu workers %d)",
        (unsigned long long)units, (unsigned long long)split, (unsigned long long)cap, workers);
  for (size_t i = 0; i < m.visits.size(); ++i)
    CHECK(m.visits[i].load() == 1, "child %zu visited %u times (split %llu cap %llu)", i,
          m.visits[i].load(), (unsigned long long)split, (unsigned long long)cap);
  CHECK(m.ctl.q_head.load() == std::min<uint64_t>(m.ctl.q_tail.load(), cap) || split == 0,
        "published tasks left unclaimed");
  std::printf("units %llu split %llu cap %llu workers %d: %zu children once each, %llu tasks\n",
              (unsigned long long)units, (unsigned long long)split, (unsigned long long)cap,
              workers, m.visits.size(), (unsigned long long)m.ctl.pushed.load());
}

int main() {
  check_split_math();
  const int hw = (int)std::max(2u, std::min(16u, std::thread::hardware_concurrency()));
  run_case(2000, 0, 16, hw, 1);        // no sharing
  run_case(2000, 1024, 8192, hw, 2);   // the ALFTJ default
  run_case(2000, 64, 8192, hw, 3);     // smallest tasks
  run_case(2000, 128, 7, hw, 4);       // queue fills: the rest is kept by its owner
  run_case(3000, 256, 1, 3, 5);        // one slot
  run_case(500, 64, 4096, 1, 6);       // one worker serves its own tasks
  run_case(4000, 192, 100000, 64, 7);  // many more workers than cores
  run_case(12000, 0, 16, hw, 9);       // 64 or more units per worker: claims of at least 4
  run_case(12000, 1024, 8192, hw, 10);  // ... with sharing
  run_case(12000, 0, 16, hw, 11, 2);   // claims capped below the minimum of 4
  run_case(12000, 256, 8192, hw, 12, 1);  // single-unit claims with sharing
  // a reserved slot whose flag never comes: the taker gives up with the error flag
  Model m(200, 64, 4096, hw, 8, true);
  m.stall_ns = 200'000'000ull;
  const uint32_t flags = m.run();
  CHECK(m.withheld.load() == 1, "nothing was withheld");
  CHECK(flags == 1, "a missing record did not raise the error flag");
  std::printf("withheld record: every worker left, error flag raised\n");
  std::printf("OK\n");
  return 0;
}
