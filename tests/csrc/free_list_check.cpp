// Host check of free_list.hpp against a unit-by-unit model: random first-fit
// allocations and releases over small arenas.  Every allocation must be the
// lowest-addressed run of free units that fits (first fit), never overlap a
// live one, and fail only when no such run exists; `used` must equal the live
// units, and the free blocks must be exactly the maximal free runs (merged on
// release).  Prints a summary, exits non-zero on a failure.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../hippolbm_amd/csrc/free_list.hpp"

struct Live {
  uint64_t off, n;
};

int main() {
  int bad = 0;
  long allocs = 0, fails = 0, releases = 0;
  std::mt19937_64 rng(12345);
  for (int trial = 0; trial < 400 && bad < 10; trial++) {
    const uint64_t size = 1 + rng() % 200;
    hlbm::FreeList fl;
    fl.init(size);
    std::vector<char> busy(size, 0);
    std::vector<Live> live;
    for (int op = 0; op < 300 && bad < 10; op++) {
      if (live.empty() || rng() % 3 != 0) {
        const uint64_t n = 1 + rng() % (size / 4 + 1);
        // the model: lowest start of n free units
        int64_t want = -1;
        for (uint64_t s = 0, run = 0; s < size; s++) {
          run = busy[s] ? 0 : run + 1;
          if (run == n) {
            want = (int64_t)(s + 1 - n);
            break;
          }
        }
        uint64_t off = 0;
        const bool ok = fl.alloc(n, &off);
        if (ok != (want >= 0) || (ok && (int64_t)off != want)) {
          std::printf("trial %d op %d: alloc(%llu) gave %d at %llu, model %lld\n", trial, op, (unsigned long long)n,
                      (int)ok, (unsigned long long)off, (long long)want);
          bad++;
          continue;
        }
        if (ok) {
          for (uint64_t k = off; k < off + n; k++) busy[k] = 1;
          live.push_back({off, n});
          allocs++;
        } else {
          fails++;
        }
      } else {
        const size_t i = rng() % live.size();
        fl.release(live[i].off, live[i].n);
        for (uint64_t k = live[i].off; k < live[i].off + live[i].n; k++) busy[k] = 0;
        live.erase(live.begin() + (long)i);
        releases++;
      }
      uint64_t used = 0;
      for (char c : busy) used += c ? 1 : 0;
      // the free blocks are exactly the maximal runs of free units
      std::vector<std::pair<uint64_t, uint64_t>> runs;
      for (uint64_t s = 0; s < size;) {
        if (busy[s]) {
          s++;
          continue;
        }
        uint64_t e = s;
        while (e < size && !busy[e]) e++;
        runs.push_back({s, e - s});
        s = e;
      }
      std::vector<std::pair<uint64_t, uint64_t>> blocks(fl.free.begin(), fl.free.end());
      if (fl.used != used || blocks != runs) {
        std::printf("trial %d op %d: used %llu model %llu, %zu blocks vs %zu runs\n", trial, op,
                    (unsigned long long)fl.used, (unsigned long long)used, blocks.size(), runs.size());
        bad++;
      }
    }
  }
  std::printf("failures %d allocs %ld refused %ld releases %ld\n", bad, allocs, fails, releases);
  return bad == 0 ? 0 : 1;
}
