// free_list.hpp -- first-fit allocator over [0, size) in abstract units, the
// batch's arena, flag-arena and read-back-pool bookkeeping (host code only;
// tests/test_free_list.py checks it against a unit-by-unit model with g++).
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace hlbm {

struct FreeList {
  std::map<uint64_t, uint64_t> free;  // offset -> length of each free block, no two adjacent
  uint64_t size = 0, used = 0;
  void init(uint64_t n) {
    size = n;
    used = 0;
    free.clear();
    if (n) free[0] = n;
  }
  // the lowest-addressed free block of at least n units
  bool alloc(uint64_t n, uint64_t* off) {
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const uint64_t rest = it->second - n, roff = it->first + n;
        free.erase(it);
        if (rest) free[roff] = rest;
        used += n;
        return true;
      }
    }
    return false;
  }
  // returns [off, off + n) and merges it with the free blocks on either side
  void release(uint64_t off, uint64_t n) {
    used -= n;
    auto it = free.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
};

}  // namespace hlbm
