// CPU check of the engine's parallel host copier (minhashcuda_amd/csrc/
// host_copy.h): copies of many shapes -- empty, one byte, around the chunk
// size, tens of MiB, many small tasks at odd offsets -- must reproduce their
// sources byte for byte, call after call (the workers' job hand-off), from
// several threads calling one copier at once (the serving loop's slots), and
// from a copier created and destroyed per call.  Built with g++ (also under
// ThreadSanitizer) by tests/test_host_copy.py; prints "ok" and exits 0 on
// success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <random>
#include <thread>
#include <vector>

#include "host_copy.h"

namespace {

bool run_case(wmh::HostCopier &copier, std::mt19937_64 &rng, size_t ntasks, size_t max_bytes) {
  std::vector<std::vector<unsigned char>> src(ntasks), dst(ntasks);
  std::vector<wmh::CopyTask> tasks;
  std::vector<size_t> offs(ntasks);
  for (size_t i = 0; i < ntasks; i++) {
    const size_t n = max_bytes ? rng() % (max_bytes + 1) : 0;
    offs[i] = rng() % 8;  // unaligned starts
    src[i].resize(n + offs[i]);
    dst[i].assign(n + offs[i] + 8, 0xA5);
    for (auto &b : src[i]) b = (unsigned char)rng();
    tasks.push_back(wmh::CopyTask{dst[i].data() + offs[i], src[i].data() + offs[i], n});
  }
  copier.copy(tasks.data(), tasks.size());
  for (size_t i = 0; i < ntasks; i++) {
    const size_t n = tasks[i].bytes;
    if (std::memcmp(dst[i].data() + offs[i], src[i].data() + offs[i], n) != 0) return false;
    for (size_t k = 0; k < offs[i]; k++)
      if (dst[i][k] != 0xA5) return false;  // nothing written before the range
    for (size_t k = offs[i] + n; k < dst[i].size(); k++)
      if (dst[i][k] != 0xA5) return false;  // nor after it
  }
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 40;
  std::mt19937_64 rng(7);
  wmh::HostCopier copier;
  const size_t shapes[][2] = {{0, 0},  {1, 1},        {1, 2u << 20},  {3, (2u << 20) + 1},
                              {2, 37u << 20}, {200, 5000}, {17, 3u << 20}, {1, 24u << 20}};
  for (int r = 0; r < rounds; r++) {
    for (const auto &sh : shapes) {
      if (!run_case(copier, rng, sh[0], sh[1])) {
        std::printf("mismatch: round %d, %zu tasks of up to %zu bytes\n", r, sh[0], sh[1]);
        return 1;
      }
    }
  }
  // four callers at once on one copier
  std::atomic<int> bad{0};
  std::vector<std::thread> callers;
  for (int t = 0; t < 4; t++) {
    callers.emplace_back([&, t] {
      std::mt19937_64 trng(100 + t);
      for (int r = 0; r < rounds; r++) {
        if (!run_case(copier, trng, 1 + trng() % 12, 3u << 20)) bad++;
        if (!run_case(copier, trng, 50, 3000)) bad++;
      }
    });
  }
  for (auto &c : callers) c.join();
  if (bad.load() != 0) {
    std::printf("mismatch: %d concurrent calls\n", bad.load());
    return 1;
  }
  for (int r = 0; r < 8; r++) {  // copiers created and destroyed around one call
    wmh::HostCopier once;
    if (!run_case(once, rng, 5, 6u << 20)) {
      std::printf("mismatch: short-lived copier %d\n", r);
      return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
