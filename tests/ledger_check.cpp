// Feeds fabricated control readings (status word, step count) to the
// library's step ledger (bowshock_amd/csrc/bs_ledger.hpp) and prints one
// line per reading: the code returned, the steps accounted after it, and
// the message. Built and run by tests/test_ledger.py with g++ (no GPU).
#include <cstdio>

#include "../bowshock_amd/csrc/bs_ledger.hpp"

int main() {
  struct Reading {
    unsigned status;
    long long steps, submitted;
  };
  // One solve: three good batches, then a short one, then one flagged.
  const Reading rs[] = {{0, 256, 256}, {0, 512, 256}, {0, 612, 100}, {0, 700, 256}, {1, 956, 256},
                        {0, 956, -1}, {0, 1000, -1}};
  long long accounted = 0;
  for (const Reading& r : rs) {
    char msg[256] = "";
    const bs_status st = bs::ledger_account(accounted, r.status, r.steps, r.submitted, msg, sizeof msg);
    std::printf("%d %lld %s\n", (int)st, accounted, msg);
  }
  return 0;
}
