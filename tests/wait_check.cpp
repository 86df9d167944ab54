// CPU check of the bounded wait (pac_wait.hpp) with stand-in completion queries: no GPU,
// nothing that hangs.  Prints one line per case and exits non-zero on a failure.
#include <chrono>
#include <cstdio>
#include <thread>

#include "pac_wait.hpp"

namespace {
int failures = 0;
void expect(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
  if (!ok) ++failures;
}
void nap(unsigned us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); }
}  // namespace

int main() {
  using clock = std::chrono::steady_clock;
  // completes on the third poll
  int polls = 0;
  int rc = pac::wait_bounded([&] { return ++polls >= 3 ? 1 : 0; }, 50.0, nap);
  expect(rc == PAC_OK && polls == 3, "completion before the deadline returns PAC_OK");
  // never completes: PAC_ERR_TIMEOUT after about the deadline
  const auto t0 = clock::now();
  rc = pac::wait_bounded([] { return 0; }, 20.0, nap);
  const double ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  expect(rc == PAC_ERR_TIMEOUT, "a query that never completes returns PAC_ERR_TIMEOUT");
  expect(ms >= 20.0 && ms < 1000.0, "the timeout comes after the deadline, not long after");
  // a failing query returns its code at once
  polls = 0;
  rc = pac::wait_bounded([&] { ++polls; return PAC_ERR_HIP; }, 20.0, nap);
  expect(rc == PAC_ERR_HIP && polls == 1, "a failure is returned as is");
  // no bound: waits until completion
  polls = 0;
  rc = pac::wait_bounded([&] { return ++polls >= 40 ? 1 : 0; }, 0.0, nap);
  expect(rc == PAC_OK && polls == 40, "timeout 0 waits without bound");
  // the pause doubles from 1 us and stops at 1 ms
  unsigned last = 0, first = 0, n = 0;
  polls = 0;
  rc = pac::wait_bounded([&] { return ++polls >= 15 ? 1 : 0; }, 0.0, [&](unsigned us) {
    if (n++ == 0) first = us;
    last = us;
  });
  expect(first == 1 && last == 1000, "pauses grow from 1 us to at most 1 ms");
  return failures ? 1 : 0;
}
