// pac_wait.hpp -- the host-side bounded wait for a solve's completion.
//
// No reference counterpart (the reference has no failure detection: a stuck loop just
// runs).  Header-only and free of HIP types so that its logic is tested on the CPU with a
// stand-in completion query.
#pragma once
#include <algorithm>
#include <chrono>

#include "../../include/pac.h"

namespace pac {

// Polls done() -- 1 complete, 0 still running, a negative PAC_ERR_* on failure -- until it
// reports completion or failure, or until timeout_ms of wall time have passed
// (timeout_ms <= 0: no bound).  Between polls it calls pause(us), starting at 1 us and
// doubling up to 1 ms.  Returns PAC_OK, the failure code, or PAC_ERR_TIMEOUT.
template <class Done, class Pause>
int wait_bounded(Done done, double timeout_ms, Pause pause) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  unsigned us = 1;
  for (;;) {
    const int d = done();
    if (d != 0) return d > 0 ? PAC_OK : d;
    if (timeout_ms > 0 &&
        std::chrono::duration<double, std::milli>(clock::now() - t0).count() >= timeout_ms)
      return PAC_ERR_TIMEOUT;
    pause(us);
    us = std::min(2 * us, 1000u);
  }
}

}  // namespace pac
