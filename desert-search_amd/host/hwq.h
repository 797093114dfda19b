// Hardware queues of the process for the HIP runtime, set before the runtime starts.
#pragma once
#include <cstdlib>
#include <string>

// 20 queues serve the pipeline worker streams (20 with the tuned settings), the gate and the server's streams;
// streams beyond the queue count share queues. With blocking host waits and 20-24 workers, 16 or 20 queues
// measured the same and 24 or 28 queues far slower (about two thirds of the rate): keep 20.
// DSX_HW_QUEUES overrides the count, clamped to [1, 28]: 28 is the most measured since the stream-ordering
// fix of the flood engine's initialisation; the time-limit stalls once seen at 32 queues came before that
// fix, and 32 has not been measured since.
inline void dsx_set_hw_queues() {
    int q = 20;
    if (const char *e = std::getenv("DSX_HW_QUEUES"))
        if (*e) q = std::atoi(e);
    q = q < 1 ? 1 : (q > 28 ? 28 : q);
    setenv("GPU_MAX_HW_QUEUES", std::to_string(q).c_str(), 1);
}
