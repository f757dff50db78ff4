// Optional per-stage timing with HIP events recorded on the stream each stage runs on.
#pragma once
#include "common.hpp"

namespace pyl {

bool profiling_enabled();
// Records a start event in the constructor and a stop event in stop() on `st`.
struct StageTimer {
    StageTimer(const char* name, hipStream_t st);
    void stop();
    ~StageTimer() { stop(); }
    const char* name;
    hipStream_t st;
    hipEvent_t e0 = nullptr;
    bool live = false;
};

}  // namespace pyl
