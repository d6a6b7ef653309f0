// profile.hpp -- optional per-kernel timing with HIP events.
#pragma once

#include "common.hpp"

namespace at
{
namespace prof
{
bool enabled();
void enable(bool on);
void reset();
// Sum of the recorded durations (ms) and launch count of one kernel label.
bool query(const char *label, double *ms, unsigned long long *launches);

// Records a start/stop event pair around the launches in its lifetime when
// timing is enabled, `active` is set and the stream is not being captured
// into a graph. An inactive scope makes no HIP call and no allocation.
class Scope
{
public:
    Scope(const char *label, hipStream_t s, bool active = true);
    ~Scope();
    Scope(const Scope &) = delete;
    Scope &operator=(const Scope &) = delete;

private:
    int slot_ = -1;
    hipStream_t s_ = nullptr;
    hipEvent_t stop_ = nullptr;
};
} // namespace prof
} // namespace at
