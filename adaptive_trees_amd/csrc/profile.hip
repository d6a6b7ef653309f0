// profile.hip -- event timing of labelled kernel launches, plus the copy
// kernel used to calibrate hardware byte counters against known traffic.
#include "profile.hpp"

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/adaptive_trees.h"

namespace at
{
namespace prof
{
namespace
{
struct Record
{
    std::string label;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double ms = 0;
    unsigned long long launches = 0;
};
std::mutex g_mu;
std::vector<Record> g_records;
bool g_on = false;

int slotOf(const char *label)
{
    for (std::size_t i = 0; i < g_records.size(); ++i)
        if (g_records[i].label == label) return int(i);
    g_records.push_back(Record{label, {}, 0, 0});
    return int(g_records.size() - 1);
}

void drain(Record &r)
{
    for (auto &p : r.pending)
    {
        float ms = 0;
        if (hipEventSynchronize(p.second) == hipSuccess && hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess)
        {
            r.ms += ms;
            ++r.launches;
        }
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    r.pending.clear();
}
} // namespace

bool enabled() { return g_on; }
void enable(bool on) { g_on = on; }
void reset()
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &r : g_records)
    {
        drain(r);
        r.ms = 0;
        r.launches = 0;
    }
}
bool query(const char *label, double *ms, unsigned long long *launches)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &r : g_records)
        if (r.label == label)
        {
            drain(r);
            *ms = r.ms;
            *launches = r.launches;
            return true;
        }
    *ms = 0;
    *launches = 0;
    return false;
}

Scope::Scope(const char *label, hipStream_t s, bool active) : s_(s)
{
    if (!g_on || !active) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    hipEvent_t start = nullptr;
    if (hipEventCreate(&start) != hipSuccess) return;
    if (hipEventCreate(&stop_) != hipSuccess)
    {
        (void)hipEventDestroy(start);
        stop_ = nullptr;
        return;
    }
    (void)hipEventRecord(start, s);
    std::lock_guard<std::mutex> lk(g_mu);
    slot_ = slotOf(label);
    g_records[std::size_t(slot_)].pending.push_back({start, stop_});
}

Scope::~Scope()
{
    if (slot_ < 0) return;
    (void)hipEventRecord(stop_, s_);
}
} // namespace prof

namespace
{
__global__ void copyKernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, std::size_t n16)
{
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
         i += std::size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
} // namespace
} // namespace at

extern "C" {
int at_profile_enable(int on)
{
    at::prof::enable(on != 0);
    return AT_OK;
}
int at_profile_reset(void)
{
    at::prof::reset();
    return AT_OK;
}
int at_profile_query(const char *label, double *ms, unsigned long long *launches)
{
    return at::prof::query(label, ms, launches) ? AT_OK : AT_EINVAL;
}
// Streaming copy of `bytes` (a multiple of 16) from src to dst: the known
// traffic that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE readings.
int at_bench_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes % 16) return AT_EINVAL;
    const std::size_t n16 = bytes / 16;
    at::prof::Scope scope("copy", hipStream_t(stream));
    hipLaunchKernelGGL(at::copyKernel, dim3(2048), dim3(256), 0, hipStream_t(stream),
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    return hipGetLastError() == hipSuccess ? AT_OK : AT_EHIP;
}
}
