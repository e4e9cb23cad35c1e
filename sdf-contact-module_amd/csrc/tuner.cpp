// Autotuner: launch variants of a kernel (brick shapes and step budgets of
// the dense fill, points per lane and block sizes of the queries, occupancy of
// the contact forward, persistent grid of the contact backward) are timed on the caller's own work the
// first time a size class is seen, when the tuner is on, and the fastest is
// kept for the rest of the process.  Every variant computes the same result;
// only the schedule differs.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <sstream>

#include "sdfo_internal.hpp"

namespace sdfo {

namespace {

struct Entry
{
    int value = 0;
    std::vector<std::pair<int, double>> timings;
    std::vector<std::string> names;
};

std::atomic<int> gMode{SDFO_TUNE_OFF};
std::mutex gLock;
std::map<std::string, Entry>& table()
{
    static std::map<std::string, Entry> t;
    return t;
}

} // namespace

bool tunerEnabled() { return gMode.load() == SDFO_TUNE_ON_FIRST_USE; }

bool tunerLookup(const std::string& key, int& value)
{
    std::lock_guard<std::mutex> lk(gLock);
    auto it = table().find(key);
    if (it == table().end()) return false;
    value = it->second.value;
    return true;
}

void tunerRecord(const std::string& key, const std::vector<std::pair<int, double>>& timings,
                 const char* const* names)
{
    if (timings.empty()) return;
    Entry e;
    e.timings = timings;
    const auto best = std::min_element(timings.begin(), timings.end(),
                                       [](const auto& a, const auto& b) { return a.second < b.second; });
    e.value = best->first;
    for (const auto& t : timings) e.names.push_back(names ? names[t.first] : std::to_string(t.first));
    std::lock_guard<std::mutex> lk(gLock);
    table()[key] = e;
}

std::string tunerReport()
{
    std::lock_guard<std::mutex> lk(gLock);
    std::ostringstream os;
    for (const auto& kv : table())
    {
        const Entry& e = kv.second;
        os << kv.first << ": ";
        for (size_t q = 0; q < e.timings.size(); ++q)
        {
            if (q) os << ", ";
            os << e.names[q] << '=' << e.timings[q].second << "ms";
            if (e.timings[q].first == e.value) os << '*';
        }
        os << '\n';
    }
    return os.str();
}

int log2Bucket(int64_t n)
{
    int b = 0;
    while ((int64_t(1) << (b + 1)) <= n && b < 62) ++b;
    return b;
}

bool streamCapturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return true;
    return st != hipStreamCaptureStatusNone;
}

double timeOnStream(hipStream_t s, const std::function<void()>& f)
{
    hipEvent_t a = nullptr, b = nullptr;
    SDFO_HIP(hipEventCreate(&a));
    SDFO_HIP(hipEventCreate(&b));
    float ms = 0.0f;
    try
    {
        SDFO_HIP(hipEventRecord(a, s));
        f();
        SDFO_HIP(hipEventRecord(b, s));
        SDFO_HIP(hipEventSynchronize(b));
        SDFO_HIP(hipEventElapsedTime(&ms, a, b));
    }
    catch (...)
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms);
}

std::vector<std::pair<int, double>> timeVariants(hipStream_t s, int n,
                                                 const std::function<void(int)>& run, int rounds)
{
    std::vector<std::pair<int, double>> t;
    for (int v = 0; v < n; ++v) t.emplace_back(v, 1e30);
    for (int r = 0; r < rounds; ++r)
        for (int v = 0; v < n; ++v)
            t[size_t(v)].second = std::min(t[size_t(v)].second, timeOnStream(s, [&] { run(v); }));
    return t;
}

int tunedVariant(const std::string& key, hipStream_t s, int n, const char* const* names,
                 const std::function<void(int)>& run)
{
    int v = 0;
    if (tunerLookup(key, v)) return v;
    if (!tunerEnabled() || streamCapturing(s)) return 0;
    run(0);
    tunerRecord(key, timeVariants(s, n, run), names);
    tunerLookup(key, v);
    return v;
}

void tunerSetMode(int mode) { gMode.store(mode); }

void tunerReset()
{
    std::lock_guard<std::mutex> lk(gLock);
    table().clear();
}

} // namespace sdfo
