// harness.hpp -- pieces shared by the two benchmark programs of this
// directory (bench_driver: the reference's per-frame driver protocol;
// bench_snapshots: its direct-vs-graph snapshot benchmark). Host C++ over
// the library's C ABI (include/adaptive_trees.h) and the HIP runtime.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "adaptive_trees.h"

namespace harness
{
using Key = std::uint64_t;
constexpr Key kKeyEnd = Key(1) << 63;                                  // K ends at the root's end
constexpr double kEscapedMarker = 1.7976931348623157e308;             // the escaped-slot marker (DBL_MAX)

// ---- errors ---------------------------------------------------------------------
struct OutOfMemory : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

inline void hipOrThrow(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory) throw OutOfMemory(std::string(what) + ": out of memory");
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
inline void libOrThrow(int status, const char *what)
{
    if (status == AT_OK) return;
    const std::string msg = std::string(what) + ": " + at_last_error();
    if (status == AT_ENOMEM) throw OutOfMemory(msg);
    if (status == AT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
#define HARNESS_HIP(x) ::harness::hipOrThrow((x), #x)
#define HARNESS_LIB(x) ::harness::libOrThrow((x), #x)

// ---- command lines --------------------------------------------------------------
// Options are registered with their target; parse() fills them from
// "--name value" pairs and bare switches, and rejects anything unknown.
class Flags
{
public:
    void text(const std::string &name, std::string *dst)
    {
        withValue_[name] = [dst](const std::string &v) { *dst = v; };
    }
    void integer(const std::string &name, int *dst)
    {
        withValue_[name] = [dst, name](const std::string &v) {
            std::size_t used = 0;
            const int x = std::stoi(v, &used);
            if (used != v.size()) throw std::invalid_argument(name + " needs an integer, got " + v);
            *dst = x;
        };
    }
    void toggle(const std::string &name, bool *dst, bool value) { switches_[name].push_back({dst, value}); }
    void parse(int argc, char **argv, const std::string &usage) const
    {
        for (int i = 1; i < argc; ++i)
        {
            const std::string a = argv[i];
            if (auto s = switches_.find(a); s != switches_.end())
            {
                for (const auto &[dst, v] : s->second) *dst = v;
                continue;
            }
            auto w = withValue_.find(a);
            if (w == withValue_.end() || i + 1 >= argc) throw std::invalid_argument(usage);
            w->second(argv[++i]);
        }
    }

private:
    std::map<std::string, std::function<void(const std::string &)>> withValue_;
    std::map<std::string, std::vector<std::pair<bool *, bool>>> switches_;
};

// ---- CSV ----------------------------------------------------------------------------
// A table of named columns; each row fills every column by name and is
// written in the header's order (a missing or unknown name is an error).
class CsvTable
{
public:
    CsvTable(std::FILE *out, std::vector<std::string> columns) : out_(out), cols_(std::move(columns))
    {
        for (std::size_t i = 0; i < cols_.size(); ++i) std::fprintf(out_, "%s%s", i ? "," : "", cols_[i].c_str());
        std::fputc('\n', out_);
        std::fflush(out_);
    }
    // A row is written by write(); one dropped unwritten is a logic error
    // of the caller (checked in debug builds).
    class Row
    {
    public:
        explicit Row(CsvTable &t) : t_(t), cells_(t.cols_.size()), set_(t.cols_.size(), false) {}
        Row &put(const std::string &name, const std::string &value)
        {
            const auto it = std::find(t_.cols_.begin(), t_.cols_.end(), name);
            if (it == t_.cols_.end()) throw std::logic_error("no CSV column " + name);
            const std::size_t i = std::size_t(it - t_.cols_.begin());
            cells_[i] = value;
            set_[i] = true;
            return *this;
        }
        Row &put(const std::string &name, long long v) { return put(name, std::to_string(v)); }
        void write()
        {
            for (std::size_t i = 0; i < cells_.size(); ++i)
            {
                if (!set_[i]) throw std::logic_error("CSV column " + t_.cols_[i] + " left empty");
                std::fprintf(t_.out_, "%s%s", i ? "," : "", cells_[i].c_str());
            }
            std::fputc('\n', t_.out_);
            std::fflush(t_.out_);
        }

    private:
        CsvTable &t_;
        std::vector<std::string> cells_;
        std::vector<bool> set_;
    };
    Row row() { return Row(*this); }

private:
    std::FILE *out_;
    std::vector<std::string> cols_;
};

// A double with `digits` significant digits as printf's %g writes it ("nan"
// for NaN): 10 digits is the driver CSV's format, 6 the snapshot CSV's.
inline std::string gfmt(double v, int digits)
{
    if (std::isnan(v)) return "nan";
    char b[64];
    std::snprintf(b, sizeof b, "%.*g", digits, v);
    return b;
}

// ---- timing -------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;
inline double msSince(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
// Wall milliseconds of f() between two device-wide synchronisations.
template <class F>
double wallMs(F &&f)
{
    HARNESS_HIP(hipDeviceSynchronize());
    const auto t0 = Clock::now();
    f();
    HARNESS_HIP(hipDeviceSynchronize());
    return msSince(t0);
}
// Named HIP events; elapsed(a, b) in milliseconds.
class EventSet
{
public:
    explicit EventSet(std::initializer_list<const char *> names)
    {
        for (const char *n : names) HARNESS_HIP(hipEventCreate(&ev_[n]));
    }
    ~EventSet()
    {
        for (auto &[n, e] : ev_) (void)hipEventDestroy(e);
    }
    EventSet(const EventSet &) = delete;
    EventSet &operator=(const EventSet &) = delete;
    hipEvent_t operator[](const std::string &n) const { return ev_.at(n); }
    double elapsed(const std::string &a, const std::string &b) const
    {
        float ms = 0;
        HARNESS_HIP(hipEventElapsedTime(&ms, ev_.at(a), ev_.at(b)));
        return ms;
    }

private:
    std::map<std::string, hipEvent_t> ev_;
};

// ---- device coordinates ---------------------------------------------------------------
class DeviceXYZ
{
public:
    explicit DeviceXYZ(std::size_t n) : n_(n)
    {
        for (auto &p : d_) HARNESS_HIP(hipMalloc(&p, std::max<std::size_t>(n, 1) * sizeof(double)));
    }
    ~DeviceXYZ()
    {
        for (auto p : d_) (void)hipFree(p);
    }
    DeviceXYZ(const DeviceXYZ &) = delete;
    DeviceXYZ &operator=(const DeviceXYZ &) = delete;
    void upload(const std::vector<double> *xyz)
    {
        for (int a = 0; a < 3; ++a)
        {
            if (xyz[a].size() != n_) throw std::runtime_error("coordinate count differs from the slot count");
            HARNESS_HIP(hipMemcpy(d_[a], xyz[a].data(), n_ * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    double *axis(int a) const { return d_[a]; }
    std::size_t size() const { return n_; }

private:
    std::size_t n_;
    double *d_[3] = {nullptr, nullptr, nullptr};
};

// ---- a tree through the C ABI ------------------------------------------------------------
class TreeHandle
{
public:
    TreeHandle(int treeType, const DeviceXYZ &xyz, const at_config &cfg) : type_(treeType)
    {
        HARNESS_LIB(at_tree_create(8, treeType, xyz.axis(0), xyz.axis(1), xyz.axis(2), xyz.size(), &cfg, &t_));
    }
    ~TreeHandle()
    {
        if (t_) at_tree_destroy(t_);
    }
    TreeHandle(const TreeHandle &) = delete;
    TreeHandle &operator=(const TreeHandle &) = delete;
    at_tree *get() const { return t_; }
    int type() const { return type_; }
    void build() { HARNESS_LIB(at_tree_build(t_)); }
    void update() { HARNESS_LIB(at_tree_update(t_)); }
    void allViews()
    {
        const std::uint8_t all[3] = {1, 1, 1};
        HARNESS_LIB(at_tree_compute_views(t_, all));
    }
    bool needsResize()
    {
        int r = 0;
        HARNESS_LIB(at_tree_needs_resize(t_, &r));
        return r != 0;
    }
    void resizeBuffers() { HARNESS_LIB(at_tree_resize_buffers(t_)); }
    at_scalars scalars() const
    {
        at_scalars s{};
        HARNESS_LIB(at_tree_scalars(t_, &s));
        return s;
    }
    // A host copy of one getter's buffer (its logical size).
    template <class T>
    std::vector<T> fetch(int id) const
    {
        const void *p = nullptr;
        std::size_t size = 0, cap = 0;
        HARNESS_LIB(at_tree_buffer(t_, id, &p, &size, &cap));
        std::vector<T> h(size);
        if (size && p) HARNESS_HIP(hipMemcpy(h.data(), p, size * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
    // Views (all three kinds), computed again after each capacity growth.
    void viewsUntilSized()
    {
        allViews();
        while (needsResize())
        {
            resizeBuffers();
            allViews();
        }
    }

private:
    at_tree *t_ = nullptr;
    int type_;
};

// Depth of a leaf from its key range: the octree level (3 key bits per
// level below the 63-bit root) or the binary depth (leading zeros).
inline unsigned leafDepth(Key range, bool binary)
{
    return binary ? unsigned(std::countl_zero(range)) : unsigned(std::countl_zero(range - 1) - 1) / 3u;
}

// Leaf structure recounted from the sorted keys P and the leaf boundaries K
// (not from the tree's own counts): one forward merge of the two sorted
// arrays gives lower_bound(P, K[i]) for every i.
struct LeafCensus
{
    std::size_t leaves = 0, nodes = 0, empty = 0, active = 0, slots = 0;
    unsigned maxDepth = 0;
};
inline LeafCensus census(const std::vector<Key> &P, const std::vector<Key> &K, bool binary)
{
    if (K.size() < 2 || K.front() != 0 || K.back() != kKeyEnd) throw std::runtime_error("K does not span the root");
    LeafCensus c;
    c.slots = P.size();
    c.leaves = K.size() - 1;
    std::size_t at = 0, first = 0, prev = 0;
    for (std::size_t i = 0; i < K.size(); ++i)
    {
        while (at < P.size() && P[at] < K[i]) ++at;
        if (i == 0) first = at;
        else
        {
            c.empty += at == prev;
            c.maxDepth = std::max(c.maxDepth, leafDepth(K[i] - K[i - 1], binary));
        }
        prev = at;
    }
    c.active = prev - first;
    const std::size_t children = binary ? 2 : 8;
    c.nodes = c.leaves + (c.leaves - 1) / (children - 1);
    return c;
}
} // namespace harness
