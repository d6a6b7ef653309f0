// print.cpp -- the trees' print() diagnostics (src/octree/octree_print.cu:
// 16-382, src/kdtree3d/kdtree3d_print.cu:11-43): the same sections, lines
// and sample counts, built from host copies of the getters' buffers. The
// text is formatted on a stream in its default state (precision 6,
// defaultfloat), as the reference's print on a fresh std::cout. Host only.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <iomanip>
#include <limits>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/adaptive_trees.h"
#include "engine.hpp"

namespace
{
using at::TreeBase;
using Key = std::uint64_t;

constexpr std::size_t kPrintSamples = 4;       // defaults::printSamples (config.hpp:73)
constexpr std::size_t kVectorPrintSamples = 3; // defaults::octreeVectorPrintSamples (config.hpp:74)

template <class T>
std::vector<T> fetch(TreeBase &t, int id)
{
    const void *d = nullptr;
    std::size_t n = 0, cap = 0;
    t.buffer(id, &d, &n, &cap);
    std::vector<T> h(n); // zeros where a buffer has a size but no storage (NF of a LeafCount octree)
    if (n && d) AT_HIP(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

// one value as operator<< prints it after unary + (uint8 -> int)
template <class T>
auto promoted(T v)
{
    return +v;
}

// printArrayLine (octree_print.cu:20-54)
template <class T>
void arrayLine(std::ostream &out, const char *name, const std::vector<T> &v)
{
    out << std::left << std::setw(18) << name << ": size=" << v.size();
    if (!v.empty())
    {
        const auto [lo, hi] = std::minmax_element(v.begin(), v.end());
        out << "  min=" << promoted(*lo) << "  max=" << promoted(*hi);
    }
    out << "  samples=[";
    const std::size_t head = std::min(kVectorPrintSamples, v.size());
    for (std::size_t i = 0; i < head; ++i) out << (i ? ", " : "") << promoted(v[i]);
    if (v.size() > 2 * kVectorPrintSamples) out << ", ...";
    if (v.size() > kVectorPrintSamples)
        for (std::size_t i = std::max(head, v.size() - kVectorPrintSamples); i < v.size(); ++i)
            out << ", " << promoted(v[i]);
    out << "]\n";
}

// printSample (octree_print.cu:97-117)
template <class T>
void sample(std::ostream &out, const std::vector<T> &v, const char *name)
{
    out << name << " size = " << v.size() << '\n';
    if (v.empty()) return;
    const std::size_t head = std::min(kPrintSamples, v.size());
    for (std::size_t i = 0; i < head; ++i) out << "  " << name << "[" << i << "] = " << v[i] << '\n';
    if (v.size() > 2 * kPrintSamples) out << "  ...\n";
    const std::size_t tail = v.size() > kPrintSamples ? std::max(head, v.size() - kPrintSamples) : head;
    for (std::size_t i = tail; i < v.size(); ++i) out << "  " << name << "[" << i << "] = " << v[i] << '\n';
}

void printOctree(std::ostream &out, TreeBase &t, const at_scalars &sc)
{
    const auto P = fetch<Key>(t, AT_BUF_PARTICLE_KEYS);
    const auto perm32 = fetch<std::uint32_t>(t, AT_BUF_PERM);
    const std::vector<Key> Perm(perm32.begin(), perm32.end()); // printed as keys, like the reference
    const auto K = fetch<Key>(t, AT_BUF_CORNERSTONE);
    const auto N = fetch<std::uint32_t>(t, AT_BUF_LEAF_COUNTS);
    const auto NF = fetch<std::uint32_t>(t, AT_BUF_NEAR_FIELD_COUNTS);
    auto separator = [&] { out << "----------------------------------------------------------------\n"; };

    separator();
    out << "Adaptive Octree Summary\n";
    separator();
    const std::size_t nLeaves = K.empty() ? 0 : K.size() - 1;
    out << "Split criterion: " << (sc.split_criterion == AT_NF_COUNT ? "NFCount" : "LeafCount") << '\n';
    out << "SFC kind       : " << (sc.sfc_kind == AT_HILBERT ? "Hilbert" : "Morton") << '\n';
    out << "Particles      : " << sc.num_active_particles << '\n';
    out << "Input slots    : " << sc.num_particles << '\n';
    out << "Leaves         : " << nLeaves << '\n';
    out << '\n';

    separator();
    out << "Leaf occupancy\n";
    separator();
    if (!N.empty())
    {
        const auto [lo, hi] = std::minmax_element(N.begin(), N.end());
        const std::uint64_t total = std::accumulate(N.begin(), N.end(), std::uint64_t{0});
        const double avg = double(total) / double(N.size());
        const std::size_t empty = std::size_t(std::count(N.begin(), N.end(), 0u));
        out << "Min particles/leaf : " << *lo << '\n';
        out << "Max particles/leaf : " << *hi << '\n';
        out << "Avg particles/leaf : " << std::fixed << std::setprecision(2) << avg << '\n';
        out << "Empty leaves        : " << empty << " (" << std::fixed << std::setprecision(2)
            << 100.0 * double(empty) / double(N.size()) << "%)\n";
        out << std::defaultfloat;
    }
    out << '\n';

    separator();
    out << "Nearfield\n";
    separator();
    if (!NF.empty())
    {
        const auto [lo, hi] = std::minmax_element(NF.begin(), NF.end());
        const std::uint64_t total = std::accumulate(NF.begin(), NF.end(), std::uint64_t{0});
        const double avg = double(total) / double(NF.size());
        const std::size_t empty = std::size_t(std::count(NF.begin(), NF.end(), 0u));
        out << "Min interactions/leaf : " << *lo << '\n';
        out << "Max interactions/leaf : " << *hi << '\n';
        out << "Total number interactions   : " << total << '\n';
        out << "Avg interactions/leaf : " << std::fixed << std::setprecision(2) << avg << '\n';
        out << "Empty leaves           : " << empty << " (" << std::fixed << std::setprecision(2)
            << 100.0 * double(empty) / double(NF.size()) << "%)\n";
        out << std::defaultfloat;
    }
    out << '\n';

    separator();
    out << "Cornerstone leaf ranges\n";
    separator();
    if (K.size() >= 2)
    {
        Key lo = std::numeric_limits<Key>::max(), hi = 0;
        for (std::size_t i = 0; i + 1 < K.size(); ++i)
        {
            lo = std::min(lo, K[i + 1] - K[i]);
            hi = std::max(hi, K[i + 1] - K[i]);
        }
        out << "Smallest key range : " << lo << '\n';
        out << "Largest key range  : " << hi << '\n';
    }
    out << '\n';

    separator();
    out << "Array samples\n";
    separator();
    sample(out, P, "P");
    out << '\n';
    sample(out, Perm, "Perm");
    out << '\n';
    sample(out, K, "K");
    out << '\n';
    sample(out, N, "N");
    out << '\n';
    sample(out, NF, "NF");
    out << '\n';
    out << '\n';

    separator();
    out << "Sample leaves\n";
    separator();
    auto leaf = [&](std::size_t i) {
        if (i >= N.size() || i + 1 >= K.size()) return;
        out << "leaf[" << i << "]"
            << "  K=[" << K[i] << ", " << K[i + 1] << ")"
            << "  range=" << (K[i + 1] - K[i]) << "  particles=" << N[i] << '\n';
    };
    const std::size_t leafHead = std::min(kPrintSamples, nLeaves);
    for (std::size_t i = 0; i < leafHead; ++i) leaf(i);
    if (nLeaves > 2 * kPrintSamples) out << "...\n";
    if (nLeaves > kPrintSamples)
        for (std::size_t i = std::max(leafHead, nLeaves - kPrintSamples); i < nLeaves; ++i) leaf(i);
    out << '\n';
    separator();

    const auto box = fetch<Key>(t, AT_BUF_SFC_BOX_INDEX);
    if (!box.empty())
    {
        out << "Tree-structure view\n";
        separator();
        out << "maxAchievedDepth  : " << t.maxAchievedDepth() << '\n';
        arrayLine(out, "SFCBoxIndex", box);
        arrayLine(out, "hasBoxSplit", fetch<std::uint8_t>(t, AT_BUF_HAS_BOX_SPLIT));
        arrayLine(out, "boxDepth", fetch<std::uint8_t>(t, AT_BUF_BOX_DEPTH));
        arrayLine(out, "parentIndex", fetch<std::int32_t>(t, AT_BUF_PARENT_INDEX));
        arrayLine(out, "childIndex", fetch<std::int32_t>(t, AT_BUF_CHILD_INDEX));
        arrayLine(out, "levelOffset", fetch<std::int32_t>(t, AT_BUF_LEVEL_OFFSET));
        out << '\n';
    }
    const auto counts = fetch<std::uint32_t>(t, AT_BUF_PARTICLE_COUNTS);
    if (!counts.empty())
    {
        separator();
        out << "Per-box particle counts\n";
        separator();
        arrayLine(out, "particleBeginIdx", fetch<std::uint32_t>(t, AT_BUF_PARTICLE_BEGIN_INDEX));
        arrayLine(out, "particleCounts", counts);
        out << '\n';
    }
    const auto mapping = fetch<Key>(t, AT_BUF_SFC_PARTICLE_INDEX);
    if (!mapping.empty())
    {
        separator();
        out << "Particle <-> leaf mapping\n";
        separator();
        arrayLine(out, "SFCParticleIndex", mapping);
        out << '\n';
    }
    separator();
}

void printBinary(std::ostream &out, TreeBase &t, const at_scalars &sc)
{
    const auto K = fetch<Key>(t, AT_BUF_CORNERSTONE);
    const auto P = fetch<Key>(t, AT_BUF_PARTICLE_KEYS);
    const auto N = fetch<std::uint32_t>(t, AT_BUF_LEAF_COUNTS);
    const auto perm = fetch<std::uint32_t>(t, AT_BUF_PERM);
    out << "Adaptive 3D Binary Tree\n"
        << "LeafCount / longest-side binary keys\nParticles: " << sc.num_active_particles << " / "
        << sc.num_particles << " input slots\nLeaves: " << N.size() << "\nNodes: " << sc.num_nodes
        << "\nMaximum binary depth: " << t.maxAchievedDepth() << '\n';
    if (!N.empty())
    {
        const auto [lo, hi] = std::minmax_element(N.begin(), N.end());
        const auto sum = std::accumulate(N.begin(), N.end(), std::uint64_t(0));
        out << "N: min=" << *lo << " max=" << *hi << " sum=" << sum << " mean=" << double(sum) / double(N.size())
            << " empty=" << std::count(N.begin(), N.end(), 0u) << '\n';
    }
    auto samples = [&](const char *name, const auto &values) {
        out << name << " size=" << values.size() << " samples=[";
        for (std::size_t i = 0; i < std::min(kPrintSamples, values.size()); ++i) out << (i ? ", " : "") << values[i];
        out << (values.size() > kPrintSamples ? ", ...]\n" : "]\n");
    };
    samples("P", P);
    samples("Perm", perm);
    out << "K: leaf boundary array (" << K.size() << " keys)\n";
    auto leaf = [&](std::size_t i) {
        out << "  leaf[" << i << "] depth=" << unsigned(__builtin_clzll(K[i + 1] - K[i])) << " K=[" << K[i] << ','
            << K[i + 1] << ") N=" << N[i] << '\n';
    };
    const std::size_t head = std::min(kPrintSamples, N.size());
    for (std::size_t i = 0; i < head; ++i) leaf(i);
    if (N.size() > 2 * kPrintSamples) out << "  ...\n";
    for (std::size_t i = std::max(head, N.size() > kPrintSamples ? N.size() - kPrintSamples : 0); i < N.size(); ++i)
        leaf(i);
}
} // namespace

namespace at
{
std::string printTree(TreeBase &t)
{
    at_scalars sc{};
    t.scalars(&sc);
    std::ostringstream out;
    if (sc.tree_type == AT_TREE_BINARY) printBinary(out, t, sc);
    else printOctree(out, t, sc);
    return out.str();
}
} // namespace at
