// adaptive_octree/config.hpp -- public types of the C++ facade over
// libadaptive_trees_hip.so (header-only; see adaptive_trees.h for the ABI).
//
// Names, defaults and meanings follow the reference's public header
// include/adaptive_octree/config.hpp:38-160 so that callers of
// adaptive_octree::Tree compile against this facade; the stream type is
// hipStream_t instead of cudaStream_t.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime_api.h>

#include "../adaptive_trees.h"

namespace adaptive_octree
{
using KeyType = unsigned long long;
using NodeIndex = int;
using PermIndex = std::uint32_t;

inline constexpr NodeIndex invalidNodeIndex = -1;
inline constexpr KeyType invalidParticleKey = std::numeric_limits<KeyType>::max();
inline constexpr KeyType invalidParticleNodeIndex = static_cast<KeyType>(invalidNodeIndex);
inline constexpr unsigned maxOctreeDepth = std::numeric_limits<KeyType>::digits / 3;  // 21
inline constexpr unsigned maxBinaryDepth = std::numeric_limits<KeyType>::digits - 1;  // 63
inline constexpr KeyType maxKey = KeyType(1) << maxBinaryDepth;

enum class TreeType { Octree, Binary };
enum class SplitCriterion { LeafCount, NFCount };
enum class SfcKind { Morton, Hilbert };

namespace defaults
{
inline constexpr unsigned maxParticlesPerLeaf = 64;
inline constexpr unsigned octreeUpdateIterations = 2;
inline constexpr unsigned binaryUpdateIterations = 6;
inline constexpr unsigned long long nearFieldLimit = 1ULL * maxParticlesPerLeaf * maxParticlesPerLeaf * 27;
inline constexpr std::size_t bufferGrowthDivisor = 4; // 25% capacity headroom
} // namespace defaults

template <class Real>
struct Box
{
    Real xmin = Real{0}, xmax = Real{1};
    Real ymin = Real{0}, ymax = Real{1};
    Real zmin = Real{0}, zmax = Real{1};
};

// Escaped particles keep their input slot and sort after the active keys.
template <class Real>
inline constexpr Real escapedParticleCoordinate = std::numeric_limits<Real>::max();

template <class Real>
constexpr bool isEscapedParticle(Real x, Real y, Real z)
{
    return x == escapedParticleCoordinate<Real> && y == escapedParticleCoordinate<Real> &&
           z == escapedParticleCoordinate<Real>;
}

struct ViewConfig
{
    bool treeStructure = false;
    bool particleCountsPerBox = false;
    bool particleMapping = false;
};

template <class Real>
struct OctreeConfig
{
    unsigned maxDepth = maxOctreeDepth; // a build-pass cap, not a depth limit
    unsigned maxParticlesPerLeaf = defaults::maxParticlesPerLeaf;
    SplitCriterion splitCriterion = SplitCriterion::LeafCount;
    SfcKind sfcKind = SfcKind::Morton;
    Box<Real> box{};
    ViewConfig views{};
    unsigned updateIterations = defaults::octreeUpdateIterations;
    hipStream_t stream = nullptr;
    bool interactive = true;
};

template <class Real>
struct KDTree3DConfig
{
    unsigned maxDepth = maxBinaryDepth;
    unsigned maxParticlesPerLeaf = defaults::maxParticlesPerLeaf;
    Box<Real> box{};
    ViewConfig views{};
    unsigned updateIterations = defaults::binaryUpdateIterations;
    hipStream_t stream = nullptr;
};

template <class Real>
struct TreeConfig
{
    int maxDepth = -1; // -1: backend default; explicit zero is valid
    unsigned maxParticlesPerLeaf = defaults::maxParticlesPerLeaf;
    int updateIterations = -1;
    Box<Real> box{};
    hipStream_t stream = nullptr;
    SplitCriterion splitCriterion = SplitCriterion::LeafCount;
    SfcKind sfcKind = SfcKind::Morton; // NFCount and Hilbert: octree only
    ViewConfig views{};
};

namespace detail
{
// Status code of the C ABI -> the exception type the reference throws.
inline void check(int status)
{
    if (status == AT_OK) return;
    const char *m = at_last_error();
    const std::string msg = m ? m : "adaptive_trees error";
    switch (status)
    {
    case AT_EINVAL: throw std::invalid_argument(msg);
    case AT_ELOGIC: throw std::logic_error(msg);
    case AT_ELENGTH: throw std::length_error(msg);
    case AT_ENOMEM: throw std::bad_alloc();
    default: throw std::runtime_error(msg); // HIP failure (the reference exits)
    }
}

template <class Real>
at_box toBox(const Box<Real> &b)
{
    return at_box{double(b.xmin), double(b.xmax), double(b.ymin), double(b.ymax), double(b.zmin), double(b.zmax)};
}
} // namespace detail
} // namespace adaptive_octree
