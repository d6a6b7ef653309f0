// views.hpp -- breadth-first node arrays of a cornerstone tree (TreeView).
#pragma once

#include "common.hpp"
#include "sort.hpp"

namespace at
{
struct ViewFlags;
struct VecShape;

// Logical size / capacity of one view output (the reference's thrust vector).
struct ViewShape
{
    std::size_t size = 0, cap = 0;
};

// Device view arrays of TreeView<b> (src/common/tree_view.hpp:20-115), b = 1
// (binary) or 3 (octree). All stages read the live leaf count from the
// device state, so a view build is capturable inside update().
class ViewEngine
{
public:
    explicit ViewEngine(unsigned levelBits);
    ~ViewEngine();
    ViewEngine(const ViewEngine &) = delete;
    ViewEngine &operator=(const ViewEngine &) = delete;

    // TreeView::prepare (tree_view.cu:755-793): allocate for leafCapacity
    // leaves and numParticles slots; no-op when no view is requested.
    void prepare(std::size_t leafCapacity, std::size_t numParticles, bool structure, bool counts, bool mapping,
                 hipStream_t s);
    // TreeView::syncSizes (tree_view.cu:795-812)
    void syncSizes(std::size_t numLeaves, std::size_t numActive);
    // TreeView::build (tree_view.cu:814-1134), state mode.
    void build(hipStream_t s, Key *const K[2], std::uint32_t *const N[2], const Key *P, std::size_t n,
               bool structure, bool counts, bool mapping, DevState *st);
    unsigned maxAchievedDepth(hipStream_t s) const;
    bool prepared() const { return scratch_; }
    bool depthAllocated() const { return depth_ != nullptr; }

    enum Out { BoxIndex, HasSplit, Depth, Parent, Child, LevelOffset, Begin, Counts, Mapping, kOuts };
    void buffer(Out o, const void **d, std::size_t *size, std::size_t *cap) const;
    std::size_t numNodes() const { return numNodes_; }

private:
    struct Impl;
    Impl *impl_;
    unsigned bits_;
    unsigned maxLevel_;
    std::size_t leafCap_ = 0, particleCap_ = 0;
    bool prepStructure_ = false, prepCounts_ = false, prepMapping_ = false;
    bool scratch_ = false;
    std::size_t numNodes_ = 0;
    std::uint8_t *depth_ = nullptr;
};
} // namespace at
