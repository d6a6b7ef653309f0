// adaptive_octree/tree.hpp -- header-only C++ facade with the reference's
// class and method names (include/adaptive_octree/tree.hpp:11-134,
// octree.hpp, kdtree3d.hpp) over the C ABI of libadaptive_trees_hip.so.
//
//   adaptive_octree::Tree<Real, TreeType::Octree | TreeType::Binary>
//   adaptive_octree::Octree<Real>    (OctreeConfig<Real>)
//   adaptive_octree::KDTree3D<Real>  (KDTree3DConfig<Real>)
//
// Semantics are the reference's: coordinates are borrowed device pointers
// re-read by every update(); buffers are owned by the tree and keep their
// addresses between build()/resizeBuffers() calls; update() and
// computeViews() may be captured into a HIP graph on the configured stream.
// Errors throw the reference's exception types (std::invalid_argument,
// std::logic_error, std::length_error); a HIP failure throws
// std::runtime_error where the reference exits the process. Getters return
// `const DeviceVector<T>&` instead of `const thrust::device_vector<T>&`.
// OctreeConfig::interactive (a stdin prompt in the reference) is ignored.
#pragma once

#include <algorithm>
#include <climits>
#include <ostream>
#include <string>
#include <type_traits>
#include <utility>

#include "config.hpp"
#include "device_vector.hpp"

namespace adaptive_octree
{
template <class Real, TreeType Kind>
class Tree
{
    static_assert(std::is_same_v<Real, float> || std::is_same_v<Real, double>, "Real must be float or double");
    static_assert(Kind == TreeType::Octree || Kind == TreeType::Binary, "TreeType must be Octree or Binary");

public:
    using KeyType = adaptive_octree::KeyType;
    using NodeIndex = adaptive_octree::NodeIndex;
    using SplitCriterion = adaptive_octree::SplitCriterion;
    using SfcKind = adaptive_octree::SfcKind;
    using ViewConfig = adaptive_octree::ViewConfig;
    using Config = TreeConfig<Real>;

    Tree(const Real *x, const Real *y, const Real *z, std::size_t count, Config config = {})
    {
        if (config.maxDepth < -1 || config.updateIterations < -1)
            throw std::invalid_argument("Tree: use -1 for backend defaults");
        at_config c{};
        c.max_depth = config.maxDepth;
        c.max_particles_per_leaf = config.maxParticlesPerLeaf;
        c.update_iterations = config.updateIterations;
        c.split_criterion = config.splitCriterion == SplitCriterion::NFCount ? AT_NF_COUNT : AT_LEAF_COUNT;
        c.sfc_kind = config.sfcKind == SfcKind::Hilbert ? AT_HILBERT : AT_MORTON;
        setViews(c, config.views);
        c.box = detail::toBox(config.box);
        c.stream = config.stream;
        create(x, y, z, count, c);
    }

    Tree(Tree &&o) noexcept { swap(o); }
    Tree &operator=(Tree &&o) noexcept
    {
        if (this != &o)
        {
            Tree tmp(std::move(o));
            swap(tmp);
        }
        return *this;
    }
    Tree(const Tree &) = delete;
    Tree &operator=(const Tree &) = delete;
    ~Tree()
    {
        if (t_) at_tree_destroy(t_);
    }

    static constexpr TreeType type() { return Kind; }
    void build() { detail::check(at_tree_build(t_)); }
    void update() { detail::check(at_tree_update(t_)); }
    void resizeBuffers() { detail::check(at_tree_resize_buffers(t_)); }
    void computeViews() { detail::check(at_tree_compute_views(t_, nullptr)); }
    void computeViews(ViewConfig views)
    {
        const std::uint8_t v[3] = {std::uint8_t(views.treeStructure), std::uint8_t(views.particleCountsPerBox),
                                   std::uint8_t(views.particleMapping)};
        detail::check(at_tree_compute_views(t_, v));
    }
    // Summary of the host-visible state (the reference's print() layout is
    // not reproduced).
    // the reference's diagnostic dump (octree_print.cu, kdtree3d_print.cu)
    void print(std::ostream &out) const
    {
        std::size_t need = 0;
        detail::check(at_tree_print(t_, nullptr, 0, &need));
        std::string text(need, '\0');
        detail::check(at_tree_print(t_, text.data(), text.size(), &need));
        text.resize(need ? need - 1 : 0);
        out << text;
    }

    bool doBuffersNeedResize() const
    {
        int r = 0;
        detail::check(at_tree_needs_resize(t_, &r));
        return r != 0;
    }
    const NodeIndex *numLeaves_d() const { return counters().l; }
    const NodeIndex *numNodes_d() const { return counters().n; }
    const unsigned *numActiveParticles_d() const { return counters().a; }
    std::size_t numParticles() const { return std::size_t(scalars().num_particles); }
    std::size_t numActiveParticles() const { return std::size_t(scalars().num_active_particles); }
    std::size_t numNodes() const { return std::size_t(scalars().num_nodes); }
    unsigned maxAchievedDepth() const
    {
        std::uint32_t d = 0;
        detail::check(at_tree_max_achieved_depth(t_, &d));
        return d;
    }

    const DeviceVector<KeyType> &particleKeys_d() const { return get(AT_BUF_PARTICLE_KEYS, keys_); }
    const DeviceVector<PermIndex> &perm_d() const { return get(AT_BUF_PERM, perm_); }
    const DeviceVector<KeyType> &cornerstone_d() const { return get(AT_BUF_CORNERSTONE, cornerstone_); }
    const DeviceVector<unsigned> &leafCounts_d() const { return get(AT_BUF_LEAF_COUNTS, leafCounts_); }
    const DeviceVector<KeyType> &sfcBoxIndex_d() const { return get(AT_BUF_SFC_BOX_INDEX, boxIndex_); }
    const DeviceVector<std::uint8_t> &hasBoxSplit_d() const { return get(AT_BUF_HAS_BOX_SPLIT, hasSplit_); }
    const DeviceVector<std::uint8_t> &boxDepth_d() const { return get(AT_BUF_BOX_DEPTH, boxDepth_); }
    const DeviceVector<NodeIndex> &parentIndex_d() const { return get(AT_BUF_PARENT_INDEX, parent_); }
    const DeviceVector<NodeIndex> &childIndex_d() const { return get(AT_BUF_CHILD_INDEX, child_); }
    const DeviceVector<NodeIndex> &levelOffset_d() const { return get(AT_BUF_LEVEL_OFFSET, levelOffset_); }
    const DeviceVector<unsigned> &particleBeginIndex_d() const { return get(AT_BUF_PARTICLE_BEGIN_INDEX, begin_); }
    const DeviceVector<unsigned> &particleCounts_d() const { return get(AT_BUF_PARTICLE_COUNTS, counts_); }
    const DeviceVector<KeyType> &sfcParticleIndex_d() const { return get(AT_BUF_SFC_PARTICLE_INDEX, mapping_); }

    SplitCriterion splitCriterion() const
    {
        if constexpr (Kind == TreeType::Octree)
            return scalars().split_criterion == AT_NF_COUNT ? SplitCriterion::NFCount : SplitCriterion::LeafCount;
        else
            return SplitCriterion::LeafCount;
    }
    SfcKind sfcKind() const
    {
        if constexpr (Kind == TreeType::Octree)
            return scalars().sfc_kind == AT_HILBERT ? SfcKind::Hilbert : SfcKind::Morton;
        else
            throw std::logic_error("Tree: SFC selection requires Octree mode");
    }
    const DeviceVector<unsigned> &nearFieldCounts_d() const
    {
        if constexpr (Kind == TreeType::Octree)
            return get(AT_BUF_NEAR_FIELD_COUNTS, nearField_);
        else
            throw std::logic_error("Tree: Binary NFCount is not implemented yet");
    }

    // The C handle, for entry points beyond the reference surface
    // (at_graph_capture_update, at_tree_scalars).
    at_tree *handle() const { return t_; }
    at_scalars scalars() const
    {
        at_scalars s{};
        detail::check(at_tree_scalars(t_, &s));
        return s;
    }

protected:
    Tree() = default;
    static void setViews(at_config &c, const ViewConfig &v)
    {
        c.views[0] = v.treeStructure;
        c.views[1] = v.particleCountsPerBox;
        c.views[2] = v.particleMapping;
    }
    void create(const Real *x, const Real *y, const Real *z, std::size_t count, const at_config &c)
    {
        detail::check(at_tree_create(int(sizeof(Real)), Kind == TreeType::Octree ? AT_TREE_OCTREE : AT_TREE_BINARY,
                                     x, y, z, count, &c, &t_));
    }

private:
    struct Counters
    {
        const NodeIndex *l;
        const NodeIndex *n;
        const unsigned *a;
    };
    Counters counters() const
    {
        const std::int32_t *l = nullptr, *n = nullptr;
        const std::uint32_t *a = nullptr;
        detail::check(at_tree_device_counters(t_, &l, &n, &a));
        return {l, n, a};
    }
    template <class T>
    const DeviceVector<T> &get(int id, DeviceVector<T> &v) const
    {
        const void *d = nullptr;
        std::size_t s = 0, c = 0;
        detail::check(at_tree_buffer(t_, id, &d, &s, &c));
        v.reset(d, s, c);
        return v;
    }
    void swap(Tree &o) noexcept { std::swap(t_, o.t_); }

    at_tree *t_ = nullptr;
    mutable DeviceVector<KeyType> keys_, cornerstone_, boxIndex_, mapping_;
    mutable DeviceVector<PermIndex> perm_;
    mutable DeviceVector<unsigned> leafCounts_, nearField_, begin_, counts_;
    mutable DeviceVector<std::uint8_t> hasSplit_, boxDepth_;
    mutable DeviceVector<NodeIndex> parent_, child_, levelOffset_;
};

// Octree<Real> with its own configuration (include/adaptive_octree/octree.hpp).
template <class Real>
class Octree : public Tree<Real, TreeType::Octree>
{
public:
    using Config = OctreeConfig<Real>;
    Octree(const Real *x, const Real *y, const Real *z, std::size_t count, Config config = {})
    {
        at_config c{};
        c.max_depth = int(std::min<unsigned>(config.maxDepth, INT_MAX));
        c.max_particles_per_leaf = config.maxParticlesPerLeaf;
        c.update_iterations = int(std::min<unsigned>(config.updateIterations, INT_MAX));
        c.split_criterion = config.splitCriterion == SplitCriterion::NFCount ? AT_NF_COUNT : AT_LEAF_COUNT;
        c.sfc_kind = config.sfcKind == SfcKind::Hilbert ? AT_HILBERT : AT_MORTON;
        this->setViews(c, config.views);
        c.box = detail::toBox(config.box);
        c.stream = config.stream;
        this->create(x, y, z, count, c);
    }
};

// KDTree3D<Real>: the binary longest-side tree (include/adaptive_octree/kdtree3d.hpp).
template <class Real>
class KDTree3D : public Tree<Real, TreeType::Binary>
{
public:
    using Config = KDTree3DConfig<Real>;
    KDTree3D(const Real *x, const Real *y, const Real *z, std::size_t count, Config config = {})
    {
        at_config c{};
        c.max_depth = int(std::min<unsigned>(config.maxDepth, INT_MAX));
        c.max_particles_per_leaf = config.maxParticlesPerLeaf;
        c.update_iterations = int(std::min<unsigned>(config.updateIterations, INT_MAX));
        c.split_criterion = AT_LEAF_COUNT;
        c.sfc_kind = AT_MORTON;
        this->setViews(c, config.views);
        c.box = detail::toBox(config.box);
        c.stream = config.stream;
        this->create(x, y, z, count, c);
    }
};
} // namespace adaptive_octree
