// engine.hpp -- host-side tree objects behind the C ABI.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/adaptive_trees.h"
#include "sort.hpp"
#include "tree_kernels.hpp"
#include "buffers.hpp"
#include "views.hpp"

namespace at
{
struct ViewFlags
{
    bool structure = false, counts = false, mapping = false;
    bool any() const { return structure || counts || mapping; }
};

class TreeBase
{
public:
    virtual ~TreeBase();
    virtual void build() = 0;
    virtual void update() = 0;
    virtual void computeViews(const ViewFlags *v) = 0;
    virtual bool needsResize() = 0;
    virtual void resizeBuffers() = 0;
    virtual void buffer(int id, const void **d, std::size_t *s, std::size_t *c) = 0;
    virtual void scalars(at_scalars *o) = 0;
    virtual unsigned maxAchievedDepth() = 0;
    hipStream_t stream() const { return stream_; }
    DevState *deviceState() const { return st_; }

protected:
    // st/sth: device and pinned host state slots owned by a batch (the tree
    // then allocates none of its own).
    explicit TreeBase(hipStream_t s, DevState *st = nullptr, DevState *sth = nullptr);
    void readState(); // synchronises the stream and copies the state to the host
    hipStream_t stream_ = nullptr;
    DevState *st_ = nullptr;  // device
    DevState *sth_ = nullptr; // pinned host copy
    bool ownsState_ = true;
    SortWorkspace sort_;
};

// Resources a batch lends to each member tree: its state slots and its
// segments of the batch's sorted keys and permutation.
struct MemberSlot
{
    DevState *st = nullptr;
    DevState *sth = nullptr;
    Key *P = nullptr;
    std::uint32_t *perm = nullptr;
};

// Message of the last failing C-ABI call on this thread.
std::string &lastError();

// The tree's print() text (print.cpp).
std::string printTree(TreeBase &t);

TreeBase *createTree(int realBytes, int type, const void *x, const void *y, const void *z, std::size_t n,
                     const at_config &c);
} // namespace at
