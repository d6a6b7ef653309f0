// batch.hpp -- many independent trees built and updated with shared launches.
#pragma once

#include "engine.hpp"

namespace at
{
class BatchBase
{
public:
    virtual ~BatchBase();
    virtual void build() = 0;
    virtual void update() = 0; // capturable into a HIP graph
    // Reads every member's state; flags[k] (if given) = member k needs a
    // resize. Returns the number of such members.
    virtual std::size_t needsResize(int *flags) = 0;
    // Grows the members that need it (the single-tree resizeBuffers rule).
    virtual void resizeBuffers() = 0;
    virtual std::size_t size() const = 0;
    virtual TreeBase &tree(std::size_t k) = 0;
    virtual hipStream_t stream() const = 0;
};

BatchBase *createBatch(int realBytes, int type, std::size_t T, const void *const *x, const void *const *y,
                       const void *const *z, const std::size_t *counts, const at_config &c);
} // namespace at
