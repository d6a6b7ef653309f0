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

// Shared launches for independent octrees that each keep their own buffers
// (the trees of a server worker). Members of one call are octrees of one
// class (precision, split criterion, leaf size, key kind, box, depth and
// iteration limits) made by createTree on the runner's stream, with at least
// one particle each. Results are those of each member's own build() and
// update(). One runner per stream; its workspace grows to the largest call.
class GroupRunner
{
public:
    explicit GroupRunner(hipStream_t s);
    ~GroupRunner();
    GroupRunner(const GroupRunner &) = delete;
    GroupRunner &operator=(const GroupRunner &) = delete;
    // Members' build() with the device-driven passes of a batch; synchronises.
    // Not while updates are pending.
    void build(const std::vector<TreeBase *> &members);
    // Members' update() launched on the stream (any number of groups in
    // turn); readBack queues the copy of every launched member's state;
    // after the caller has synchronised the stream, finishUpdates applies
    // them: needsResize[g][k] = member k of the g-th launched group needs
    // the resize protocol. discardUpdates forgets the launched groups (after
    // a failed synchronisation).
    void launchUpdate(const std::vector<TreeBase *> &members);
    void readBack();
    void finishUpdates(std::vector<std::vector<char>> &needsResize);
    void discardUpdates();
    // Whether a tree of this type, configuration and size can join a group,
    // and the particles one group may hold.
    static bool groupable(int type, const at_config &c, std::size_t n);
    static std::size_t maxParticles();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};
} // namespace at
