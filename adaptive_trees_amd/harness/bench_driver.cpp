// bench_driver.cpp -- runs one dataset of the reference's benchmark harness
// through one tree of this library and writes the harness's CSV rows, so
// the reference's analysis scripts read the results like their own
// (protocol and columns: benchmarks/drivers/common.hpp:182-489 with the v3
// adapter of benchmarks/drivers/driver_v3.cu).
//
//   bench_driver --manifest M --config octree_leafcount|octree_nfcount|kdtree3d_leafcount
//                --mode steps|cold --version V --commit H --run R --out FILE [--max-snapshots N]
//
// Frame 0 is built from the root and given all views ("cold" writes one row
// with those times and stops). "steps" then loads each further frame and
// runs one update + views: replayed from a captured HIP graph (with event
// nodes around the two phases), or enqueued directly when
// BENCH_V3_PATH=direct. A step whose tree reports needsResize grows the
// buffers (timed as maintenance), is captured again and is repeated, at
// most 64 attempts. Structure columns are recounted from P and K.
// Exit status: 0 done, 3 out of device/host memory, 1 anything else.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "harness.hpp"

namespace
{
using namespace harness;

// The three configurations of the harness (common.hpp:196-202): tree type,
// criterion and bucket (LeafCount 64, NFCount 64*64*27).
struct Setup
{
    const char *name;
    int tree, criterion;
    unsigned bucket;
};
constexpr Setup kSetups[] = {
    {"octree_leafcount", AT_TREE_OCTREE, AT_LEAF_COUNT, 64},
    {"octree_nfcount", AT_TREE_OCTREE, AT_NF_COUNT, 64u * 64u * 27u},
    {"kdtree3d_leafcount", AT_TREE_BINARY, AT_LEAF_COUNT, 64},
};

const Setup &setupNamed(const std::string &name)
{
    for (const Setup &s : kSetups)
        if (name == s.name) return s;
    throw std::invalid_argument("unknown config " + name);
}

struct FrameRef
{
    int index = 0;
    long id = 0;
    std::string format, path;
};
struct Dataset
{
    std::string name;
    std::size_t slots = 0;
    double box[6]{};
    std::vector<FrameRef> frames;
};

Dataset openManifest(const std::string &path)
{
    void *m = nullptr;
    HARNESS_LIB(at_io_manifest_open(path.c_str(), &m));
    std::unique_ptr<void, void (*)(void *)> hold(m, at_io_manifest_close);
    Dataset d;
    std::vector<char> name(4096), fmt(64), file(8192);
    std::size_t frames = 0;
    HARNESS_LIB(at_io_manifest_info(m, name.data(), name.size(), &d.slots, d.box, &frames));
    d.name = name.data();
    d.frames.resize(frames);
    for (std::size_t i = 0; i < frames; ++i)
    {
        FrameRef &f = d.frames[i];
        HARNESS_LIB(at_io_manifest_frame(m, i, &f.index, &f.id, fmt.data(), fmt.size(), file.data(), file.size()));
        f.format = fmt.data();
        f.path = file.data();
    }
    return d;
}

void loadFrame(const FrameRef &f, std::vector<double> (&xyz)[3])
{
    double *cols = nullptr;
    std::size_t n = 0;
    HARNESS_LIB(at_io_read_frame(f.path.c_str(), f.format.c_str(), &cols, &n));
    std::unique_ptr<double, void (*)(void *)> hold(cols, at_io_free);
    for (int a = 0; a < 3; ++a) xyz[a].assign(cols + std::size_t(a) * n, cols + std::size_t(a + 1) * n);
}

// One step of the protocol: update() then all views, bracketed by the
// "begin" / "tree" / "views" events -- captured once into a graph (event
// nodes record when the graph runs) or enqueued directly on every step.
class Stepper
{
public:
    Stepper(TreeHandle &tree, hipStream_t s, bool graph, EventSet &ev) : tree_(tree), s_(s), graph_(graph), ev_(ev) {}
    ~Stepper() { drop(); }
    void capture()
    {
        if (!graph_) return;
        drop();
        HARNESS_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeGlobal));
        try
        {
            enqueue(hipEventRecordExternal);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        HARNESS_HIP(hipStreamEndCapture(s_, &g));
        const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HARNESS_HIP(e);
        HARNESS_HIP(hipGraphUpload(exec_, s_));
        HARNESS_HIP(hipStreamSynchronize(s_));
    }
    void step()
    {
        if (graph_) HARNESS_HIP(hipGraphLaunch(exec_, s_));
        else enqueue(hipEventRecordDefault);
    }
    void grow()
    {
        drop();
        tree_.resizeBuffers();
        capture();
    }

private:
    void enqueue(unsigned flags)
    {
        HARNESS_HIP(hipEventRecordWithFlags(ev_["begin"], s_, flags));
        tree_.update();
        HARNESS_HIP(hipEventRecordWithFlags(ev_["tree"], s_, flags));
        tree_.allViews();
        HARNESS_HIP(hipEventRecordWithFlags(ev_["views"], s_, flags));
    }
    void drop()
    {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    TreeHandle &tree_;
    hipStream_t s_;
    bool graph_;
    EventSet &ev_;
    hipGraphExec_t exec_ = nullptr;
};

const std::vector<std::string> kStepColumns = {
    "version", "commit", "path", "config", "dataset", "run", "snapshot", "frame_id", "tree_update_ms", "view_ms",
    "total_ms", "wall_ms", "attempts", "maintenance_ms", "iterations", "num_nodes", "num_leaves", "max_depth",
    "max_depth_octree_equiv", "num_empty_leaves", "empty_leaf_pct", "avg_particles_per_leaf", "particles_active",
    "particles_escaped", "particle_slots", "lib_num_nodes"};
const std::vector<std::string> kColdColumns = {
    "version", "commit", "path", "config", "dataset", "process_rep", "initial_build_ms", "initial_tree_gpu_ms",
    "initial_view_gpu_ms", "num_leaves", "num_nodes", "max_depth"};

struct Options
{
    std::string manifest, config, mode = "steps", version, commit, out;
    int run = 0, maxSnapshots = -1;
};

struct StepTimes
{
    double tree = NAN, view = NAN, total = NAN, wall = NAN, maintenance = 0;
    int attempts = 0;
};

int drive(int argc, char **argv)
{
    Options o;
    Flags flags;
    flags.text("--manifest", &o.manifest);
    flags.text("--config", &o.config);
    flags.text("--mode", &o.mode);
    flags.text("--version", &o.version);
    flags.text("--commit", &o.commit);
    flags.text("--out", &o.out);
    flags.integer("--run", &o.run);
    flags.integer("--max-snapshots", &o.maxSnapshots);
    const std::string usage = "usage: --manifest M --config C --mode steps|cold --version V --commit H --run R "
                              "--out FILE [--max-snapshots N]";
    flags.parse(argc, argv, usage);
    if (o.manifest.empty() || o.config.empty() || o.out.empty() || (o.mode != "steps" && o.mode != "cold"))
        throw std::invalid_argument(usage);
    const Setup &setup = setupNamed(o.config);
    const bool binary = setup.tree == AT_TREE_BINARY;
    const char *env = std::getenv("BENCH_V3_PATH");
    const bool graph = !(env && std::string(env) == "direct");
    const char *pathName = graph ? "graph" : "direct";

    HARNESS_HIP(hipFree(nullptr)); // runtime start-up, outside the measurements
    const Dataset d = openManifest(o.manifest);
    std::size_t freeB = 0, totalB = 0;
    HARNESS_HIP(hipMemGetInfo(&freeB, &totalB));
    std::cerr << "GPU memory free/total [MiB]: " << (freeB >> 20) << '/' << (totalB >> 20) << '\n';

    std::vector<double> xyz[3];
    loadFrame(d.frames.front(), xyz);
    if (xyz[0].size() != d.slots) throw std::runtime_error("frame 0 has a slot count other than the manifest's");
    DeviceXYZ coords(d.slots);
    coords.upload(xyz);

    hipStream_t s = nullptr;
    HARNESS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<std::FILE, int (*)(std::FILE *)> out(std::fopen(o.out.c_str(), "w"), std::fclose);
    if (!out) throw std::runtime_error("cannot open " + o.out);
    EventSet ev{"outer_begin", "begin", "tree", "views", "outer_end"};

    at_config cfg{};
    cfg.max_depth = -1;
    cfg.update_iterations = -1;
    cfg.max_particles_per_leaf = setup.bucket;
    cfg.split_criterion = setup.criterion;
    cfg.sfc_kind = AT_MORTON;
    cfg.box = at_box{d.box[0], d.box[1], d.box[2], d.box[3], d.box[4], d.box[5]};
    cfg.stream = s;

    // frame 0: construction + build from the root + first views
    HARNESS_HIP(hipDeviceSynchronize());
    const auto t0 = Clock::now();
    HARNESS_HIP(hipEventRecord(ev["begin"], s));
    TreeHandle tree(setup.tree, coords, cfg);
    tree.build();
    HARNESS_HIP(hipEventRecord(ev["tree"], s));
    tree.viewsUntilSized();
    HARNESS_HIP(hipEventRecord(ev["views"], s));
    HARNESS_HIP(hipDeviceSynchronize());
    const double coldWall = msSince(t0);

    auto measure = [&] { return census(tree.fetch<Key>(AT_BUF_PARTICLE_KEYS), tree.fetch<Key>(AT_BUF_CORNERSTONE), binary); };
    const LeafCensus first = measure();
    if (first.slots != d.slots) throw std::runtime_error("P does not hold one key per slot");

    if (o.mode == "cold")
    {
        CsvTable cold(out.get(), kColdColumns);
        cold.row()
            .put("version", o.version)
            .put("commit", o.commit)
            .put("path", pathName)
            .put("config", setup.name)
            .put("dataset", d.name)
            .put("process_rep", o.run)
            .put("initial_build_ms", gfmt(coldWall, 10))
            .put("initial_tree_gpu_ms", gfmt(ev.elapsed("begin", "tree"), 10))
            .put("initial_view_gpu_ms", gfmt(ev.elapsed("tree", "views"), 10))
            .put("num_leaves", (long long)first.leaves)
            .put("num_nodes", (long long)first.nodes)
            .put("max_depth", (long long)first.maxDepth)
            .write();
        return 0;
    }

    CsvTable steps(out.get(), kStepColumns);
    auto emit = [&](const FrameRef &f, const StepTimes &t, const LeafCensus &c) {
        const double lib = double(tree.scalars().num_nodes);
        steps.row()
            .put("version", o.version)
            .put("commit", o.commit)
            .put("path", pathName)
            .put("config", setup.name)
            .put("dataset", d.name)
            .put("run", o.run)
            .put("snapshot", f.index)
            .put("frame_id", f.id)
            .put("tree_update_ms", gfmt(t.tree, 10))
            .put("view_ms", gfmt(t.view, 10))
            .put("total_ms", gfmt(t.total, 10))
            .put("wall_ms", gfmt(t.wall, 10))
            .put("attempts", t.attempts)
            .put("maintenance_ms", gfmt(t.maintenance, 10))
            .put("iterations", -1) // fixed pass counts: the library reports no iteration count
            .put("num_nodes", (long long)c.nodes)
            .put("num_leaves", (long long)c.leaves)
            .put("max_depth", (long long)c.maxDepth)
            .put("max_depth_octree_equiv", gfmt(binary ? c.maxDepth / 3.0 : double(c.maxDepth), 10))
            .put("num_empty_leaves", (long long)c.empty)
            .put("empty_leaf_pct", gfmt(100.0 * double(c.empty) / double(c.leaves), 10))
            .put("avg_particles_per_leaf", gfmt(double(c.active) / double(c.leaves), 10))
            .put("particles_active", (long long)c.active)
            .put("particles_escaped", (long long)(c.slots - c.active))
            .put("particle_slots", (long long)c.slots)
            .put("lib_num_nodes", gfmt(lib, 10))
            .write();
    };
    emit(d.frames.front(), StepTimes{}, first);

    Stepper stepper(tree, s, graph, ev);
    stepper.capture();
    const std::size_t frames = o.maxSnapshots > 0 ? std::min<std::size_t>(std::size_t(o.maxSnapshots), d.frames.size())
                                                  : d.frames.size();
    for (std::size_t k = 1; k < frames; ++k)
    {
        const FrameRef &f = d.frames[k];
        loadFrame(f, xyz);
        if (xyz[0].size() != d.slots) throw std::runtime_error("frame " + f.path + " changes the slot count");
        coords.upload(xyz);
        StepTimes t{0, 0, 0, 0, 0, 0};
        for (;;)
        {
            HARNESS_HIP(hipDeviceSynchronize());
            const auto w0 = Clock::now();
            HARNESS_HIP(hipEventRecord(ev["outer_begin"], s));
            stepper.step();
            HARNESS_HIP(hipEventRecord(ev["outer_end"], s));
            const bool grow = tree.needsResize();
            HARNESS_HIP(hipDeviceSynchronize());
            t.wall += msSince(w0);
            t.tree += ev.elapsed("begin", "tree");
            t.view += ev.elapsed("tree", "views");
            t.total += ev.elapsed("outer_begin", "outer_end");
            ++t.attempts;
            if (!grow) break;
            if (t.attempts >= 64) throw std::runtime_error("frame " + f.path + ": capacity still short after 64 attempts");
            t.maintenance += wallMs([&] { stepper.grow(); });
        }
        emit(f, t, measure());
    }
    HARNESS_HIP(hipStreamDestroy(s));
    return 0;
}
} // namespace

int main(int argc, char **argv)
{
    try
    {
        return drive(argc, argv);
    }
    catch (const harness::OutOfMemory &e)
    {
        std::cerr << "bench_driver: " << e.what() << '\n';
        return 3;
    }
    catch (const std::bad_alloc &)
    {
        std::cerr << "bench_driver: out of host memory\n";
        return 3;
    }
    catch (const std::exception &e)
    {
        std::cerr << "bench_driver: " << e.what() << '\n';
        return 1;
    }
}
