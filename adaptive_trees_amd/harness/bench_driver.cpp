// bench_driver.cpp -- the reference's benchmark driver protocol and CSV
// schema (benchmarks/drivers/common.hpp:182-510 with the v3 adapter of
// benchmarks/drivers/driver_v3.cu) over this library's C++ facade:
//
//   bench_driver --manifest M --config octree_leafcount|octree_nfcount|kdtree3d_leafcount
//                --mode steps|cold --version V --commit H --run R --out FILE [--max-snapshots N]
//
// Frame 0 is built from the root (plus views); "cold" writes one row with
// its times. "steps" writes a row for frame 0 and then, for every further
// frame, uploads the coordinates and runs one update + views -- replayed
// from a captured HIP graph, or enqueued directly when BENCH_V3_PATH=direct
// -- repeating it after a capacity growth (timed as maintenance) while the
// tree reports needsResize. Structure columns are recounted from P and K.
// Exit status 3 = out of memory, 1 = any other failure.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "adaptive_octree/tree.hpp"

namespace ao = adaptive_octree;

namespace
{
using Key = ao::KeyType;
using Clock = std::chrono::steady_clock;
constexpr Key kRootEnd = Key(1) << 63;
constexpr int kOomExit = 3;

void hipOk(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define BENCH_HIP(x) hipOk((x), #x)

void ioOk(int status)
{
    if (status == AT_ENOMEM) throw std::bad_alloc();
    if (status != AT_OK) throw std::runtime_error(at_last_error());
}

// ---- input: manifest and frames through the library's readers ----
struct Frame
{
    int index;
    long id;
    std::string format, path;
};
struct Manifest
{
    std::string dataset;
    std::size_t slots = 0;
    double box[6]{};
    std::vector<Frame> frames;
};

Manifest readManifest(const std::string &path)
{
    void *h = nullptr;
    ioOk(at_io_manifest_open(path.c_str(), &h));
    std::unique_ptr<void, void (*)(void *)> guard(h, at_io_manifest_close);
    Manifest m;
    std::vector<char> name(4096), fmt(64), fpath(8192);
    std::size_t frames = 0;
    ioOk(at_io_manifest_info(h, name.data(), name.size(), &m.slots, m.box, &frames));
    m.dataset = name.data();
    for (std::size_t i = 0; i < frames; ++i)
    {
        Frame f{};
        ioOk(at_io_manifest_frame(h, i, &f.index, &f.id, fmt.data(), fmt.size(), fpath.data(), fpath.size()));
        f.format = fmt.data();
        f.path = fpath.data();
        m.frames.push_back(f);
    }
    return m;
}

void readFrame(const Frame &f, std::vector<double> &x, std::vector<double> &y, std::vector<double> &z)
{
    double *xyz = nullptr;
    std::size_t n = 0;
    ioOk(at_io_read_frame(f.path.c_str(), f.format.c_str(), &xyz, &n));
    std::unique_ptr<double, void (*)(void *)> guard(xyz, at_io_free);
    x.assign(xyz, xyz + n);
    y.assign(xyz + n, xyz + 2 * n);
    z.assign(xyz + 2 * n, xyz + 3 * n);
}

// ---- configuration and arguments (common.hpp:184-231) ----
enum class Kind { Octree, KDTree3D };
struct RunConfig
{
    std::string name;
    Kind tree;
    bool nf;
    unsigned limit;
    unsigned levelBits() const { return tree == Kind::Octree ? 3u : 1u; }
};

RunConfig parseConfig(const std::string &name)
{
    if (name == "octree_leafcount") return {name, Kind::Octree, false, 64};
    if (name == "octree_nfcount") return {name, Kind::Octree, true, 64u * 64u * 27u};
    if (name == "kdtree3d_leafcount") return {name, Kind::KDTree3D, false, 64};
    throw std::invalid_argument("unknown config " + name);
}

struct Args
{
    std::string manifest, config, mode = "steps", version, commit, out;
    int run = 0;
    int maxSnapshots = -1;
};

Args parseArgs(int argc, char **argv)
{
    Args a;
    for (int i = 1; i + 1 < argc; i += 2)
    {
        const std::string k = argv[i], v = argv[i + 1];
        if (k == "--manifest") a.manifest = v;
        else if (k == "--config") a.config = v;
        else if (k == "--mode") a.mode = v;
        else if (k == "--version") a.version = v;
        else if (k == "--commit") a.commit = v;
        else if (k == "--run") a.run = std::stoi(v);
        else if (k == "--out") a.out = v;
        else if (k == "--max-snapshots") a.maxSnapshots = std::stoi(v);
        else throw std::invalid_argument("unknown option " + k);
    }
    if (a.manifest.empty() || a.config.empty() || a.out.empty() || (a.mode != "steps" && a.mode != "cold"))
        throw std::invalid_argument("usage: --manifest M --config C --mode steps|cold --version V --commit H "
                                    "--run R --out FILE [--max-snapshots N]");
    return a;
}

// ---- structure, recounted from P and K (common.hpp:233-264) ----
struct Structure
{
    std::size_t leaves = 0, nodes = 0, emptyLeaves = 0, active = 0, slots = 0;
    unsigned maxDepth = 0;
};

Structure structure(const std::vector<Key> &P, const std::vector<Key> &K, unsigned levelBits)
{
    Structure s;
    s.slots = P.size();
    s.leaves = K.size() - 1;
    if (K.front() != 0 || K.back() != kRootEnd) throw std::runtime_error("K does not span the root");
    std::vector<std::size_t> pos(K.size());
    for (std::size_t i = 0; i < K.size(); ++i)
        pos[i] = std::size_t(std::lower_bound(P.begin(), P.end(), K[i]) - P.begin());
    s.active = pos.back() - pos.front();
    for (std::size_t i = 0; i + 1 < K.size(); ++i)
    {
        s.emptyLeaves += pos[i + 1] == pos[i];
        const unsigned rangeBits = unsigned(std::bit_width(K[i + 1] - K[i]) - 1);
        s.maxDepth = std::max(s.maxDepth, (63u - rangeBits) / levelBits);
    }
    s.nodes = s.leaves + (s.leaves - 1) / ((1u << levelBits) - 1);
    return s;
}

// ---- timing (common.hpp:268-298) ----
struct Events
{
    // begin/treeEnd/end bracket the phases (inside the graph on the graph
    // path); outerBegin/outerEnd bracket the whole submission
    hipEvent_t outerBegin{}, begin{}, treeEnd{}, end{}, outerEnd{};
    Events()
    {
        for (auto *e : {&outerBegin, &begin, &treeEnd, &end, &outerEnd}) BENCH_HIP(hipEventCreate(e));
    }
    ~Events()
    {
        for (auto e : {outerBegin, begin, treeEnd, end, outerEnd}) (void)hipEventDestroy(e);
    }
};

double elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    BENCH_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}
double msSince(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

// ---- CSV (common.hpp:300-336) ----
std::string num(double v)
{
    if (std::isnan(v)) return "nan";
    std::ostringstream s;
    s << std::setprecision(10) << v;
    return s.str();
}

constexpr const char *kStepHeader =
    "version,commit,path,config,dataset,run,snapshot,frame_id,tree_update_ms,view_ms,total_ms,wall_ms,"
    "attempts,maintenance_ms,iterations,num_nodes,num_leaves,max_depth,max_depth_octree_equiv,"
    "num_empty_leaves,empty_leaf_pct,avg_particles_per_leaf,particles_active,particles_escaped,particle_slots,"
    "lib_num_nodes";
constexpr const char *kColdHeader =
    "version,commit,path,config,dataset,process_rep,initial_build_ms,initial_tree_gpu_ms,initial_view_gpu_ms,"
    "num_leaves,num_nodes,max_depth";

struct StepTiming
{
    double tree = NAN, view = NAN, total = NAN, wall = NAN, maintenance = 0;
    int attempts = 0, iterations = -1;
};

void writeStep(std::ostream &o, const Args &a, const char *path, const RunConfig &c, const Manifest &m,
               const Frame &f, const StepTiming &t, const Structure &s, double libNodes)
{
    const double maxDepthEquiv = c.tree == Kind::KDTree3D ? s.maxDepth / 3.0 : double(s.maxDepth);
    o << a.version << ',' << a.commit << ',' << path << ',' << c.name << ',' << m.dataset << ',' << a.run << ','
      << f.index << ',' << f.id << ',' << num(t.tree) << ',' << num(t.view) << ',' << num(t.total) << ','
      << num(t.wall) << ',' << t.attempts << ',' << num(t.maintenance) << ',' << t.iterations << ',' << s.nodes
      << ',' << s.leaves << ',' << s.maxDepth << ',' << num(maxDepthEquiv) << ',' << s.emptyLeaves << ','
      << num(100.0 * double(s.emptyLeaves) / double(s.leaves)) << ',' << num(double(s.active) / double(s.leaves))
      << ',' << s.active << ',' << s.slots - s.active << ',' << s.slots << ',' << num(libNodes) << '\n';
    o.flush();
}

// ---- the v3 adapter (driver_v3.cu:14-121) on Tree<double, Octree|Binary> ----
template <ao::TreeType TK>
class V3
{
public:
    using Tree = ao::Tree<double, TK>;
    static bool useGraph()
    {
        static const bool g = !(std::getenv("BENCH_V3_PATH") && std::string(std::getenv("BENCH_V3_PATH")) == "direct");
        return g;
    }
    static const char *pathName() { return useGraph() ? "graph" : "direct"; }
    static bool supports(const RunConfig &c) { return (c.tree == Kind::KDTree3D) == (TK == ao::TreeType::Binary); }
    static hipStream_t stream()
    {
        static hipStream_t s = [] {
            hipStream_t t;
            BENCH_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
            return t;
        }();
        return s;
    }

    V3(const RunConfig &c, const Manifest &m, double *x, double *y, double *z, std::size_t n, hipStream_t s) : s_(s)
    {
        typename Tree::Config cfg;
        cfg.maxParticlesPerLeaf = c.limit;
        cfg.splitCriterion = c.nf ? ao::SplitCriterion::NFCount : ao::SplitCriterion::LeafCount;
        cfg.sfcKind = ao::SfcKind::Morton;
        cfg.box = {m.box[0], m.box[1], m.box[2], m.box[3], m.box[4], m.box[5]};
        cfg.stream = s;
        tree_ = std::make_unique<Tree>(x, y, z, n, cfg);
    }
    ~V3() { discard(); }

    void build() { tree_->build(); }
    void views()
    {
        tree_->computeViews(kAllViews);
        while (tree_->doBuffersNeedResize())
        {
            tree_->resizeBuffers();
            tree_->computeViews(kAllViews);
        }
    }
    void prepare(Events &ev)
    {
        if (!useGraph()) return;
        BENCH_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeGlobal));
        enqueue(ev, hipEventRecordExternal);
        BENCH_HIP(hipStreamEndCapture(s_, &graph_));
        BENCH_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        BENCH_HIP(hipGraphUpload(exec_, s_));
        BENCH_HIP(hipStreamSynchronize(s_));
    }
    void step(Events &ev)
    {
        if (useGraph()) BENCH_HIP(hipGraphLaunch(exec_, s_));
        else enqueue(ev, hipEventRecordDefault);
    }
    bool needsResize() { return tree_->doBuffersNeedResize(); }
    void resize(Events &ev)
    {
        discard();
        tree_->resizeBuffers();
        prepare(ev);
    }
    std::vector<Key> P() const { return tree_->particleKeys_d().to_host(); }
    std::vector<Key> K() const { return tree_->cornerstone_d().to_host(); }
    double libNodes() const { return double(tree_->numNodes()); }
    int iterations() const { return -1; }

private:
    static constexpr ao::ViewConfig kAllViews{true, true, true};
    // the same event placement on both paths (inside the graph on the graph path)
    void enqueue(Events &ev, unsigned flags)
    {
        BENCH_HIP(hipEventRecordWithFlags(ev.begin, s_, flags));
        tree_->update();
        BENCH_HIP(hipEventRecordWithFlags(ev.treeEnd, s_, flags));
        tree_->computeViews(kAllViews);
        BENCH_HIP(hipEventRecordWithFlags(ev.end, s_, flags));
    }
    void discard()
    {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
    }
    hipStream_t s_;
    std::unique_ptr<Tree> tree_;
    hipGraph_t graph_{};
    hipGraphExec_t exec_{};
};

// common.hpp:349-489
template <class Adapter>
int runDriver(int argc, char **argv)
{
    try
    {
        const Args args = parseArgs(argc, argv);
        const RunConfig config = parseConfig(args.config);
        if (!Adapter::supports(config)) throw std::invalid_argument("config not supported by this version");
        BENCH_HIP(hipFree(nullptr)); // context creation, outside every measurement

        const Manifest m = readManifest(args.manifest);
        std::size_t freeB = 0, totalB = 0;
        BENCH_HIP(hipMemGetInfo(&freeB, &totalB));
        std::cerr << "GPU memory free/total [MiB]: " << (freeB >> 20) << '/' << (totalB >> 20) << '\n';

        std::vector<double> x, y, z;
        readFrame(m.frames.front(), x, y, z);
        if (x.size() != m.slots) throw std::runtime_error("slot count differs from manifest");
        const std::size_t n = x.size();
        double *x_d = nullptr, *y_d = nullptr, *z_d = nullptr;
        BENCH_HIP(hipMalloc(&x_d, std::max<std::size_t>(1, n) * sizeof(double)));
        BENCH_HIP(hipMalloc(&y_d, std::max<std::size_t>(1, n) * sizeof(double)));
        BENCH_HIP(hipMalloc(&z_d, std::max<std::size_t>(1, n) * sizeof(double)));
        auto upload = [&] {
            BENCH_HIP(hipMemcpy(x_d, x.data(), n * sizeof(double), hipMemcpyHostToDevice));
            BENCH_HIP(hipMemcpy(y_d, y.data(), n * sizeof(double), hipMemcpyHostToDevice));
            BENCH_HIP(hipMemcpy(z_d, z.data(), n * sizeof(double), hipMemcpyHostToDevice));
        };
        upload();

        const hipStream_t s = Adapter::stream();
        Events ev;
        std::ofstream out(args.out);
        if (!out) throw std::runtime_error("cannot open " + args.out);

        // initial build: allocation + build from the root + first views
        BENCH_HIP(hipDeviceSynchronize());
        const auto t0 = Clock::now();
        BENCH_HIP(hipEventRecord(ev.begin, s));
        Adapter tree(config, m, x_d, y_d, z_d, n, s);
        tree.build();
        BENCH_HIP(hipEventRecord(ev.treeEnd, s));
        tree.views();
        BENCH_HIP(hipEventRecord(ev.end, s));
        BENCH_HIP(hipDeviceSynchronize());
        const double buildWall = msSince(t0);
        const Structure s0 = structure(tree.P(), tree.K(), config.levelBits());
        if (s0.slots != n) throw std::runtime_error("P size differs from slot count");

        if (args.mode == "cold")
        {
            out << kColdHeader << '\n'
                << args.version << ',' << args.commit << ',' << Adapter::pathName() << ',' << config.name << ','
                << m.dataset << ',' << args.run << ',' << num(buildWall) << ',' << num(elapsed(ev.begin, ev.treeEnd))
                << ',' << num(elapsed(ev.treeEnd, ev.end)) << ',' << s0.leaves << ',' << s0.nodes << ','
                << s0.maxDepth << '\n';
            return 0;
        }

        out << kStepHeader << '\n';
        StepTiming none;
        none.iterations = tree.iterations();
        writeStep(out, args, Adapter::pathName(), config, m, m.frames.front(), none, s0, tree.libNodes());
        tree.prepare(ev);

        const std::size_t last = args.maxSnapshots > 0 ? std::min<std::size_t>(args.maxSnapshots, m.frames.size())
                                                       : m.frames.size();
        for (std::size_t k = 1; k < last; ++k)
        {
            const Frame &f = m.frames[k];
            readFrame(f, x, y, z);
            if (x.size() != n) throw std::runtime_error("slot count changed in " + f.path);
            upload();

            StepTiming t{0, 0, 0, 0, 0, 0, -1};
            bool resize = false;
            do
            {
                BENCH_HIP(hipDeviceSynchronize());
                const auto w0 = Clock::now();
                BENCH_HIP(hipEventRecord(ev.outerBegin, s));
                tree.step(ev);
                BENCH_HIP(hipEventRecord(ev.outerEnd, s));
                resize = tree.needsResize();
                BENCH_HIP(hipDeviceSynchronize());
                t.wall += msSince(w0);
                t.tree += elapsed(ev.begin, ev.treeEnd);
                t.view += elapsed(ev.treeEnd, ev.end);
                t.total += elapsed(ev.outerBegin, ev.outerEnd);
                ++t.attempts;
                if (resize)
                {
                    const auto r0 = Clock::now();
                    tree.resize(ev);
                    BENCH_HIP(hipDeviceSynchronize());
                    t.maintenance += msSince(r0);
                }
            } while (resize && t.attempts < 64);
            if (resize) throw std::runtime_error("buffers still need resizing after 64 attempts");
            t.iterations = tree.iterations();
            const Structure st = structure(tree.P(), tree.K(), config.levelBits());
            writeStep(out, args, Adapter::pathName(), config, m, f, t, st, tree.libNodes());
        }
        BENCH_HIP(hipFree(x_d));
        BENCH_HIP(hipFree(y_d));
        BENCH_HIP(hipFree(z_d));
        return 0;
    }
    catch (const std::bad_alloc &e)
    {
        std::cerr << "bench_driver: out of memory: " << e.what() << '\n';
        return kOomExit;
    }
    catch (const std::exception &e)
    {
        const std::string what = e.what();
        std::cerr << "bench_driver: " << what << '\n';
        return what.find("out of memory") != std::string::npos ? kOomExit : 1;
    }
}
} // namespace

int main(int argc, char **argv)
{
    for (int i = 1; i + 1 < argc; ++i)
        if (std::string(argv[i]) == "--config" && std::string(argv[i + 1]) == "kdtree3d_leafcount")
            return runDriver<V3<ao::TreeType::Binary>>(argc, argv);
    return runDriver<V3<ao::TreeType::Octree>>(argc, argv);
}
