// bench_snapshots.cpp -- the reference's snapshot benchmark
// (tests/bench_snapshots.cu:449-625 and its command line :768-812): the same
// updates through a directly enqueued tree and through a tree whose
// update() + views are replayed from a captured HIP graph, timed in one run,
// with both trees checked for equality after every comparison and one CSV
// row per snapshot (the reference's bench_snapshots.csv columns).
//
//   bench_snapshots --dataset FILE_OR_FOLDER [--series|--halo] [--last-snapshot N]
//                   [--stride N] [--repeats N] [--leaf-limit N] [--nf-limit N]
//                   [--tree all|binary|octree] [--criterion all|leafcount|nfcount]
//                   [--csv PATH]   (default bench_snapshots.csv)
//
// A file is one static snapshot ("x y z charge" rows): after the build, the
// CSV gets one row with the mean of --repeats warm comparisons. A folder is a
// series (xyz_N: "x y z charge"; snapshot_NNNNNN.dat and cluster-NNNN.dat:
// "charge x y z"): frames up to --last-snapshot whose number is a multiple of
// --stride; the first frame's row has the build and view times, every later
// frame one comparison. The box is the bounding box of the active particles
// of every selected frame, padded by 5% per axis. Runs: binary LeafCount,
// octree LeafCount, octree NFCount (as --tree / --criterion select). The
// per-leaf frame exports of the reference (tree_frames/, for its plots) are
// not written.
#include <filesystem>
#include <future>
#include <iostream>
#include <limits>
#include <regex>
#include <string>
#include <vector>

#include "harness.hpp"

namespace
{
using namespace harness;
namespace fs = std::filesystem;

// ---- inputs ------------------------------------------------------------------------------
struct Series
{
    std::string path;
    bool folder = false;
    bool chargeFirst = false; // "charge x y z" rows (else "x y z charge")
    std::vector<std::pair<int, std::string>> frames; // (snapshot number, file), ascending
    double box[6]{};
};

// The snapshot naming schemes of the reference's datasets and their row layouts.
struct Scheme
{
    std::regex re;
    const char *name;
    bool chargeFirst;
};

Series selectSeries(const std::string &path, int lastSnapshot, int stride)
{
    Series s;
    s.path = fs::absolute(path).string();
    s.folder = fs::is_directory(path);
    if (!s.folder)
    {
        if (!fs::is_regular_file(path)) throw std::invalid_argument("Dataset does not exist: " + path);
        s.frames.push_back({0, s.path});
        return s;
    }
    const Scheme schemes[] = {{std::regex(R"(xyz_([0-9]+))"), "xyz_{snapshot}", false},
                              {std::regex(R"(snapshot_([0-9]{6})\.dat)"), "snapshot_{snapshot:06d}.dat", true},
                              {std::regex(R"(cluster-([0-9]{4})\.dat)"), "cluster-{snapshot:04d}.dat", true}};
    const char *seen = nullptr;
    for (const auto &entry : fs::directory_iterator(s.path))
    {
        if (!entry.is_regular_file()) continue;
        const std::string name = entry.path().filename().string();
        for (const Scheme &sc : schemes)
        {
            std::smatch m;
            if (!std::regex_match(name, m, sc.re)) continue;
            if (seen && seen != sc.name) throw std::invalid_argument("Mixed snapshot formats in: " + path);
            seen = sc.name;
            s.chargeFirst = sc.chargeFirst;
            const int k = std::stoi(m[1].str());
            if (k <= lastSnapshot && k % stride == 0) s.frames.push_back({k, entry.path().string()});
        }
    }
    if (s.frames.empty()) throw std::invalid_argument("No selected snapshot frames in: " + path);
    std::sort(s.frames.begin(), s.frames.end());
    return s;
}

struct Particles
{
    std::vector<double> xyz[3];
};

Particles readSnapshot(const std::string &file, bool chargeFirst)
{
    double *cols = nullptr;
    std::size_t rows = 0;
    HARNESS_LIB(at_io_read_columns(file.c_str(), 4, &cols, &rows));
    std::unique_ptr<double, void (*)(void *)> hold(cols, at_io_free);
    Particles p;
    const int first = chargeFirst ? 1 : 0;
    for (int a = 0; a < 3; ++a)
        p.xyz[a].assign(cols + std::size_t(first + a) * rows, cols + std::size_t(first + a + 1) * rows);
    return p;
}

// Bounds of the active particles of every selected frame, each axis padded
// by 5% of its extent (at least a few ulps, so flat clouds keep a positive
// extent). The box is fixed for the whole run, as a captured update needs.
void boundingBox(Series &s, const Particles &first, std::size_t slots)
{
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a)
    {
        lo[a] = std::numeric_limits<double>::max();
        hi[a] = -std::numeric_limits<double>::max();
    }
    bool any = false;
    auto extend = [&](const Particles &p) {
        if (p.xyz[0].empty()) throw std::invalid_argument("Cannot compute bounds for empty or mismatched coordinates");
        for (std::size_t i = 0; i < p.xyz[0].size(); ++i)
        {
            if (p.xyz[0][i] == kEscapedMarker && p.xyz[1][i] == kEscapedMarker && p.xyz[2][i] == kEscapedMarker)
                continue;
            any = true;
            for (int a = 0; a < 3; ++a)
            {
                lo[a] = std::min(lo[a], p.xyz[a][i]);
                hi[a] = std::max(hi[a], p.xyz[a][i]);
            }
        }
    };
    extend(first);
    std::cout << "Scanning bounds for " << s.frames.size() << " snapshots (outside benchmark timings)...\n";
    for (std::size_t k = 1; k < s.frames.size(); ++k)
    {
        const Particles p = readSnapshot(s.frames[k].second, s.chargeFirst);
        if (p.xyz[0].size() != slots)
            throw std::runtime_error("Particle slot count changed in snapshot " + std::to_string(s.frames[k].first) +
                                     "; preserve escaped slots with ESC rows");
        extend(p);
    }
    if (!any) throw std::invalid_argument("Cannot infer bounds: selected snapshots contain no active particles");
    for (int a = 0; a < 3; ++a)
    {
        const double margin = std::max(0.05 * (hi[a] - lo[a]), 8 * std::numeric_limits<double>::epsilon() *
                                                                    std::max({1.0, std::abs(lo[a]), std::abs(hi[a])}));
        s.box[2 * a] = lo[a] - margin;
        s.box[2 * a + 1] = hi[a] + margin;
    }
}

// ---- the two update paths --------------------------------------------------------------
struct Phases
{
    double tree = 0, view = 0, total = 0, maintenance = 0;
    bool grow = false;
    Phases &operator+=(const Phases &o)
    {
        tree += o.tree;
        view += o.view;
        total += o.total;
        maintenance += o.maintenance;
        return *this;
    }
};

// update() + all views of one tree on the run's stream, between three
// events; the graph path records them once (event nodes) and replays.
class UpdatePath
{
public:
    UpdatePath(TreeHandle &tree, hipStream_t s, bool graph)
        : tree_(tree), s_(s), graph_(graph), ev_{"begin", "tree", "views"}
    {
    }
    ~UpdatePath() { forget(); }

    double capture()
    {
        return wallMs([&] {
            HARNESS_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeGlobal));
            work(hipEventRecordExternal);
            hipGraph_t g = nullptr;
            HARNESS_HIP(hipStreamEndCapture(s_, &g));
            const hipError_t e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HARNESS_HIP(e);
            HARNESS_HIP(hipGraphUpload(exec_, s_));
        });
    }
    // One update with the status check; wall time around launch, completion
    // and check, GPU times of the two phases.
    Phases once()
    {
        Phases p;
        p.total = wallMs([&] {
            if (graph_) HARNESS_HIP(hipGraphLaunch(exec_, s_));
            else work(hipEventRecordDefault);
            p.grow = tree_.needsResize();
        });
        p.tree = ev_.elapsed("begin", "tree");
        p.view = ev_.elapsed("tree", "views");
        return p;
    }
    // Capacity growth (and a new capture on the graph path), wall time.
    double grow()
    {
        const double ms = wallMs([&] {
            forget();
            tree_.resizeBuffers();
        });
        return ms + (graph_ ? capture() : 0.0);
    }

private:
    void work(unsigned flags)
    {
        HARNESS_HIP(hipEventRecordWithFlags(ev_["begin"], s_, flags));
        tree_.update();
        HARNESS_HIP(hipEventRecordWithFlags(ev_["tree"], s_, flags));
        tree_.allViews();
        HARNESS_HIP(hipEventRecordWithFlags(ev_["views"], s_, flags));
    }
    void forget()
    {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    TreeHandle &tree_;
    hipStream_t s_;
    bool graph_;
    EventSet ev_;
    hipGraphExec_t exec_ = nullptr;
};

struct Pair
{
    Phases direct, graph;
};

// Both paths until neither needs more capacity; the order alternates between
// calls so that neither path always runs on warm caches.
Pair comparePaths(UpdatePath &direct, UpdatePath &graph, bool graphFirst)
{
    Pair sum;
    for (;;)
    {
        Phases a, b;
        if (graphFirst)
        {
            b = graph.once();
            a = direct.once();
        }
        else
        {
            a = direct.once();
            b = graph.once();
        }
        sum.direct += a;
        sum.graph += b;
        if (a.grow != b.grow) throw std::runtime_error("Graph and ordinary update disagree on capacity");
        if (!a.grow) return sum;
        sum.direct.maintenance += direct.grow();
        sum.graph.maintenance += graph.grow();
    }
}

// Exact equality of the two trees: scalars, P, Perm, K, N (NF for NFCount
// octrees) and every view array.
void requireSameTrees(const TreeHandle &a, const TreeHandle &b)
{
    const at_scalars sa = a.scalars(), sb = b.scalars();
    auto same = [&](int id, auto tag) {
        using T = decltype(tag);
        return a.fetch<T>(id) == b.fetch<T>(id);
    };
    if (sa.num_particles != sb.num_particles || sa.num_active_particles != sb.num_active_particles ||
        !same(AT_BUF_PARTICLE_KEYS, Key{}) || !same(AT_BUF_PERM, std::uint32_t{}) ||
        !same(AT_BUF_CORNERSTONE, Key{}) || !same(AT_BUF_LEAF_COUNTS, std::uint32_t{}))
        throw std::runtime_error("Graph and ordinary update produced different trees");
    if (a.type() == AT_TREE_OCTREE && sa.split_criterion == AT_NF_COUNT && !same(AT_BUF_NEAR_FIELD_COUNTS, std::uint32_t{}))
        throw std::runtime_error("Graph and ordinary update produced different near-field counts");
    if (sa.num_nodes != sb.num_nodes || !same(AT_BUF_SFC_BOX_INDEX, Key{}) ||
        !same(AT_BUF_HAS_BOX_SPLIT, std::uint8_t{}) || !same(AT_BUF_BOX_DEPTH, std::uint8_t{}) ||
        !same(AT_BUF_PARENT_INDEX, std::int32_t{}) || !same(AT_BUF_CHILD_INDEX, std::int32_t{}) ||
        !same(AT_BUF_LEVEL_OFFSET, std::int32_t{}) || !same(AT_BUF_PARTICLE_BEGIN_INDEX, std::uint32_t{}) ||
        !same(AT_BUF_PARTICLE_COUNTS, std::uint32_t{}) || !same(AT_BUF_SFC_PARTICLE_INDEX, Key{}))
        throw std::runtime_error("Graph and ordinary update produced different views");
}

// Snapshot statistics from the tree's own getters: nodes of the full tree,
// empty leaves and mean population from N (whose sum must be the active
// count), deepest leaf from K.
struct Stats
{
    std::size_t slots = 0, active = 0, nodes = 0, leaves = 0, empty = 0;
    double avg = 0;
    unsigned maxDepth = 0;
};
Stats statsOf(const TreeHandle &t)
{
    Stats st;
    const at_scalars sc = t.scalars();
    const bool binary = t.type() == AT_TREE_BINARY;
    st.slots = sc.num_particles;
    st.active = sc.num_active_particles;
    const std::vector<Key> K = t.fetch<Key>(AT_BUF_CORNERSTONE);
    const std::vector<std::uint32_t> N = t.fetch<std::uint32_t>(AT_BUF_LEAF_COUNTS);
    st.leaves = N.size();
    st.nodes = st.leaves + (st.leaves - 1) / (binary ? 1 : 7);
    std::uint64_t sum = 0;
    for (std::size_t i = 0; i < N.size(); ++i)
    {
        sum += N[i];
        st.empty += N[i] == 0;
        st.maxDepth = std::max(st.maxDepth, leafDepth(K[i + 1] - K[i], binary));
    }
    if (sum != st.active) throw std::runtime_error("Leaf counts do not match the active particle count");
    if (st.leaves) st.avg = double(sum) / double(st.leaves);
    return st;
}

const std::vector<std::string> kColumns = {
    "snapshot", "h2d_ms", "init_tree_ms", "tree_ms", "view_ms", "compute_total_ms", "total_with_h2d_ms",
    "num_nodes", "num_leaves", "num_empty_leaves", "empty_leaf_percent", "avg_particles_per_leaf", "max_depth",
    "graph_tree_ms", "graph_view_ms", "graph_compute_total_ms", "maintenance_ms", "graph_maintenance_ms",
    "graph_initial_capture_ms", "num_particle_slots", "num_active_particles", "num_escaped_particles", "tree_type",
    "split_criterion"};

struct RowTimes
{
    double h2d = 0, init = 0, tree = 0, view = 0, total = 0, withH2d = 0, capture = 0;
};

void putRow(CsvTable &csv, int snapshot, const RowTimes &t, const Stats &st, const Pair *paths, bool binary,
            bool nf)
{
    const double none = std::numeric_limits<double>::quiet_NaN();
    csv.row()
        .put("snapshot", snapshot)
        .put("h2d_ms", gfmt(t.h2d, 6))
        .put("init_tree_ms", gfmt(t.init, 6))
        .put("tree_ms", gfmt(t.tree, 6))
        .put("view_ms", gfmt(t.view, 6))
        .put("compute_total_ms", gfmt(t.total, 6))
        .put("total_with_h2d_ms", gfmt(t.withH2d, 6))
        .put("num_nodes", (long long)st.nodes)
        .put("num_leaves", (long long)st.leaves)
        .put("num_empty_leaves", (long long)st.empty)
        .put("empty_leaf_percent", gfmt(st.leaves ? 100.0 * double(st.empty) / double(st.leaves) : 0.0, 6))
        .put("avg_particles_per_leaf", gfmt(st.avg, 6))
        .put("max_depth", (long long)st.maxDepth)
        .put("graph_tree_ms", gfmt(paths ? paths->graph.tree : none, 6))
        .put("graph_view_ms", gfmt(paths ? paths->graph.view : none, 6))
        .put("graph_compute_total_ms", gfmt(paths ? paths->graph.total : none, 6))
        .put("maintenance_ms", gfmt(paths ? paths->direct.maintenance : 0.0, 6))
        .put("graph_maintenance_ms", gfmt(paths ? paths->graph.maintenance : 0.0, 6))
        .put("graph_initial_capture_ms", gfmt(t.capture, 6))
        .put("num_particle_slots", (long long)st.slots)
        .put("num_active_particles", (long long)st.active)
        .put("num_escaped_particles", (long long)(st.slots - st.active))
        .put("tree_type", binary ? "binary" : "octree")
        .put("split_criterion", nf ? "nfcount" : "leafcount")
        .write();
}

void printPaths(const Pair &p)
{
    std::cout << "  Ordinary: tree " << p.direct.tree << " ms, view " << p.direct.view << " ms, total "
              << p.direct.total << " ms\n  Graph:    tree " << p.graph.tree << " ms, view " << p.graph.view
              << " ms, total " << p.graph.total << " ms\n  Maintenance (ordinary / graph): " << p.direct.maintenance
              << " / " << p.graph.maintenance << " ms\n";
}

void benchmark(const Series &series, int treeType, bool nf, unsigned limit, int repeats, CsvTable &csv)
{
    const bool binary = treeType == AT_TREE_BINARY;
    const std::string run = std::string(binary ? "binary" : "octree") + "_" + (nf ? "nfcount" : "leafcount");
    Particles p = readSnapshot(series.frames.front().second, series.chargeFirst);
    const std::size_t n = p.xyz[0].size();
    std::cout << "Benchmark: " << run << " (limit " << limit << ")\nDataset: " << series.path
              << "\nParticle slots: " << n << '\n';
    DeviceXYZ coords(n);
    RowTimes first;
    first.h2d = wallMs([&] { coords.upload(p.xyz); });
    hipStream_t s = nullptr;
    HARNESS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    at_config cfg{};
    cfg.max_depth = -1;
    cfg.update_iterations = -1;
    cfg.max_particles_per_leaf = limit;
    cfg.split_criterion = nf ? AT_NF_COUNT : AT_LEAF_COUNT;
    cfg.sfc_kind = AT_MORTON;
    cfg.box = at_box{series.box[0], series.box[1], series.box[2], series.box[3], series.box[4], series.box[5]};
    cfg.stream = s;
    {
        std::unique_ptr<TreeHandle> direct, graphed;
        first.init = wallMs([&] { direct = std::make_unique<TreeHandle>(treeType, coords, cfg); });
        graphed = std::make_unique<TreeHandle>(treeType, coords, cfg);
        first.tree = wallMs([&] { direct->build(); });
        graphed->build();
        // every view buffer is prepared before the capture; the first views'
        // time includes that preparation
        first.view = wallMs([&] { direct->viewsUntilSized(); });
        graphed->viewsUntilSized();
        requireSameTrees(*direct, *graphed);

        UpdatePath pd(*direct, s, false), pg(*graphed, s, true);
        first.capture = pg.capture();
        comparePaths(pd, pg, false); // warm-up of both paths, the first graph launch included
        requireSameTrees(*direct, *graphed);
        std::cout << "Initial " << (binary ? "binary" : "octree") << " build: " << first.tree
                  << " ms; views: " << first.view << " ms (wall, including preparation)\nGraph capture/instantiate: "
                  << first.capture << " ms\n";
        const int firstSnapshot = series.frames.front().first;
        if (!series.folder)
        {
            Pair mean;
            for (int r = 0; r < repeats; ++r)
            {
                const Pair one = comparePaths(pd, pg, r % 2 != 0);
                mean.direct += one.direct;
                mean.graph += one.graph;
                requireSameTrees(*direct, *graphed);
            }
            for (Phases *ph : {&mean.direct, &mean.graph})
            {
                ph->tree /= repeats;
                ph->view /= repeats;
                ph->total /= repeats;
                ph->maintenance /= repeats;
            }
            std::cout << "Static dataset, mean of " << repeats << " warm updates:\n";
            printPaths(mean);
            RowTimes t = first;
            t.tree = mean.direct.tree;
            t.view = mean.direct.view;
            t.total = mean.direct.total;
            t.withH2d = first.h2d + mean.direct.total;
            putRow(csv, firstSnapshot, t, statsOf(*direct), &mean, binary, nf);
        }
        else
        {
            RowTimes t = first;
            t.total = first.tree + first.view;
            t.withH2d = first.h2d + first.init + first.tree + first.view;
            putRow(csv, firstSnapshot, t, statsOf(*direct), nullptr, binary, nf);
            // the next frame is read while the current one is compared
            std::future<Particles> next;
            auto fetch = [&](std::size_t k) {
                return std::async(std::launch::async, readSnapshot, series.frames[k].second, series.chargeFirst);
            };
            if (series.frames.size() > 1) next = fetch(1);
            for (std::size_t k = 1; k < series.frames.size(); ++k)
            {
                p = next.get();
                if (k + 1 < series.frames.size()) next = fetch(k + 1);
                const int snapshot = series.frames[k].first;
                if (p.xyz[0].size() != n)
                    throw std::runtime_error("Particle slot count changed in snapshot " + std::to_string(snapshot) +
                                             "; preserve escaped slots with ESC rows");
                RowTimes rt;
                rt.h2d = wallMs([&] { coords.upload(p.xyz); });
                const Pair r = comparePaths(pd, pg, (k - 1) % 2 != 0);
                requireSameTrees(*direct, *graphed);
                rt.tree = r.direct.tree;
                rt.view = r.direct.view;
                rt.total = r.direct.total;
                rt.withH2d = rt.h2d + r.direct.total;
                std::cout << "Snapshot " << snapshot << ":\n";
                printPaths(r);
                putRow(csv, snapshot, rt, statsOf(*direct), &r, binary, nf);
            }
        }
    }
    HARNESS_HIP(hipStreamDestroy(s));
}

int mainBody(int argc, char **argv)
{
    std::string datasetPath, csvPath = "bench_snapshots.csv", treeChoice = "all", criterionChoice = "all";
    bool series = false;
    int lastSnapshot = std::numeric_limits<int>::max(), stride = 1, repeats = 20, leafLimit = 64,
        nfLimit = 64 * 64 * 27;
    const std::string usage = "Usage: bench_snapshots --dataset FILE_OR_FOLDER [--series|--halo] [--last-snapshot N] "
                              "[--stride N] [--repeats N] [--leaf-limit N] [--nf-limit N] [--tree all|binary|octree] "
                              "[--criterion all|leafcount|nfcount] [--csv PATH]";
    for (int i = 1; i < argc; ++i)
        if (std::string(argv[i]) == "--help")
        {
            std::cout << usage << '\n';
            return 0;
        }
    Flags f;
    f.text("--dataset", &datasetPath);
    f.text("--csv", &csvPath);
    f.text("--tree", &treeChoice);
    f.text("--criterion", &criterionChoice);
    f.toggle("--series", &series, true);
    f.toggle("--halo", &series, false);
    f.integer("--last-snapshot", &lastSnapshot);
    f.integer("--stride", &stride);
    f.integer("--repeats", &repeats);
    f.integer("--leaf-limit", &leafLimit);
    f.integer("--nf-limit", &nfLimit);
    f.parse(argc, argv, usage);
    if (treeChoice != "all" && treeChoice != "binary" && treeChoice != "octree")
        throw std::invalid_argument("--tree must be all, binary, or octree");
    if (criterionChoice != "all" && criterionChoice != "leafcount" && criterionChoice != "nfcount")
        throw std::invalid_argument("--criterion must be all, leafcount, or nfcount");
    if (stride <= 0 || lastSnapshot < 0 || repeats <= 0 || leafLimit <= 0 || nfLimit <= 0)
        throw std::invalid_argument("Require positive stride/repeats/leaf-limit/nf-limit and nonnegative last-snapshot");
    if (treeChoice == "binary" && criterionChoice == "nfcount")
        throw std::invalid_argument("Binary NFCount is not implemented; use octree or LeafCount");
    if (datasetPath.empty()) throw std::invalid_argument("--dataset is required (this build ships no test data)");
    (void)series; // a folder is always read as a series, a file as one snapshot

    Series s = selectSeries(datasetPath, lastSnapshot, stride);
    {
        const Particles first = readSnapshot(s.frames.front().second, s.chargeFirst);
        boundingBox(s, first, first.xyz[0].size());
    }
    std::unique_ptr<std::FILE, int (*)(std::FILE *)> out(std::fopen(csvPath.c_str(), "w"), std::fclose);
    if (!out) throw std::runtime_error("Could not open " + csvPath);
    CsvTable csv(out.get(), kColumns);
    if (treeChoice != "octree" && criterionChoice != "nfcount")
        benchmark(s, AT_TREE_BINARY, false, unsigned(leafLimit), repeats, csv);
    if (treeChoice != "binary")
        for (const bool nf : {false, true})
        {
            if (criterionChoice != "all" && criterionChoice != (nf ? "nfcount" : "leafcount")) continue;
            benchmark(s, AT_TREE_OCTREE, nf, unsigned(nf ? nfLimit : leafLimit), repeats, csv);
        }
    return 0;
}
} // namespace

int main(int argc, char **argv)
{
    try
    {
        return mainBody(argc, argv);
    }
    catch (const std::exception &e)
    {
        std::cerr << "bench_snapshots: " << e.what() << '\n';
        return 1;
    }
}
