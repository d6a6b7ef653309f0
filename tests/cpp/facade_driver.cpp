// facade_driver.cpp -- exercises the C++ facade (include/adaptive_octree/
// tree.hpp) the way the reference's tests/tree_example.cu drives
// adaptive_octree::Tree: device coordinates, build(), update() after the
// caller moves the particles in place, the resize protocol, views, and the
// exception types. Dumps every getter to <out>/<tree>_<getter>.bin for the
// parity test, which checks them against the oracle.
//
// usage: facade_driver <input.bin> <outdir>
//   input.bin: uint64 n, then double x0[n], y0[n], z0[n], x1[n], y1[n], z1[n]
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "adaptive_octree/tree.hpp"

namespace ao = adaptive_octree;

#define CHECK_HIP(x)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        if ((x) != hipSuccess)                                                                                         \
        {                                                                                                              \
            std::fprintf(stderr, "HIP error at %s:%d\n", __FILE__, __LINE__);                                          \
            return 3;                                                                                                  \
        }                                                                                                              \
    } while (0)

template <class T>
void dump(const std::string &path, const ao::DeviceVector<T> &v)
{
    const std::vector<T> h = v.to_host();
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(h.data()), std::streamsize(h.size() * sizeof(T)));
}

template <class TreeT>
void dumpAll(const std::string &dir, const std::string &name, const TreeT &t, bool nf)
{
    const std::string p = dir + "/" + name + "_";
    dump(p + "particleKeys_d.bin", t.particleKeys_d());
    dump(p + "perm_d.bin", t.perm_d());
    dump(p + "cornerstone_d.bin", t.cornerstone_d());
    dump(p + "leafCounts_d.bin", t.leafCounts_d());
    if (nf) dump(p + "nearFieldCounts_d.bin", t.nearFieldCounts_d());
    dump(p + "sfcBoxIndex_d.bin", t.sfcBoxIndex_d());
    dump(p + "hasBoxSplit_d.bin", t.hasBoxSplit_d());
    dump(p + "boxDepth_d.bin", t.boxDepth_d());
    dump(p + "parentIndex_d.bin", t.parentIndex_d());
    dump(p + "childIndex_d.bin", t.childIndex_d());
    dump(p + "levelOffset_d.bin", t.levelOffset_d());
    dump(p + "particleBeginIndex_d.bin", t.particleBeginIndex_d());
    dump(p + "particleCounts_d.bin", t.particleCounts_d());
    dump(p + "sfcParticleIndex_d.bin", t.sfcParticleIndex_d());
    std::ofstream s(p + "scalars.txt");
    s << t.numParticles() << " " << t.numActiveParticles() << " " << t.numNodes() << " " << t.maxAchievedDepth()
      << "\n";
}

template <class Real>
int runTrees(const std::vector<double> &in, std::size_t n, const std::string &out, const char *tag)
{
    std::vector<Real> h(6 * n);
    for (std::size_t i = 0; i < 6 * n; ++i) h[i] = Real(in[i]);
    Real *d = nullptr;
    CHECK_HIP(hipMalloc(&d, 3 * n * sizeof(Real) + 1));
    CHECK_HIP(hipMemcpy(d, h.data(), 3 * n * sizeof(Real), hipMemcpyHostToDevice));
    const Real *x = d, *y = d + n, *z = d + 2 * n;

    // octree: Hilbert, NFCount, all views, through the generic facade
    ao::TreeConfig<Real> oc;
    oc.maxParticlesPerLeaf = 4096;
    oc.splitCriterion = ao::SplitCriterion::NFCount;
    oc.sfcKind = ao::SfcKind::Hilbert;
    oc.views = {true, true, true};
    ao::Tree<Real, ao::TreeType::Octree> oct(x, y, z, n, oc);
    // binary: LeafCount 16 through the KDTree3D backend class
    ao::KDTree3DConfig<Real> bc;
    bc.maxParticlesPerLeaf = 16;
    bc.views = {true, true, true};
    ao::KDTree3D<Real> kd(x, y, z, n, bc);
    ao::Tree<Real, ao::TreeType::Binary> bin = std::move(kd); // movable, not copyable

    oct.build();
    bin.build();
    // the caller moves the particles in place, then update()
    CHECK_HIP(hipMemcpy(d, h.data() + 3 * n, 3 * n * sizeof(Real), hipMemcpyHostToDevice));
    oct.update();
    while (oct.doBuffersNeedResize())
    {
        oct.resizeBuffers();
        oct.update();
    }
    bin.update();
    while (bin.doBuffersNeedResize())
    {
        bin.resizeBuffers();
        bin.update();
    }
    oct.computeViews();
    bin.computeViews();
    (void)oct.doBuffersNeedResize();
    (void)bin.doBuffersNeedResize();
    dumpAll(out, std::string(tag) + "_octree", oct, true);
    dumpAll(out, std::string(tag) + "_binary", bin, false);
    oct.print(std::cout);
    bin.print(std::cout);

    // device counters are device pointers to int / unsigned
    int leaves = -1;
    CHECK_HIP(hipMemcpy(&leaves, oct.numLeaves_d(), sizeof(int), hipMemcpyDeviceToHost));
    if (std::size_t(leaves) + 1 != oct.cornerstone_d().size())
    {
        std::fprintf(stderr, "numLeaves_d %d disagrees with cornerstone size %zu\n", leaves,
                     oct.cornerstone_d().size());
        return 4;
    }
    CHECK_HIP(hipFree(d));
    return 0;
}

int exceptions()
{
    float *d = nullptr;
    if (hipMalloc(&d, 3 * 64 * sizeof(float)) != hipSuccess) return 3;
    int ok = 0;
    try
    {
        ao::TreeConfig<float> c;
        c.splitCriterion = ao::SplitCriterion::NFCount;
        ao::Tree<float, ao::TreeType::Binary> t(d, d + 64, d + 128, 64, c);
    }
    catch (const std::invalid_argument &) { ++ok; }
    try
    {
        ao::TreeConfig<float> c;
        c.maxDepth = -2;
        ao::Tree<float, ao::TreeType::Octree> t(d, d + 64, d + 128, 64, c);
    }
    catch (const std::invalid_argument &) { ++ok; }
    try
    {
        ao::Tree<float, ao::TreeType::Binary> t(d, d + 64, d + 128, 64);
        (void)t.sfcKind();
    }
    catch (const std::logic_error &) { ++ok; }
    try
    {
        ao::Tree<float, ao::TreeType::Binary> t(d, d + 64, d + 128, 64);
        (void)t.nearFieldCounts_d();
    }
    catch (const std::logic_error &) { ++ok; }
    try
    {
        ao::Tree<float, ao::TreeType::Octree> t(d, d + 64, d + 128, 64);
        t.update(); // before build()
    }
    catch (const std::logic_error &) { ++ok; }
    try
    {
        ao::KDTree3DConfig<float> c;
        c.updateIterations = 0;
        ao::KDTree3D<float> t(d, d + 64, d + 128, 64, c);
    }
    catch (const std::invalid_argument &) { ++ok; }
    (void)hipFree(d);
    std::cout << "exceptions " << ok << "/6\n";
    return ok == 6 ? 0 : 5;
}

int main(int argc, char **argv)
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s input.bin outdir\n", argv[0]);
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    std::uint64_t n = 0;
    f.read(reinterpret_cast<char *>(&n), sizeof(n));
    std::vector<double> in(6 * n);
    f.read(reinterpret_cast<char *>(in.data()), std::streamsize(in.size() * sizeof(double)));
    if (!f)
    {
        std::fprintf(stderr, "short input file\n");
        return 2;
    }
    try
    {
        if (int r = runTrees<double>(in, n, argv[2], "f64")) return r;
        if (int r = runTrees<float>(in, n, argv[2], "f32")) return r;
        return exceptions();
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 1;
    }
}
