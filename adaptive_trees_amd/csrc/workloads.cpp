// workloads.cpp -- seeded synthetic particle sets (host code, C ABI at_workload_*).
//
// Inputs of the benchmark and test workloads are drawn from one
// xoshiro256** stream per call, its state seeded by four splitmix64 outputs
// of the seed. Uniform doubles are (u >> 11) * 2^-53; Gaussians come from
// Box-Muller in double, two per pair of uniforms, used in order; FP32 inputs
// are cast by the caller after generation. Coordinates are written
// column-major (x[0..n), y[0..n), z[0..n)). Draw order, per particle:
//   uniform box      x, y, z
//   Plummer          u for the radius (redrawn while r > 20 a or u = 0),
//                    then cos(theta), then phi; centred at (0.5, 0.5, 0.5),
//                    clipped to the unit box
//   two clusters     the first floor(0.4 n) particles ~ N(c1, sigma^2 I),
//                    the next floor(0.4 n) ~ N(c2, sigma^2 I) (three
//                    Gaussians x, y, z each), the rest uniform (x, y, z);
//                    c1, c2 at 1/4 and 3/4 of x and the middle of y and z;
//                    clipped to the box
//   reflected step   x, y, z each += N(0, (sigma * side)^2), reflected at
//                    the walls (lo - d -> lo + d, hi + d -> hi - d), clamped
//   sizes            round(exp(U(ln lo, ln hi))) per entry
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <new>

#include "../../include/adaptive_trees.h"

namespace
{
std::uint64_t splitmix64(std::uint64_t &s)
{
    std::uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256** (Blackman and Vigna)
class Stream
{
public:
    explicit Stream(std::uint64_t seed)
    {
        for (auto &w : s_) w = splitmix64(seed);
    }
    std::uint64_t next()
    {
        const std::uint64_t r = rotl(s_[1] * 5, 7) * 9;
        const std::uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    double gaussian()
    {
        if (hasSpare_)
        {
            hasSpare_ = false;
            return spare_;
        }
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double phi = 2.0 * M_PI * u2;
        spare_ = r * std::sin(phi);
        hasSpare_ = true;
        return r * std::cos(phi);
    }

private:
    std::uint64_t s_[4];
    double spare_ = 0.0;
    bool hasSpare_ = false;
};

inline double clampTo(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class F>
int guard(F &&f)
{
    try
    {
        f();
        return AT_OK;
    }
    catch (const std::bad_alloc &)
    {
        return AT_ENOMEM;
    }
    catch (...)
    {
        return AT_EINVAL;
    }
}
} // namespace

extern "C" {
int at_workload_uniform(uint64_t seed, size_t n, const double *box, double *xyz)
{
    if ((n && !xyz) || !box) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        for (std::size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) xyz[a * n + i] = box[2 * a] + (box[2 * a + 1] - box[2 * a]) * g.uniform();
    });
}

int at_workload_plummer(uint64_t seed, size_t n, double a, double *xyz)
{
    if ((n && !xyz) || !(a > 0)) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        for (std::size_t i = 0; i < n; ++i)
        {
            double r;
            for (;;)
            {
                const double u = g.uniform();
                if (u <= 0.0) continue;
                r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
                if (r <= 20.0 * a) break;
            }
            const double cz = 2.0 * g.uniform() - 1.0;
            const double phi = 2.0 * M_PI * g.uniform();
            const double s = std::sqrt(1.0 - cz * cz);
            xyz[i] = clampTo(0.5 + r * s * std::cos(phi), 0.0, 1.0);
            xyz[n + i] = clampTo(0.5 + r * s * std::sin(phi), 0.0, 1.0);
            xyz[2 * n + i] = clampTo(0.5 + r * cz, 0.0, 1.0);
        }
    });
}

int at_workload_two_clusters(uint64_t seed, size_t n, const double *box, double sigma_of_long_side, double *xyz)
{
    if ((n && !xyz) || !box) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        const double lx = box[1] - box[0];
        const double sigma = sigma_of_long_side * lx;
        const std::size_t n1 = std::size_t(0.4 * double(n)), n2 = std::size_t(0.4 * double(n));
        const double cy = 0.5 * (box[2] + box[3]), cz = 0.5 * (box[4] + box[5]);
        const double cx[2] = {box[0] + 0.25 * lx, box[0] + 0.75 * lx};
        for (std::size_t i = 0; i < n; ++i)
        {
            double p[3];
            if (i < n1 + n2)
            {
                const double c[3] = {cx[i < n1 ? 0 : 1], cy, cz};
                for (int a = 0; a < 3; ++a) p[a] = c[a] + sigma * g.gaussian();
            }
            else
                for (int a = 0; a < 3; ++a) p[a] = box[2 * a] + (box[2 * a + 1] - box[2 * a]) * g.uniform();
            for (int a = 0; a < 3; ++a) xyz[a * n + i] = clampTo(p[a], box[2 * a], box[2 * a + 1]);
        }
    });
}

int at_workload_reflect_step(uint64_t seed, size_t n, const double *box, double sigma_of_side, const double *in,
                             double *out)
{
    if ((n && (!in || !out)) || !box) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        for (std::size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a)
            {
                const double lo = box[2 * a], hi = box[2 * a + 1];
                double v = in[a * n + i] + g.gaussian() * sigma_of_side * (hi - lo);
                if (v < lo) v = 2.0 * lo - v;
                if (v > hi) v = 2.0 * hi - v;
                out[a * n + i] = clampTo(v, lo, hi);
            }
    });
}

int at_workload_log_uniform_sizes(uint64_t seed, size_t count, double lo, double hi, int64_t *out)
{
    if ((count && !out) || !(lo > 0) || !(hi >= lo)) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        const double l0 = std::log(lo), l1 = std::log(hi);
        for (std::size_t i = 0; i < count; ++i) out[i] = std::int64_t(std::nearbyint(std::exp(l0 + (l1 - l0) * g.uniform())));
    });
}

int at_workload_bits(uint64_t seed, size_t count, uint64_t *out)
{
    if (count && !out) return AT_EINVAL;
    return guard([&] {
        Stream g(seed);
        for (std::size_t i = 0; i < count; ++i) out[i] = g.next();
    });
}
}
