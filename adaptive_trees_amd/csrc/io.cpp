// io.cpp -- particle input formats: the test reader's text rows
// (tests/particle_io.hpp:35-139) and the benchmark harness's manifest and
// frames (benchmarks/drivers/common.hpp:54-175). Host code only.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/adaptive_trees.h"
#include "common.hpp"
#include "engine.hpp"

namespace
{
using at::Error;

// escapedParticleCoordinate<double> (include/adaptive_octree/config.hpp:103-113)
constexpr double kEscaped = std::numeric_limits<double>::max();

// Read-only view of a whole file (the harness maps frames; the test reader
// reads them into a buffer -- the bytes seen are the same).
class FileBytes
{
public:
    explicit FileBytes(const std::string &path)
    {
        fd_ = ::open(path.c_str(), O_RDONLY);
        if (fd_ < 0) throw Error(AT_EIO, "Could not open input file: " + path);
        struct stat st{};
        if (::fstat(fd_, &st) != 0) throw Error(AT_EIO, "Could not stat input file: " + path);
        size_ = std::size_t(st.st_size);
        if (size_)
        {
            void *p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
            if (p == MAP_FAILED) throw Error(AT_EIO, "cannot mmap " + path);
            data_ = static_cast<const char *>(p);
            ::madvise(p, size_, MADV_SEQUENTIAL);
        }
    }
    ~FileBytes()
    {
        if (data_) ::munmap(const_cast<char *>(data_), size_);
        if (fd_ >= 0) ::close(fd_);
    }
    FileBytes(const FileBytes &) = delete;
    FileBytes &operator=(const FileBytes &) = delete;
    const char *data() const { return data_; }
    std::size_t size() const { return size_; }

private:
    int fd_ = -1;
    std::size_t size_ = 0;
    const char *data_ = nullptr;
};

bool isSpace(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

// particle_io::readColumns: `numColumns` numbers per row parsed as strtod
// does, each followed by whitespace, '#' or the end of the row; blank and
// '#' rows skipped; an ESC row (nothing but whitespace or '#' after it)
// fills every column with the escaped marker. Columns past numColumns are
// not read.
std::vector<std::vector<double>> readColumns(const std::string &path, int numColumns)
{
    if (numColumns <= 0 || numColumns > 16)
        throw Error(AT_EINVAL, "particle_io::readColumns: expected 1 to 16 columns");
    FileBytes file(path);
    // strtod needs a terminated row: each row is copied into a small buffer
    const char *p = file.data(), *end = p + file.size();
    const std::size_t numLines = std::size_t(std::count(p, end, '\n')) + 1;
    std::vector<std::vector<double>> columns(static_cast<std::size_t>(numColumns));
    for (auto &c : columns) c.reserve(numLines);
    std::string row;
    std::size_t lineNumber = 0;
    while (p < end)
    {
        const char *rowEnd = static_cast<const char *>(std::memchr(p, '\n', std::size_t(end - p)));
        if (!rowEnd) rowEnd = end;
        row.assign(p, rowEnd);
        p = rowEnd < end ? rowEnd + 1 : end;
        ++lineNumber;
        auto invalidRow = [&] {
            return Error(AT_EIO, "Invalid particle row in " + path + ":" + std::to_string(lineNumber) +
                                     "; expected numeric columns or ESC");
        };
        // a NUL byte ends the row, as it ends the reference's C string
        const char *r = row.c_str();
        while (isSpace(*r)) ++r;
        if (*r == '\0' || *r == '#') continue;
        double values[16];
        if (std::strncmp(r, "ESC", 3) == 0)
        {
            r += 3;
            while (isSpace(*r)) ++r;
            if (*r != '\0' && *r != '#') throw invalidRow();
            std::fill_n(values, numColumns, kEscaped);
        }
        else
        {
            for (int col = 0; col < numColumns; ++col)
            {
                char *next = nullptr;
                values[col] = std::strtod(r, &next);
                if (next == r || (*next != '\0' && *next != '#' && !isSpace(*next))) throw invalidRow();
                r = next;
                while (isSpace(*r)) ++r;
            }
        }
        for (int col = 0; col < numColumns; ++col) columns[std::size_t(col)].push_back(values[col]);
    }
    return columns;
}

// bench::readFrame for the text formats: rows of four numbers (from_chars,
// one leading '+' allowed), or a row starting with "ESC"; blank and '#'
// rows skipped; columns 1..3 (cxyz) or 0..2 (xyzc) are the coordinates.
void readTextFrame(const std::string &path, int offset, std::vector<double> &x, std::vector<double> &y,
                   std::vector<double> &z)
{
    FileBytes file(path);
    const char *p = file.data(), *end = p + file.size();
    std::size_t line = 0;
    while (p < end)
    {
        const char *eol = static_cast<const char *>(std::memchr(p, '\n', std::size_t(end - p)));
        if (!eol) eol = end;
        ++line;
        const char *q = p;
        while (q < eol && isSpace(*q)) ++q;
        if (q < eol && *q != '#')
        {
            if (eol - q >= 3 && std::strncmp(q, "ESC", 3) == 0)
            {
                x.push_back(kEscaped);
                y.push_back(kEscaped);
                z.push_back(kEscaped);
            }
            else
            {
                double v[4];
                for (double &c : v)
                {
                    while (q < eol && isSpace(*q)) ++q;
                    if (q < eol && *q == '+') ++q;
                    const auto r = std::from_chars(q, eol, c);
                    if (r.ec != std::errc()) throw Error(AT_EIO, path + ":" + std::to_string(line) + ": bad row");
                    q = r.ptr;
                }
                x.push_back(v[offset]);
                y.push_back(v[offset + 1]);
                z.push_back(v[offset + 2]);
            }
        }
        p = eol + 1;
    }
}

// bench::readFrame for "bin": x y z doubles per slot; NaN in x = escaped.
void readBinFrame(const std::string &path, std::vector<double> &x, std::vector<double> &y, std::vector<double> &z)
{
    FileBytes file(path);
    const std::size_t n = file.size() / (3 * sizeof(double));
    if (n * 3 * sizeof(double) != file.size()) throw Error(AT_EIO, "bad binary frame " + path);
    x.resize(n);
    y.resize(n);
    z.resize(n);
    for (std::size_t i = 0; i < n; ++i)
    {
        double v[3];
        std::memcpy(v, file.data() + i * sizeof(v), sizeof(v));
        const bool esc = std::isnan(v[0]);
        x[i] = esc ? kEscaped : v[0];
        y[i] = esc ? kEscaped : v[1];
        z[i] = esc ? kEscaped : v[2];
    }
}

struct Frame
{
    int index = 0;
    long id = 0;
    std::string format, path;
};
struct Manifest
{
    std::string dataset;
    std::size_t slots = 0;
    double box[6]{};
    std::vector<Frame> frames;
};

// bench::readManifest: "dataset", "slots", "box" (six numbers) and "frame
// index id format path" lines (relative paths are taken from the manifest's
// directory); other lines are ignored. At least one frame and slots > 0.
Manifest readManifest(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw Error(AT_EIO, "cannot open manifest " + path);
    Manifest m;
    std::string line;
    while (std::getline(in, line))
    {
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "dataset") ss >> m.dataset;
        else if (tag == "slots") ss >> m.slots;
        else if (tag == "box")
            for (double &b : m.box) ss >> b;
        else if (tag == "frame")
        {
            Frame f;
            ss >> f.index >> f.id >> f.format;
            std::getline(ss >> std::ws, f.path);
            const std::filesystem::path framePath(f.path);
            if (framePath.is_relative())
                f.path = (std::filesystem::path(path).parent_path() / framePath).lexically_normal().string();
            m.frames.push_back(f);
        }
    }
    if (m.frames.empty() || m.slots == 0) throw Error(AT_EIO, "invalid manifest " + path);
    return m;
}

template <class F>
int guard(F &&f)
{
    try
    {
        f();
        return AT_OK;
    }
    catch (const Error &e)
    {
        at::lastError() = e.what();
        return e.code;
    }
    catch (const std::bad_alloc &e)
    {
        at::lastError() = e.what();
        return AT_ENOMEM;
    }
    catch (const std::exception &e)
    {
        at::lastError() = e.what();
        return AT_EIO;
    }
}

double *copyOut(const std::vector<std::vector<double>> &cols, std::size_t rows)
{
    double *out = static_cast<double *>(std::malloc(std::max<std::size_t>(1, cols.size() * rows) * sizeof(double)));
    if (!out) throw std::bad_alloc();
    for (std::size_t c = 0; c < cols.size(); ++c)
        if (rows) std::memcpy(out + c * rows, cols[c].data(), rows * sizeof(double));
    return out;
}

void copyString(const std::string &s, char *buf, std::size_t len)
{
    if (!buf || len == 0) return;
    const std::size_t k = std::min(s.size(), len - 1);
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
}
} // namespace

extern "C" {

int at_io_read_columns(const char *path, int columns, double **data, size_t *rows)
{
    return guard([&] {
        if (!path || !data || !rows) throw Error(AT_EINVAL, "null argument");
        const auto cols = readColumns(path, columns);
        const std::size_t n = cols.empty() ? 0 : cols[0].size();
        *data = copyOut(cols, n);
        *rows = n;
    });
}

int at_io_read_frame(const char *path, const char *format, double **xyz, size_t *n)
{
    return guard([&] {
        if (!path || !format || !xyz || !n) throw Error(AT_EINVAL, "null argument");
        const std::string f(format);
        std::vector<std::vector<double>> c(3);
        if (f == "bin") readBinFrame(path, c[0], c[1], c[2]);
        else if (f == "cxyz" || f == "xyzc") readTextFrame(path, f == "cxyz" ? 1 : 0, c[0], c[1], c[2]);
        else throw Error(AT_EIO, "unknown frame format " + f);
        *xyz = copyOut(c, c[0].size());
        *n = c[0].size();
    });
}

void at_io_free(void *p) { std::free(p); }

int at_io_manifest_open(const char *path, void **manifest)
{
    return guard([&] {
        if (!path || !manifest) throw Error(AT_EINVAL, "null argument");
        *manifest = new Manifest(readManifest(path));
    });
}

int at_io_manifest_info(const void *manifest, char *dataset, size_t dataset_len, size_t *slots, double *box,
                        size_t *frames)
{
    return guard([&] {
        if (!manifest) throw Error(AT_EINVAL, "null manifest");
        const auto &m = *static_cast<const Manifest *>(manifest);
        copyString(m.dataset, dataset, dataset_len);
        if (slots) *slots = m.slots;
        if (box) std::copy(m.box, m.box + 6, box);
        if (frames) *frames = m.frames.size();
    });
}

int at_io_manifest_frame(const void *manifest, size_t i, int *index, long *id, char *format, size_t format_len,
                         char *path, size_t path_len)
{
    return guard([&] {
        if (!manifest) throw Error(AT_EINVAL, "null manifest");
        const auto &m = *static_cast<const Manifest *>(manifest);
        if (i >= m.frames.size()) throw Error(AT_EINVAL, "frame index out of range");
        const Frame &f = m.frames[i];
        if (index) *index = f.index;
        if (id) *id = f.id;
        copyString(f.format, format, format_len);
        copyString(f.path, path, path_len);
    });
}

void at_io_manifest_close(void *manifest) { delete static_cast<Manifest *>(manifest); }

} // extern "C"
