// io.cpp -- particle files on the host: the text rows of the reference's
// tests (tests/particle_io.hpp:35-139), and the benchmark harness's dataset
// manifests and frames (benchmarks/drivers/common.hpp:54-175).
//
// The accepted formats are fixed by those readers; the reading is organised
// for large files on a many-core host. A file is mapped once, cut into
// chunks that end at row boundaries, and the chunks are parsed by one thread
// each into their own row-major values; the rows are then copied straight
// into the column-major array the C ABI returns, each chunk at its row
// offset. Line numbers in messages are the global ones (each chunk counts
// its lines; a chunk's first line is the sum over the chunks before it).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <limits>
#include <new>
#include <string>
#include <string_view>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/adaptive_trees.h"
#include "common.hpp"
#include "engine.hpp"

namespace
{
using at::Error;

// The escaped-slot marker of double trees (include/adaptive_octree/config.hpp:103-113).
constexpr double kEscapedMarker = std::numeric_limits<double>::max();

// isspace of the "C" locale, without the locale lookup.
inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

// ---- the bytes of a file ------------------------------------------------------
class Mapping
{
public:
    explicit Mapping(const std::string &path, const char *what = "input file")
    {
        fd_ = ::open(path.c_str(), O_RDONLY);
        if (fd_ < 0) throw Error(AT_EIO, std::string("Could not open ") + what + ": " + path);
        struct stat st{};
        if (::fstat(fd_, &st) != 0) throw Error(AT_EIO, std::string("Could not stat ") + what + ": " + path);
        bytes_ = std::size_t(st.st_size);
        if (bytes_)
        {
            void *p = ::mmap(nullptr, bytes_, PROT_READ, MAP_PRIVATE, fd_, 0);
            if (p == MAP_FAILED) throw Error(AT_EIO, "Could not map input file: " + path);
            base_ = static_cast<const char *>(p);
            ::madvise(p, bytes_, MADV_WILLNEED);
        }
    }
    ~Mapping()
    {
        if (base_) ::munmap(const_cast<char *>(base_), bytes_);
        if (fd_ >= 0) ::close(fd_);
    }
    Mapping(const Mapping &) = delete;
    Mapping &operator=(const Mapping &) = delete;
    std::string_view text() const { return {base_, bytes_}; }

private:
    int fd_ = -1;
    std::size_t bytes_ = 0;
    const char *base_ = nullptr;
};

// ---- chunked parallel parsing ----------------------------------------------------
// Byte offsets that cut the text into at most `parts` pieces, each ending
// just after a '\n' (the last one at the end of the text).
std::vector<std::size_t> rowAlignedCuts(std::string_view t, unsigned parts)
{
    std::vector<std::size_t> cuts{0};
    for (unsigned k = 1; k < parts; ++k)
    {
        std::size_t at = t.size() / parts * k;
        if (at <= cuts.back()) continue;
        const std::size_t nl = t.find('\n', at);
        if (nl == std::string_view::npos) break;
        if (nl + 1 > cuts.back() && nl + 1 < t.size()) cuts.push_back(nl + 1);
    }
    cuts.push_back(t.size());
    return cuts;
}

unsigned parserThreads(std::size_t bytes)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t byBytes = bytes / (std::size_t(4) << 20) + 1; // ~4 MiB per thread at least
    return unsigned(std::min<std::size_t>({hw, 32, byBytes}));
}

// What one chunk produced: `width` values per accepted row, the number of
// lines it spans, and the first bad line (1-based within the chunk, 0 = none).
struct Piece
{
    std::vector<double> values;
    std::size_t lines = 0;
    std::size_t badLine = 0;
};

// Runs parseChunk(text of chunk, Piece&) on every chunk in parallel, then
// either reports the first bad row (file order) or gathers all rows into a
// column-major malloc'd array [width][rows].
template <class ParseChunk>
double *parseRows(std::string_view t, int width, const std::string &path, const char *badRowText,
                  std::size_t *rowsOut, ParseChunk parseChunk)
{
    const std::vector<std::size_t> cuts = rowAlignedCuts(t, parserThreads(t.size()));
    const std::size_t chunks = cuts.size() - 1;
    std::vector<Piece> pieces(chunks);
    {
        std::vector<std::thread> pool;
        for (std::size_t c = 1; c < chunks; ++c)
            pool.emplace_back([&, c] { parseChunk(t.substr(cuts[c], cuts[c + 1] - cuts[c]), pieces[c]); });
        parseChunk(t.substr(cuts[0], cuts[1] - cuts[0]), pieces[0]);
        for (auto &th : pool) th.join();
    }
    std::size_t lineBase = 0, rows = 0;
    std::vector<std::size_t> rowBase(chunks);
    for (std::size_t c = 0; c < chunks; ++c)
    {
        if (pieces[c].badLine) throw Error(AT_EIO, path + ":" + std::to_string(lineBase + pieces[c].badLine) + badRowText);
        lineBase += pieces[c].lines;
        rowBase[c] = rows;
        rows += pieces[c].values.size() / std::size_t(width);
    }
    auto *out = static_cast<double *>(std::malloc(std::max<std::size_t>(1, rows * std::size_t(width)) * sizeof(double)));
    if (!out) throw std::bad_alloc();
    std::vector<std::thread> pool;
    for (std::size_t c = 0; c < chunks; ++c)
        pool.emplace_back([&, c] {
            const Piece &p = pieces[c];
            const std::size_t n = p.values.size() / std::size_t(width);
            for (int col = 0; col < width; ++col)
            {
                double *dst = out + std::size_t(col) * rows + rowBase[c];
                for (std::size_t r = 0; r < n; ++r) dst[r] = p.values[r * std::size_t(width) + std::size_t(col)];
            }
        });
    for (auto &th : pool) th.join();
    *rowsOut = rows;
    return out;
}

// Calls visit(row, lineNumber) for the rows of a chunk (split at '\n'; a
// final row without '\n' counts) and returns the number of rows; visit
// returns false to stop at a bad row.
template <class Visit>
std::size_t forEachRow(std::string_view chunk, Visit visit)
{
    std::size_t line = 0, at = 0;
    while (at < chunk.size())
    {
        std::size_t nl = chunk.find('\n', at);
        if (nl == std::string_view::npos) nl = chunk.size();
        ++line;
        if (!visit(chunk.substr(at, nl - at), line)) return line;
        at = nl + 1;
    }
    return line;
}

// ---- the test reader (particle_io::readColumns) -----------------------------------
// A row is blank or a comment when only whitespace precedes its end or a
// '#'. "ESC" (then only whitespace up to the end or a '#') keeps an escaped
// slot. Otherwise the first `width` numbers are read as strtod reads them;
// each must be followed by whitespace, '#' or the end of the row; text past
// them is not looked at. A NUL byte ends a row's text, as it ends a C string.
class ColumnRow
{
public:
    explicit ColumnRow(int width) : width_(width) {}

    // Appends the row's values to out; false for a malformed row.
    bool take(std::string_view row, std::vector<double> &out)
    {
        // strtod needs a terminated string that ends with the row
        buf_.assign(row.data(), row.size());
        const char *p = buf_.c_str();
        while (blank(*p)) ++p;
        if (*p == '\0' || *p == '#') return true;
        if (p[0] == 'E' && p[1] == 'S' && p[2] == 'C')
        {
            p += 3;
            while (blank(*p)) ++p;
            if (*p != '\0' && *p != '#') return false;
            out.insert(out.end(), std::size_t(width_), kEscapedMarker);
            return true;
        }
        double v[16];
        for (int c = 0; c < width_; ++c)
        {
            char *stop = nullptr;
            v[c] = std::strtod(p, &stop);
            if (stop == p || !(*stop == '\0' || *stop == '#' || blank(*stop))) return false;
            p = stop;
            while (blank(*p)) ++p;
        }
        out.insert(out.end(), v, v + width_);
        return true;
    }

private:
    int width_;
    std::string buf_;
};

double *readColumnFile(const std::string &path, int width, std::size_t *rows)
{
    if (width <= 0 || width > 16) throw Error(AT_EINVAL, "particle_io::readColumns: expected 1 to 16 columns");
    Mapping file(path);
    return parseRows(file.text(), width, path, ": invalid particle row; expected numeric columns or ESC", rows,
                     [width](std::string_view chunk, Piece &piece) {
                         ColumnRow reader(width);
                         piece.lines = forEachRow(chunk, [&](std::string_view row, std::size_t line) {
                             if (reader.take(row, piece.values)) return true;
                             piece.badLine = line;
                             return false;
                         });
                     });
}

// ---- harness frames (bench::readFrame) -------------------------------------------
// Text frames: rows of four numbers (leading whitespace, then per number
// optional whitespace and at most one '+', parsed by from_chars inside the
// row), "ESC" at the start of a row (rest ignored) for an escaped slot,
// blank and '#' rows skipped. `first` is the column of x (1 for "cxyz",
// 0 for "xyzc").
bool takeFrameRow(std::string_view row, int first, std::vector<double> &out)
{
    const char *p = row.data(), *end = p + row.size();
    while (p < end && blank(*p)) ++p;
    if (p == end || *p == '#') return true;
    if (end - p >= 3 && p[0] == 'E' && p[1] == 'S' && p[2] == 'C')
    {
        out.insert(out.end(), 3, kEscapedMarker);
        return true;
    }
    double v[4];
    for (double &value : v)
    {
        while (p < end && blank(*p)) ++p;
        if (p < end && *p == '+') ++p;
        const auto r = std::from_chars(p, end, value);
        if (r.ec != std::errc()) return false;
        p = r.ptr;
    }
    out.insert(out.end(), v + first, v + first + 3);
    return true;
}

double *readTextFrame(const std::string &path, int first, std::size_t *n)
{
    Mapping file(path);
    return parseRows(file.text(), 3, path, ": bad row", n, [first](std::string_view chunk, Piece &piece) {
        piece.lines = forEachRow(chunk, [&](std::string_view row, std::size_t line) {
            if (takeFrameRow(row, first, piece.values)) return true;
            piece.badLine = line;
            return false;
        });
    });
}

// Binary frames: x y z doubles per slot; a NaN x marks an escaped slot.
double *readBinaryFrame(const std::string &path, std::size_t *n)
{
    Mapping file(path);
    const std::string_view t = file.text();
    constexpr std::size_t kSlot = 3 * sizeof(double);
    if (t.size() % kSlot) throw Error(AT_EIO, "bad binary frame " + path + " (size not a multiple of 24 bytes)");
    const std::size_t slots = t.size() / kSlot;
    auto *out = static_cast<double *>(std::malloc(std::max<std::size_t>(1, 3 * slots) * sizeof(double)));
    if (!out) throw std::bad_alloc();
    const unsigned parts = parserThreads(t.size());
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < parts; ++k)
        pool.emplace_back([&, k] {
            const std::size_t lo = slots * k / parts, hi = slots * (k + 1) / parts;
            for (std::size_t i = lo; i < hi; ++i)
            {
                double v[3];
                std::memcpy(v, t.data() + i * kSlot, kSlot);
                const bool esc = std::isnan(v[0]);
                for (int a = 0; a < 3; ++a) out[std::size_t(a) * slots + i] = esc ? kEscapedMarker : v[a];
            }
        });
    for (auto &th : pool) th.join();
    *n = slots;
    return out;
}

// ---- manifests (bench::readManifest) ----------------------------------------------
struct FrameEntry
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
    std::vector<FrameEntry> frames;
};

// Whitespace-separated fields of one manifest line; rest() is what follows
// the fields taken so far, after its leading whitespace.
class Fields
{
public:
    explicit Fields(std::string_view line) : s_(line) {}
    std::string_view next()
    {
        while (at_ < s_.size() && blank(s_[at_])) ++at_;
        const std::size_t b = at_;
        while (at_ < s_.size() && !blank(s_[at_])) ++at_;
        return s_.substr(b, at_ - b);
    }
    std::string_view rest()
    {
        while (at_ < s_.size() && blank(s_[at_])) ++at_;
        return s_.substr(at_);
    }

private:
    std::string_view s_;
    std::size_t at_ = 0;
};

template <class T>
T number(std::string_view f)
{
    const std::string s(f);
    char *stop = nullptr;
    if constexpr (std::is_floating_point_v<T>) return T(std::strtod(s.c_str(), &stop));
    else if constexpr (std::is_unsigned_v<T>) return s.empty() || s[0] == '-' ? T(0) : T(std::strtoull(s.c_str(), &stop, 10));
    else return T(std::strtoll(s.c_str(), &stop, 10));
}

Manifest readManifestFile(const std::string &path)
{
    Mapping file(path, "manifest");
    Manifest m;
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    forEachRow(file.text(), [&](std::string_view line, std::size_t) {
        Fields f(line);
        const std::string_view tag = f.next();
        if (tag == "dataset")
        {
            const std::string_view name = f.next();
            if (!name.empty()) m.dataset = std::string(name);
        }
        else if (tag == "slots") m.slots = number<std::size_t>(f.next());
        else if (tag == "box")
            for (double &b : m.box) b = number<double>(f.next());
        else if (tag == "frame")
        {
            FrameEntry e;
            e.index = number<int>(f.next());
            e.id = number<long>(f.next());
            e.format = std::string(f.next());
            const std::filesystem::path p{std::string(f.rest())};
            e.path = p.is_relative() ? (dir / p).lexically_normal().string() : p.string();
            m.frames.push_back(std::move(e));
        }
        return true; // other tags are ignored
    });
    if (m.frames.empty() || m.slots == 0) throw Error(AT_EIO, "invalid manifest " + path + " (no frames or no slots)");
    return m;
}

// ---- C ABI plumbing ----------------------------------------------------------------
template <class F>
int withStatus(F &&f)
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
    catch (const std::bad_alloc &)
    {
        at::lastError() = "out of host memory";
        return AT_ENOMEM;
    }
    catch (const std::exception &e)
    {
        at::lastError() = e.what();
        return AT_EIO;
    }
}

void putString(const std::string &s, char *buf, std::size_t len)
{
    if (!buf || !len) return;
    const std::size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
}
} // namespace

extern "C" {

int at_io_read_columns(const char *path, int columns, double **data, size_t *rows)
{
    return withStatus([&] {
        if (!path || !data || !rows) throw Error(AT_EINVAL, "null argument");
        *data = readColumnFile(path, columns, rows);
    });
}

int at_io_read_frame(const char *path, const char *format, double **xyz, size_t *n)
{
    return withStatus([&] {
        if (!path || !format || !xyz || !n) throw Error(AT_EINVAL, "null argument");
        const std::string f(format);
        if (f == "bin") *xyz = readBinaryFrame(path, n);
        else if (f == "cxyz" || f == "xyzc") *xyz = readTextFrame(path, f == "cxyz" ? 1 : 0, n);
        else throw Error(AT_EIO, "unknown frame format " + f);
    });
}

void at_io_free(void *p) { std::free(p); }

int at_io_manifest_open(const char *path, void **manifest)
{
    return withStatus([&] {
        if (!path || !manifest) throw Error(AT_EINVAL, "null argument");
        *manifest = new Manifest(readManifestFile(path));
    });
}

int at_io_manifest_info(const void *manifest, char *dataset, size_t dataset_len, size_t *slots, double *box,
                        size_t *frames)
{
    return withStatus([&] {
        if (!manifest) throw Error(AT_EINVAL, "null manifest");
        const auto &m = *static_cast<const Manifest *>(manifest);
        putString(m.dataset, dataset, dataset_len);
        if (slots) *slots = m.slots;
        if (box) std::copy(m.box, m.box + 6, box);
        if (frames) *frames = m.frames.size();
    });
}

int at_io_manifest_frame(const void *manifest, size_t i, int *index, long *id, char *format, size_t format_len,
                         char *path, size_t path_len)
{
    return withStatus([&] {
        if (!manifest) throw Error(AT_EINVAL, "null manifest");
        const auto &m = *static_cast<const Manifest *>(manifest);
        if (i >= m.frames.size()) throw Error(AT_EINVAL, "frame index out of range");
        const FrameEntry &f = m.frames[i];
        if (index) *index = f.index;
        if (id) *id = f.id;
        putString(f.format, format, format_len);
        putString(f.path, path, path_len);
    });
}

void at_io_manifest_close(void *manifest) { delete static_cast<Manifest *>(manifest); }

} // extern "C"
