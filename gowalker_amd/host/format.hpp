// format.hpp -- GOWalker's on-disk graph format and block table.
//
//   <base>.beg   u64 edge offsets indexed by vertex id (ids 0..max_vert, plus the end)
//   <base>.csr   u32 targets, each adjacency list sorted
//   <base>.meta  {u32 nvertices, u64 nedges, u32 min_vert}, written field by field
//                (preprocess/graph_converter.hpp:258-275, read by util/io.hpp:29-42)
//   <dir>/sowalker_<MB>/<name>_<MB>MB.vert.blocks   u32 vertex cut points
//   <dir>/sowalker_<MB>/<name>_<MB>MB.edge.blocks   u64 edge cut points
//                (names util/util.hpp:128-165, cut rule preprocess/split.hpp:15-87)
// The block table applies the giant-vertex truncation of engine/cache.hpp:170-172.
// All IO is checked: a short read is an error (the reference's assert(ret > 0) on a
// size_t lets -1 through, util/io.hpp:51).
#pragma once
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gw.h"

namespace gwh {

struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline std::string remove_extension(const std::string& f)
{
    size_t p = f.find_last_of('.');
    return (p > 0 && p != std::string::npos) ? f.substr(0, p) : f;
}
inline std::string path_dir(const std::string& s)
{
    size_t p = s.rfind('/');
    return p == std::string::npos ? "./" : s.substr(0, p + 1);
}
inline std::string path_file(const std::string& s)
{
    size_t p = s.rfind('/');
    return p == std::string::npos ? s : s.substr(p + 1);
}
inline std::string beg_name(const std::string& base) { return base + ".beg"; }
inline std::string csr_name(const std::string& base) { return base + ".csr"; }
inline std::string meta_name(const std::string& base) { return base + ".meta"; }
inline std::string block_folder(const std::string& base, uint64_t block_bytes)
{
    return path_dir(base) + "sowalker_" + std::to_string(block_bytes / (1024 * 1024));
}
inline std::string vert_blocks_name(const std::string& base, uint64_t block_bytes)
{
    return block_folder(base, block_bytes) + "/" + path_file(base) + "_" +
           std::to_string(block_bytes / (1024 * 1024)) + "MB.vert.blocks";
}
inline std::string edge_blocks_name(const std::string& base, uint64_t block_bytes)
{
    return block_folder(base, block_bytes) + "/" + path_file(base) + "_" +
           std::to_string(block_bytes / (1024 * 1024)) + "MB.edge.blocks";
}
inline bool file_exists(const std::string& f)
{
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

inline uint64_t file_size(const std::string& f)
{
    struct stat st;
    if (stat(f.c_str(), &st) != 0) throw IoError("cannot stat " + f + ": " + std::strerror(errno));
    return (uint64_t)st.st_size;
}

inline void pread_all(int fd, void* buf, uint64_t bytes, uint64_t off, const std::string& what)
{
    char* p = (char*)buf;
    while (bytes) {
        ssize_t r = ::pread(fd, p, bytes > (1u << 30) ? (1u << 30) : bytes, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) throw IoError("short read of " + what + (r < 0 ? std::string(": ") + std::strerror(errno) : ""));
        p += r;
        bytes -= (uint64_t)r;
        off += (uint64_t)r;
    }
}

inline void write_all(const std::string& name, const void* buf, uint64_t bytes)
{
    int fd = ::open(name.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) throw IoError("cannot create " + name + ": " + std::strerror(errno));
    const char* p = (const char*)buf;
    while (bytes) {
        ssize_t r = ::write(fd, p, bytes > (1u << 30) ? (1u << 30) : bytes);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) {
            ::close(fd);
            throw IoError("short write of " + name);
        }
        p += r;
        bytes -= (uint64_t)r;
    }
    if (::close(fd) != 0) throw IoError("close failed for " + name);
}

template <typename T>
std::vector<T> read_vector(const std::string& name)
{
    uint64_t sz = file_size(name);
    if (sz % sizeof(T)) throw IoError(name + " has a size that is not a multiple of the element size");
    std::vector<T> v(sz / sizeof(T));
    int fd = ::open(name.c_str(), O_RDONLY);
    if (fd < 0) throw IoError("cannot open " + name);
    try {
        pread_all(fd, v.data(), sz, 0, name);
    } catch (...) {
        ::close(fd);
        throw;
    }
    ::close(fd);
    return v;
}

struct Meta {
    uint32_t nvertices = 0;
    uint64_t nedges = 0;
    uint32_t min_vert = 0;
};

inline void write_meta(const std::string& base, const Meta& m)
{
    char buf[16];
    std::memcpy(buf, &m.nvertices, 4);
    std::memcpy(buf + 4, &m.nedges, 8);
    std::memcpy(buf + 12, &m.min_vert, 4);
    write_all(meta_name(base), buf, 16);
}

inline Meta read_meta(const std::string& base)
{
    std::string f = meta_name(base);
    if (file_size(f) != 16) throw IoError(f + " is not a 16-byte meta file");
    auto v = read_vector<char>(f);
    Meta m;
    std::memcpy(&m.nvertices, v.data(), 4);
    std::memcpy(&m.nedges, v.data() + 4, 8);
    std::memcpy(&m.min_vert, v.data() + 12, 4);
    return m;
}

// Block table from cut points, with engine/cache.hpp:170-172's truncation of a giant
// single-vertex block to blocksize/4 edges.
inline std::vector<gw_block_desc> make_block_table(const std::vector<uint32_t>& vb, const std::vector<uint64_t>& eb,
                                                   uint64_t block_bytes)
{
    if (vb.size() < 2 || vb.size() != eb.size()) throw IoError("malformed block files");
    const uint64_t maxedges = block_bytes / 4;
    std::vector<gw_block_desc> t(vb.size() - 1);
    for (size_t b = 0; b + 1 < vb.size(); ++b) {
        if (vb[b + 1] < vb[b] || eb[b + 1] < eb[b]) throw IoError("block cut points are not monotone");
        gw_block_desc& d = t[b];
        d.blk = (uint32_t)b;
        d.start_vert = vb[b];
        d.nverts = vb[b + 1] - vb[b];
        d.pad = 0;
        d.start_edge = eb[b];
        d.nedges = eb[b + 1] - eb[b];
        if (d.nverts == 1 && d.nedges > maxedges) d.nedges = maxedges;
    }
    return t;
}

// Text edge list -> .beg/.csr/.meta with the reference converter's rules
// (preprocess/graph_converter.hpp:286-384, :193-262): lines starting with '#' or '%' are
// comments; fields are split on tabs, commas and spaces; ids below the first line's
// source get empty lists (dealfirstline); self-loops are dropped; sources must not
// decrease (the reference asserts; here an IoError names the line); duplicates are kept;
// every id up to the largest one seen gets a (possibly empty) list, so .beg has
// max_id + 2 entries; .meta holds max - min + 1, the edge count and the smallest id.
// Lists are sorted, because the membership search needs sorted lists (the reference
// sorts only with its `sorted` option, :444-447; its EDGE_SIZE of 0, api/constants.hpp:5,
// makes the original stop before writing anything).
inline Meta convert_text(const std::string& txt)
{
    FILE* f = std::fopen(txt.c_str(), "r");
    if (!f) throw IoError("cannot open " + txt);
    std::vector<uint64_t> beg{0};
    std::vector<uint32_t> csr;
    uint64_t cur = 0, lineno = 0, start = 0;
    uint32_t maxv = 0, minv = 0xFFFFFFFFu;
    bool first = true;
    char line[1024];
    auto close_list = [&]() {   // sort the current source's list and record its end
        std::sort(csr.begin() + (ptrdiff_t)start, csr.end());
        beg.push_back(csr.size());
        start = csr.size();
    };
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        if (line[0] == '#' || line[0] == '%') continue;
        char* save = nullptr;
        char* t1 = strtok_r(line, "\t, \r\n", &save);
        char* t2 = t1 ? strtok_r(nullptr, "\t, \r\n", &save) : nullptr;
        if (!t1 || !t2) {
            std::fclose(f);
            throw IoError(txt + ":" + std::to_string(lineno) + ": expected <from> <to>");
        }
        const uint32_t from = (uint32_t)std::strtoul(t1, nullptr, 10), to = (uint32_t)std::strtoul(t2, nullptr, 10);
        if (first) {   // ids below the first source have empty lists
            for (uint32_t v = 0; v < from; ++v) beg.push_back(0);
            cur = from;
            first = false;
        }
        if (from == to) continue;
        if (from < cur) {
            std::fclose(f);
            throw IoError(txt + ":" + std::to_string(lineno) + ": sources must not decrease (sort the edge list)");
        }
        while (cur < from) {   // finish the current list and any empty ones in between
            close_list();
            ++cur;
        }
        csr.push_back(to);
        maxv = std::max(maxv, std::max(from, to));
        minv = std::min(minv, std::min(from, to));
    }
    std::fclose(f);
    if (csr.empty()) throw IoError(txt + ": no edges");
    close_list();
    while (cur < maxv) {
        ++cur;
        beg.push_back(csr.size());
    }
    const std::string base = remove_extension(txt);
    write_all(beg_name(base), beg.data(), beg.size() * 8);
    write_all(csr_name(base), csr.data(), csr.size() * 4);
    Meta m;
    m.nvertices = maxv - minv + 1;
    m.nedges = csr.size();
    m.min_vert = minv;
    write_meta(base, m);
    return m;
}

// split_blocks over an in-memory .beg and write both block files.
inline void write_block_files(const std::string& base, const uint64_t* beg, uint64_t entries, uint64_t block_bytes)
{
    std::vector<uint32_t> vb(entries + 2);
    std::vector<uint64_t> eb(entries + 2);
    uint64_t n = 0;
    if (gw_split_blocks(beg, entries, block_bytes, vb.data(), eb.data(), vb.size(), &n) != GW_OK)
        throw IoError(std::string("split_blocks failed: ") + gw_last_error(nullptr));
    ::mkdir(block_folder(base, block_bytes).c_str(), 0755);
    write_all(vert_blocks_name(base, block_bytes), vb.data(), n * 4);
    write_all(edge_blocks_name(base, block_bytes), eb.data(), n * 8);
}

}  // namespace gwh
