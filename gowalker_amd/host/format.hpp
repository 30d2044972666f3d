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

#include <cerrno>
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
