// tblcache.cpp - .tbl cache files: Table::from_cache (src/build_table.cpp:147-203) maps a file
// written by DumpTable::dump (include/table.h:128-212). Layout: one 8192-byte header page with
// TableMeta {u64 num_rows, u64 num_cols, DataType types[16], u64 num_pages[16]} (208 bytes, the
// rest zero), then the pages of column 0, column 1, ... in order. The reader here also checks
// what the reference takes on trust: the column count, the types and that the file holds
// exactly the pages its header counts.
#include "tblcache.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <memory>

#include "hj_internal.hpp"

namespace hj {

hj_cache* open_cache(const std::string& path) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd == -1) throw Error(HJ_E_INVALID, "Failed to open file: " + path);
    struct stat sb;
    if (fstat(fd, &sb) == -1) {
        ::close(fd);
        throw Error(HJ_E_INVALID, "Failed to stat file: " + path);
    }
    if (sb.st_size == 0) {
        ::close(fd);
        throw Error(HJ_E_INVALID, "File is empty: " + path);
    }
    const size_t bytes = (size_t)sb.st_size;
    void* map = mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED) throw Error(HJ_E_INVALID, "Failed to mmap file: " + path);
    std::unique_ptr<hj_cache> c(new hj_cache());
    c->map = map;
    c->bytes = bytes;
    try {
        if (bytes < HJ_PAGE_SIZE)
            throw Error(HJ_E_INVALID, "cache file shorter than its header page: " + path);
        CacheMeta m;
        std::memcpy(&m, map, sizeof m);
        if (m.num_cols > kCacheMaxCols)
            throw Error(HJ_E_INVALID, "cache file has more than 16 columns: " + path);
        uint64_t total = 0;
        for (uint64_t i = 0; i < m.num_cols; ++i) {
            if (m.types[i] < HJ_INT32 || m.types[i] > HJ_VARCHAR)
                throw Error(HJ_E_INVALID, "cache file has an unknown column type: " + path);
            if (m.num_pages[i] >= (1ull << 32))
                throw Error(HJ_E_INVALID, "cache file column has too many pages: " + path);
            total += m.num_pages[i];
        }
        if ((total + 1) * HJ_PAGE_SIZE != bytes)
            throw Error(HJ_E_INVALID, "cache file size does not match its page counts: " + path);
        const uint8_t* data = (const uint8_t*)map + HJ_PAGE_SIZE;
        c->pages.resize(m.num_cols);
        c->cols.resize(m.num_cols);
        for (uint64_t i = 0; i < m.num_cols; ++i) {
            auto& pv = c->pages[i];
            pv.resize(m.num_pages[i]);
            for (uint64_t j = 0; j < m.num_pages[i]; ++j, data += HJ_PAGE_SIZE) pv[j] = data;
            c->cols[i].type = m.types[i];
            c->cols[i].num_pages = (uint32_t)m.num_pages[i];
            c->cols[i].pages = pv.empty() ? nullptr : pv.data();
        }
        c->table.num_rows = m.num_rows;
        c->table.num_columns = (uint32_t)m.num_cols;
        c->table.columns = c->cols.empty() ? nullptr : c->cols.data();
    } catch (...) {
        munmap(map, bytes);
        c->map = nullptr;
        throw;
    }
    return c.release();
}

void close_cache(hj_cache* c) {
    if (!c) return;
    if (c->map) munmap(c->map, c->bytes);
    delete c;
}

void write_cache(const hj_table& t, const std::string& path) {
    if (t.num_columns > kCacheMaxCols)
        throw Error(HJ_E_INVALID, "a cache file holds at most 16 columns");
    if (t.num_columns && !t.columns) throw Error(HJ_E_INVALID, "columns missing");
    std::vector<uint8_t> head(HJ_PAGE_SIZE, 0);
    CacheMeta m;
    std::memset(&m, 0, sizeof m);
    m.num_rows = t.num_rows;
    m.num_cols = t.num_columns;
    for (uint32_t i = 0; i < t.num_columns; ++i) {
        const hj_column& col = t.columns[i];
        if (col.num_pages && !col.pages) throw Error(HJ_E_INVALID, "column pages missing");
        m.types[i] = col.type;
        m.num_pages[i] = col.num_pages;
    }
    std::memcpy(head.data(), &m, sizeof m);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(HJ_E_INVALID, "Failed to open file: " + path);
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    for (uint32_t i = 0; ok && i < t.num_columns; ++i)
        for (uint32_t j = 0; ok && j < t.columns[i].num_pages; ++j) {
            const uint8_t* pg = t.columns[i].pages[j];
            ok = pg && std::fwrite(pg, 1, HJ_PAGE_SIZE, f) == HJ_PAGE_SIZE;
        }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw Error(HJ_E_INVALID, "Failed to write file: " + path);
}

}  // namespace hj
