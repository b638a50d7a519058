// tblcache.hpp - the reference's .tbl table cache files, read and written on the host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/hjoin.h"

// A cache file as mapped: the table view's page pointers point into the read-only mapping.
struct hj_cache {
    void* map = nullptr;
    size_t bytes = 0;
    std::vector<std::vector<const uint8_t*>> pages;  // per column
    std::vector<hj_column> cols;
    hj_table table{};
};

namespace hj {

// Header page of a cache file (include/table.h:30-37): TableMeta, then zeros to 8192 bytes.
struct CacheMeta {
    uint64_t num_rows;
    uint64_t num_cols;
    int32_t types[16];  // DataType, an int-sized enum
    uint64_t num_pages[16];
};
static_assert(sizeof(CacheMeta) == 208, "TableMeta is 208 bytes");
constexpr int kCacheMaxCols = 16;

// maps and checks the file (throws hj::Error)
hj_cache* open_cache(const std::string& path);
void close_cache(hj_cache* c);
// writes the header page and every column's pages (throws hj::Error)
void write_cache(const hj_table& t, const std::string& path);

}  // namespace hj
