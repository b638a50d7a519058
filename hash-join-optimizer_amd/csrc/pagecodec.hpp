// pagecodec.hpp - host page writer/reader (reference page format, include/plan.h:100-396).
#pragma once
#include <cstdint>
#include <functional>

#include "../../include/hjoin.h"

namespace hj {

using ValidFn = std::function<bool(uint64_t)>;

void encode_int32(const int32_t* values, const ValidFn& valid, uint64_t n, hj_out_column* out);
void encode_varchar(const uint64_t* offsets, const char* chars, const ValidFn& valid, uint64_t n,
                    hj_out_column* out);
uint64_t column_rows(const hj_column* col);
uint64_t column_chars(const hj_column* col);
int decode_int32(const hj_column* col, int32_t* values, uint8_t* validity);
int decode_varchar(const hj_column* col, uint64_t* offsets, char* chars, uint8_t* validity);

}  // namespace hj
