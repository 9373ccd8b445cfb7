// sell.hpp -- SELL-64 gather tables for one sparse operator (A or A^T).
//
// Rows are cut into slices of 64 (one wavefront, one lane per row).  A
// slice stores width(slice) x 64 packed entries, k-major, so that step k of
// all 64 lanes reads 256 contiguous bytes.  An entry is (code << 24) | col:
// col < 2^24 indexes the input vector, code indexes a per-operator table of
// at most 254 distinct coefficients (code 0 = 0.0 is padding).  Rows longer
// than a threshold are stored separately and reduced by a whole wavefront;
// their slice lane carries the marker code 255 at k = 0 and writes nothing.
#pragma once

#include <cstdint>
#include <vector>

#include "constraints.hpp"

namespace v2h {

constexpr uint32_t kCodeLongRow = 255;

struct SellHost {
    int64_t nrows = 0;
    int nslices = 0;
    std::vector<uint32_t> slice_start;  // [nslices + 1], in units of 64 entries
    std::vector<uint32_t> ent;
    std::vector<int32_t> long_row;      // [nlong]
    std::vector<uint32_t> long_start;   // [nlong + 1]
    std::vector<uint32_t> long_ent;
    std::vector<double> coef;           // [256]
    int nlong() const { return (int)long_row.size(); }
};

SellHost build_sell(const Csr& M, int long_threshold);

}  // namespace v2h
