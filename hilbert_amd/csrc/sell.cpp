// sell.cpp -- pack a CSR operator into SELL-64 gather tables.
#include "sell.hpp"

#include <map>
#include <stdexcept>

namespace v2h {

SellHost build_sell(const Csr& M, int long_threshold) {
    SellHost S;
    S.nrows = M.nrows;
    if (M.ncols >= ((int64_t)1 << 24) || M.nrows >= ((int64_t)1 << 31))
        throw std::invalid_argument("operator too large for 24-bit column indices");
    S.coef.assign(256, 0.0);
    std::map<double, uint32_t> code_of;
    auto code = [&](double v) -> uint32_t {
        auto it = code_of.find(v);
        if (it != code_of.end()) return it->second;
        uint32_t c = (uint32_t)code_of.size() + 1;
        if (c >= kCodeLongRow) throw std::runtime_error("more than 254 distinct coefficients");
        code_of[v] = c;
        S.coef[c] = v;
        return c;
    };
    auto pack = [&](int64_t e) { return (code(M.val[e]) << 24) | (uint32_t)M.col[e]; };

    S.nslices = (int)((M.nrows + 63) / 64);
    S.slice_start.assign(S.nslices + 1, 0);
    S.long_start.assign(1, 0);
    uint64_t total = 0;
    for (int s = 0; s < S.nslices; s++) {
        int width = 0;
        for (int lane = 0; lane < 64; lane++) {
            int64_t r = (int64_t)s * 64 + lane;
            if (r >= M.nrows) break;
            int len = (int)(M.ptr[r + 1] - M.ptr[r]);
            width = std::max(width, len > long_threshold ? 1 : len);
        }
        S.slice_start[s] = (uint32_t)total;
        total += (uint64_t)width;
        if (total * 64 >= ((uint64_t)1 << 32)) throw std::runtime_error("gather table too large");
        S.slice_start[s + 1] = (uint32_t)total;
    }
    S.ent.assign(total * 64, 0u);
    for (int s = 0; s < S.nslices; s++) {
        uint64_t base = (uint64_t)S.slice_start[s] * 64;
        for (int lane = 0; lane < 64; lane++) {
            int64_t r = (int64_t)s * 64 + lane;
            if (r >= M.nrows) break;
            int64_t len = M.ptr[r + 1] - M.ptr[r];
            if (len > long_threshold) {
                S.ent[base + lane] = kCodeLongRow << 24;
                S.long_row.push_back((int32_t)r);
                for (int64_t e = M.ptr[r]; e < M.ptr[r + 1]; e++) S.long_ent.push_back(pack(e));
                S.long_start.push_back((uint32_t)S.long_ent.size());
                continue;
            }
            for (int64_t k = 0; k < len; k++) S.ent[base + (uint64_t)k * 64 + lane] = pack(M.ptr[r] + k);
        }
    }
    return S;
}

}  // namespace v2h
