// ASCII numbers of OpenFOAM's field files and of the single-cell trace files, formatted
// and parsed on the host.  OpenFOAM writes scalars through std::ostream: general format
// with writePrecision significant digits (the printf %g form), or std::ios::fixed with a
// set number of decimals (singleCellSolver.C:117-118).  std::to_chars gives exactly those
// forms; several threads format disjoint slices that are then concatenated in order.
#pragma once
#include <algorithm>
#include <charconv>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace cg {

// one value and its separator; returns the end, or nullptr if [p, end) is too short
inline char* format_one(char* p, char* end, double v, int precision, bool fixed, char sep)
{
    const auto r = fixed ? std::to_chars(p, end, v, std::chars_format::fixed, precision)
                         : std::to_chars(p, end, v, std::chars_format::general, precision);
    if (r.ec != std::errc() || r.ptr >= end) return nullptr;
    *r.ptr = sep;
    return r.ptr + 1;
}

// formats n values into buf (capacity cap), `row` values to a line separated by single
// spaces; *len gets the bytes written.  Returns false when cap is too small (nothing
// useful in buf then).
inline bool format_scalars(const double* v, int64_t n, int precision, bool fixed, int64_t row,
                           char* buf, int64_t cap, int64_t* len)
{
    auto sep = [row](int64_t i) { return (i + 1) % row == 0 ? '\n' : ' '; };
    constexpr int64_t kSlice = 1 << 16;
    const int64_t want = (n + kSlice - 1) / kSlice;
    const int nt = (int)std::max<int64_t>(
        1, std::min<int64_t>(want, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    if (nt == 1) {
        char* p = buf;
        char* end = buf + cap;
        for (int64_t i = 0; i < n; ++i)
            if (!(p = format_one(p, end, v[i], precision, fixed, sep(i)))) return false;
        *len = p - buf;
        return true;
    }
    // each thread formats its contiguous range into its own buffer, then the pieces are
    // copied to buf in order
    std::vector<std::vector<char>> part(nt);
    std::vector<char> ok(nt, 1);
    std::vector<std::thread> th;
    const int64_t per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            const int64_t a = std::min(n, t * per), b = std::min(n, a + per);
            std::vector<char>& out = part[t];
            out.resize((size_t)(b - a) * 32 + 64);
            char* p = out.data();
            char* end = out.data() + out.size();
            for (int64_t i = a; i < b; ++i) {
                if (end - p < 400) {  // fixed format of a large value needs room
                    const size_t used = p - out.data();
                    out.resize(out.size() * 2 + 800);
                    p = out.data() + used;
                    end = out.data() + out.size();
                }
                if (!(p = format_one(p, end, v[i], precision, fixed, sep(i)))) {
                    ok[t] = 0;
                    return;
                }
            }
            out.resize(p - out.data());
        });
    }
    for (auto& x : th) x.join();
    int64_t total = 0;
    for (int t = 0; t < nt; ++t) {
        if (!ok[t]) return false;
        total += (int64_t)part[t].size();
    }
    if (total > cap) return false;
    char* p = buf;
    for (int t = 0; t < nt; ++t) {
        std::memcpy(p, part[t].data(), part[t].size());
        p += part[t].size();
    }
    *len = total;
    return true;
}

// parses n numbers from text; whitespace and the list brackets '(' ')' separate them
// (List<scalar>, List<vector>, List<tensor> bodies).  *consumed gets the offset just past
// the n-th number.  Returns 0, or 1 on a malformed token, 2 when the text ends early.
inline int parse_scalars(const char* text, int64_t len, double* out, int64_t n,
                         int64_t* consumed)
{
    const char* p = text;
    const char* end = text + len;
    for (int64_t i = 0; i < n; ++i) {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '(' ||
                           *p == ')'))
            ++p;
        if (p == end) return 2;
        const char* q = p + (*p == '+');  // from_chars takes no leading '+'
        const auto r = std::from_chars(q, end, out[i]);
        if (r.ec != std::errc()) return 1;
        p = r.ptr;
        if (p < end && !(*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '(' ||
                         *p == ')' || *p == ';'))
            return 1;
    }
    *consumed = p - text;
    return 0;
}

} // namespace cg
