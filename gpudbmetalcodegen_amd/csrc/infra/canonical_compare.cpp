#include "canonical_compare.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace gdb {
namespace infra {

namespace {

std::vector<std::string> lines(const std::string& s) {
    std::vector<std::string> out;
    size_t p = 0;
    while (p <= s.size()) {
        size_t e = s.find('\n', p);
        if (e == std::string::npos) e = s.size();
        std::string ln = s.substr(p, e - p);
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        out.push_back(std::move(ln));
        if (e == s.size()) break;
        p = e + 1;
    }
    while (!out.empty() && out.back().empty()) out.pop_back();
    return out;
}

std::vector<std::string> fields(const std::string& ln) {
    std::vector<std::string> out(1);
    bool quoted = false;
    for (size_t i = 0; i < ln.size(); ++i) {
        const char c = ln[i];
        if (quoted) {
            if (c != '"') {
                out.back() += c;
            } else if (i + 1 < ln.size() && ln[i + 1] == '"') {
                out.back() += '"';
                ++i;
            } else {
                quoted = false;
            }
        } else if (c == '"') {
            quoted = true;
        } else if (c == ',') {
            out.emplace_back();
        } else {
            out.back() += c;
        }
    }
    return out;
}

bool number(const std::string& s, double& v) {
    if (s.empty()) return false;
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end != s.c_str() && *end == '\0';
}

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    std::vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

}  // namespace

std::string compareCanonical(const std::string& got, const std::string& expected, double absTol, double relTol) {
    const auto a = lines(got), b = lines(expected);
    if (a.empty() && b.empty()) return "";
    if (a.empty()) return fmt("got 0 rows, expected %zu data rows", b.size() - 1);
    if (b.empty()) return fmt("got %zu data rows, expected 0", a.size() - 1);
    const auto ha = fields(a[0]), hb = fields(b[0]);
    std::vector<std::pair<size_t, size_t>> shared;
    for (size_t i = 0; i < ha.size(); ++i) {
        const auto it = std::find(hb.begin(), hb.end(), ha[i]);
        if (it != hb.end()) shared.push_back({i, static_cast<size_t>(it - hb.begin())});
    }
    if (shared.empty()) return "schema mismatch: got cols=[" + a[0] + "] expected cols=[" + b[0] + "]";
    if (a.size() != b.size()) return fmt("row count mismatch: got=%zu expected=%zu", a.size() - 1, b.size() - 1);
    for (size_t r = 1; r < a.size(); ++r) {
        const auto ra = fields(a[r]), rb = fields(b[r]);
        for (const auto& [ia, ib] : shared) {
            if (ia >= ra.size() || ib >= rb.size()) continue;
            const std::string &va = ra[ia], &vb = rb[ib];
            if (va == vb) continue;
            double x = 0, y = 0;
            if (number(va, x) && number(vb, y)) {
                const double diff = std::fabs(x - y), tol = absTol + relTol * std::max(std::fabs(x), std::fabs(y));
                if (diff <= tol) continue;
                return fmt("row %zu col '%s': %s vs %s (diff=%.6g tol=%.6g)", r, ha[ia].c_str(), va.c_str(),
                           vb.c_str(), diff, tol);
            }
            return fmt("row %zu col '%s': '%s' vs '%s'", r, ha[ia].c_str(), va.c_str(), vb.c_str());
        }
    }
    return "";
}

}  // namespace infra
}  // namespace gdb
