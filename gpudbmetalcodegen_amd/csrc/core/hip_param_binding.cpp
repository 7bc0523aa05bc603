#include "hip_param_binding.h"

#include <cctype>

namespace gdb {
namespace codegen {

namespace {

// Recursive-descent evaluator over unsigned 64-bit integers:
//   expr := term (('+'|'-') term)*
//   term := atom (('*'|'/') atom)*
//   atom := number | symbol | '(' expr ')' | ('min'|'max') '(' expr ',' expr ')'
//         | 'pow2' '(' expr ')'      (the smallest power of two >= expr, 1 for 0)
class SizeParser {
public:
    SizeParser(const std::string& s, const HipSizeResolver& r) : s_(s), r_(r) {}
    uint64_t parse() {
        uint64_t v = expr();
        skip();
        if (pos_ != s_.size()) throw std::runtime_error("HipSizeResolver: trailing input in '" + s_ + "'");
        return v;
    }

private:
    void skip() {
        while (pos_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[pos_]))) ++pos_;
    }
    uint64_t expr() {
        uint64_t v = term();
        for (;;) {
            skip();
            if (pos_ < s_.size() && (s_[pos_] == '+' || s_[pos_] == '-')) {
                char op = s_[pos_++];
                uint64_t rhs = term();
                if (op == '+') {
                    v += rhs;
                } else {
                    if (rhs > v) throw std::runtime_error("HipSizeResolver: negative size in '" + s_ + "'");
                    v -= rhs;
                }
            } else {
                return v;
            }
        }
    }
    uint64_t term() {
        uint64_t v = atom();
        for (;;) {
            skip();
            if (pos_ < s_.size() && (s_[pos_] == '*' || s_[pos_] == '/')) {
                char op = s_[pos_++];
                uint64_t rhs = atom();
                if (op == '*') {
                    v *= rhs;
                } else {
                    if (rhs == 0) throw std::runtime_error("HipSizeResolver: division by zero in '" + s_ + "'");
                    v /= rhs;
                }
            } else {
                return v;
            }
        }
    }
    uint64_t atom() {
        skip();
        if (pos_ >= s_.size()) throw std::runtime_error("HipSizeResolver: empty expression '" + s_ + "'");
        char c = s_[pos_];
        if (c == '(') {
            ++pos_;
            uint64_t v = expr();
            skip();
            if (pos_ >= s_.size() || s_[pos_] != ')')
                throw std::runtime_error("HipSizeResolver: unbalanced parentheses in '" + s_ + "'");
            ++pos_;
            return v;
        }
        if (std::isdigit(static_cast<unsigned char>(c))) {
            uint64_t v = 0;
            while (pos_ < s_.size() && std::isdigit(static_cast<unsigned char>(s_[pos_])))
                v = v * 10 + static_cast<uint64_t>(s_[pos_++] - '0');
            return v;
        }
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            size_t b = pos_;
            while (pos_ < s_.size() &&
                   (std::isalnum(static_cast<unsigned char>(s_[pos_])) || s_[pos_] == '_'))
                ++pos_;
            const std::string name = s_.substr(b, pos_ - b);
            skip();
            if (name == "pow2" && pos_ < s_.size() && s_[pos_] == '(') {
                ++pos_;
                const uint64_t a = expr();
                skip();
                if (pos_ >= s_.size() || s_[pos_] != ')')
                    throw std::runtime_error("HipSizeResolver: unbalanced parentheses in '" + s_ + "'");
                ++pos_;
                uint64_t p = 1;
                while (p < a) {
                    if (p >> 63) throw std::runtime_error("HipSizeResolver: pow2 overflow in '" + s_ + "'");
                    p <<= 1;
                }
                return p;
            }
            if ((name == "min" || name == "max") && pos_ < s_.size() && s_[pos_] == '(') {
                ++pos_;
                const uint64_t a = expr();
                skip();
                if (pos_ >= s_.size() || s_[pos_] != ',')
                    throw std::runtime_error("HipSizeResolver: expected ',' in '" + s_ + "'");
                ++pos_;
                const uint64_t c = expr();
                skip();
                if (pos_ >= s_.size() || s_[pos_] != ')')
                    throw std::runtime_error("HipSizeResolver: unbalanced parentheses in '" + s_ + "'");
                ++pos_;
                return name == "min" ? (a < c ? a : c) : (a > c ? a : c);
            }
            return r_.getSymbol(name);
        }
        throw std::runtime_error("HipSizeResolver: unexpected '" + std::string(1, c) + "' in '" + s_ + "'");
    }

    const std::string& s_;
    const HipSizeResolver& r_;
    size_t pos_ = 0;
};

}  // namespace

uint64_t HipSizeResolver::resolve(const std::string& expr) const {
    if (expr.empty()) throw std::runtime_error("HipSizeResolver: empty expression");
    return SizeParser(expr, *this).parse();
}

}  // namespace codegen
}  // namespace gdb
