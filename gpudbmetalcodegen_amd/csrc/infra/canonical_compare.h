#pragma once
// Golden-file check of a canonical result CSV.
//
// Same verdicts as the reference's compareCanonical
// (codegen/codegen_main.cpp:63-186), kept alike on purpose so a result that
// passes --check there passes here: lines split on '\n' with a trailing '\r'
// and trailing empty lines dropped; fields split on ',' with "..." quoting
// and "" escapes; columns matched by header name (columns only one side has
// are ignored; none shared is a schema mismatch); equal row counts; per
// shared field, equal text passes, two numbers pass when
// |a - b| <= absTol + relTol * max(|a|, |b|), anything else fails. The first
// difference is reported.

#include <string>

namespace gdb {
namespace infra {

// Empty string on a match, otherwise a one-line description of the first
// difference.
std::string compareCanonical(const std::string& got, const std::string& expected, double absTol, double relTol);

}  // namespace infra
}  // namespace gdb
