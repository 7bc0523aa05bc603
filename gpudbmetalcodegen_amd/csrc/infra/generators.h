#pragma once
// Seeded synthetic column generators.
//
// The reference loads TPC-H .tbl/.colbin files (src/infra.h:141-782); the
// workloads here are synthetic, so columns come from counter-based
// generators: value(seed, stream, row) is a pure function, which makes every
// column reproducible in any order, on any number of threads, and identical
// for the device path and the CPU oracle. Each generator fills host memory;
// loops are split over threads.

#include <cstdint>
#include <string>
#include <vector>

namespace gdb {
namespace infra {

// splitmix64 finalizer and a counter-based stream: rng(seed, stream, i) is
// element i of the splitmix64 sequence whose state is derived from
// (seed, stream).
uint64_t mix64(uint64_t x);
uint64_t rng(uint64_t seed, uint64_t stream, uint64_t counter);
// Uniform integer in [lo, hi] from a 64-bit random value (multiply-shift).
int64_t uniformInt(uint64_t r, int64_t lo, int64_t hi);
// Uniform double in [0, 1) from the top 53 bits.
double uniform01(uint64_t r);

int defaultThreads();

void genUniformI32(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int32_t* out, int threads = 0);
void genUniformI64(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int64_t* out, int threads = 0);
// Values drawn with integer weights.
void genCategoricalU8(uint64_t seed, uint64_t stream, uint64_t n, const uint8_t* values, const uint32_t* weights,
                      int nvalues, uint8_t* out, int threads = 0);
// out[i] = base + perm(i), perm a seeded bijection of [0, n) (Feistel network
// with cycle walking).
void genPermutationI32(uint64_t seed, uint64_t stream, uint64_t n, int64_t base, int32_t* out, int threads = 0);
uint64_t permute(uint64_t seed, uint64_t stream, uint64_t n, uint64_t i);
// Parent key k in [1, nparent] gets c(k) ~ U{cmin..cmax} children; returns
// the total number of children.
uint64_t fanoutTotal(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax, int threads = 0);
// Child foreign keys: the multiset {k repeated c(k) times}, in a seeded
// permuted order. nchild must equal fanoutTotal(...).
void genFanoutFkeyI32(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax, uint64_t nchild,
                      int32_t* out, int threads = 0);
// Zipf(s) over 1..nkeys by rejection-inversion (Hoermann & Derflinger 1996).
void genZipfI32(uint64_t seed, uint64_t stream, uint64_t n, double s, int64_t nkeys, int32_t* out, int threads = 0);
// Vocabulary of nwords lowercase pseudo-words (length minlen..maxlen), each
// stored NUL-padded in 16 bytes.
void genVocabulary(uint64_t seed, int nwords, int minlen, int maxlen, char* out);
// Fixed-width NUL-padded strings of space-separated vocabulary words.
void genText(uint64_t seed, uint64_t stream, uint64_t n, int width, const char* vocab, int nwords, char* out,
             int threads = 0);

}  // namespace infra
}  // namespace gdb
