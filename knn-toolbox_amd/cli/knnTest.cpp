// knnTest: file-driven driver of the drop-in C API, the counterpart of src/knnTest.cu:65-110.
//
//   knnTest       <corpus.bin> <queries.bin> N Q D k     (float knntype,  links libcuknn.so)
//   knnTestDouble <corpus.bin> <queries.bin> N Q D k     (double knntype, links libcuknn_double.so)
//
// Inputs are raw little-endian knntype arrays, D leading: the corpus is N rows of D values, the
// queries Q rows of D values (MATLAB's column-major d x N written with fwrite, demo/exportData.m).
// Outputs KNNdist.bin and KNNidx.bin in the working directory: Q rows of k knntype values, k
// leading (read back in MATLAB with fread(fid, [k Q], type), demo/importData.m:3-14); indices are
// 0-based row numbers stored as knntype, sorted by (distance, index).
//
// Unlike the reference driver, arguments and file sizes are checked and failures end the program
// with a non-zero status: 2 for usage / input errors, 1 for a failed search.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cuknns.h"

namespace {

bool parse_long(const char* s, long* out) {
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s, &end, 10);
  if (errno != 0 || end == s || *end != '\0') return false;
  *out = v;
  return true;
}

bool load(std::vector<knntype>& a, const char* file, long count) {
  std::printf("Loading data from file: %s\n", file);
  FILE* f = std::fopen(file, "rb");
  if (!f) {
    std::fprintf(stderr, "knnTest: cannot open input file %s: %s\n", file, std::strerror(errno));
    return false;
  }
  a.resize((size_t)count);
  const size_t got = std::fread(a.data(), sizeof(knntype), (size_t)count, f);
  std::fclose(f);
  if (got != (size_t)count) {
    std::fprintf(stderr, "knnTest: %s holds %zu values of %zu bytes, expected %ld\n", file, got,
                 sizeof(knntype), count);
    return false;
  }
  return true;
}

bool save(const std::vector<knntype>& a, const char* file) {
  std::printf("Saving data to file: %s\n", file);
  FILE* f = std::fopen(file, "wb");
  if (!f) {
    std::fprintf(stderr, "knnTest: cannot open output file %s: %s\n", file, std::strerror(errno));
    return false;
  }
  const size_t put = std::fwrite(a.data(), sizeof(knntype), a.size(), f);
  const bool ok = std::fclose(f) == 0 && put == a.size();
  if (!ok) std::fprintf(stderr, "knnTest: short write to %s\n", file);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s <corpus.bin> <queries.bin> N Q D k\n", argc > 0 ? argv[0] : "knnTest");
    return 2;
  }
  long N, Q, D, k;
  if (!parse_long(argv[3], &N) || !parse_long(argv[4], &Q) || !parse_long(argv[5], &D) ||
      !parse_long(argv[6], &k) || N < 1 || Q < 1 || D < 1 || k < 1 || N > 0x7fffff80L ||
      Q > 0x7fffffffL || D > 65536 || k > 1024) {
    std::fprintf(stderr, "knnTest: N, Q, D, k must be integers with N, Q >= 1, 1 <= D <= 65536, "
                         "1 <= k <= 1024\n");
    return 2;
  }
  const char* datafile = argv[1];
  const char* queryfile = argv[2];
  std::printf("files: %s %s\n", datafile, queryfile);
  std::vector<knntype> data, queries;
  if (!load(data, datafile, N * D) || !load(queries, queryfile, Q * D)) return 2;
  std::vector<knntype> dist((size_t)(Q * k)), idx((size_t)(Q * k));

  knnplan plan;
  knnsplan(&plan, N, Q, D, k);
  knnsexecute(plan, data.data(), queries.data(), dist.data(), idx.data());
  const char* err = knnsGetLastError();
  if (err && err[0]) {
    std::fprintf(stderr, "knnTest: search failed: %s\n", err);
    return 1;
  }
  if (!save(dist, "KNNdist.bin") || !save(idx, "KNNidx.bin")) return 1;
  return 0;
}
