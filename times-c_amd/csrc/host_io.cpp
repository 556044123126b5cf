// CSV readers/writer with the reference's file format (host code).
//
// Format (Code/include/readfile.cuh:12-142): one series per line, fields
// separated by ','; the labeled form has an integer label first (parsed like
// std::stoi), values are parsed like atof and rounded to float; a trailing
// comma is ignored.  Deviations, each an error instead of undefined
// behaviour in the reference: a missing or empty file (data.at(0) throws,
// readfile.cuh:56), ragged rows (out-of-range reads, readfile.cuh:59-64) and
// a non-integer label (stoi throws) return TCX_ERR_IO.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tcx_host.hpp"
#include "timesc.h"

namespace {

bool slurp(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)sz);
  size_t got = sz ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == (size_t)sz;
}

// Parses [b, e) as comma-separated fields the way getline(ss, s, ',') splits
// them: an empty trailing field after the last comma is not a field.
void split_fields(const char* b, const char* e, std::vector<std::string>& f) {
  f.clear();
  const char* p = b;
  while (p < e) {
    const char* q = p;
    while (q < e && *q != ',') ++q;
    f.emplace_back(p, q);
    if (q == e) break;
    p = q + 1;
  }
}

// std::stoi semantics: optional leading whitespace, sign, digits; anything
// after the digits is ignored; no digits is an error.
bool parse_label(const std::string& s, int* out) {
  const char* c = s.c_str();
  char* end = nullptr;
  errno = 0;
  long v = std::strtol(c, &end, 10);
  if (end == c || errno == ERANGE || v < -2147483648L || v > 2147483647L)
    return false;
  *out = (int)v;
  return true;
}

int read_csv(const char* path, bool labeled, float** data, int** labels,
             int* rows, int* cols) {
  if (!path || !data || !rows || !cols || (labeled && !labels))
    return tcx::fail(TCX_ERR_ARG, "read_csv: null argument");
  std::string text;
  if (!slurp(path, text))
    return tcx::fail(TCX_ERR_IO, std::string("cannot read file: ") + path);
  std::vector<float> vals;
  std::vector<int> labs;
  std::vector<std::string> fields;
  long nrows = 0;
  long ncols = -1;
  const char* p = text.data();
  const char* end = p + text.size();
  while (p < end) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    const char* le = eol ? eol : end;
    split_fields(p, le, fields);
    p = eol ? eol + 1 : end;
    long nval = (long)fields.size() - (labeled ? 1 : 0);
    if (nval < 0) nval = 0;
    if (ncols < 0) ncols = nval;
    if (nval != ncols)
      return tcx::fail(TCX_ERR_IO, "ragged CSV: line " +
                                       std::to_string(nrows + 1) + " has " +
                                       std::to_string(nval) + " values, expected " +
                                       std::to_string(ncols));
    size_t f0 = 0;
    if (labeled) {
      int lab;
      if (!parse_label(fields[0], &lab))
        return tcx::fail(TCX_ERR_IO, "non-integer label on line " +
                                         std::to_string(nrows + 1));
      labs.push_back(lab);
      f0 = 1;
    }
    for (size_t j = f0; j < fields.size(); ++j)
      vals.push_back((float)std::atof(fields[j].c_str()));
    ++nrows;
  }
  if (nrows == 0 || ncols <= 0)
    return tcx::fail(TCX_ERR_IO, std::string("no data in file: ") + path);
  float* d = (float*)std::malloc(sizeof(float) * vals.size());
  if (!d) return tcx::fail(TCX_ERR_NOMEM, "read_csv: host allocation");
  std::memcpy(d, vals.data(), sizeof(float) * vals.size());
  *data = d;
  if (labeled) {
    int* l = (int*)std::malloc(sizeof(int) * labs.size());
    if (!l) {
      std::free(d);
      return tcx::fail(TCX_ERR_NOMEM, "read_csv: host allocation");
    }
    std::memcpy(l, labs.data(), sizeof(int) * labs.size());
    *labels = l;
  }
  *rows = (int)nrows;
  *cols = (int)ncols;
  return TCX_OK;
}

}  // namespace

extern "C" int tcx_read_labeled_csv(const char* path, float** data,
                                    int** labels, int* rows, int* cols) {
  return read_csv(path, true, data, labels, rows, cols);
}

extern "C" int tcx_read_unlabeled_csv(const char* path, float** data,
                                      int* rows, int* cols) {
  return read_csv(path, false, data, nullptr, rows, cols);
}

extern "C" int tcx_write_labels(const char* path, const int* labels, int n) {
  if (!path || (!labels && n > 0) || n < 0)
    return tcx::fail(TCX_ERR_ARG, "write_labels: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return tcx::fail(TCX_ERR_IO, std::string("cannot write ") + path);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%d,", labels[i]);
  bool ok = std::fclose(f) == 0;
  return ok ? TCX_OK : tcx::fail(TCX_ERR_IO, std::string("write failed: ") + path);
}

extern "C" void tcx_free_host(void* p) { std::free(p); }
