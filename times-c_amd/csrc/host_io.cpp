// Host I/O with the reference's file formats, plus a binary matrix format.
//
// CSV (Code/include/readfile.cuh:12-142): one series per line, fields
// separated by ','; the labeled form has an integer label first (parsed like
// std::stoi), values are parsed like atof and rounded to float; a trailing
// comma is ignored (getline(ss, s, ',') yields no empty last field).
// Deviations, each an error instead of undefined behaviour in the
// reference: a missing or empty file (data.at(0) throws, readfile.cuh:56),
// ragged rows (out-of-range reads, readfile.cuh:59-64) and a non-integer
// label (stoi throws) return TCX_ERR_IO.
//
// The reader is multithreaded: the file is memory-mapped and split at line
// boundaries into one chunk per thread; a first pass counts each chunk's
// lines, a second parses every chunk straight into its rows of the output,
// so the result does not depend on the thread count.  Values go through std::from_chars (correctly rounded, like
// strtod), with strtod itself for the forms from_chars does not take (hex
// floats, out-of-range magnitudes), then are rounded to float -- the same
// double-then-float rounding as atof into a float.
//
// Binary (".tcxb"): the 8 bytes "TCXB0001", int64 rows, int64 cols, int64
// labeled (0/1), then rows int32 labels when labeled, then rows*cols
// float32 values, all little-endian.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "tcx_host.hpp"
#include "timesc.h"

namespace {

constexpr char kMagic[8] = {'T', 'C', 'X', 'B', '0', '0', '0', '1'};

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

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// atof on the field [b, e): leading white space, optional sign, the longest
// valid prefix; no digits gives 0.
double parse_value(const char* b, const char* e) {
  while (b < e && is_space(*b)) ++b;
  const char* p = b;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
    if (p < e && (*p == '+' || *p == '-')) return 0.0;  // a second sign: no number
  }
  const bool hex = e - p >= 2 && p[0] == '0' && (p[1] == 'x' || p[1] == 'X');
  if (!hex) {
    double v = 0.0;
    const std::from_chars_result r = std::from_chars(p, e, v);
    if (r.ec == std::errc()) return neg ? -v : v;
    if (r.ec == std::errc::invalid_argument) return 0.0;
  }
  // hex floats and out-of-range magnitudes: strtod on a bounded copy
  char buf[128];
  const size_t n = std::min<size_t>((size_t)(e - b), sizeof(buf) - 1);
  std::memcpy(buf, b, n);
  buf[n] = '\0';
  return std::strtod(buf, nullptr);
}

// std::stoi on [b, e): false when there are no digits or the value
// overflows int.
bool parse_label(const char* b, const char* e, int* out) {
  char buf[64];
  size_t n = std::min<size_t>((size_t)(e - b), sizeof(buf) - 1);
  std::memcpy(buf, b, n);
  buf[n] = '\0';
  char* end = nullptr;
  errno = 0;
  long v = std::strtol(buf, &end, 10);
  if (end == buf || errno == ERANGE || v < -2147483648L || v > 2147483647L) return false;
  *out = (int)v;
  return true;
}

// Number of values of the line [b, e) (fields minus the label).
long count_values(const char* b, const char* e, bool labeled) {
  long f = 0;
  const char* p = b;
  while (p < e) {
    const char* q = (const char*)std::memchr(p, ',', (size_t)(e - p));
    ++f;
    if (!q) break;
    p = q + 1;
  }
  return std::max(0L, f - (labeled ? 1 : 0));
}

struct Chunk {
  const char* b;
  const char* e;
  long rows = 0;      // lines in [b, e)
  long row0 = 0;      // first row of the chunk in the file
  int err = TCX_OK;
  long err_line = 0;  // line within the chunk (1-based)
  std::string err_what;
};

long count_lines(const char* b, const char* e) {
  long n = 0;
  const char* p = b;
  while (p < e) {
    const char* q = (const char*)std::memchr(p, '\n', (size_t)(e - p));
    ++n;
    if (!q) break;
    p = q + 1;
  }
  return n;
}

// Parses the lines of one chunk straight into rows row0.. of the output.
void parse_chunk(Chunk& c, bool labeled, long ncols, float* data, int* labels) {
  float* out = data + (size_t)c.row0 * ncols;
  int* lout = labeled ? labels + c.row0 : nullptr;
  long row = 0;
  const char* p = c.b;
  while (p < c.e) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(c.e - p));
    const char* le = eol ? eol : c.e;
    long nval = 0;
    const char* f = p;
    bool first = true;
    while (f < le) {
      const char* q = (const char*)std::memchr(f, ',', (size_t)(le - f));
      const char* fe = q ? q : le;
      if (labeled && first) {
        if (!parse_label(f, fe, lout + row)) {
          c.err = TCX_ERR_IO;
          c.err_line = row + 1;
          c.err_what = "non-integer label";
          return;
        }
      } else {
        if (nval < ncols) out[(size_t)row * ncols + nval] = (float)parse_value(f, fe);
        ++nval;
      }
      first = false;
      if (!q) break;
      f = q + 1;
    }
    if (nval != ncols) {
      c.err = TCX_ERR_IO;
      c.err_line = row + 1;
      c.err_what = "ragged CSV: line has " + std::to_string(nval) + " values, expected " +
                   std::to_string(ncols);
      return;
    }
    ++row;
    p = eol ? eol + 1 : c.e;
  }
}

// The file's bytes: memory-mapped (read-only), or read into `copy` when
// mapping is not possible.
struct FileView {
  const char* data = nullptr;
  size_t size = 0;
  void* map = nullptr;
  std::string copy;
  ~FileView() {
    if (map) munmap(map, size);
  }
  bool open(const char* path) {
    int fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
      void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m != MAP_FAILED) {
        map = m;
        size = (size_t)st.st_size;
        data = (const char*)m;
        ::close(fd);
        return true;
      }
    }
    ::close(fd);
    if (!slurp(path, copy)) return false;
    data = copy.data();
    size = copy.size();
    return true;
  }
};

template <class F>
void parallel_for(size_t n, F fn) {
  if (n == 1) {
    fn(0);
    return;
  }
  std::vector<std::thread> th;
  for (size_t t = 0; t < n; ++t) th.emplace_back(fn, t);
  for (std::thread& x : th) x.join();
}

int read_csv(const char* path, bool labeled, float** data, int** labels, int* rows, int* cols) {
  if (!path || !data || !rows || !cols || (labeled && !labels))
    return tcx::fail(TCX_ERR_ARG, "read_csv: null argument");
  FileView fv;
  if (!fv.open(path)) return tcx::fail(TCX_ERR_IO, std::string("cannot read file: ") + path);
  const char* b = fv.data;
  const char* e = b + fv.size;
  if (b == e) return tcx::fail(TCX_ERR_IO, std::string("no data in file: ") + path);
  const char* eol0 = (const char*)std::memchr(b, '\n', fv.size);
  const long ncols = count_values(b, eol0 ? eol0 : e, labeled);
  if (ncols <= 0) return tcx::fail(TCX_ERR_IO, std::string("no data in file: ") + path);
  // chunks at line boundaries, about 4 MB or more each (hardware_concurrency
  // may exceed the CPUs this process may use; it only bounds the count)
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nthreads = std::max<size_t>(
      1, std::min<size_t>({hw ? hw : 1, 16, fv.size / (4u << 20) + 1}));
  std::vector<Chunk> ch(nthreads);
  const char* p = b;
  for (size_t t = 0; t < nthreads; ++t) {
    const char* target = t + 1 == nthreads ? e : b + fv.size * (t + 1) / nthreads;
    if (target < p) target = p;
    const char* nl = target < e ? (const char*)std::memchr(target, '\n', (size_t)(e - target))
                                : nullptr;
    const char* ce = (t + 1 == nthreads || !nl) ? e : nl + 1;
    ch[t].b = p;
    ch[t].e = ce;
    p = ce;
  }
  parallel_for(nthreads, [&](size_t t) { ch[t].rows = count_lines(ch[t].b, ch[t].e); });
  long nrows = 0;
  for (Chunk& c : ch) {
    c.row0 = nrows;
    nrows += c.rows;
  }
  if (nrows > 0x7fffffffL) return tcx::fail(TCX_ERR_IO, "CSV has more than 2^31 rows");
  float* d = (float*)std::malloc(sizeof(float) * (size_t)nrows * ncols);
  int* l = labeled ? (int*)std::malloc(sizeof(int) * (size_t)nrows) : nullptr;
  if (!d || (labeled && !l)) {
    std::free(d);
    std::free(l);
    return tcx::fail(TCX_ERR_NOMEM, "read_csv: host allocation");
  }
  parallel_for(nthreads, [&](size_t t) { parse_chunk(ch[t], labeled, ncols, d, l); });
  for (const Chunk& c : ch) {
    if (c.err != TCX_OK) {
      std::free(d);
      std::free(l);
      return tcx::fail(c.err,
                       c.err_what + " (line " + std::to_string(c.row0 + c.err_line) + ")");
    }
  }
  *data = d;
  if (labeled) *labels = l;
  *rows = (int)nrows;
  *cols = (int)ncols;
  return TCX_OK;
}

}  // namespace

extern "C" {

int tcx_read_labeled_csv(const char* path, float** data, int** labels, int* rows, int* cols) {
  tcx::clear_error();
  return read_csv(path, true, data, labels, rows, cols);
}

int tcx_read_unlabeled_csv(const char* path, float** data, int* rows, int* cols) {
  tcx::clear_error();
  return read_csv(path, false, data, nullptr, rows, cols);
}

int tcx_write_labels(const char* path, const int* labels, int n) {
  tcx::clear_error();
  if (!path || (!labels && n > 0) || n < 0)
    return tcx::fail(TCX_ERR_ARG, "write_labels: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return tcx::fail(TCX_ERR_IO, std::string("cannot write ") + path);
  std::string buf;
  buf.reserve((size_t)n * 4);
  char tmp[16];
  for (int i = 0; i < n; ++i) {
    std::to_chars_result r = std::to_chars(tmp, tmp + sizeof(tmp), labels[i]);
    buf.append(tmp, r.ptr);
    buf.push_back(',');
  }
  bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TCX_OK : tcx::fail(TCX_ERR_IO, std::string("write failed: ") + path);
}

int tcx_write_labels_lines(const char* path, const int* labels, int n) {
  tcx::clear_error();
  if (!path || (!labels && n > 0) || n < 0)
    return tcx::fail(TCX_ERR_ARG, "write_labels_lines: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return tcx::fail(TCX_ERR_IO, std::string("cannot write ") + path);
  std::string buf;
  buf.reserve((size_t)n * 4);
  char tmp[16];
  for (int i = 0; i < n; ++i) {
    std::to_chars_result r = std::to_chars(tmp, tmp + sizeof(tmp), labels[i]);
    buf.append(tmp, r.ptr);
    buf.push_back('\n');
  }
  bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TCX_OK : tcx::fail(TCX_ERR_IO, std::string("write failed: ") + path);
}

int tcx_write_binary(const char* path, const float* data, const int* labels, int rows,
                     int cols) {
  tcx::clear_error();
  if (!path || !data || rows < 1 || cols < 1)
    return tcx::fail(TCX_ERR_ARG, "write_binary: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return tcx::fail(TCX_ERR_IO, std::string("cannot write ") + path);
  const int64_t hdr[3] = {rows, cols, labels ? 1 : 0};
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
  if (ok && labels) ok = std::fwrite(labels, sizeof(int), (size_t)rows, f) == (size_t)rows;
  if (ok) ok = std::fwrite(data, sizeof(float), (size_t)rows * cols, f) == (size_t)rows * cols;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TCX_OK : tcx::fail(TCX_ERR_IO, std::string("write failed: ") + path);
}

int tcx_read_binary(const char* path, float** data, int** labels, int* rows, int* cols) {
  tcx::clear_error();
  if (!path || !data || !rows || !cols) return tcx::fail(TCX_ERR_ARG, "read_binary: null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return tcx::fail(TCX_ERR_IO, std::string("cannot read file: ") + path);
  char magic[8];
  int64_t hdr[3];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
            std::fread(hdr, sizeof(hdr), 1, f) == 1;
  if (!ok || hdr[0] < 1 || hdr[1] < 1 || hdr[0] > 0x7fffffff || hdr[1] > 0x7fffffff ||
      (hdr[2] != 0 && hdr[2] != 1)) {
    std::fclose(f);
    return tcx::fail(TCX_ERR_IO, std::string("not a timesc binary matrix: ") + path);
  }
  const size_t r = (size_t)hdr[0], c = (size_t)hdr[1];
  float* d = (float*)std::malloc(sizeof(float) * r * c);
  int* l = hdr[2] ? (int*)std::malloc(sizeof(int) * r) : nullptr;
  if (!d || (hdr[2] && !l)) {
    std::fclose(f);
    std::free(d);
    std::free(l);
    return tcx::fail(TCX_ERR_NOMEM, "read_binary: host allocation");
  }
  ok = (!l || std::fread(l, sizeof(int), r, f) == r) && std::fread(d, sizeof(float), r * c, f) == r * c;
  std::fclose(f);
  if (!ok) {
    std::free(d);
    std::free(l);
    return tcx::fail(TCX_ERR_IO, std::string("truncated binary matrix: ") + path);
  }
  *data = d;
  if (labels)
    *labels = l;
  else
    std::free(l);
  *rows = (int)r;
  *cols = (int)c;
  return TCX_OK;
}

void tcx_free_host(void* p) { std::free(p); }

}  // extern "C"
