// A caller written against the reference's C++ entry points (the
// float instantiations of Code/include/{normalization,ncc,SBD,shapeExtract,
// kshape,readfile}.cuh), built on include/timesc_ref.hpp.  It runs every
// forward once on the rows of a labeled CSV and writes each result as raw
// little-endian arrays into <outdir>/<name>.bin for tests/test_gpu_forwards.py.
//
//   ref_forwards <labeled.csv> <unlabeled.csv> <outdir> <k>
#include <cstdio>
#include <string>
#include <vector>

#include "timesc_ref.hpp"

namespace {

std::string g_out;

template <typename T>
void dump(const char* name, const T* p, size_t count) {
  const std::string path = g_out + "/" + name + ".bin";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, sizeof(T), count, f) != count) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  std::printf("wrote %s (%zu values)\n", name, count);
  std::fflush(stdout);
}

float* device_copy(const float* h, size_t count) {
  tcx_ctx* c = timesc_ref::context();
  float* d = nullptr;
  timesc_ref::check(tcx_malloc(c, (void**)&d, sizeof(float) * count), "malloc");
  timesc_ref::check(tcx_memcpy_h2d(c, d, h, sizeof(float) * count), "h2d");
  return d;
}

template <typename T>
std::vector<T> host_copy(const T* d, size_t count) {
  std::vector<T> h(count);
  timesc_ref::check(tcx_memcpy_d2h(timesc_ref::context(), h.data(), d, sizeof(T) * count), "d2h");
  return h;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s labeled.csv unlabeled.csv outdir k\n", argv[0]);
    return 2;
  }
  g_out = argv[3];
  const int k = std::atoi(argv[4]);
  tcx_ctx* c = timesc_ref::context();

  // readFile / read_unlabeled_File (readfile.cuh:12-116)
  std::vector<int> label;
  int rows = 0, cols = 0;
  float* h = readFile<float>(argv[1], label, rows, cols);
  dump("read_values", h, (size_t)rows * cols);
  dump("read_labels", label.data(), label.size());
  int urows = 0, ucols = 0;
  float* hu = read_unlabeled_File<float>(argv[2], urows, ucols);
  dump("read_unlabeled", hu, (size_t)urows * ucols);
  delete[] hu;

  // z_norm_gpu_x with ddof 1, as test_kshape.cu:37 calls it
  float* d = device_copy(h, (size_t)rows * cols);
  z_norm_gpu_x<float>(d, rows, cols, 1);
  const std::vector<float> z = host_copy(d, (size_t)rows * cols);
  dump("znorm", z.data(), z.size());

  // NCC_3D of rows 0..4 against rows 5..7 -> [5][3][2m-1]
  const int xr = 5, yr = 3, L = 2 * cols - 1;
  float* d_ncc = nullptr;
  timesc_ref::check(tcx_malloc(c, (void**)&d_ncc, sizeof(float) * xr * yr * L), "malloc");
  NCC_3D<float>(d, d + (size_t)xr * cols, d_ncc, xr, yr, cols);
  const std::vector<float> ncc = host_copy(d_ncc, (size_t)xr * yr * L);
  dump("ncc", ncc.data(), ncc.size());

  // sbd3D: rows 8..15 aligned in place to row 0
  float* d_mem = device_copy(z.data() + (size_t)8 * cols, (size_t)8 * cols);
  sbd3D<float>(d_mem, d, nullptr, 8, cols);
  const std::vector<float> aligned = host_copy(d_mem, (size_t)8 * cols);
  dump("sbd3d", aligned.data(), aligned.size());

  // extract_shape from the file's labels (one-based on the device, as the
  // reference's k_shape keeps them), old centroids = rows 0..k-1
  std::vector<int> one(rows);
  for (int i = 0; i < rows; ++i) one[i] = label[i] % k + 1;
  int* d_idx = nullptr;
  int* d_cnt = nullptr;
  timesc_ref::check(tcx_malloc(c, (void**)&d_idx, sizeof(int) * rows), "malloc");
  timesc_ref::check(tcx_malloc(c, (void**)&d_cnt, sizeof(int) * k), "malloc");
  timesc_ref::check(tcx_memcpy_h2d(c, d_idx, one.data(), sizeof(int) * rows), "h2d");
  float* d_cent = device_copy(z.data(), (size_t)k * cols);
  int sum_class = -1;
  extract_shape<float>(d_idx, d, d_cnt, d_cent, k, rows, rows, cols, sum_class, 0);
  const std::vector<float> cent = host_copy(d_cent, (size_t)k * cols);
  const std::vector<int> cnt = host_copy(d_cnt, (size_t)k);
  dump("extract_centroids", cent.data(), cent.size());
  dump("extract_counts", cnt.data(), cnt.size());
  dump("extract_sum_class", &sum_class, 1);

  // k_shape and k_shape_tsl with the CLI's caps (test_kshape.cu:21, 52, 47)
  std::vector<int> idx(rows);
  int iter_num = 10;
  sum_class = -1;
  k_shape<float>(d, idx.data(), rows, cols, k, iter_num, sum_class, 0);
  dump("kshape_labels", idx.data(), idx.size());
  int info[2] = {iter_num, sum_class};
  dump("kshape_info", info, 2);
  iter_num = 10;
  sum_class = -1;
  k_shape_tsl<float>(d, h, idx.data(), rows, cols, k, iter_num, sum_class, 0);
  dump("kshape_tsl_labels", idx.data(), idx.size());
  info[0] = iter_num;
  info[1] = sum_class;
  dump("kshape_tsl_info", info, 2);

  // writeFile (readfile.cuh:120-129)
  writeFile<int>(g_out + "/labels.txt", idx.data(), rows);

  for (void* p : {(void*)d, (void*)d_ncc, (void*)d_mem, (void*)d_idx, (void*)d_cnt, (void*)d_cent})
    timesc_ref::check(tcx_free(c, p), "free");
  delete[] h;
  std::printf("ref_forwards done\n");
  return 0;
}
