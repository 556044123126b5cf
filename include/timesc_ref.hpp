/*
 * timesc_ref.hpp -- the reference's C++ entry points (float instantiations)
 * forwarded to the timesc C ABI, for callers written against Times-C's
 * headers.  Each forward keeps the reference signature; on a non-zero
 * status it prints the error and calls exit(-1), as the reference's CHECK
 * does (Code/include/check.cuh:7-19).  double instantiations are not
 * provided (the reference CLI only uses float, test_kshape.cu:9).
 *
 * The forwards share one process-wide context on the device chosen with
 * timesc_ref::set_device (default 0), the counterpart of cudaSetDevice in
 * test_kshape.cu:178.
 */
#ifndef TIMESC_REF_HPP
#define TIMESC_REF_HPP

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "timesc.h"

namespace timesc_ref {

inline int& device_slot() {
  static int dev = 0;
  return dev;
}
inline void set_device(int dev) { device_slot() = dev; }

inline void check(int rc, const char* what) {
  if (rc != TCX_OK) {
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, rc, tcx_last_error());
    std::exit(-1);
  }
}

inline tcx_ctx* context() {
  static tcx_ctx* ctx = nullptr;
  if (!ctx) check(tcx_create(device_slot(), 0, &ctx), "tcx_create");
  return ctx;
}

}  // namespace timesc_ref

/* z_norm_gpu_x<float>: Code/include/normalization.cuh:149-177 */
template <typename T>
void z_norm_gpu_x(T* d_Mat, const int MAT_ROWS, const int MAT_COLS, int ddof);
template <>
inline void z_norm_gpu_x<float>(float* d_Mat, const int MAT_ROWS, const int MAT_COLS, int ddof) {
  timesc_ref::check(tcx_znorm(timesc_ref::context(), d_Mat, MAT_ROWS, MAT_COLS, ddof),
                    "z_norm_gpu_x");
}

/* NCC_3D<float>: Code/include/ncc.cuh:467-517 (linear, non-aliased lags) */
template <typename T>
void NCC_3D(const T* d_mat_x, const T* d_mat_y, T* d_mat_out, const int mat_x_row,
            const int mat_y_row, const int mat_col);
template <>
inline void NCC_3D<float>(const float* d_mat_x, const float* d_mat_y, float* d_mat_out,
                          const int mat_x_row, const int mat_y_row, const int mat_col) {
  timesc_ref::check(tcx_ncc(timesc_ref::context(), d_mat_x, d_mat_y, d_mat_out, mat_x_row,
                            mat_y_row, mat_col),
                    "NCC_3D");
}

/* sbd3D<float>: Code/include/SBD.cuh:68-85 -- aligns members to the centre
 * in place with the reference GPU path's sign (NCC_3D(center, member)). */
template <typename T>
void sbd3D(T* d_mat, T* d_center, T* d_mat_shift, const int mat_row, const int mat_col);
template <>
inline void sbd3D<float>(float* d_mat, float* d_center, float* d_mat_shift, const int mat_row,
                         const int mat_col) {
  (void)d_mat_shift;  // unused by the reference as well (SBD.cuh:80-81)
  timesc_ref::check(tcx_sbd_align(timesc_ref::context(), d_mat, d_center, mat_row, mat_col,
                                  TCX_CONV_KSHAPE),
                    "sbd3D");
}

/* extract_shape<float>: Code/include/shapeExtract.cuh:1057-1181.  d_idx
 * holds one-based labels like the reference GPU path (kshape.cuh:91);
 * d_k_count receives the cluster sizes. */
template <typename T>
void extract_shape(int* d_idx, const T* d_mat, int* d_k_count, T* d_centers, const int k,
                   const int idx_len, const int mat_row, const int mat_col, int& sum_class,
                   const int cal_select);
template <>
inline void extract_shape<float>(int* d_idx, const float* d_mat, int* d_k_count,
                                 float* d_centers, const int k, const int idx_len,
                                 const int mat_row, const int mat_col, int& sum_class,
                                 const int cal_select) {
  (void)mat_row;
  tcx_ctx* c = timesc_ref::context();
  std::vector<int> lab(idx_len), cnt(k, 0);
  timesc_ref::check(tcx_memcpy_d2h(c, lab.data(), d_idx, sizeof(int) * idx_len), "extract_shape");
  for (int& l : lab) {
    l -= 1;
    if (l >= 0 && l < k) ++cnt[l];
  }
  int* d_lab = nullptr;
  timesc_ref::check(tcx_malloc(c, (void**)&d_lab, sizeof(int) * idx_len), "extract_shape");
  timesc_ref::check(tcx_memcpy_h2d(c, d_lab, lab.data(), sizeof(int) * idx_len), "extract_shape");
  timesc_ref::check(tcx_extract_shape(c, d_lab, d_mat, nullptr, d_centers, idx_len, mat_col, k,
                                      0, cal_select, &sum_class),
                    "extract_shape");
  timesc_ref::check(tcx_free(c, d_lab), "extract_shape");
  if (d_k_count)
    timesc_ref::check(tcx_memcpy_h2d(c, d_k_count, cnt.data(), sizeof(int) * k), "extract_shape");
}

/* k_shape<float>: Code/include/kshape.cuh:197-324 (one-based labels) */
template <typename T>
void k_shape(const T* d_mat, int* idx, const int mat_row, const int mat_col, const int k,
             int& iter_num, int& sum_class, const int cal_select);
template <>
inline void k_shape<float>(const float* d_mat, int* idx, const int mat_row, const int mat_col,
                           const int k, int& iter_num, int& sum_class, const int cal_select) {
  timesc_ref::check(tcx_kshape(timesc_ref::context(), d_mat, idx, mat_row, mat_col, k, &iter_num,
                               &sum_class, cal_select),
                    "k_shape");
}

/* k_shape_tsl<float>: Code/include/kshape.cuh:332-474 (one-based labels) */
template <typename T>
void k_shape_tsl(const T* d_mat, const T* h_mat, int* idx, const int mat_row, const int mat_col,
                 const int k, int& iter_num, int& sum_class, const int cal_select);
template <>
inline void k_shape_tsl<float>(const float* d_mat, const float* h_mat, int* idx,
                               const int mat_row, const int mat_col, const int k, int& iter_num,
                               int& sum_class, const int cal_select) {
  timesc_ref::check(tcx_kshape_tsl(timesc_ref::context(), d_mat, h_mat, idx, mat_row, mat_col, k,
                                   &iter_num, &sum_class, cal_select),
                    "k_shape_tsl");
}

/* readFile<float>: Code/include/readfile.cuh:12-67 (returns new[]) */
template <typename T>
T* readFile(std::string filePath, std::vector<int>& label, int& mat_rows, int& mat_cols);
template <>
inline float* readFile<float>(std::string filePath, std::vector<int>& label, int& mat_rows,
                              int& mat_cols) {
  float* data = nullptr;
  int* lab = nullptr;
  timesc_ref::check(tcx_read_labeled_csv(filePath.c_str(), &data, &lab, &mat_rows, &mat_cols),
                    "readFile");
  float* out = new float[(size_t)mat_rows * mat_cols];
  std::memcpy(out, data, sizeof(float) * (size_t)mat_rows * mat_cols);
  label.assign(lab, lab + mat_rows);
  tcx_free_host(data);
  tcx_free_host(lab);
  return out;
}

/* read_unlabeled_File<float>: Code/include/readfile.cuh:72-116 */
template <typename T>
T* read_unlabeled_File(std::string filePath, int& mat_rows, int& mat_cols);
template <>
inline float* read_unlabeled_File<float>(std::string filePath, int& mat_rows, int& mat_cols) {
  float* data = nullptr;
  timesc_ref::check(tcx_read_unlabeled_csv(filePath.c_str(), &data, &mat_rows, &mat_cols),
                    "read_unlabeled_File");
  float* out = new float[(size_t)mat_rows * mat_cols];
  std::memcpy(out, data, sizeof(float) * (size_t)mat_rows * mat_cols);
  tcx_free_host(data);
  return out;
}

/* writeFile<int>: Code/include/readfile.cuh:120-129 */
template <typename T>
void writeFile(std::string filePath, T* data, const int dataSize);
template <>
inline void writeFile<int>(std::string filePath, int* data, const int dataSize) {
  timesc_ref::check(tcx_write_labels(filePath.c_str(), data, dataSize), "writeFile");
}

#endif
