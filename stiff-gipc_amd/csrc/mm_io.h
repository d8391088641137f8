// mm_io.h -- MatrixMarket I/O of block systems and the upper-BCOO to full-BSR
// expansion used at the linear-system boundary (host code, mm_io.hip).
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace sgipc
{
// A (full BSR, both triangles; only blocks with col >= row are read, each
// written with its transpose) and b to the given paths
void mm_write_system(const std::string& A_path, const std::string& b_path, int nb,
                     const std::vector<int>& rowptr, const std::vector<int>& cols,
                     const std::vector<double>& vals, const std::vector<double>& b);
// <dir>/linear_system/frame_<frame:05d>_ns_<newton_step>_A.mtx (directories
// created); *b_path receives the _b.mtx path
std::string mm_system_paths(const std::string& dir, int frame, int newton_step,
                            std::string* b_path);
// unique upper 3x3 blocks (row <= col, (row, col) order, row-major values)
// of a coordinate real general/symmetric file; *nb = block rows
void mm_read_upper_blocks(const std::string& path, int* nb, std::vector<int>& br,
                          std::vector<int>& bc, std::vector<double>& bv);
// array real M x 1
std::vector<double> mm_read_vector(const std::string& path);
// upper BCOO (duplicates summed) -> full BSR, columns sorted, every diagonal
// block required
void upper_bcoo_to_full_bsr(int nb, int nnz, const int* rows, const int* cols, const double* vals,
                            std::vector<int>& rowptr, std::vector<int>& out_cols,
                            std::vector<double>& out_vals);
}  // namespace sgipc
