// MatrixMarket coordinate files and plain vector files for the lmp classes
// (lMatrix::ReadFile / WriteFile, lmatrix.h:344-348, implemented for CSR
// through COO in lmatrix_csr_cpu.cc:1781-1810 and lmatrix_coo_cpu.cc:
// 123-279; lVector::ReadFile / WriteFile, lvector_cpu.cc:496-565).
//
// Reading follows the reference: "%%MatrixMarket matrix coordinate" with
// field real or integer, 1-based "row col value" lines, and for symmetry
// "symmetric" every off-diagonal entry (i, j) followed by its mirror (j, i)
// in the entry order.  Other fields (pattern, complex) and the dense array
// format are refused, as the reference refuses them.  Two differences, on
// purpose: "skew-symmetric" adds the mirror with the opposite sign (the
// reference keeps only the stored triangle), and a file with fewer entries
// than its size line announces is an error (the reference stops reading
// and keeps uninitialised entries).
//
// Writing produces "coordinate real general" with "row col value" lines
// in CSR order and values at 17 significant digits, so that reading the
// file back gives the same matrix bit for bit.  (The reference writes
// "col row" with 14 digits, which transposes the matrix on a round trip.)
#pragma once

#include <string>
#include <vector>

namespace hiflow {
namespace la {

struct MmCoo {
  int nrow = 0, ncol = 0;
  std::vector<int> row, col;  // 0-based, in file order (mirrors after their entry)
  std::vector<double> val;
};

// Returns an empty string on success, else the reason.
std::string mm_read_coo(const std::string &path, MmCoo &out);
std::string mm_write_csr(const std::string &path, int nrow, int ncol, const int *row_ptr,
                         const int *col, const double *val);
// One value per line (lVector::ReadFile reads values until the end).
std::string vec_read(const std::string &path, std::vector<double> &out);
std::string vec_write(const std::string &path, const double *v, int n);

}  // namespace la
}  // namespace hiflow
