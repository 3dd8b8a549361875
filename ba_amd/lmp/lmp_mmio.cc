// MatrixMarket coordinate files and plain vector files (see lmp_mmio.h).
#include "lmp_mmio.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace hiflow {
namespace la {

namespace {

struct File {
  FILE *f;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// next line that is neither blank nor a comment; false at the end
bool data_line(FILE *f, std::string &line) {
  static thread_local std::vector<char> buf(1 << 16);
  while (fgets(buf.data(), (int)buf.size(), f)) {
    line.assign(buf.data());
    while (!line.empty() && line.back() != '\n' && !feof(f)) {  // long line
      if (!fgets(buf.data(), (int)buf.size(), f)) break;
      line.append(buf.data());
    }
    size_t k = 0;
    while (k < line.size() && std::isspace((unsigned char)line[k])) ++k;
    if (k == line.size() || line[k] == '%') continue;
    return true;
  }
  return false;
}

}  // namespace

std::string mm_read_coo(const std::string &path, MmCoo &out) {
  File fp(fopen(path.c_str(), "r"));
  if (!fp.f) return "cannot open " + path;
  char head[1024];
  if (!fgets(head, sizeof(head), fp.f)) return "empty file";
  char t0[64] = {0}, t1[64] = {0}, t2[64] = {0}, t3[64] = {0}, t4[64] = {0};
  if (sscanf(head, "%63s %63s %63s %63s %63s", t0, t1, t2, t3, t4) != 5 ||
      lower(t0) != "%%matrixmarket")
    return "no MatrixMarket banner";
  const std::string object = lower(t1), format = lower(t2), field = lower(t3),
                    symmetry = lower(t4);
  if (object != "matrix") return "object '" + object + "' is not a matrix";
  if (format != "coordinate") return "only the sparse coordinate format is read";
  if (field != "real" && field != "integer" && field != "double")
    return "field '" + field + "' is not supported (real or integer)";
  const bool sym = symmetry == "symmetric", skew = symmetry == "skew-symmetric";
  if (!sym && !skew && symmetry != "general")
    return "symmetry '" + symmetry + "' is not supported";
  std::string line;
  if (!data_line(fp.f, line)) return "missing size line";
  long long m = -1, n = -1, nz = -1;
  if (sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &nz) != 3 || m < 0 || n < 0 || nz < 0 ||
      m > 0x7fffffffLL || n > 0x7fffffffLL)
    return "bad size line";
  if ((sym || skew) && m != n) return "a symmetric matrix must be square";
  out = MmCoo();
  out.nrow = (int)m;
  out.ncol = (int)n;
  const long long cap = (sym || skew) ? 2 * nz : nz;
  if (cap > 0x7fffffffLL) return "more than 2^31-1 entries";
  out.row.reserve((size_t)cap);
  out.col.reserve((size_t)cap);
  out.val.reserve((size_t)cap);
  for (long long k = 0; k < nz; ++k) {
    if (!data_line(fp.f, line))
      return "file ends after " + std::to_string(k) + " of " + std::to_string(nz) + " entries";
    const char *p = line.c_str();
    char *e = nullptr;
    errno = 0;
    const long long i = strtoll(p, &e, 10);
    if (e == p) return "bad entry line " + std::to_string(k + 1);
    p = e;
    const long long j = strtoll(p, &e, 10);
    if (e == p) return "bad entry line " + std::to_string(k + 1);
    p = e;
    const double v = strtod(p, &e);
    if (e == p) return "missing value on entry line " + std::to_string(k + 1);
    if (i < 1 || i > m || j < 1 || j > n)
      return "entry " + std::to_string(k + 1) + " out of range";
    out.row.push_back((int)(i - 1));
    out.col.push_back((int)(j - 1));
    out.val.push_back(v);
    if ((sym || skew) && i != j) {
      out.row.push_back((int)(j - 1));
      out.col.push_back((int)(i - 1));
      out.val.push_back(skew ? -v : v);
    }
  }
  return "";
}

std::string mm_write_csr(const std::string &path, int nrow, int ncol, const int *row_ptr,
                         const int *col, const double *val) {
  File fp(fopen(path.c_str(), "w"));
  if (!fp.f) return "cannot create " + path;
  const int nnz = row_ptr[nrow];
  if (fprintf(fp.f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", nrow, ncol,
              nnz) < 0)
    return "write failed";
  for (int i = 0; i < nrow; ++i)
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      if (fprintf(fp.f, "%d %d %.17g\n", i + 1, col[k] + 1, val[k]) < 0) return "write failed";
  return "";
}

std::string vec_read(const std::string &path, std::vector<double> &out) {
  File fp(fopen(path.c_str(), "r"));
  if (!fp.f) return "cannot open " + path;
  out.clear();
  double v;
  int r;
  while ((r = fscanf(fp.f, "%lf", &v)) == 1) out.push_back(v);
  if (r != EOF) return "bad value after " + std::to_string(out.size()) + " entries";
  return "";
}

std::string vec_write(const std::string &path, const double *v, int n) {
  File fp(fopen(path.c_str(), "w"));
  if (!fp.f) return "cannot create " + path;
  for (int i = 0; i < n; ++i)
    if (fprintf(fp.f, "%.17g\n", v[i]) < 0) return "write failed";
  return "";
}

}  // namespace la
}  // namespace hiflow
