// Host-side file formats of the linear-tetmesh path (no device code): TetGen .node / .ele and MEDIT
// .mesh tetrahedral meshes, and sparse matrices in the MatrixMarket coordinate format that Eigen's
// saveMarket / loadMarket read and write (the reference's fixtures, data/matrices/*.txt, and its tests'
// loadMarket calls, tests/test_assemble.cpp:33, tests/test_constitutive_model.cpp:88).
//
// Every reader follows the two-call convention of the C ABI: called with NULL output arrays it only
// returns the sizes; called again with arrays of those sizes it fills them.  Arrays use the reference's
// Eigen layouts: V nV x 3 double column-major, E nT x 4 int32 column-major, sparse matrices as CSC.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bartels_hip.h"

namespace {

// TetGen body: after the header, up to `count` records of `idx v0 v1 ...`; lines shorter than 3
// characters and lines starting with '#' are skipped, reading stops after `count` records -- the
// rules of sim::loadTet (src/read_tetgen.cpp:32-99).  The leading index is read and discarded.
template <typename T>
bool tetgen_records(std::istream& in, int count, int width, int keep, std::vector<T>& out) {
  std::string line;
  std::getline(in, line);  // rest of the header line
  out.clear();
  out.reserve((size_t)count * keep);
  int cnt = 0;
  while (cnt < count && std::getline(in, line)) {
    if (line.size() < 3 || line[0] == '#') continue;
    std::stringstream ss(line);
    long long idx;
    ss >> idx;
    for (int i = 0; i < width; ++i) {
      T v{};
      ss >> v;
      if (ss.fail()) return false;
      if (i < keep) out.push_back(v);
    }
    ++cnt;
  }
  return cnt == count;
}

struct MeshData {
  std::vector<double> V;  // row-major nV x 3 while reading
  std::vector<int32_t> E; // row-major nT x 4
};

int copy_out(const MeshData& d, int* nV, int* nT, double* V, int32_t* E) {
  const int nv = (int)(d.V.size() / 3), nt = (int)(d.E.size() / 4);
  if (nV) *nV = nv;
  if (nT) *nT = nt;
  if (V)
    for (int i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) V[i + (int64_t)c * nv] = d.V[3ll * i + c];
  if (E)
    for (int i = 0; i < nt; ++i)
      for (int k = 0; k < 4; ++k) E[i + (int64_t)k * nt] = d.E[4ll * i + k];
  return BH_OK;
}

// lower-case copy of a MEDIT keyword
std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

}  // namespace

extern "C" {

int bh_read_tetgen(const char* node_file, const char* ele_file, int* nV, int* nT, double* V, int32_t* E) {
  if (!node_file || !ele_file) return BH_BAD_ARG;
  std::ifstream nodeIn(node_file), eleIn(ele_file);
  if (!nodeIn.good() || !eleIn.good()) {
    std::fprintf(stderr, "bartels_hip: cannot read %s / %s\n", node_file, ele_file);
    return BH_BAD_ARG;
  }
  MeshData d;
  int nv = 0, nt = 0, per = 0;
  nodeIn >> nv;
  if (nodeIn.fail() || nv < 0 || !tetgen_records(nodeIn, nv, 3, 3, d.V)) return BH_BAD_ARG;
  eleIn >> nt >> per;
  if (eleIn.fail() || nt < 0 || per < 4 || !tetgen_records(eleIn, nt, per, 4, d.E)) return BH_BAD_ARG;
  return copy_out(d, nV, nT, V, E);
}

int bh_read_medit(const char* path, int* nV, int* nT, double* V, int32_t* E) {
  if (!path) return BH_BAD_ARG;
  std::ifstream in(path);
  if (!in.good()) {
    std::fprintf(stderr, "bartels_hip: cannot read %s\n", path);
    return BH_BAD_ARG;
  }
  MeshData d;
  std::string tok;
  bool haveV = false, haveT = false;
  while (in >> tok) {
    const std::string k = lower(tok);
    if (k == "end") break;
    if (k == "vertices" || k == "tetrahedra" || k == "triangles" || k == "edges") {
      long long n = 0;
      in >> n;
      if (in.fail() || n < 0) return BH_BAD_ARG;
      const int cols = (k == "vertices") ? 3 : (k == "tetrahedra") ? 4 : (k == "triangles") ? 3 : 2;
      for (long long i = 0; i < n; ++i) {
        if (k == "vertices") {
          double x[3];
          int ref;
          in >> x[0] >> x[1] >> x[2] >> ref;
          d.V.insert(d.V.end(), x, x + 3);
        } else {
          long long idx[4] = {0, 0, 0, 0};
          long long ref;
          for (int c = 0; c < cols; ++c) in >> idx[c];
          in >> ref;
          if (k == "tetrahedra")
            for (int c = 0; c < 4; ++c) d.E.push_back((int32_t)(idx[c] - 1));  // MEDIT is 1-based
        }
        if (in.fail()) return BH_BAD_ARG;
      }
      haveV |= k == "vertices";
      haveT |= k == "tetrahedra";
    }
  }
  if (!haveV || !haveT) return BH_BAD_ARG;
  for (int32_t v : d.E)
    if (v < 0 || 3ll * v >= (long long)d.V.size()) return BH_BAD_ARG;
  return copy_out(d, nV, nT, V, E);
}

int bh_save_market(const char* path, int64_t rows, int64_t cols, const int32_t* colptr, const int32_t* rowind,
                   const double* vals) {
  if (!path || rows < 0 || cols < 0 || !colptr || (colptr[cols] > 0 && (!rowind || !vals))) return BH_BAD_ARG;
  std::FILE* f = std::fopen(path, "w");
  if (!f) {
    std::fprintf(stderr, "bartels_hip: cannot write %s\n", path);
    return BH_BAD_ARG;
  }
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate  real general\n%lld %lld %lld\n", (long long)rows,
               (long long)cols, (long long)colptr[cols]);
  for (int64_t c = 0; c < cols; ++c)
    for (int32_t k = colptr[c]; k < colptr[c + 1]; ++k)
      std::fprintf(f, "%d %lld %.64e\n", rowind[k] + 1, (long long)(c + 1), vals[k]);
  const bool ok = std::ferror(f) == 0;
  return (std::fclose(f) == 0 && ok) ? BH_OK : BH_DEVICE_ERROR;
}

int bh_load_market(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int32_t* colptr, int32_t* rowind,
                   double* vals) {
  if (!path) return BH_BAD_ARG;
  std::ifstream in(path);
  if (!in.good()) {
    std::fprintf(stderr, "bartels_hip: cannot read %s\n", path);
    return BH_BAD_ARG;
  }
  std::string line;
  if (!std::getline(in, line) || line.rfind("%%MatrixMarket", 0) != 0) return BH_BAD_ARG;
  std::string banner = lower(line);
  if (banner.find("coordinate") == std::string::npos || banner.find("complex") != std::string::npos ||
      banner.find("pattern") != std::string::npos)
    return BH_UNSUPPORTED;
  const bool sym = banner.find("symmetric") != std::string::npos;
  while (std::getline(in, line))
    if (!line.empty() && line[0] != '%') break;
  long long m = 0, n = 0, ne = 0;
  {
    std::stringstream ss(line);
    ss >> m >> n >> ne;
    if (ss.fail() || m < 0 || n < 0 || ne < 0) return BH_BAD_ARG;
  }
  struct Trip {
    int32_t r, c;
    double v;
  };
  std::vector<Trip> t;
  t.reserve((size_t)(sym ? 2 : 1) * ne);
  for (long long i = 0; i < ne; ++i) {
    long long r, c;
    double v;
    if (!(in >> r >> c >> v) || r < 1 || c < 1 || r > m || c > n) return BH_BAD_ARG;
    t.push_back({(int32_t)(r - 1), (int32_t)(c - 1), v});
    if (sym && r != c) t.push_back({(int32_t)(c - 1), (int32_t)(r - 1), v});
  }
  // setFromTriplets semantics: sorted rows within each column, duplicates summed in file order
  std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) { return a.c != b.c ? a.c < b.c : a.r < b.r; });
  std::vector<int32_t> cp(n + 1, 0), ri;
  std::vector<double> vv;
  for (size_t k = 0; k < t.size();) {
    size_t k2 = k + 1;
    double s = t[k].v;
    while (k2 < t.size() && t[k2].c == t[k].c && t[k2].r == t[k].r) s += t[k2++].v;
    ri.push_back(t[k].r);
    vv.push_back(s);
    cp[t[k].c + 1]++;
    k = k2;
  }
  for (long long c = 0; c < n; ++c) cp[c + 1] += cp[c];
  if (rows) *rows = m;
  if (cols) *cols = n;
  if (nnz) *nnz = (int64_t)ri.size();
  if (colptr) std::copy(cp.begin(), cp.end(), colptr);
  if (rowind) std::copy(ri.begin(), ri.end(), rowind);
  if (vals) std::copy(vv.begin(), vv.end(), vals);
  return BH_OK;
}

}  // extern "C"
