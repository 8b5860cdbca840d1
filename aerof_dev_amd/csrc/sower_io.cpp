// sower_io.cpp -- AERO-F's binary mesh and solution files on the host.
//
// Every file starts with the BinFileHandler header (BinFileHandler.h:167-196):
// int 1 (a reader that sees another value swaps the bytes of everything it
// reads) and a double version; offsets inside the file (OffType, 8 bytes)
// count from the end of that header.
//
// Mesh: one cluster of a SOWER decomposition is a geometry file (.geo) and a
// decomposition file (.dec).  GeoSource::getSubDomain (GeoSource.C:246-356)
// reads one subdomain: its node, element and face ranges from the .dec
// (getRangeInfo, GeoSource.C:204-240: nr ranges then ns singles), then the
// nodes (NodeSet::read, Node.C:7-50: global number + 3 coordinates per
// cluster node, fixed stride), tets (ElemSet::read, ElemCore.C:135-200: a
// table of contents of int offsets, then type 5, global number, volume id,
// 4 cluster node numbers per element) and boundary triangles
// (FaceSet::read, FaceCore.C:270-347: type 4, global number, 3 nodes, code
// with 6 read as BC_SYMMETRY, surface id) from the .geo, and renumbers
// element and face nodes from cluster to subdomain numbering
// (GeoSource.C:360-379).
//
// Solutions: SubDomain::openFileForWriting / writeTagToFile /
// writeVectorToFile / readVectorFromFile (SubDomain.C:5609-5810): a header
// {cluster nodes, dim, solutions} and per solution a double tag followed by
// [cluster nodes][dim] values.
#include <sys/types.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "aerofx.h"
#include "host_internal.h"

namespace aerofx {
namespace {

using Off = int64_t;
constexpr int kTet = 5;           // Elem::TET (Elem.h:535)
constexpr int kTria = 4;          // Face::TRIA (Face.h:432)
constexpr int kSymmetry = 11;     // BC_SYMMETRY (BcDef.h:4)
constexpr Off kHeader = sizeof(int) + sizeof(double);

template <class T>
void swap_bytes(T *p, size_t n) {
  for (size_t k = 0; k < n; ++k) {
    unsigned char *b = reinterpret_cast<unsigned char *>(p + k);
    for (size_t c = 0; c < sizeof(T) / 2; ++c) std::swap(b[c], b[sizeof(T) - 1 - c]);
  }
}

// A BinFileHandler-format file: reads and seeks raise (as a C++ exception
// caught at the C boundary) on a short read or a failed seek.
class BinFile {
 public:
  struct Error {
    std::string msg;
  };
  // mode "r": read, detect byte order; "w": create/truncate, write header;
  // "u": update an existing file in place (header kept)
  BinFile(const char *path, char mode) : path_(path) {
    f_ = std::fopen(path, mode == 'r' ? "rb" : mode == 'w' ? "wb" : "r+b");
    if (!f_) throw Error{std::string("cannot open '") + path + "'"};
    if (mode == 'r') {
      int one = 0;
      raw_read(&one, 1);
      swap_ = one != 1;
      if (swap_) {
        swap_bytes(&one, 1);
        if (one != 1) throw Error{std::string("'") + path + "' is not a BinFileHandler file"};
      }
      raw_read(&version_, 1);
    } else if (mode == 'w') {
      int one = 1;
      raw_write(&one, 1);
      raw_write(&version_, 1);
    }
    seek(0);
  }
  ~BinFile() {
    if (f_) std::fclose(f_);
  }
  BinFile(const BinFile &) = delete;
  BinFile &operator=(const BinFile &) = delete;

  template <class T>
  void read(T *p, size_t n) {
    raw_read(p, n);
  }
  template <class T>
  T get() {
    T v;
    raw_read(&v, 1);
    return v;
  }
  template <class T>
  void write(const T *p, size_t n) {
    if (n && std::fwrite(p, sizeof(T), n, f_) != n) throw Error{"write failed on '" + path_ + "'"};
  }
  void seek(Off pos) {
    if (fseeko(f_, (off_t)(pos + kHeader), SEEK_SET) != 0) throw Error{"seek failed on '" + path_ + "'"};
  }
  Off tell() { return (Off)ftello(f_) - kHeader; }
  double version() const { return version_; }

 private:
  template <class T>
  void raw_read(T *p, size_t n) {
    if (n && std::fread(p, sizeof(T), n, f_) != n) throw Error{"short read on '" + path_ + "'"};
    if (swap_) swap_bytes(p, n);
  }
  template <class T>
  void raw_write(const T *p, size_t n) {
    if (std::fwrite(p, sizeof(T), n, f_) != n) throw Error{"write failed on '" + path_ + "'"};
  }
  std::FILE *f_ = nullptr;
  std::string path_;
  bool swap_ = false;
  double version_ = 0.0;
};

// getRangeInfo: the ranges first, then each single as a range of one
struct Ranges {
  std::vector<int> lo, hi;
  long long count() const {
    long long n = 0;
    for (size_t k = 0; k < lo.size(); ++k) n += (long long)hi[k] - lo[k] + 1;
    return n;
  }
};

Ranges read_ranges(BinFile &f) {
  const int nr = f.get<int>(), ns = f.get<int>();
  if (nr < 0 || ns < 0) throw BinFile::Error{"negative range count in the decomposition file"};
  Ranges r;
  std::vector<int> buf(2 * (size_t)nr), singles((size_t)ns);
  f.read(buf.data(), buf.size());
  f.read(singles.data(), singles.size());
  for (int k = 0; k < nr; ++k) {
    if (buf[2 * k + 1] < buf[2 * k]) throw BinFile::Error{"empty range in the decomposition file"};
    r.lo.push_back(buf[2 * k]);
    r.hi.push_back(buf[2 * k + 1]);
  }
  for (int s : singles) {
    r.lo.push_back(s);
    r.hi.push_back(s);
  }
  return r;
}

struct Decomposition {
  int num_sub = 0;
  Ranges nodes, elems, faces;
  int num_neighb = 0;
};

Decomposition read_dec(const char *dec, int sub) {
  BinFile f(dec, 'r');
  Decomposition d;
  d.num_sub = f.get<int>();
  if (sub < 0 || sub >= d.num_sub) throw BinFile::Error{"subdomain index outside the cluster"};
  f.seek((Off)sizeof(int) + (Off)sub * (Off)sizeof(Off));
  f.seek(f.get<Off>());
  d.nodes = read_ranges(f);
  d.elems = read_ranges(f);
  d.faces = read_ranges(f);
  d.num_neighb = f.get<int>();
  return d;
}

struct GeoToc {
  Off nodes, elems, faces, bcs;
};

GeoToc read_toc(BinFile &g) {
  Off toc[3];
  g.read(toc, 3);
  return GeoToc{toc[0], toc[1], toc[2], g.get<Off>()};
}

}  // namespace
}  // namespace aerofx

using namespace aerofx;

#define SOWER_TRY try {
#define SOWER_CATCH(who)                                                            \
  }                                                                                 \
  catch (const BinFile::Error &e) {                                                 \
    return set_error(AEROFX_ERR_ARG, std::string(who) + ": " + e.msg);              \
  }                                                                                 \
  catch (const std::bad_alloc &) {                                                  \
    return set_error(AEROFX_ERR_NOMEM, std::string(who) + ": out of host memory");  \
  }

extern "C" int aerofx_sower_info(const char *geo, const char *dec, int sub, aerofx_sower_header *info) {
  if (!geo || !dec || !info) return set_error(AEROFX_ERR_ARG, "sower_info: null argument");
  SOWER_TRY
  const Decomposition d = read_dec(dec, sub);
  BinFile g(geo, 'r');
  const GeoToc toc = read_toc(g);
  g.seek(toc.nodes);
  info->num_clus_nodes = g.get<int>();
  g.seek(toc.bcs);
  info->num_bcs = g.get<int>();
  const long long n = d.nodes.count(), t = d.elems.count(), fc = d.faces.count();
  if (n > INT32_MAX || t > INT32_MAX || fc > INT32_MAX) throw BinFile::Error{"subdomain too large"};
  info->num_nodes = (int)n;
  info->num_tets = (int)t;
  info->num_faces = (int)fc;
  info->num_sub_cluster = d.num_sub;
  info->num_neighbours = d.num_neighb;
  info->version = g.version();
  return AEROFX_OK;
  SOWER_CATCH("sower_info")
}

extern "C" int aerofx_sower_read(const char *geo, const char *dec, int sub, double *X, int *tets, int *faces,
                                 int *codes, int *node_glob, int *node_clus) {
  if (!geo || !dec || !X || !tets || !faces || !codes)
    return set_error(AEROFX_ERR_ARG, "sower_read: null argument");
  SOWER_TRY
  const Decomposition d = read_dec(dec, sub);
  BinFile g(geo, 'r');
  const GeoToc toc = read_toc(g);
  // nodes: fixed records (global number, x, y, z) at start + clus * 28
  g.seek(toc.nodes);
  const int num_clus = g.get<int>();
  const Off nstart = g.get<Off>();
  std::map<int, int> clus_to_loc;
  int count = 0;
  for (size_t r = 0; r < d.nodes.lo.size(); ++r) {
    if (d.nodes.lo[r] < 0 || d.nodes.hi[r] >= num_clus) throw BinFile::Error{"node range outside the cluster"};
    g.seek(nstart + (Off)d.nodes.lo[r] * (Off)(sizeof(int) + 3 * sizeof(double)));
    for (int c = d.nodes.lo[r]; c <= d.nodes.hi[r]; ++c, ++count) {
      const int glob = g.get<int>();
      g.read(X + 3 * (size_t)count, 3);
      if (node_glob) node_glob[count] = glob;
      if (node_clus) node_clus[count] = c;
      clus_to_loc[c] = count;
    }
  }
  auto local = [&](int c) {
    auto it = clus_to_loc.find(c);
    if (it == clus_to_loc.end()) throw BinFile::Error{"element or face node outside the subdomain"};
    return it->second;
  };
  // tets: a table of contents (one int offset per cluster element) after the
  // start offset; each range is read sequentially from its first element
  g.seek(toc.elems);
  (void)g.get<int>();
  const Off estart = g.get<Off>(), etoc = g.tell();
  count = 0;
  for (size_t r = 0; r < d.elems.lo.size(); ++r) {
    g.seek(etoc + (Off)sizeof(int) * d.elems.lo[r]);
    g.seek(estart + (Off)sizeof(int) * g.get<int>());
    for (int e = d.elems.lo[r]; e <= d.elems.hi[r]; ++e, ++count) {
      int rec[3];  // type, global number, volume id
      g.read(rec, 3);
      if (rec[0] != kTet) throw BinFile::Error{"element type " + std::to_string(rec[0]) + " (only tets, type 5)"};
      int nodes[4];
      g.read(nodes, 4);
      for (int k = 0; k < 4; ++k) tets[4 * (size_t)count + k] = local(nodes[k]);
    }
  }
  // boundary triangles, same table-of-contents scheme
  g.seek(toc.faces);
  (void)g.get<int>();
  const Off fstart = g.get<Off>(), ftoc = g.tell();
  count = 0;
  for (size_t r = 0; r < d.faces.lo.size(); ++r) {
    g.seek(ftoc + (Off)sizeof(int) * d.faces.lo[r]);
    g.seek(fstart + (Off)sizeof(int) * g.get<int>());
    for (int q = d.faces.lo[r]; q <= d.faces.hi[r]; ++q, ++count) {
      int rec[2];  // type, global number
      g.read(rec, 2);
      if (rec[0] != kTria) throw BinFile::Error{"face type " + std::to_string(rec[0]) + " (only triangles, type 4)"};
      int nodes[3];
      g.read(nodes, 3);
      for (int k = 0; k < 3; ++k) faces[3 * (size_t)count + k] = local(nodes[k]);
      int code = g.get<int>();
      (void)g.get<int>();  // surface id
      codes[count] = code == 6 ? kSymmetry : code;
    }
  }
  return AEROFX_OK;
  SOWER_CATCH("sower_read")
}

// One cluster holding one subdomain (the whole mesh): what a one-GPU run
// reads back with sub = 0.  Global numbers are the indices; volume and
// surface ids 0; no boundary-condition records, no neighbours.
extern "C" int aerofx_sower_write(const char *geo, const char *dec, int N, int T, int F, const double *X,
                                  const int *tets, const int *faces, const int *codes) {
  if (!geo || !dec || !X || !tets || (F > 0 && (!faces || !codes)))
    return set_error(AEROFX_ERR_ARG, "sower_write: null argument");
  if (N < 1 || T < 1 || F < 0) return set_error(AEROFX_ERR_ARG, "sower_write: N, T >= 1, F >= 0");
  for (size_t k = 0; k < 4 * (size_t)T; ++k)
    if (tets[k] < 0 || tets[k] >= N) return set_error(AEROFX_ERR_ARG, "sower_write: tet node out of range");
  for (size_t k = 0; k < 3 * (size_t)F; ++k)
    if (faces[k] < 0 || faces[k] >= N) return set_error(AEROFX_ERR_ARG, "sower_write: face node out of range");
  SOWER_TRY
  {
    BinFile d(dec, 'w');
    const int one_sub = 1;
    d.write(&one_sub, 1);
    const Off at = (Off)sizeof(int) + (Off)sizeof(Off);
    d.write(&at, 1);
    const int ranges[3][4] = {{1, 0, 0, N - 1}, {1, 0, 0, T - 1}, {F > 0 ? 1 : 0, 0, 0, F - 1}};
    for (int k = 0; k < 3; ++k) d.write(ranges[k], ranges[k][0] ? 4 : 2);
    const int no_neighb = 0;
    d.write(&no_neighb, 1);
  }
  BinFile g(geo, 'w');
  const Off toc_size = 4 * (Off)sizeof(Off);
  const Off nodes_at = toc_size;
  const Off nodes_start = nodes_at + (Off)sizeof(int) + (Off)sizeof(Off);
  const Off elems_at = nodes_start + (Off)N * (Off)(sizeof(int) + 3 * sizeof(double));
  const Off elems_start = elems_at + (Off)sizeof(int) + (Off)sizeof(Off) + (Off)T * (Off)sizeof(int);
  const Off faces_at = elems_start + (Off)T * 7 * (Off)sizeof(int);
  const Off faces_start = faces_at + (Off)sizeof(int) + (Off)sizeof(Off) + (Off)F * (Off)sizeof(int);
  const Off bcs_at = faces_start + (Off)F * 7 * (Off)sizeof(int);
  const Off toc[4] = {nodes_at, elems_at, faces_at, bcs_at};
  g.write(toc, 4);
  g.write(&N, 1);
  g.write(&nodes_start, 1);
  for (int i = 0; i < N; ++i) {
    g.write(&i, 1);
    g.write(X + 3 * (size_t)i, 3);
  }
  g.write(&T, 1);
  g.write(&elems_start, 1);
  for (int e = 0; e < T; ++e) {
    const int off = 7 * e;
    g.write(&off, 1);
  }
  for (int e = 0; e < T; ++e) {
    const int rec[3] = {kTet, e, 0};
    g.write(rec, 3);
    g.write(tets + 4 * (size_t)e, 4);
  }
  g.write(&F, 1);
  g.write(&faces_start, 1);
  for (int q = 0; q < F; ++q) {
    const int off = 7 * q;
    g.write(&off, 1);
  }
  for (int q = 0; q < F; ++q) {
    const int rec[2] = {kTria, q};
    g.write(rec, 2);
    g.write(faces + 3 * (size_t)q, 3);
    const int tail[2] = {codes[q] == kSymmetry ? 6 : codes[q], 0};
    g.write(tail, 2);
  }
  const int no_bcs = 0;
  g.write(&no_bcs, 1);
  return AEROFX_OK;
  SOWER_CATCH("sower_write")
}

extern "C" int aerofx_solution_info(const char *path, int *num_clus_nodes, int *dim, int *nsol) {
  if (!path || !num_clus_nodes || !dim || !nsol) return set_error(AEROFX_ERR_ARG, "solution_info: null argument");
  SOWER_TRY
  BinFile f(path, 'r');
  int info[3];
  f.read(info, 3);
  *num_clus_nodes = info[0];
  *dim = info[1];
  *nsol = info[2];
  return AEROFX_OK;
  SOWER_CATCH("solution_info")
}

extern "C" int aerofx_solution_write(const char *path, int step, double tag, int n, int dim, const double *U) {
  if (!path || !U) return set_error(AEROFX_ERR_ARG, "solution_write: null argument");
  if (step < 0 || n < 1 || dim < 1) return set_error(AEROFX_ERR_ARG, "solution_write: step >= 0, n, dim >= 1");
  SOWER_TRY
  // step 0 creates the file; later steps update it in place and set the
  // solution count to step + 1 (openFileForWriting)
  BinFile f(path, step == 0 ? 'w' : 'u');
  const int info[3] = {n, dim, step + 1};
  f.write(info, 3);
  const Off unit = (Off)dim * (Off)sizeof(double);
  f.seek(3 * (Off)sizeof(int) + (Off)step * ((Off)sizeof(double) + (Off)n * unit));
  f.write(&tag, 1);
  f.write(U, (size_t)n * dim);
  return AEROFX_OK;
  SOWER_CATCH("solution_write")
}

extern "C" int aerofx_solution_read(const char *path, int step, double *tag, int n, int dim, const int *node_clus,
                                    double *U) {
  if (!path || !U) return set_error(AEROFX_ERR_ARG, "solution_read: null argument");
  if (n < 1 || dim < 1) return set_error(AEROFX_ERR_ARG, "solution_read: n, dim >= 1");
  SOWER_TRY
  BinFile f(path, 'r');
  int info[3];
  f.read(info, 3);
  const int num_clus = info[0], neq = info[1], nsol = info[2];
  if (step < 0 || step >= nsol) throw BinFile::Error{"solution " + std::to_string(step) + " not in the file"};
  if (neq < 1) throw BinFile::Error{"bad solution width"};
  if (!node_clus && n != num_clus)
    throw BinFile::Error{"size mismatch (" + std::to_string(num_clus) + " cluster nodes in the file)"};
  const Off unit = (Off)neq * (Off)sizeof(double);
  const Off rec = 3 * (Off)sizeof(int) + (Off)step * ((Off)sizeof(double) + (Off)num_clus * unit);
  f.seek(rec);
  const double t = f.get<double>();
  if (tag) *tag = t;
  std::vector<double> row((size_t)neq);
  const int m = neq < dim ? neq : dim;
  // rows in file order where they are consecutive (one seek per run)
  int prev = -2;
  for (int i = 0; i < n; ++i) {
    const int c = node_clus ? node_clus[i] : i;
    if (c < 0 || c >= num_clus) throw BinFile::Error{"cluster node outside the file"};
    if (c != prev + 1) f.seek(rec + (Off)sizeof(double) + (Off)c * unit);
    f.read(row.data(), (size_t)neq);
    for (int k = 0; k < m; ++k) U[(size_t)i * dim + k] = row[k];
    prev = c;
  }
  return AEROFX_OK;
  SOWER_CATCH("solution_read")
}
