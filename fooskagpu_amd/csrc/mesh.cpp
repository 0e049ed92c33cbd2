// Host mesh and case library: box generator (tetrahedra, or hexahedra,
// prisms, pyramids and tetrahedra mixed), Gambit neutral file I/O, face
// numbering and geometry in the reference solver's conventions, and initial
// conditions.  C ABI in include/fooska_mesh.h.
//
// Reference conventions followed here (one rank, so no METIS/halo maps):
//  * element types by node count, 4 tet, 5 pyramid, 6 prism, 8 hex, with the
//    face -> element-node tables of meshProperties.cu:4-49 (fcNd2ElNd,
//    h_nFaceNodes); element nodes are stored 8 per cell as elNd2GlbNd
//  * face centres summed in translator.cu's node order per element type and
//    divided by the face's node count, cell centre = mean of the element's
//    nodes (translator.cu:122-293)
//  * unit normal from the first three fcNd2ElNd nodes of the face
//    (meshProperties.cu:51-63), tangent t1 along the first edge and
//    t2 = n x t1 (:65-82), area as the fan of triangles from the first node
//    (:84-100), volume as the sum of the sub-tets of the element that share
//    its node 0 (:102-137, tables :168-171)
//  * faces are numbered by walking cells in order and their local faces,
//    creating a face when the neighbour is a boundary or not yet visited, so
//    the owner is always the lower cell index (buildLocalMesh.cu:244-343)
//  * BC tags 45/26/13/42/51 -> codes 1..5, element face numbers of the
//    boundary sets = local face + 1 (translator.cu:295-349)
// Neighbours are matched by exact sorted node ids instead of the reference's
// face-centre tolerance search (meshConnectivity.cu:97-128).

#include "../../include/fooska_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

constexpr int kNd = 8;  // node slots per cell (elNd2GlbNd stride)
constexpr int kFc = 6;  // face slots per cell

// One element type of the reference.
struct ElemType {
  int nnode, nface;
  int fnn[kFc];      // nodes per face (h_nFaceNodes, meshProperties.cu:3-12)
  int fn[kFc][4];    // fcNd2ElNd (meshProperties.cu:15-49): normal, area, tangent order
  int fc[kFc][4];    // node order of the face-centre sums (translator.cu:122-293)
  int ntet;          // sub-tets of the volume, all with node 0 (meshProperties.cu:168-171)
  int vn[6][3];
};
const ElemType kTet = {4, 4, {3, 3, 3, 3, 0, 0},
                       {{1, 0, 2}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}},
                       {{1, 0, 2}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}},
                       1, {{1, 2, 3}}};
const ElemType kPyramid = {5, 5, {4, 3, 3, 3, 3, 0},
                           {{0, 2, 3, 1}, {0, 1, 4}, {1, 3, 4}, {3, 2, 4}, {2, 0, 4}},
                           {{0, 2, 3, 1}, {0, 1, 4}, {1, 3, 4}, {3, 2, 4}, {2, 0, 4}},
                           2, {{1, 3, 4}, {3, 2, 4}}};
const ElemType kPrism = {6, 5, {4, 4, 4, 3, 3, 0},
                         {{0, 1, 4, 3}, {2, 5, 4, 1}, {0, 3, 5, 2}, {0, 2, 1}, {3, 4, 5}},
                         {{0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}, {0, 2, 1}, {3, 4, 5}},
                         3, {{1, 2, 5}, {1, 5, 4}, {4, 5, 3}}};
const ElemType kHex = {8, 6, {4, 4, 4, 4, 4, 4},
                       {{0, 1, 5, 4}, {3, 7, 5, 1}, {2, 6, 7, 3}, {2, 0, 4, 6}, {2, 3, 1, 0}, {6, 4, 5, 7}},
                       {{0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}, {1, 0, 2, 3}, {4, 5, 7, 6}},
                       6, {{1, 3, 5}, {5, 3, 7}, {5, 7, 4}, {3, 2, 7}, {7, 2, 6}, {7, 6, 4}}};

inline const ElemType &etype(int nnode) {
  return nnode == 4 ? kTet : nnode == 5 ? kPyramid : nnode == 6 ? kPrism : kHex;
}
inline bool valid_nnode(long n) { return n == 4 || n == 5 || n == 6 || n == 8; }

}  // namespace

struct fkm_mesh {
  int64_t nc = 0, nf = 0, nn = 0, nbf = 0;
  std::vector<double> nodes;     // [nn][3]
  std::vector<int64_t> topo;     // node -> topological id (periodic merge); empty = identity
  std::vector<int32_t> el2nd;    // [nc][8] element nodes (elNd2GlbNd), -1 past the element's nodes
  std::vector<int32_t> nnode;    // [nc] 4 tet, 5 pyramid, 6 prism, 8 hex
  std::vector<int32_t> elbc;     // [nc][6] boundary code per local face (0 = none)
  std::vector<int32_t> nb;       // [nc][6] neighbour cell per local face, -1 boundary or no face
  std::vector<int8_t> nbface;    // [nc][6] local face index in the neighbour
  std::vector<double> elshift;   // [nc][6][3] periodic shift per local face (empty if none)
  // reference MESH arrays
  std::vector<int32_t> fc2el, fc2elfc, bc;
  std::vector<double> n, nt1, nt2, area, fc_center, shift, cell_center, volume;
  bool periodic = false;
};

namespace {

int64_t topo_of(const fkm_mesh &m, int64_t node) {
  return m.topo.empty() ? node : m.topo[(size_t)node];
}

inline int nface_of(const fkm_mesh &m, int64_t c) { return etype(m.nnode[(size_t)c]).nface; }

// Match faces by their sorted topological node ids (3 for triangles, 4 for
// quadrilaterals; a triangle never matches a quadrilateral).  Faces are
// bucketed by the smallest id (counting sort), so the whole match is
// O(faces).
int match_neighbours(fkm_mesh &m) {
  const int64_t nc = m.nc;
  int64_t ntopo = 0;
  for (int64_t i = 0; i < m.nn; ++i) ntopo = std::max(ntopo, topo_of(m, i) + 1);
  struct Key {
    int64_t b, c, d;
    int64_t cell;
    int8_t face;
  };
  std::vector<int64_t> start((size_t)ntopo + 1, 0);
  auto face_key = [&](int64_t c, int j, int64_t v[4]) {
    const ElemType &t = etype(m.nnode[(size_t)c]);
    const int k = t.fnn[j];
    for (int q = 0; q < k; ++q) v[q] = topo_of(m, m.el2nd[(size_t)(c * kNd + t.fn[j][q])]);
    std::sort(v, v + k);
    if (k == 3) v[3] = -1;
  };
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < nface_of(m, c); ++j) {
      int64_t v[4];
      face_key(c, j, v);
      start[(size_t)v[0] + 1]++;
    }
  for (int64_t i = 0; i < ntopo; ++i) start[(size_t)i + 1] += start[(size_t)i];
  std::vector<Key> keys((size_t)start[(size_t)ntopo]);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < nface_of(m, c); ++j) {
      int64_t v[4];
      face_key(c, j, v);
      keys[(size_t)fill[(size_t)v[0]]++] = Key{v[1], v[2], v[3], c, (int8_t)j};
    }
  m.nb.assign((size_t)(nc * kFc), -1);
  m.nbface.assign((size_t)(nc * kFc), -1);
  for (int64_t a = 0; a < ntopo; ++a) {
    Key *lo = keys.data() + start[(size_t)a], *hi = keys.data() + start[(size_t)a + 1];
    std::sort(lo, hi, [](const Key &x, const Key &y) {
      if (x.b != y.b) return x.b < y.b;
      if (x.c != y.c) return x.c < y.c;
      if (x.d != y.d) return x.d < y.d;
      return x.cell < y.cell;
    });
    for (Key *p = lo; p < hi;) {
      Key *q = p + 1;
      while (q < hi && q->b == p->b && q->c == p->c && q->d == p->d) ++q;
      if (q - p == 2) {
        m.nb[(size_t)(p->cell * kFc + p->face)] = (int32_t)p[1].cell;
        m.nbface[(size_t)(p->cell * kFc + p->face)] = p[1].face;
        m.nb[(size_t)(p[1].cell * kFc + p[1].face)] = (int32_t)p->cell;
        m.nbface[(size_t)(p[1].cell * kFc + p[1].face)] = p->face;
      } else if (q - p > 2) {
        return fail(FK_EINVAL, "non-manifold mesh: a face is shared by more than two cells");
      }
      p = q;
    }
  }
  return FK_OK;
}

// Reverse Cuthill-McKee on the cell adjacency graph.  Start cell: a
// pseudo-peripheral cell found by repeated BFS (George-Liu), neighbours
// queued by increasing degree (ties by index).  Returns perm[new] = old.
std::vector<int64_t> rcm_order(const fkm_mesh &m) {
  const int64_t nc = m.nc > 0 ? m.nc : 0;
  std::vector<int8_t> deg((size_t)nc, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < kFc; ++j) deg[(size_t)c] += (m.nb[(size_t)(c * kFc + j)] >= 0);
  std::vector<int64_t> level((size_t)nc, -1);
  std::vector<int64_t> queue;
  queue.reserve((size_t)nc);
  auto bfs_last = [&](int64_t s, int64_t &depth) {
    std::fill(level.begin(), level.end(), -1);
    queue.clear();
    queue.push_back(s);
    level[(size_t)s] = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
      int64_t c = queue[h];
      for (int j = 0; j < kFc; ++j) {
        int64_t d = m.nb[(size_t)(c * kFc + j)];
        if (d >= 0 && level[(size_t)d] < 0) {
          level[(size_t)d] = level[(size_t)c] + 1;
          queue.push_back(d);
        }
      }
    }
    depth = level[(size_t)queue.back()];
    int64_t best = queue.back();
    for (size_t h = queue.size(); h-- > 0;) {
      int64_t c = queue[h];
      if (level[(size_t)c] != depth) break;
      if (deg[(size_t)c] < deg[(size_t)best] || (deg[(size_t)c] == deg[(size_t)best] && c < best)) best = c;
    }
    return best;
  };
  std::vector<int64_t> perm;
  perm.reserve((size_t)nc);
  std::vector<int8_t> done((size_t)nc, 0);
  for (int64_t seed = 0; seed < nc; ++seed) {
    if (done[(size_t)seed]) continue;
    // pseudo-peripheral start within this component
    int64_t s = seed, depth = -1;
    for (int it = 0; it < 4; ++it) {
      int64_t d;
      int64_t t = bfs_last(s, d);
      if (d <= depth) break;
      depth = d;
      s = t;
    }
    size_t head = perm.size();
    perm.push_back(s);
    done[(size_t)s] = 1;
    for (size_t h = head; h < perm.size(); ++h) {
      int64_t c = perm[h];
      int64_t nbs[kFc];
      int k = 0;
      for (int j = 0; j < kFc; ++j) {
        int64_t d = m.nb[(size_t)(c * kFc + j)];
        if (d >= 0 && !done[(size_t)d]) nbs[k++] = d;
      }
      std::sort(nbs, nbs + k, [&](int64_t x, int64_t y) {
        if (deg[(size_t)x] != deg[(size_t)y]) return deg[(size_t)x] < deg[(size_t)y];
        return x < y;
      });
      for (int i = 0; i < k; ++i) {
        done[(size_t)nbs[i]] = 1;
        perm.push_back(nbs[i]);
      }
    }
  }
  std::reverse(perm.begin(), perm.end());
  return perm;
}

// Morton (Z-order) of the cell centroids quantised to 21 bits per axis over
// the bounding box; ties keep the previous order.  Returns perm[new] = old.
std::vector<int64_t> morton_order(const fkm_mesh &m) {
  const int64_t nc = m.nc;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < m.nn; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], m.nodes[(size_t)(i * 3 + a)]);
      hi[a] = std::max(hi[a], m.nodes[(size_t)(i * 3 + a)]);
    }
  auto spread = [](uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffffULL;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  std::vector<std::pair<uint64_t, int64_t>> key((size_t)nc);
  for (int64_t c = 0; c < nc; ++c) {
    uint64_t code = 0;
    for (int a = 0; a < 3; ++a) {
      const int nn = m.nnode[(size_t)c];
      double x = 0;
      for (int q = 0; q < nn; ++q) x += m.nodes[(size_t)(m.el2nd[(size_t)(c * kNd + q)] * 3 + a)];
      x *= 1.0 / nn;
      const double span = hi[a] > lo[a] ? hi[a] - lo[a] : 1.0;
      double f = (x - lo[a]) / span;
      f = std::min(std::max(f, 0.0), 1.0);
      code |= spread((uint64_t)(f * 2097151.0)) << a;
    }
    key[(size_t)c] = {code, c};
  }
  std::sort(key.begin(), key.end());
  std::vector<int64_t> perm((size_t)nc);
  for (int64_t i = 0; i < nc; ++i) perm[(size_t)i] = key[(size_t)i].second;
  return perm;
}

void apply_order(fkm_mesh &m, const std::vector<int64_t> &perm) {
  const int64_t nc = m.nc;
  std::vector<int64_t> inv((size_t)nc);
  for (int64_t i = 0; i < nc; ++i) inv[(size_t)perm[(size_t)i]] = i;
  std::vector<int32_t> el2nd((size_t)(nc * kNd)), nnode((size_t)nc), elbc((size_t)(nc * kFc)), nb((size_t)(nc * kFc));
  std::vector<int8_t> nbface((size_t)(nc * kFc));
  std::vector<double> elshift;
  if (!m.elshift.empty()) elshift.resize((size_t)(nc * kFc * 3));
  for (int64_t i = 0; i < nc; ++i) {
    int64_t o = perm[(size_t)i];
    nnode[(size_t)i] = m.nnode[(size_t)o];
    for (int q = 0; q < kNd; ++q) el2nd[(size_t)(i * kNd + q)] = m.el2nd[(size_t)(o * kNd + q)];
    for (int j = 0; j < kFc; ++j) {
      elbc[(size_t)(i * kFc + j)] = m.elbc[(size_t)(o * kFc + j)];
      int32_t d = m.nb[(size_t)(o * kFc + j)];
      nb[(size_t)(i * kFc + j)] = d >= 0 ? (int32_t)inv[(size_t)d] : -1;
      nbface[(size_t)(i * kFc + j)] = m.nbface[(size_t)(o * kFc + j)];
      if (!elshift.empty())
        for (int k = 0; k < 3; ++k)
          elshift[(size_t)((i * kFc + j) * 3 + k)] = m.elshift[(size_t)((o * kFc + j) * 3 + k)];
    }
  }
  m.el2nd.swap(el2nd);
  m.nnode.swap(nnode);
  m.elbc.swap(elbc);
  m.nb.swap(nb);
  m.nbface.swap(nbface);
  m.elshift.swap(elshift);
}

// Face numbering (buildLocalMesh.cu:244-343, single rank) and geometry
// (meshProperties.cu:139-180 + translator.cu face/cell centres).
int build_faces_and_geometry(fkm_mesh &m) {
  const int64_t nc = m.nc;
  int64_t nf = 0, nbf = 0;
  for (int64_t c = 0; c < nc; ++c)
    for (int j = 0; j < nface_of(m, c); ++j) {
      int32_t d = m.nb[(size_t)(c * kFc + j)];
      if (d < 0) {
        ++nf;
        ++nbf;
      } else if (d > c) {
        ++nf;
      }
    }
  if (nf > INT32_MAX) return fail(FK_ERANGE, "mesh too large for int32 face ids");
  m.nf = nf;
  m.nbf = nbf;
  try {
    m.fc2el.resize((size_t)(nf * 2));
    m.fc2elfc.resize((size_t)(nf * 2));
    m.bc.resize((size_t)nf);
    m.n.resize((size_t)(nf * 3));
    m.area.resize((size_t)nf);
    m.fc_center.resize((size_t)(nf * 3));
    m.shift.assign((size_t)(nf * 3), 0.0);
    m.cell_center.resize((size_t)(nc * 3));
    m.volume.resize((size_t)nc);
  } catch (const std::bad_alloc &) {
    return fail(FK_ENOMEM, "out of host memory building faces");
  }
  m.nt1.clear();
  m.nt2.clear();
  const double *X = m.nodes.data();
  int64_t f = 0;
  for (int64_t c = 0; c < nc; ++c) {
    const ElemType &t = etype(m.nnode[(size_t)c]);
    const int32_t *e = &m.el2nd[(size_t)(c * kNd)];
    for (int j = 0; j < t.nface; ++j) {
      int32_t d = m.nb[(size_t)(c * kFc + j)];
      if (!(d < 0 || d > c)) continue;
      m.fc2el[(size_t)(f * 2 + 0)] = (int32_t)c;
      m.fc2el[(size_t)(f * 2 + 1)] = d;
      m.fc2elfc[(size_t)(f * 2 + 0)] = j;
      m.fc2elfc[(size_t)(f * 2 + 1)] = d >= 0 ? m.nbface[(size_t)(c * kFc + j)] : -1;
      m.bc[(size_t)f] = d >= 0 ? 0 : m.elbc[(size_t)(c * kFc + j)];
      const int k = t.fnn[j];
      double x[4], y[4], z[4];
      for (int q = 0; q < k; ++q) {
        int64_t nd = e[t.fn[j][q]];
        x[q] = X[nd * 3 + 0];
        y[q] = X[nd * 3 + 1];
        z[q] = X[nd * 3 + 2];
      }
      // normalVector (meshProperties.cu:51-63), first three face nodes
      double c0 = (y[1] - y[0]) * (z[2] - z[0]) - (y[2] - y[0]) * (z[1] - z[0]);
      double c1 = (x[1] - x[0]) * (z[2] - z[0]) - (x[2] - x[0]) * (z[1] - z[0]);
      double c2 = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]);
      double norm = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
      m.n[(size_t)(f * 3 + 0)] = c0 / norm;
      m.n[(size_t)(f * 3 + 1)] = -c1 / norm;
      m.n[(size_t)(f * 3 + 2)] = c2 / norm;
      // computeArea (meshProperties.cu:84-100): triangles (0, 1+i, 2+i)
      double ar = 0.0;
      for (int it = 0; it < k - 2; ++it) {
        double dx1 = x[1 + it] - x[0], dx2 = x[2 + it] - x[0];
        double dy1 = y[1 + it] - y[0], dy2 = y[2 + it] - y[0];
        double dz1 = z[1 + it] - z[0], dz2 = z[2 + it] - z[0];
        double a0 = dy1 * dz2 - dz1 * dy2, a1 = dz1 * dx2 - dx1 * dz2, a2 = dx1 * dy2 - dy1 * dx2;
        ar += std::fabs(0.5 * std::sqrt(a0 * a0 + a1 * a1 + a2 * a2));
      }
      m.area[(size_t)f] = ar;
      // face centre in translator order
      for (int a = 0; a < 3; ++a) {
        double sum = X[(int64_t)e[t.fc[j][0]] * 3 + a];
        for (int q = 1; q < k; ++q) sum = sum + X[(int64_t)e[t.fc[j][q]] * 3 + a];
        m.fc_center[(size_t)(f * 3 + a)] = sum / k;
      }
      if (!m.elshift.empty() && d >= 0)
        for (int a = 0; a < 3; ++a) m.shift[(size_t)(f * 3 + a)] = m.elshift[(size_t)((c * kFc + j) * 3 + a)];
      ++f;
    }
    // cell centre (translator.cu, mean of the element nodes in order) and
    // volume (meshProperties.cu:102-137: sub-tets sharing node 0)
    for (int a = 0; a < 3; ++a) {
      double sum = X[(int64_t)e[0] * 3 + a];
      for (int q = 1; q < t.nnode; ++q) sum = sum + X[(int64_t)e[q] * 3 + a];
      m.cell_center[(size_t)(c * 3 + a)] = sum / t.nnode;
    }
    double vol = 0.0;
    for (int it = 0; it < t.ntet; ++it) {
      double dx[3], dy[3], dz[3];
      for (int in = 0; in < 3; ++in) {
        dx[in] = X[e[t.vn[it][in]] * 3 + 0] - X[e[0] * 3 + 0];
        dy[in] = X[e[t.vn[it][in]] * 3 + 1] - X[e[0] * 3 + 1];
        dz[in] = X[e[t.vn[it][in]] * 3 + 2] - X[e[0] * 3 + 2];
      }
      double cr0 = dy[0] * dz[1] - dz[0] * dy[1];
      double cr1 = dz[0] * dx[1] - dx[0] * dz[1];
      double cr2 = dx[0] * dy[1] - dy[0] * dx[1];
      vol += (cr0 * dx[2] + cr1 * dy[2] + cr2 * dz[2]) / 6.0;
    }
    m.volume[(size_t)c] = vol;
    if (!(vol > 0.0)) return fail(FK_EINVAL, "cell with non-positive volume (inverted element)");
  }
  return FK_OK;
}

// tangentVector (meshProperties.cu:65-82), computed on request only.
void build_tangents(fkm_mesh &m) {
  if (!m.nt1.empty()) return;
  m.nt1.resize((size_t)(m.nf * 3));
  m.nt2.resize((size_t)(m.nf * 3));
  const double *X = m.nodes.data();
  for (int64_t f = 0; f < m.nf; ++f) {
    int64_t c = m.fc2el[(size_t)(f * 2)];
    int j = m.fc2elfc[(size_t)(f * 2)];
    const ElemType &t = etype(m.nnode[(size_t)c]);
    double x[2], y[2], z[2];
    for (int k = 0; k < 2; ++k) {
      int64_t nd = m.el2nd[(size_t)(c * kNd + t.fn[j][k])];
      x[k] = X[nd * 3 + 0];
      y[k] = X[nd * 3 + 1];
      z[k] = X[nd * 3 + 2];
    }
    const double *nv = &m.n[(size_t)(f * 3)];
    double norm = std::sqrt((x[1] - x[0]) * (x[1] - x[0]) + (y[1] - y[0]) * (y[1] - y[0]) + (z[1] - z[0]) * (z[1] - z[0]));
    double t1[3] = {(x[1] - x[0]) / norm, (y[1] - y[0]) / norm, (z[1] - z[0]) / norm};
    double a = nv[1] * t1[2] - t1[1] * nv[2], b = nv[0] * t1[2] - t1[0] * nv[2], cc = nv[0] * t1[1] - t1[0] * nv[1];
    double norm1 = std::sqrt(a * a + b * b + cc * cc);
    for (int k = 0; k < 3; ++k) m.nt1[(size_t)(f * 3 + k)] = t1[k];
    m.nt2[(size_t)(f * 3 + 0)] = a / norm1;
    m.nt2[(size_t)(f * 3 + 1)] = -b / norm1;
    m.nt2[(size_t)(f * 3 + 2)] = cc / norm1;
  }
}

int finish_mesh(fkm_mesh &m, int order) {
  int rc = match_neighbours(m);
  if (rc) return rc;
  // boundary faces must carry a code
  for (int64_t c = 0; c < m.nc; ++c)
    for (int j = 0; j < nface_of(m, c); ++j)
      if (m.nb[(size_t)(c * kFc + j)] < 0 && m.elbc[(size_t)(c * kFc + j)] == 0)
        m.elbc[(size_t)(c * kFc + j)] = 2;  // untagged boundary faces: zero-gradient outlet
  if (order == 1) apply_order(m, rcm_order(m));
  else if (order == 2) apply_order(m, morton_order(m));
  return build_faces_and_geometry(m);
}

}  // namespace

extern "C" {

const char *fkm_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// Elements of one hex of the box (hex vertex v = a + 2b + 4c of the lattice
// offsets; 8 = the hex's centre node).
struct HexElems {
  int n = 0;
  int nnode[12];
  int v[12][8];
  void add(int nn, const int *vv) {
    nnode[n] = nn;
    for (int q = 0; q < nn; ++q) v[n][q] = vv[q];
    ++n;
  }
};

// Hex face s (0/1: x = 0/1, 2/3: y, 4/5: z): its four vertices as a pyramid
// base (b0, b1, b2, b3 with cyclic order b0 b1 b3 b2 and (b1 - b0) x (b2 - b0)
// pointing into the hex), and its two triangles along the diagonal of the
// 6-tet split (the diagonal through hex vertex 0 or 7).
const int kHexFaceBase[6][4] = {{0, 2, 4, 6}, {1, 5, 3, 7}, {0, 4, 1, 5}, {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 6, 5, 7}};
const int kHexFaceTri[6][2][3] = {{{0, 2, 6}, {0, 6, 4}}, {{1, 3, 7}, {1, 7, 5}}, {{0, 1, 5}, {0, 5, 4}},
                                  {{2, 3, 7}, {2, 7, 6}}, {{0, 1, 3}, {0, 3, 2}}, {{4, 5, 7}, {4, 7, 6}}};

// Face shape a hex of kind `kind` presents on its face s: 0 triangles (along
// the 6-tet diagonal), 1 quadrilateral, -1 adapts to its neighbour.
int face_shape(int kind, int s) {
  switch (kind) {
    case 0: return 0;
    case 1: return 1;
    case 2: return s >= 4 ? 0 : 1;
    default: return -1;
  }
}

// lattice offset (in half spacings) of hex vertex v (8 = centre) along axis a
inline int half_off(int v, int a) { return v == 8 ? 1 : 2 * ((v >> a) & 1); }

void hex_elems(int kind, const int shape[6], HexElems &h) {
  // 6 tets per hex sharing the diagonal 0-7
  static const int kTet6[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
  auto add_tet = [&](int a, int b, int c, int d) {
    int v[4] = {a, b, c, d};
    // orient positively on the undistorted lattice
    double e[3][3];
    for (int q = 0; q < 3; ++q)
      for (int x = 0; x < 3; ++x) e[q][x] = half_off(v[q + 1], x) - half_off(v[0], x);
    double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) - e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                 e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    if (det < 0) std::swap(v[1], v[2]);
    h.add(4, v);
  };
  if (kind == 0) {
    for (int t = 0; t < 6; ++t) add_tet(kTet6[t][0], kTet6[t][1], kTet6[t][2], kTet6[t][3]);
  } else if (kind == 1) {
    const int v[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    h.add(8, v);
  } else if (kind == 2) {
    // two prisms along z, split by the plane through vertices 0, 3, 4, 7
    const int p1[6] = {0, 1, 3, 4, 5, 7}, p2[6] = {0, 3, 2, 4, 7, 6};
    h.add(6, p1);
    h.add(6, p2);
  } else {
    // centre node 8: a pyramid on every quadrilateral face, two tets on every
    // triangulated one
    for (int sf = 0; sf < 6; ++sf) {
      if (shape[sf] == 1) {
        const int v[5] = {kHexFaceBase[sf][0], kHexFaceBase[sf][1], kHexFaceBase[sf][2], kHexFaceBase[sf][3], 8};
        h.add(5, v);
      } else {
        for (int t = 0; t < 2; ++t) add_tet(kHexFaceTri[sf][t][0], kHexFaceTri[sf][t][1], kHexFaceTri[sf][t][2], 8);
      }
    }
  }
}

}  // namespace

namespace {

int box_impl(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
             int periodic_mask, const int bc[6], int order, const int32_t *kinds, fkm_mesh **out) {
  if (!out || !lo || !hi || !bc) return fail(FK_EINVAL, "null argument");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(FK_EINVAL, "box needs at least one hex per axis");
  if (order < 0 || order > 2) return fail(FK_EINVAL, "order must be 0, 1 or 2");
  if (jitter < 0 || jitter > 0.2) return fail(FK_EINVAL, "jitter must lie in [0, 0.2]");
  const int N[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    if (!(hi[a] > lo[a])) return fail(FK_EINVAL, "box extent must be positive");
    if ((periodic_mask >> a & 1) && N[a] < 3) return fail(FK_EINVAL, "periodic axes need >= 3 hexes");
  }
  for (int s = 0; s < 6; ++s)
    if (!(periodic_mask >> (s / 2) & 1) && (bc[s] < 1 || bc[s] > 5))
      return fail(FK_EINVAL, "boundary codes must be 1..5");
  const int64_t nhex = (int64_t)nx * ny * nz;
  auto kind_of = [&](int64_t i, int64_t j, int64_t k) -> int {
    return kinds ? kinds[(size_t)(i + nx * (j + (int64_t)ny * k))] : 0;
  };
  // per hex and face: the shape it presents (conformity between neighbours)
  std::vector<int8_t> shape;
  int64_t nc = 0, ncentre = 0;
  if (kinds) {
    if (periodic_mask) return fail(FK_EINVAL, "mixed element kinds need a box without periodic axes");
    try {
      shape.assign((size_t)(nhex * 6), 1);
    } catch (const std::bad_alloc &) {
      return fail(FK_ENOMEM, "out of host memory");
    }
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
          const int kd = kind_of(i, j, k);
          if (kd < 0 || kd > 3) return fail(FK_EINVAL, "element kinds must be 0..3");
          const int64_t I[3] = {i, j, k};
          for (int sf = 0; sf < 6; ++sf) {
            const int a = sf / 2, dir = sf % 2 ? 1 : -1;
            int64_t J[3] = {I[0], I[1], I[2]};
            J[a] += dir;
            const bool inside = J[a] >= 0 && J[a] < N[a];
            int mine = face_shape(kd, sf);
            const int theirs = inside ? face_shape(kind_of(J[0], J[1], J[2]), sf ^ 1) : -1;
            if (mine < 0) mine = theirs < 0 ? 1 : theirs;
            else if (theirs >= 0 && theirs != mine)
              return fail(FK_EINVAL, "element kinds do not conform: a quadrilateral face meets triangles "
                                     "(put kind 3 between them)");
            shape[(size_t)((i + nx * (j + (int64_t)ny * k)) * 6 + sf)] = (int8_t)mine;
          }
          HexElems h;
          int sh[6];
          for (int sf = 0; sf < 6; ++sf) sh[sf] = shape[(size_t)((i + nx * (j + (int64_t)ny * k)) * 6 + sf)];
          hex_elems(kd, sh, h);
          nc += h.n;
          ncentre += kd == 3;
        }
  } else {
    nc = nhex * 6;
  }
  if (nc > INT32_MAX / kFc) return fail(FK_ERANGE, "mesh too large for int32 cell ids");
  fkm_mesh *m = new (std::nothrow) fkm_mesh;
  if (!m) return fail(FK_ENOMEM, "out of host memory");
  try {
    const int64_t px = nx + 1, py = ny + 1, pz = nz + 1;
    const int64_t nlat = px * py * pz;
    m->nn = nlat + ncentre;
    m->nc = nc;
    m->nodes.resize((size_t)(m->nn * 3));
    m->periodic = periodic_mask != 0;
    if (periodic_mask) m->topo.resize((size_t)m->nn);
    const int64_t tx = (periodic_mask & 1) ? nx : px, ty = (periodic_mask & 2) ? ny : py;
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = (hi[a] - lo[a]) / N[a];
    // jitter per topological node, drawn in topological order
    const int64_t ntz = (periodic_mask & 4) ? nz : pz;
    std::vector<double> jit;
    if (jitter > 0) {
      jit.resize((size_t)(tx * ty * ntz * 3));
      uint64_t s = seed;
      for (size_t i = 0; i < jit.size(); ++i) jit[i] = (2.0 * uniform01(s) - 1.0) * jitter;
    }
    for (int64_t k = 0; k < pz; ++k)
      for (int64_t j = 0; j < py; ++j)
        for (int64_t i = 0; i < px; ++i) {
          int64_t id = i + px * (j + py * k);
          int64_t I[3] = {i, j, k};
          int64_t T[3];
          bool interior = true;
          for (int a = 0; a < 3; ++a) {
            bool per = periodic_mask >> a & 1;
            T[a] = per ? I[a] % N[a] : I[a];
            if (!per && (I[a] == 0 || I[a] == N[a])) interior = false;
          }
          int64_t tid = T[0] + tx * (T[1] + ty * T[2]);
          if (periodic_mask) m->topo[(size_t)id] = tid;
          for (int a = 0; a < 3; ++a) {
            double x = (I[a] == N[a]) ? hi[a] : lo[a] + h[a] * (double)I[a];
            if (interior && !jit.empty()) x += jit[(size_t)(tid * 3 + a)] * h[a];
            m->nodes[(size_t)(id * 3 + a)] = x;
          }
        }
    m->el2nd.assign((size_t)(nc * kNd), -1);
    m->nnode.assign((size_t)nc, 0);
    m->elbc.assign((size_t)(nc * kFc), 0);
    if (periodic_mask) m->elshift.assign((size_t)(nc * kFc * 3), 0.0);
    int64_t c = 0, centre = nlat;
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
          const int64_t hx = i + nx * (j + (int64_t)ny * k);
          const int kd = kind_of(i, j, k);
          int sh[6] = {1, 1, 1, 1, 1, 1};
          if (kinds)
            for (int sf = 0; sf < 6; ++sf) sh[sf] = shape[(size_t)(hx * 6 + sf)];
          HexElems he;
          hex_elems(kd, sh, he);
          int64_t gid[9];
          for (int v = 0; v < 8; ++v) gid[v] = (i + (v & 1)) + px * ((j + (v >> 1 & 1)) + py * (k + (v >> 2 & 1)));
          if (kd == 3) {
            gid[8] = centre++;
            for (int a = 0; a < 3; ++a) {
              double sum = 0;
              for (int v = 0; v < 8; ++v) sum += m->nodes[(size_t)(gid[v] * 3 + a)];
              m->nodes[(size_t)(gid[8] * 3 + a)] = sum / 8;
            }
          }
          for (int e = 0; e < he.n; ++e, ++c) {
            const int nn = he.nnode[e];
            m->nnode[(size_t)c] = nn;
            for (int q = 0; q < nn; ++q) m->el2nd[(size_t)(c * kNd + q)] = (int32_t)gid[he.v[e][q]];
            // boundary code / periodic shift per local face from lattice positions
            const ElemType &t = etype(nn);
            for (int fj = 0; fj < t.nface; ++fj) {
              int64_t L[4][3];
              const int kq = t.fnn[fj];
              for (int q = 0; q < kq; ++q) {
                const int vv = he.v[e][t.fn[fj][q]];
                L[q][0] = 2 * i + half_off(vv, 0);
                L[q][1] = 2 * j + half_off(vv, 1);
                L[q][2] = 2 * k + half_off(vv, 2);
              }
              for (int a = 0; a < 3; ++a) {
                bool all0 = true, allN = true;
                for (int q = 0; q < kq; ++q) {
                  all0 = all0 && L[q][a] == 0;
                  allN = allN && L[q][a] == 2 * (int64_t)N[a];
                }
                if (periodic_mask >> a & 1) {
                  // face on the upper periodic plane: the neighbour sits at the
                  // lower end, its copy of the face is translated by -(hi-lo)
                  if (allN) m->elshift[(size_t)((c * kFc + fj) * 3 + a)] = hi[a] - lo[a];
                  if (all0) m->elshift[(size_t)((c * kFc + fj) * 3 + a)] = -(hi[a] - lo[a]);
                } else {
                  if (all0) m->elbc[(size_t)(c * kFc + fj)] = bc[2 * a];
                  if (allN) m->elbc[(size_t)(c * kFc + fj)] = bc[2 * a + 1];
                }
              }
            }
          }
        }
    int rc = finish_mesh(*m, order);
    if (rc) {
      delete m;
      return rc;
    }
  } catch (const std::bad_alloc &) {
    delete m;
    return fail(FK_ENOMEM, "out of host memory generating box mesh");
  }
  *out = m;
  return FK_OK;
}

}  // namespace

extern "C" {

int fkm_box(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
            int periodic_mask, const int bc[6], int order, fkm_mesh **out) {
  return box_impl(nx, ny, nz, lo, hi, jitter, seed, periodic_mask, bc, order, nullptr, out);
}

int fkm_box_kinds(int nx, int ny, int nz, const double lo[3], const double hi[3], double jitter, uint64_t seed,
                  const int bc[6], int order, const int32_t *kinds, fkm_mesh **out) {
  if (!kinds) return fail(FK_EINVAL, "null kinds");
  return box_impl(nx, ny, nz, lo, hi, jitter, seed, 0, bc, order, kinds, out);
}

int fkm_read_neu(const char *path, double scale, int order, fkm_mesh **out) {
  if (!out || !path) return fail(FK_EINVAL, "null argument");
  *out = nullptr;
  FILE *fp = std::fopen(path, "r");
  if (!fp) return fail(FK_EIO, std::string("mesh file not found: ") + path);
  fkm_mesh *m = new (std::nothrow) fkm_mesh;
  if (!m) {
    std::fclose(fp);
    return fail(FK_ENOMEM, "out of host memory");
  }
  char line[512];
  auto next_section = [&](const char *key) -> bool {
    while (std::fgets(line, sizeof line, fp)) {
      const char *p = line;
      while (*p == ' ' || *p == '\t') ++p;
      if (std::strncmp(p, key, std::strlen(key)) == 0) return true;
    }
    return false;
  };
  auto bail = [&](int code, const char *msg) {
    std::fclose(fp);
    delete m;
    return fail(code, msg);
  };
  long ndnum = 0, elnum = 0, ngrps = 0, bcnum = 0, ndfcd = 0, ndfvl = 0;
  if (!next_section("CONTROL INFO")) return bail(FK_EIO, "neu: no CONTROL INFO section");
  // skip to the line of sizes: first line with 6 integers after the NUMNP header
  if (!next_section("NUMNP")) return bail(FK_EIO, "neu: no NUMNP header");
  if (std::fscanf(fp, "%ld %ld %ld %ld %ld %ld", &ndnum, &elnum, &ngrps, &bcnum, &ndfcd, &ndfvl) != 6)
    return bail(FK_EIO, "neu: bad size line");
  if (ndnum <= 0 || elnum <= 0 || elnum > INT32_MAX / kFc) return bail(FK_EINVAL, "neu: bad node/element counts");
  if (!next_section("NODAL COORDINATES")) return bail(FK_EIO, "neu: no NODAL COORDINATES section");
  try {
    m->nn = ndnum;
    m->nc = elnum;
    m->nodes.resize((size_t)(ndnum * 3));
    m->el2nd.assign((size_t)(elnum * kNd), -1);
    m->nnode.assign((size_t)elnum, 0);
    m->elbc.assign((size_t)(elnum * kFc), 0);
  } catch (const std::bad_alloc &) {
    return bail(FK_ENOMEM, "neu: out of memory");
  }
  for (long i = 0; i < ndnum; ++i) {
    long id;
    double x, y, z;
    if (std::fscanf(fp, "%ld %lf %lf %lf", &id, &x, &y, &z) != 4 || id < 1 || id > ndnum)
      return bail(FK_EIO, "neu: bad node line");
    m->nodes[(size_t)((id - 1) * 3 + 0)] = x * scale;
    m->nodes[(size_t)((id - 1) * 3 + 1)] = y * scale;
    m->nodes[(size_t)((id - 1) * 3 + 2)] = z * scale;
  }
  if (!next_section("ELEMENTS/CELLS")) return bail(FK_EIO, "neu: no ELEMENTS/CELLS section");
  // element type by node count, as translator.cu:122-293: 8 hex, 6 prism,
  // 5 pyramid, 4 tet, nodes in the element's Gambit order
  for (long i = 0; i < elnum; ++i) {
    long el, typ, nnd, c[kNd];
    if (std::fscanf(fp, "%ld %ld %ld", &el, &typ, &nnd) != 3 || el < 1 || el > elnum)
      return bail(FK_EIO, "neu: bad element line");
    if (!valid_nnode(nnd)) return bail(FK_EINVAL, "neu: elements must have 4, 5, 6 or 8 nodes (linear tet, pyramid, prism, hex)");
    for (int q = 0; q < nnd; ++q)
      if (std::fscanf(fp, "%ld", &c[q]) != 1) return bail(FK_EIO, "neu: bad element nodes");
    m->nnode[(size_t)(el - 1)] = (int32_t)nnd;
    for (int q = 0; q < nnd; ++q) {
      if (c[q] < 1 || c[q] > ndnum) return bail(FK_EINVAL, "neu: node id out of range");
      m->el2nd[(size_t)((el - 1) * kNd + q)] = (int32_t)(c[q] - 1);
    }
  }
  for (long i = 0; i < elnum; ++i)
    if (m->nnode[(size_t)i] == 0) return bail(FK_EIO, "neu: element missing from ELEMENTS/CELLS");
  for (long b = 0; b < bcnum; ++b) {
    if (!next_section("BOUNDARY CONDITIONS")) return bail(FK_EIO, "neu: missing BOUNDARY CONDITIONS section");
    char name[256];
    long itype, nentry, nval, code;
    if (std::fscanf(fp, "%255s %ld %ld %ld %ld", name, &itype, &nentry, &nval, &code) != 5)
      return bail(FK_EIO, "neu: bad boundary header");
    int bcv = code == 45 ? 1 : code == 26 ? 2 : code == 13 ? 3 : code == 42 ? 4 : code == 51 ? 5 : 0;
    for (long i = 0; i < nentry; ++i) {
      long el, typ, face;
      if (std::fscanf(fp, "%ld %ld %ld", &el, &typ, &face) != 3) return bail(FK_EIO, "neu: bad boundary entry");
      if (el < 1 || el > elnum || face < 1 || face > etype(m->nnode[(size_t)(el - 1)]).nface)
        return bail(FK_EINVAL, "neu: boundary entry out of range");
      if (bcv) m->elbc[(size_t)((el - 1) * kFc + face - 1)] = bcv;
    }
  }
  std::fclose(fp);
  try {
    int rc = finish_mesh(*m, order);
    if (rc) {
      delete m;
      return rc;
    }
  } catch (const std::bad_alloc &) {
    delete m;
    return fail(FK_ENOMEM, "out of host memory building mesh");
  }
  *out = m;
  return FK_OK;
}

int fkm_write_neu(const fkm_mesh *m, const char *path) {
  if (!m || !path) return fail(FK_EINVAL, "null argument");
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(FK_EIO, std::string("cannot open for writing: ") + path);
  static const int kCode[6] = {0, 45, 26, 13, 42, 51};
  static const char *kName[6] = {"", "symmetry", "outlet", "inlet", "slip", "wall"};
  int64_t count[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t c = 0; c < m->nc; ++c)
    for (int j = 0; j < nface_of(*m, c); ++j)
      if (m->nb[(size_t)(c * kFc + j)] < 0) count[m->elbc[(size_t)(c * kFc + j)]]++;
  int nsets = 0;
  for (int b = 1; b <= 5; ++b) nsets += count[b] > 0;
  std::fprintf(fp, "CONTROL INFO         1.2.1\r\n** GAMBIT NEUTRAL FILE\r\nfooska\r\nPROGRAM:  fooska\r\n\r\n");
  std::fprintf(fp, "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL\r\n");
  std::fprintf(fp, "%10lld%10lld%10d%10d%10d%10d\r\nENDOFSECTION\r\n", (long long)m->nn, (long long)m->nc, 1, nsets, 3, 3);
  std::fprintf(fp, "NODAL COORDINATES    1.2.1\r\n");
  for (int64_t i = 0; i < m->nn; ++i)
    std::fprintf(fp, "%10lld %.17e %.17e %.17e\r\n", (long long)(i + 1), m->nodes[(size_t)(i * 3)], m->nodes[(size_t)(i * 3 + 1)],
                 m->nodes[(size_t)(i * 3 + 2)]);
  std::fprintf(fp, "ENDOFSECTION\r\nELEMENTS/CELLS       1.2.1\r\n");
  for (int64_t c = 0; c < m->nc; ++c) {
    // Gambit element types: 4 brick, 5 wedge, 6 tetrahedron, 7 pyramid
    const int nn = m->nnode[(size_t)c];
    const int gtype = nn == 8 ? 4 : nn == 6 ? 5 : nn == 4 ? 6 : 7;
    std::fprintf(fp, "%8lld %2d %2d", (long long)(c + 1), gtype, nn);
    for (int q = 0; q < nn; ++q) std::fprintf(fp, " %8d", m->el2nd[(size_t)(c * kNd + q)] + 1);
    std::fprintf(fp, "\r\n");
  }
  std::fprintf(fp, "ENDOFSECTION\r\nELEMENT GROUP        1.2.1\r\nGROUP: 1 ELEMENTS: %lld MATERIAL: 2 NFLAGS: 1\r\nfluid\r\n0\r\n",
               (long long)m->nc);
  for (int64_t c = 0; c < m->nc; ++c) std::fprintf(fp, "%8lld%s", (long long)(c + 1), (c % 10 == 9 || c == m->nc - 1) ? "\r\n" : "");
  std::fprintf(fp, "ENDOFSECTION\r\n");
  for (int b = 1; b <= 5; ++b) {
    if (!count[b]) continue;
    std::fprintf(fp, "BOUNDARY CONDITIONS  1.2.1\r\n%32s%8d%8lld%8d%8d\r\n", kName[b], 1, (long long)count[b], 0, kCode[b]);
    for (int64_t c = 0; c < m->nc; ++c) {
      const int nn = m->nnode[(size_t)c];
      const int gtype = nn == 8 ? 4 : nn == 6 ? 5 : nn == 4 ? 6 : 7;
      for (int j = 0; j < nface_of(*m, c); ++j)
        if (m->nb[(size_t)(c * kFc + j)] < 0 && m->elbc[(size_t)(c * kFc + j)] == b)
          std::fprintf(fp, "%10lld %5d %5d\r\n", (long long)(c + 1), gtype, j + 1);
    }
    std::fprintf(fp, "ENDOFSECTION\r\n");
  }
  int bad = std::ferror(fp);
  std::fclose(fp);
  return bad ? fail(FK_EIO, "write error") : FK_OK;
}

void fkm_free(fkm_mesh *m) { delete m; }

int fkm_sizes(const fkm_mesh *m, int64_t *n_cells, int64_t *n_faces, int64_t *n_nodes, int64_t *n_bfaces) {
  if (!m) return fail(FK_EINVAL, "null mesh");
  if (n_cells) *n_cells = m->nc;
  if (n_faces) *n_faces = m->nf;
  if (n_nodes) *n_nodes = m->nn;
  if (n_bfaces) *n_bfaces = m->nbf;
  return FK_OK;
}

int fkm_array(const fkm_mesh *mc, const char *name, void **ptr) {
  if (!mc || !name || !ptr) return fail(FK_EINVAL, "null argument");
  fkm_mesh *m = const_cast<fkm_mesh *>(mc);
  std::string s(name);
  if (s == "fc2el") *ptr = m->fc2el.data();
  else if (s == "bc") *ptr = m->bc.data();
  else if (s == "el2nd") *ptr = m->el2nd.data();
  else if (s == "nnode") *ptr = m->nnode.data();
  else if (s == "n") *ptr = m->n.data();
  else if (s == "nt1") { build_tangents(*m); *ptr = m->nt1.data(); }
  else if (s == "nt2") { build_tangents(*m); *ptr = m->nt2.data(); }
  else if (s == "fc_center") *ptr = m->fc_center.data();
  else if (s == "shift") *ptr = m->shift.data();
  else if (s == "area") *ptr = m->area.data();
  else if (s == "cell_center") *ptr = m->cell_center.data();
  else if (s == "volume") *ptr = m->volume.data();
  else if (s == "nodes") *ptr = m->nodes.data();
  else return fail(FK_EINVAL, "unknown mesh array: " + s);
  return FK_OK;
}

// ---- initial conditions ----

int fkm_init_sod(const fkm_mesh *m, double gamma, double pinf, double x_split, const double left[5], const double right[5],
                 double *U) {
  if (!m || !left || !right || !U) return fail(FK_EINVAL, "null argument");
  for (int64_t c = 0; c < m->nc; ++c) {
    const double *w = m->cell_center[(size_t)(c * 3)] < x_split ? left : right;
    double rho = w[0], u = w[1], v = w[2], ww = w[3], p = w[4];
    // prim2cons of the single-phase solver (SP/prim2cons.h:2-10, SP/eqos.h case 3)
    double e = (p + gamma * pinf) / (rho * (gamma - 1));
    double *q = U + c * 5;
    q[0] = rho;
    q[1] = rho * u;
    q[2] = rho * v;
    q[3] = rho * ww;
    q[4] = rho * (e + 0.5 * (u * u + v * v + ww * ww));
  }
  return FK_OK;
}

int fkm_init_tgv(const fkm_mesh *m, double gamma, double rho0, double U0, double mach, double *U) {
  if (!m || !U) return fail(FK_EINVAL, "null argument");
  double p0 = rho0 * U0 * U0 / (gamma * mach * mach);
  for (int64_t c = 0; c < m->nc; ++c) {
    double x = m->cell_center[(size_t)(c * 3)], y = m->cell_center[(size_t)(c * 3 + 1)], z = m->cell_center[(size_t)(c * 3 + 2)];
    double u = U0 * std::sin(x) * std::cos(y) * std::cos(z);
    double v = -U0 * std::cos(x) * std::sin(y) * std::cos(z);
    double w = 0.0;
    double p = p0 + rho0 * U0 * U0 / 16.0 * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
    double e = p / (rho0 * (gamma - 1));
    double *q = U + c * 5;
    q[0] = rho0;
    q[1] = rho0 * u;
    q[2] = rho0 * v;
    q[3] = rho0 * w;
    q[4] = rho0 * (e + 0.5 * (u * u + v * v + w * w));
  }
  return FK_OK;
}

int fkm_init_regions(const fkm_mesh *m, const fkm_phase_eos *eos, const fkm_region *rg, int nr, double *U) {
  if (!m || !eos || !rg || !U || nr < 1) return fail(FK_EINVAL, "null argument");
  if (rg[0].kind != 0) return fail(FK_EINVAL, "first region must be the background");
  for (int64_t c = 0; c < m->nc; ++c) {
    const double *x = &m->cell_center[(size_t)(c * 3)];
    const fkm_region *r = &rg[0];
    for (int i = 1; i < nr; ++i) {
      const fkm_region &q = rg[i];
      bool in = false;
      if (q.kind == 1) {
        double d0 = x[0] - q.c[0], d1 = x[1] - q.c[1], d2 = x[2] - q.c[2];
        in = d0 * d0 + d1 * d1 + d2 * d2 < q.r * q.r;
      } else if (q.kind == 2) {
        in = x[0] < q.c[0];
      }
      if (in) r = &q;
    }
    // initVec (initialize.cu:37-54): rho_k = eqos(k,2,p,T), e_k = eqos(k,3,p,rho_k)
    double rho[2], e[2];
    for (int k = 0; k < 2; ++k) {
      double g = eos->gamma[k], pi = eos->pinf[k], cp = eos->cp[k], et = eos->eta[k];
      rho[k] = (r->p + pi) / ((cp / g) * r->T * (g - 1));
      e[k] = (r->p + g * pi + (g - 1) * et * rho[k]) / (rho[k] * (g - 1));
    }
    double aL = r->alphaL, aG = 1 - aL;
    double rhoMix = aL * rho[0] + aG * rho[1];
    double *q = U + c * 10;
    q[0] = aL;
    q[1] = aL * rho[0];
    q[2] = aG * rho[1];
    q[3] = rhoMix * r->u[0];
    q[4] = rhoMix * r->u[1];
    q[5] = rhoMix * r->u[2];
    q[6] = aL * rho[0] * e[0];
    q[7] = aG * rho[1] * e[1];
    q[8] = aL * rho[0] * e[0] + aG * rho[1] * e[1] + 0.5 * rhoMix * (r->u[0] * r->u[0] + r->u[1] * r->u[1] + r->u[2] * r->u[2]);
    q[9] = 0;
  }
  return FK_OK;
}

int fkm_place_bubbles(int n, const double lo[3], const double hi[3], double rmin, double rmax, double xmin, uint64_t seed,
                      double *centres, double *radii) {
  if (n < 0 || !lo || !hi || (n > 0 && (!centres || !radii))) return fail(FK_EINVAL, "bad argument");
  if (!(rmax >= rmin) || !(rmin > 0)) return fail(FK_EINVAL, "bad radius range");
  uint64_t s = seed;
  for (int i = 0; i < n; ++i) {
    int tries = 0;
    for (;;) {
      if (++tries > 100000) return fail(FK_ERANGE, "could not place non-overlapping bubbles");
      double r = rmin + (rmax - rmin) * uniform01(s);
      double c[3];
      bool ok = true;
      for (int a = 0; a < 3; ++a) {
        double l = (a == 0 ? std::max(lo[0], xmin) : lo[a]) + r, h = hi[a] - r;
        if (h <= l) {
          ok = false;
          break;
        }
        c[a] = l + (h - l) * uniform01(s);
      }
      if (!ok) continue;
      for (int k = 0; k < i && ok; ++k) {
        double d0 = c[0] - centres[k * 3], d1 = c[1] - centres[k * 3 + 1], d2 = c[2] - centres[k * 3 + 2];
        double rr = r + radii[k];
        ok = d0 * d0 + d1 * d1 + d2 * d2 > rr * rr;
      }
      if (!ok) continue;
      for (int a = 0; a < 3; ++a) centres[i * 3 + a] = c[a];
      radii[i] = r;
      break;
    }
  }
  return FK_OK;
}

}  // extern "C"
