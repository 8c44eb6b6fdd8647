// Host geometry of the update and remesh steps: interior smoothing, point location in a tet mesh and the
// isosurface-stuffing mesher.
//
// Replaces TetLocator (TetMesh.cs:115-262) and IsosurfaceStuffing.Stuff / StuffLattice (IsosurfaceStuffing.cs:17-88,
// :115-294),
// the pieces of Forging3DSolver.RemeshCore (Forging3DSolver.cs:1958-2102) that are not per-node placement logic.
// Everything is plain IEEE double arithmetic in the order of the reference's Vec3 operations (no contraction), so the
// same inputs give the same mesh bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ifea_remesh.h"

#pragma clang fp contract(off)

namespace {

struct P3 {
  double x, y, z;
};
inline P3 add(const P3& a, const P3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 sub(const P3& a, const P3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 scale(double s, const P3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(const P3& a, const P3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(const P3& a, const P3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(const P3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double len2(const P3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline P3 unit(const P3& a) {
  const double l = len(a);
  return l > 0 ? P3{a.x / l, a.y / l, a.z / l} : a;
}
inline double tet_volume(const P3& a, const P3& b, const P3& c, const P3& d) {
  return dot(sub(b, a), cross(sub(c, a), sub(d, a))) / 6.0;
}

// Uniform bucket grid over the bounding boxes of a tet mesh (cell = a multiple of the element size): a point is
// tested against the tets listed in its cell, in ascending tet order.
class Locator {
 public:
  Locator(int nn, const double* x, int ne, const int* tets, double cell)
      : x_(x, x + 3 * (size_t)nn), t_(tets, tets + 4 * (size_t)ne), ne_(ne), cell_(cell) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int n = 0; n < nn; n++)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], x[3 * n + a]);
        hi[a] = std::max(hi[a], x[3 * n + a]);
      }
    for (int a = 0; a < 3; a++) {
      min_[a] = lo[a] - cell;
      dim_[a] = (int)((hi[a] - min_[a]) / cell) + 2;
    }
    const size_t ncell = (size_t)dim_[0] * dim_[1] * dim_[2];
    start_.assign(ncell + 1, 0);
    for (int e = 0; e < ne; e++) cells_of(e, [&](size_t c) { start_[c + 1]++; });
    for (size_t c = 0; c < ncell; c++) start_[c + 1] += start_[c];
    items_.resize(start_[ncell]);
    std::vector<long long> fill(start_.begin(), start_.end() - 1);
    for (int e = 0; e < ne; e++) cells_of(e, [&](size_t c) { items_[fill[c]++] = e; });
  }

  P3 node(int n) const { return {x_[3 * (size_t)n], x_[3 * (size_t)n + 1], x_[3 * (size_t)n + 2]}; }

  // containing tet (first in the cell's list with every barycentric weight >= -1e-9) and its weights, or -1
  int find(const P3& p, double w[4]) const {
    const size_t c = cell(p);
    for (long long k = start_[c]; k < start_[c + 1]; k++) {
      const int e = items_[k];
      const P3 a = node(t_[4 * (size_t)e]), b = node(t_[4 * (size_t)e + 1]);
      const P3 cc = node(t_[4 * (size_t)e + 2]), d = node(t_[4 * (size_t)e + 3]);
      const double v = tet_volume(a, b, cc, d);
      if (v <= 0) continue;
      const double b0 = tet_volume(p, b, cc, d) / v;
      const double b1 = tet_volume(a, p, cc, d) / v;
      const double b2 = tet_volume(a, b, p, d) / v;
      const double b3 = 1 - b0 - b1 - b2;
      const double tol = -1e-9;
      if (b0 >= tol && b1 >= tol && b2 >= tol && b3 >= tol) {
        if (w) { w[0] = b0; w[1] = b1; w[2] = b2; w[3] = b3; }
        return e;
      }
    }
    return -1;
  }

  bool contains(const P3& p) const { return find(p, nullptr) >= 0; }

  // the containing tet, else the tet of the point's cell with the nearest centroid, else the nearest over all tets
  int find_nearest(const P3& p) const {
    const int e = find(p, nullptr);
    if (e >= 0) return e;
    const size_t c = cell(p);
    double best = INFINITY;
    int be = -1;
    for (long long k = start_[c]; k < start_[c + 1]; k++) near_tet(items_[k], p, best, be);
    if (be >= 0) return be;
    for (int q = 0; q < ne_; q++) near_tet(q, p, best, be);
    return be;
  }

  // linear interpolation of a nodal 3-vector field; the nearest node's value outside the mesh
  P3 interpolate(const P3& p, const double* f) const {
    double w[4];
    const int e = find(p, w);
    auto fv = [&](int n) { return P3{f[3 * (size_t)n], f[3 * (size_t)n + 1], f[3 * (size_t)n + 2]}; };
    if (e >= 0) {
      const int* te = &t_[4 * (size_t)e];
      return add(add(add(scale(w[0], fv(te[0])), scale(w[1], fv(te[1]))), scale(w[2], fv(te[2]))),
                 scale(w[3], fv(te[3])));
    }
    const int nn = (int)(x_.size() / 3);
    int best = 0;
    double dmin = INFINITY;
    for (int n = 0; n < nn; n++) {
      const double dd = len2(sub(node(n), p));
      if (dd < dmin) { dmin = dd; best = n; }
    }
    return fv(best);
  }

 private:
  int ci(double v, int a) const {
    const double q = (v - min_[a]) / cell_;
    // the grid's own range: points far outside clamp to the border cells
    if (!(q >= 0)) return 0;
    if (q >= dim_[a] - 1) return dim_[a] - 1;
    return (int)q;
  }
  size_t cell(const P3& p) const {
    return ((size_t)ci(p.z, 2) * dim_[1] + ci(p.y, 1)) * dim_[0] + ci(p.x, 0);
  }
  template <typename F>
  void cells_of(int e, F&& f) const {
    P3 lo = node(t_[4 * (size_t)e]), hi = lo;
    for (int a = 1; a < 4; a++) {
      const P3 p = node(t_[4 * (size_t)e + a]);
      lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
      hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    for (int k = ci(lo.z, 2); k <= ci(hi.z, 2); k++)
      for (int j = ci(lo.y, 1); j <= ci(hi.y, 1); j++)
        for (int i = ci(lo.x, 0); i <= ci(hi.x, 0); i++) f(((size_t)k * dim_[1] + j) * dim_[0] + i);
  }
  void near_tet(int e, const P3& p, double& best, int& be) const {
    const int* te = &t_[4 * (size_t)e];
    const P3 cen = scale(0.25, add(add(add(node(te[0]), node(te[1])), node(te[2])), node(te[3])));
    const double dd = len2(sub(cen, p));
    if (dd < best) { best = dd; be = e; }
  }

  std::vector<double> x_;
  std::vector<int> t_;
  int ne_;
  double cell_;
  double min_[3];
  int dim_[3];
  std::vector<long long> start_;
  std::vector<int> items_;
};

// Isosurface stuffing (Labelle and Shewchuk, SIGGRAPH 2007) on a body-centred-cubic lattice of spacing h, the domain
// given by an inside test. Lattice vertex ids: cube corners first (x fastest), then cube centres; cut vertices are
// appended in the order the clipping first asks for them.
class Stuffer {
 public:
  Stuffer(const Locator& inside, const double* bmin, const double* bmax, double h, bool symX, bool symY)
      : in_(inside), h_(h), symX_(symX), symY_(symY), eps_(1e-7 * h) {
    const P3 off{0.31 * h, 0.37 * h, 0.17 * h};  // no lattice vertex on a symmetry plane
    const int i0 = (int)std::floor(bmin[0] / h) - 1, i1 = (int)std::ceil(bmax[0] / h) + 1;
    const int j0 = (int)std::floor(bmin[1] / h) - 1, j1 = (int)std::ceil(bmax[1] / h) + 1;
    const int k0 = (int)std::floor(bmin[2] / h) - 1, k1 = (int)std::ceil(bmax[2] / h) + 1;
    const int ni = i1 - i0, nj = j1 - j0, nk = k1 - k0;
    const long long ncorner = (long long)(ni + 1) * (nj + 1) * (nk + 1), ncentre = (long long)ni * nj * nk;
    if (ncorner + ncentre > (1LL << 30)) throw std::runtime_error("stuffing lattice too large");
    auto corner = [&](int i, int j, int k) { return (int)(((long long)(k - k0) * (nj + 1) + (j - j0)) * (ni + 1) + (i - i0)); };
    auto centre = [&](int i, int j, int k) { return (int)(ncorner + ((long long)(k - k0) * nj + (j - j0)) * ni + (i - i0)); };
    pos_.reserve(ncorner + ncentre);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) pos_.push_back(add(P3{i * h, j * h, k * h}, off));
    for (int k = k0; k < k1; k++)
      for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) pos_.push_back(add(P3{(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h}, off));
    ncorner_ = ncorner;
    // lattice tets: two face-adjacent cube centres and one edge of their shared face, four per face
    auto pair = [&](int c1, int c2, int a, int b, int c, int d) {
      const int q[4][2] = {{a, b}, {b, c}, {c, d}, {d, a}};
      for (auto& e : q) lat_.insert(lat_.end(), {c1, c2, e[0], e[1]});
    };
    for (int k = k0; k < k1; k++)
      for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) {
          const int c = centre(i, j, k);
          if (i + 1 < i1)
            pair(c, centre(i + 1, j, k), corner(i + 1, j, k), corner(i + 1, j + 1, k), corner(i + 1, j + 1, k + 1),
                 corner(i + 1, j, k + 1));
          if (j + 1 < j1)
            pair(c, centre(i, j + 1, k), corner(i, j + 1, k), corner(i + 1, j + 1, k), corner(i + 1, j + 1, k + 1),
                 corner(i, j + 1, k + 1));
          if (k + 1 < k1)
            pair(c, centre(i, j, k + 1), corner(i, j, k + 1), corner(i + 1, j, k + 1), corner(i + 1, j + 1, k + 1),
                 corner(i, j + 1, k + 1));
        }
    run();
  }

  std::vector<double> points;
  std::vector<int> tets;

 private:
  static constexpr double kAlphaLong = 0.24999, kAlphaShort = 0.41189;
  static constexpr int EA[6] = {0, 0, 0, 1, 1, 2}, EB[6] = {1, 2, 3, 2, 3, 3};

  bool inside(const P3& p) const {
    if ((symX_ && p.x < 0) || (symY_ && p.y < 0)) return false;
    return in_.contains(P3{symX_ ? std::max(p.x, eps_) : p.x, symY_ ? std::max(p.y, eps_) : p.y, p.z});
  }
  bool is_corner(int v) const { return v < ncorner_; }
  static unsigned long long key(int u, int w) {
    const unsigned a = (unsigned)std::min(u, w), b = (unsigned)std::max(u, w);
    return ((unsigned long long)a << 32) | b;
  }
  bool crosses(int u, int w) const { return (sign_[u] < 0 && sign_[w] > 0) || (sign_[u] > 0 && sign_[w] < 0); }

  int cut_vertex(int u, int w) {
    const unsigned long long k = key(u, w);
    auto it = cut_id_.find(k);
    if (it != cut_id_.end()) return it->second;
    pos_.push_back(cut_pos_[cut_index_.at(k)]);
    sign_.push_back(0);
    const int id = (int)pos_.size() - 1;
    cut_id_.emplace(k, id);
    return id;
  }

  void emit(int a, int b, int c, int d) {
    const double v = tet_volume(pos_[a], pos_[b], pos_[c], pos_[d]);
    if (std::fabs(v) < min_vol_) return;
    if (v < 0) std::swap(c, d);
    out_.insert(out_.end(), {a, b, c, d});
  }

  // cone from apex over a polygon, the polygon split into a fan from its smallest id
  void cone(const std::vector<int>& f, int apex) {
    if (f.size() < 3 || std::find(f.begin(), f.end(), apex) != f.end()) return;
    const int n = (int)f.size();
    const int s = (int)(std::min_element(f.begin(), f.end()) - f.begin());
    for (int q = 1; q < n - 1; q++) emit(apex, f[s], f[(s + q) % n], f[(s + q + 1) % n]);
  }

  // the cut polygon's vertices in angular order about its centroid, viewed along the normal away from the inside
  void order_planar(std::vector<int>& f, const std::vector<int>& all) {
    P3 c{0, 0, 0};
    for (int v : f) c = add(c, pos_[v]);
    c = scale(1.0 / (double)f.size(), c);
    P3 inner{0, 0, 0};
    int cnt = 0;
    for (int v : all)
      if (sign_[v] < 0) { inner = add(inner, pos_[v]); cnt++; }
    const P3 nrm = unit(sub(c, scale(1.0 / std::max(cnt, 1), inner)));
    const P3 u = unit(cross(std::fabs(nrm.x) < 0.9 ? P3{1, 0, 0} : P3{0, 1, 0}, nrm));
    const P3 w = cross(nrm, u);
    std::vector<std::pair<double, int>> ang;
    for (int v : f) {
      const P3 d = sub(pos_[v], c);
      ang.push_back({std::atan2(dot(d, w), dot(d, u)), v});
    }
    std::stable_sort(ang.begin(), ang.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t q = 0; q < f.size(); q++) f[q] = ang[q].second;
  }

  void run() {
    sign_.resize(pos_.size());
    for (size_t v = 0; v < pos_.size(); v++) sign_[v] = inside(pos_[v]) ? -1 : 1;
    // cut points of the lattice edges joining inside and outside vertices, by 16 bisections of the inside test
    for (size_t t = 0; t < lat_.size(); t += 4)
      for (int q = 0; q < 6; q++) {
        const int u = lat_[t + EA[q]], w = lat_[t + EB[q]];
        if (!crosses(u, w)) continue;
        const unsigned long long k = key(u, w);
        if (cut_index_.count(k)) continue;
        P3 pin = sign_[u] < 0 ? pos_[u] : pos_[w], pout = sign_[u] < 0 ? pos_[w] : pos_[u];
        for (int it = 0; it < 16; it++) {
          const P3 m = scale(0.5, add(pin, pout));
          if (inside(m)) pin = m; else pout = m;
        }
        P3 cp = scale(0.5, add(pin, pout));
        if (symX_ && std::fabs(cp.x) < 1e-4 * h_) cp.x = 0;
        if (symY_ && std::fabs(cp.y) < 1e-4 * h_) cp.y = 0;
        cut_index_.emplace(k, (int)cut_pos_.size());
        cut_pos_.push_back(cp);
        cut_u_.push_back(std::min(u, w));
        cut_w_.push_back(std::max(u, w));
      }
    // warp every lattice vertex closer than alpha * (edge length) to a cut point of its edges onto the nearest such
    // cut point (the first of equal distances, in the cuts' order)
    std::unordered_map<int, std::pair<P3, double>> warp;
    std::vector<int> warp_order;
    for (size_t q = 0; q < cut_pos_.size(); q++) {
      const int u = cut_u_[q], w = cut_w_[q];
      const double l = len(sub(pos_[w], pos_[u]));
      const double alpha = is_corner(u) == is_corner(w) ? kAlphaLong : kAlphaShort;
      for (int v : {u, w}) {
        const double d = len(sub(cut_pos_[q], pos_[v]));
        if (sign_[v] == 0 || d >= alpha * l) continue;
        auto it = warp.find(v);
        if (it == warp.end()) {
          warp.emplace(v, std::make_pair(cut_pos_[q], d));
          warp_order.push_back(v);
        } else if (d < it->second.second) {
          it->second = std::make_pair(cut_pos_[q], d);
        }
      }
    }
    for (int v : warp_order) {
      pos_[v] = warp[v].first;
      sign_[v] = 0;
    }
    // clip every lattice tet that reaches into the domain and split it into tets
    min_vol_ = 1e-6 * h_ * h_ * h_;
    static const int FACES[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    std::vector<int> poly, face, iface;
    for (size_t t = 0; t < lat_.size(); t += 4) {
      const int v[4] = {lat_[t], lat_[t + 1], lat_[t + 2], lat_[t + 3]};
      int neg = 0, posc = 0;
      for (int x : v) {
        if (sign_[x] < 0) neg++;
        else if (sign_[x] > 0) posc++;
      }
      if (neg == 0) continue;
      if (posc == 0) {
        emit(v[0], v[1], v[2], v[3]);
        continue;
      }
      poly.clear();
      iface.clear();
      for (int x : v) {
        if (sign_[x] <= 0) poly.push_back(x);
        if (sign_[x] == 0) iface.push_back(x);
      }
      for (int q = 0; q < 6; q++) {
        const int u = v[EA[q]], w = v[EB[q]];
        if (crosses(u, w)) {
          const int c = cut_vertex(u, w);
          poly.push_back(c);
          iface.push_back(c);
        }
      }
      const int apex = *std::min_element(poly.begin(), poly.end());
      for (const auto& f : FACES) {
        face.clear();
        for (int q = 0; q < 3; q++) {
          const int a = v[f[q]], b = v[f[(q + 1) % 3]];
          if (sign_[a] <= 0) face.push_back(a);
          if (crosses(a, b)) face.push_back(cut_vertex(a, b));
        }
        cone(face, apex);
      }
      if (iface.size() >= 3) {
        order_planar(iface, poly);
        cone(iface, apex);
      }
    }
    // compact: vertices numbered in the order the tets first use them
    std::unordered_map<int, int> map;
    tets.resize(out_.size());
    for (size_t q = 0; q < out_.size(); q++) {
      auto it = map.find(out_[q]);
      int id;
      if (it == map.end()) {
        id = (int)(points.size() / 3);
        map.emplace(out_[q], id);
        const P3& p = pos_[out_[q]];
        points.insert(points.end(), {p.x, p.y, p.z});
      } else {
        id = it->second;
      }
      tets[q] = id;
    }
  }

  const Locator& in_;
  double h_;
  bool symX_, symY_;
  double eps_, min_vol_ = 0;
  int ncorner_ = 0;
  std::vector<P3> pos_;
  std::vector<int> sign_;
  std::vector<int> lat_;
  std::unordered_map<unsigned long long, int> cut_index_, cut_id_;
  std::vector<P3> cut_pos_;
  std::vector<int> cut_u_, cut_w_;
  std::vector<int> out_;
};

// TetMesh.Quality (TetMesh.cs:107-112): 6 sqrt(2) V / l2^1.5, l2 the mean squared edge length; -1 when degenerate
inline double tet_quality(const P3& a, const P3& b, const P3& c, const P3& d) {
  const double l2 = (len2(sub(b, a)) + len2(sub(c, a)) + len2(sub(d, a)) + len2(sub(c, b)) + len2(sub(d, b)) +
                     len2(sub(d, c))) / 6.0;
  return l2 > 0 ? 6 * std::sqrt(2.0) * tet_volume(a, b, c, d) / std::pow(l2, 1.5) : -1;
}

// SmoothInterior (Forging3DSolver.cs:1776-1838) with the tets' qualities kept current: a sweep marks the nodes of the
// tets below qTarget, then each marked interior node (ascending) tries its neighbours' centroid at full, half and
// quarter steps and keeps the first that raises the worst quality of its tets by more than 1e-6; only the moved node's
// tets are re-evaluated, so every decision sees the qualities a full re-evaluation would give.
int smooth_interior(int nn, int ne, const int* t, const int* nts, const int* ntl, const unsigned char* surf,
                    const unsigned char* ox, const unsigned char* oy, double* xp, double qTarget, int sweeps) {
  auto X = [&](int n) { return P3{xp[3 * (size_t)n], xp[3 * (size_t)n + 1], xp[3 * (size_t)n + 2]}; };
  auto Q = [&](int e) { return tet_quality(X(t[4 * (size_t)e]), X(t[4 * (size_t)e + 1]), X(t[4 * (size_t)e + 2]),
                                           X(t[4 * (size_t)e + 3])); };
  std::vector<double> q(ne);
  {
    // every tet's quality is independent: evaluate them on a few host threads
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int w = 0; w < nt; w++)
      th.emplace_back([&, w] {
        for (int e = (int)((long long)ne * w / nt); e < (int)((long long)ne * (w + 1) / nt); e++) q[e] = Q(e);
      });
    for (auto& x : th) x.join();
  }
  std::vector<char> bad(nn);
  int moves = 0;
  for (int s = 0; s < sweeps; s++) {
    std::fill(bad.begin(), bad.end(), 0);
    bool any = false;
    for (int e = 0; e < ne; e++)
      if (q[e] < qTarget) {
        for (int a = 0; a < 4; a++) bad[t[4 * (size_t)e + a]] = 1;
        any = true;
      }
    if (!any) break;
    int sweep_moves = 0;
    for (int n = 0; n < nn; n++) {
      if (!bad[n] || surf[n]) continue;
      P3 sum{0, 0, 0};
      int cnt = 0;
      double worst = INFINITY;
      for (int k = nts[n]; k < nts[n + 1]; k++) {
        const int e = ntl[k];
        for (int a = 0; a < 4; a++) {
          const int m = t[4 * (size_t)e + a];
          if (m == n) continue;
          sum = add(sum, X(m));
          cnt++;
        }
        worst = std::min(worst, q[e]);
      }
      if (cnt == 0) continue;
      const P3 old = X(n);
      const P3 c = scale(1.0 / cnt, sum);
      for (double lam : {1.0, 0.5, 0.25}) {
        const P3 p = add(old, scale(lam, sub(c, old)));
        xp[3 * (size_t)n] = ox[n] ? 0 : p.x;
        xp[3 * (size_t)n + 1] = oy[n] ? 0 : p.y;
        xp[3 * (size_t)n + 2] = p.z;
        double qn = INFINITY;
        for (int k = nts[n]; k < nts[n + 1]; k++) qn = std::min(qn, Q(ntl[k]));
        if (qn > worst + 1e-6) {
          sweep_moves++;
          for (int k = nts[n]; k < nts[n + 1]; k++) q[ntl[k]] = Q(ntl[k]);
          break;
        }
        xp[3 * (size_t)n] = old.x;
        xp[3 * (size_t)n + 1] = old.y;
        xp[3 * (size_t)n + 2] = old.z;
      }
    }
    moves += sweep_moves;
    if (sweep_moves == 0) break;
  }
  return moves;
}

thread_local char g_err[256] = "";
void set_err(const char* m) { std::snprintf(g_err, sizeof g_err, "%s", m); }

}  // namespace

struct ifea_locator { Locator loc; };
struct ifea_stuffing { std::vector<double> points; std::vector<int> tets; };

extern "C" {

int ifea_locator_create(int nn, const double* x, int ne, const int* tets, double cell, ifea_locator_t* out) {
  if (!out || nn <= 0 || ne < 0 || !x || (ne > 0 && !tets) || !(cell > 0)) { set_err("invalid locator arguments"); return -1; }
  try {
    for (long long q = 0; q < 4LL * ne; q++)
      if (tets[q] < 0 || tets[q] >= nn) { set_err("tet node index out of range"); return -1; }
    *out = new ifea_locator{Locator(nn, x, ne, tets, cell)};
    return 0;
  } catch (const std::exception& e) {
    set_err(e.what());
    return -6;
  }
}

int ifea_locator_find(ifea_locator_t h, int np, const double* p, int* tet, double* w) {
  if (!h || np < 0 || (np > 0 && (!p || !tet))) { set_err("invalid locator query"); return -1; }
  for (int i = 0; i < np; i++) {
    const P3 q{p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
    tet[i] = h->loc.find(q, w ? w + 4 * (size_t)i : nullptr);
  }
  return 0;
}

int ifea_locator_find_nearest(ifea_locator_t h, int np, const double* p, int* tet) {
  if (!h || np < 0 || (np > 0 && (!p || !tet))) { set_err("invalid locator query"); return -1; }
  for (int i = 0; i < np; i++)
    tet[i] = h->loc.find_nearest(P3{p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]});
  return 0;
}

int ifea_locator_interpolate(ifea_locator_t h, int np, const double* p, const double* field, double* out) {
  if (!h || np < 0 || (np > 0 && (!p || !field || !out))) { set_err("invalid locator query"); return -1; }
  for (int i = 0; i < np; i++) {
    const P3 v = h->loc.interpolate(P3{p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}, field);
    out[3 * (size_t)i] = v.x;
    out[3 * (size_t)i + 1] = v.y;
    out[3 * (size_t)i + 2] = v.z;
  }
  return 0;
}

void ifea_locator_destroy(ifea_locator_t h) { delete h; }

int ifea_stuff(ifea_locator_t inside, const double* bmin, const double* bmax, double h, int symX, int symY,
               ifea_stuffing_t* out) {
  if (!inside || !bmin || !bmax || !out || !(h > 0)) { set_err("invalid stuffing arguments"); return -1; }
  try {
    Stuffer s(inside->loc, bmin, bmax, h, symX != 0, symY != 0);
    *out = new ifea_stuffing{std::move(s.points), std::move(s.tets)};
    return 0;
  } catch (const std::exception& e) {
    set_err(e.what());
    return -6;
  }
}

int ifea_stuffing_sizes(ifea_stuffing_t s, int* nn, int* ne) {
  if (!s || !nn || !ne) { set_err("invalid stuffing handle"); return -1; }
  *nn = (int)(s->points.size() / 3);
  *ne = (int)(s->tets.size() / 4);
  return 0;
}

int ifea_stuffing_get(ifea_stuffing_t s, double* x, int* tets) {
  if (!s || !x || !tets) { set_err("invalid stuffing handle"); return -1; }
  std::memcpy(x, s->points.data(), s->points.size() * sizeof(double));
  std::memcpy(tets, s->tets.data(), s->tets.size() * sizeof(int));
  return 0;
}

void ifea_stuffing_destroy(ifea_stuffing_t s) { delete s; }

int ifea_smooth_interior(int nn, int ne, const int* tets, const int* node_tet_start, const int* node_tet_list,
                         const unsigned char* is_surface, const unsigned char* on_x0, const unsigned char* on_y0,
                         double* x, double q_target, int sweeps, int* moves) {
  if (nn <= 0 || ne < 0 || !tets || !node_tet_start || !node_tet_list || !is_surface || !on_x0 || !on_y0 || !x ||
      !moves || sweeps < 0) {
    set_err("invalid smoothing arguments");
    return -1;
  }
  try {
    *moves = smooth_interior(nn, ne, tets, node_tet_start, node_tet_list, is_surface, on_x0, on_y0, x, q_target,
                             sweeps);
    return 0;
  } catch (const std::exception& e) {
    set_err(e.what());
    return -6;
  }
}

const char* ifea_remesh_last_error(void) { return g_err; }

}  // extern "C"
