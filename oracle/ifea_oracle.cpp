// TEST INFRASTRUCTURE ONLY — the FP64 CPU oracle for ImpressionFEA's 3D rigid-viscoplastic solve step.
//
// A scalar restatement of src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs (and TetMesh.cs, FlowStressLaw.cs,
// Solver/Linear/ISpdSolver.cs, Solver/Principal.cs) used as the checker by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg. It is never linked into the product library. Each function cites the reference lines
// it follows. Element work runs on OpenMP threads into per-element buffers and is summed serially in element order,
// as the reference does (Forging3DSolver.cs:1133-1171), so results do not depend on the thread count.
//
// Linear solver behind the ISpdSolver slot: an FP64 left-looking sparse Cholesky (symbolic structure from the
// elimination tree, ordering = caller permutation or an exact minimum-degree ordering on the node graph), or Intel
// MKL PARDISO through its documented C prototype (oracle/mkl_pardiso.cpp), as the reference's MklPardisoSolver does.
//
// Parity status (see the tests): every formula here is pinned by a closed form, a finite-difference check, an
// invariant or brute force, except the minimum-degree ordering, which is NOT CSparse's AMD (CSparse.NET is absent),
// so the reference's own ordering is "parity unpinned"; orderings are compared by feeding one permutation to both sides.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "pardiso_dl.h"

namespace {

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  V3 operator-() const { return {-x, -y, -z}; }
  double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
  double len() const { return std::sqrt(x * x + y * y + z * z); }
  V3 normalized() const { double l = len(); return l > 0 ? V3(x / l, y / l, z / l) : *this; }
};
inline V3 operator*(double s, const V3& v) { return {s * v.x, s * v.y, s * v.z}; }

// FlowStressLaw.Stress (Materials/FlowStressLaw.cs:8-9)
struct Law {
  double C, N, M, E0;
  double stress(double strain, double rate) const {
    return C * std::pow(E0 + std::max(strain, 0.0), N) * std::pow(std::max(rate, 1e-12), M);
  }
};

// Principal.Max3 (Solver/Principal.cs:7-19)
double max3(double xx, double yy, double zz, double xy, double yz, double zx) {
  double p1 = xy * xy + yz * yz + zx * zx;
  double q = (xx + yy + zz) / 3;
  if (p1 < 1e-30 * (xx * xx + yy * yy + zz * zz + 1e-300)) return std::max(xx, std::max(yy, zz));
  double p2 = (xx - q) * (xx - q) + (yy - q) * (yy - q) + (zz - q) * (zz - q) + 2 * p1;
  double p = std::sqrt(p2 / 6);
  double bxx = (xx - q) / p, byy = (yy - q) / p, bzz = (zz - q) / p, bxy = xy / p, byz = yz / p, bzx = zx / p;
  double detB = bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bzx) + bzx * (bxy * byz - byy * bzx);
  double r = std::min(1.0, std::max(-1.0, detB / 2));
  double phi = std::acos(r) / 3;
  return q + 2 * p * std::cos(phi);
}

// ---------------------------------------------------------------------------------------------- rigid dies
// Analytic height-field dies. side 0 = bottom (metal above), 1 = top (metal below). kind 0: flat plane at z = p[0].
// kind 1: top die with a hexagonal cavity: land at z = p[0] (the flash thickness t at closure), cavity ceiling at
// p[0] + p[1], across-flats p[2], wall transition width p[3], land width p[4], gutter depth p[5].
// kind 2: one half of a connecting-rod closed die, mirrored about the parting plane z = 0: a big boss (centre x = -42,
// radius 20) and a small boss (x = 49.5, radius 12.5) joined by an I-section shank of width p[5] (flanges p[3] and
// web p[4] deep), boss depths p[1] and p[2], flash thickness p[0] at closure, wall transition p[6], land width p[7],
// gutter depth 5; the surface sits at +-(p[0]/2 + cavity depth + gutter).
// Clearance follows HeightFieldDie.Clearance (HeightFieldDie.cs:130-134); Project is the vertical projection onto
// the surface with the normal of the analytic surface pointing out of the die (the reference searches the closest
// point on a triangulated height field, HeightFieldDie.cs:157-191; on flat regions both agree).
struct Die {
  int side = 0, kind = 0;
  double p[8] = {0};
  static double smooth(double u) { u = std::min(1.0, std::max(0.0, u)); return u * u * (3 - 2 * u); }
  double height(double x, double y) const {
    if (kind == 0) return p[0];
    if (kind == 2) {
      const double xb = -42.0, rb = 20.0, xs = 49.5, rs = 12.5, g = 5.0;
      const double t = p[0], wall = p[6], land = p[7], w = p[5], fl = 0.25 * w;
      double db = std::sqrt((x - xb) * (x - xb) + y * y) - rb, ds = std::sqrt((x - xs) * (x - xs) + y * y) - rs;
      double dsh = std::max(std::max(xb - x, x - xs), std::fabs(y) - 0.5 * w);
      double shank = p[4] + (p[3] - p[4]) * smooth((std::fabs(y) - (0.5 * w - fl) + 0.5 * wall) / wall);
      double depth = std::max(std::max(p[1] * smooth((-db + 0.5 * wall) / wall), p[2] * smooth((-ds + 0.5 * wall) / wall)),
                              shank * smooth((-dsh + 0.5 * wall) / wall));
      double h = 0.5 * t + depth + g * smooth((std::min(std::min(db, ds), dsh) - land) / wall);
      return side == 1 ? h : -h;
    }
    const double s3 = std::sqrt(3.0) / 2;
    double d = std::max(std::fabs(x), std::max(std::fabs(0.5 * x + s3 * y), std::fabs(-0.5 * x + s3 * y))) - 0.5 * p[2];
    double h = p[0] + p[1] * smooth((-d + 0.5 * p[3]) / p[3]);
    h += p[5] * smooth((d - p[4]) / p[3]);
    return h;
  }
  double clearance(const V3& q, double off) const {
    double h = height(q.x, q.y) + off;
    return side == 0 ? q.z - h : h - q.z;
  }
  V3 project(const V3& q, double off, V3* nrm) const {
    const double dd = 1e-5;
    double hx = (height(q.x + dd, q.y) - height(q.x - dd, q.y)) / (2 * dd);
    double hy = (height(q.x, q.y + dd) - height(q.x, q.y - dd)) / (2 * dd);
    V3 up = V3(-hx, -hy, 1.0).normalized();
    if (nrm) *nrm = side == 0 ? up : -up;
    return V3(q.x, q.y, height(q.x, q.y) + off);
  }
  double surface_z(double x, double y) const { return height(x, y); }
};

enum Contact : int { FREE = 0, BOTTOM = 1, TOP = 2 };

}  // namespace

extern "C" {

typedef struct OrcParams {
  double C, N, M, Eps0;          // flow law
  double friction;               // shear friction factor m (SimulationConfig3D.FrictionFactor)
  double ram;                    // ram speed, mm/s
  double penalty_factor;         // Forging3DSolver.PenaltyFactor (relative)
  double newton_blend;           // Forging3DSolver.NewtonBlend
  double tolerance;              // SimulationConfig3D.Tolerance on beta*|dv|/|v|
  double size;                   // billet size for the reference rate (H for a vertical cylinder, Forging3DSolver.cs:145)
  double stroke_increment;       // nominal stroke increment, mm
  double h;                      // nominal element size, mm (contact tolerances)
  int max_iterations;            // SimulationConfig3D.MaxIterations
  int pcg_budget;                // Forging3DSolver.PcgBudget
  int quadratic_ls;              // Forging3DSolver.QuadraticLineSearch
  int adaptive_newton;           // Forging3DSolver.AdaptiveNewton
  int extrapolate_start;         // Forging3DSolver.ExtrapolateStart
  int sym_factor;                // DieSet3D.SymmetryFactor
  int linear;                    // 0: oracle sparse Cholesky, 1: MKL PARDISO
  int bottom_kind, top_kind;     // die kinds
  double bottom_p[8], top_p[8];  // die parameters (closed position)
} OrcParams;

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------------------------- sparse Cholesky
// Symbolic: elimination tree (Liu), row structures by row-subtree traversal, column counts; numeric: left-looking
// column Cholesky in FP64. The matrix is given as a full symmetric CSC pattern (Forging3DSolver.cs:469-521), the
// solver works on the lower triangle of P A P^T.
struct SparseChol {
  int n = 0;
  std::vector<int> perm, iperm;       // perm[k] = original index of the k-th pivot
  std::vector<int> Lp, Li;            // CSC of L (lower, diagonal first)
  std::vector<double> Lx;
  std::vector<int> parent;            // etree of P A P^T
  // gather map: for each entry of L's pattern that comes from A: (A position) -> (L position)
  std::vector<int> amap_src, amap_dst;
  std::vector<double> work;

  static void etree_of(int n, const std::vector<int>& cp, const std::vector<int>& ri, std::vector<int>& parent) {
    // Liu's algorithm with path compression on the upper triangle (cs_etree): column j, rows i < j.
    parent.assign(n, -1);
    std::vector<int> anc(n, -1);
    for (int j = 0; j < n; j++) {
      for (int p = cp[j]; p < cp[j + 1]; p++) {
        int i = ri[p];
        while (i != -1 && i < j) {
          int nx = anc[i];
          anc[i] = j;
          if (nx == -1) { parent[i] = j; break; }
          i = nx;
        }
      }
    }
  }

  // Analyze: permuted upper pattern, etree, structure of L.
  int analyze(int n_, const int* colPtr, const int* rowIdx, const int* userPerm, std::vector<int>* colcount_out) {
    n = n_;
    perm.resize(n);
    iperm.resize(n);
    for (int k = 0; k < n; k++) perm[k] = userPerm ? userPerm[k] : k;
    std::vector<int> seen(n, 0);
    for (int k = 0; k < n; k++) {
      if (perm[k] < 0 || perm[k] >= n || seen[perm[k]]) return -2;
      seen[perm[k]] = 1;
      iperm[perm[k]] = k;
    }
    // upper triangle of C = P A P^T by columns: C(iperm[i], iperm[j]) = A(i, j)
    std::vector<int> cnt(n + 1, 0);
    for (int j = 0; j < n; j++)
      for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
        int i = rowIdx[p];
        int pi = iperm[i], pj = iperm[j];
        if (pi <= pj) cnt[pj + 1]++;
      }
    std::vector<int> ucp(n + 1, 0);
    for (int j = 0; j < n; j++) ucp[j + 1] = ucp[j] + cnt[j + 1];
    std::vector<int> uri(ucp[n]), usrc(ucp[n]);
    std::vector<int> fill(ucp.begin(), ucp.end() - 1);
    for (int j = 0; j < n; j++)
      for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
        int i = rowIdx[p];
        int pi = iperm[i], pj = iperm[j];
        if (pi <= pj) { uri[fill[pj]] = pi; usrc[fill[pj]] = p; fill[pj]++; }
      }
    etree_of(n, ucp, uri, parent);
    // row structure of L: row i of L = reach of the upper-column i entries in the etree (cs_ereach)
    std::vector<std::vector<int>> cols(n);
    std::vector<int> mark(n, -1);
    for (int i = 0; i < n; i++) {
      mark[i] = i;
      for (int p = ucp[i]; p < ucp[i + 1]; p++) {
        int k = uri[p];
        while (k != -1 && k < i && mark[k] != i) {
          cols[k].push_back(i);  // L(i, k) != 0
          mark[k] = i;
          k = parent[k];
        }
      }
    }
    Lp.assign(n + 1, 0);
    for (int j = 0; j < n; j++) Lp[j + 1] = Lp[j] + 1 + (int)cols[j].size();
    Li.resize(Lp[n]);
    for (int j = 0; j < n; j++) {
      Li[Lp[j]] = j;
      std::copy(cols[j].begin(), cols[j].end(), Li.begin() + Lp[j] + 1);  // rows pushed in increasing i
    }
    if (colcount_out) {
      colcount_out->resize(n);
      for (int j = 0; j < n; j++) (*colcount_out)[j] = Lp[j + 1] - Lp[j];
    }
    // value map A -> L (lower of C is the transpose of the upper part)
    amap_src.clear();
    amap_dst.clear();
    for (int j = 0; j < n; j++) {  // upper column j of C: entries (i, j), i <= j => L position of (j, i) in column i
      for (int p = ucp[j]; p < ucp[j + 1]; p++) {
        int i = uri[p];
        const int* b = &Li[Lp[i]];
        const int* e = &Li[Lp[i + 1]];
        const int* f = std::lower_bound(b, e, j);
        amap_src.push_back(usrc[p]);
        amap_dst.push_back((int)(f - &Li[0]));
      }
    }
    Lx.assign(Lp[n], 0.0);
    work.assign(n, 0.0);
    return 0;
  }

  // Left-looking numeric factorization: column j = A(:,j) - sum_k L(j,k) L(j:,k).
  int factorize(const double* full) {
    std::fill(Lx.begin(), Lx.end(), 0.0);
    for (size_t k = 0; k < amap_src.size(); k++) Lx[amap_dst[k]] += full[amap_src[k]];
    std::vector<int> next(n, 0);           // next[k] = position in column k of the first row >= current j
    std::vector<int> head(n, -1), link(n, -1);  // linked lists of columns k waiting for row j
    for (int k = 0; k < n; k++) next[k] = Lp[k] + 1;
    std::vector<double>& w = work;
    for (int j = 0; j < n; j++) {
      for (int p = Lp[j]; p < Lp[j + 1]; p++) w[Li[p]] = Lx[p];
      int k = head[j];
      while (k != -1) {
        int nk = link[k];
        int pj = next[k];  // Li[pj] == j
        double ljk = Lx[pj];
        for (int p = pj; p < Lp[k + 1]; p++) w[Li[p]] -= Lx[p] * ljk;
        next[k] = pj + 1;
        if (next[k] < Lp[k + 1]) {
          int r = Li[next[k]];
          link[k] = head[r];
          head[r] = k;
        }
        k = nk;
      }
      double d = w[j];
      if (!(d > 0) || !std::isfinite(d)) return -4;
      double ljj = std::sqrt(d);
      Lx[Lp[j]] = ljj;
      for (int p = Lp[j] + 1; p < Lp[j + 1]; p++) Lx[p] = w[Li[p]] / ljj;
      for (int p = Lp[j]; p < Lp[j + 1]; p++) w[Li[p]] = 0.0;
      if (Lp[j + 1] > Lp[j] + 1) {
        int r = Li[Lp[j] + 1];
        link[j] = head[r];
        head[r] = j;
      }
    }
    return 0;
  }

  void solve(const double* b, double* x) {
    std::vector<double>& y = work;
    for (int k = 0; k < n; k++) y[k] = b[perm[k]];
    for (int j = 0; j < n; j++) {
      y[j] /= Lx[Lp[j]];
      double yj = y[j];
      for (int p = Lp[j] + 1; p < Lp[j + 1]; p++) y[Li[p]] -= Lx[p] * yj;
    }
    for (int j = n - 1; j >= 0; j--) {
      double s = y[j];
      for (int p = Lp[j] + 1; p < Lp[j + 1]; p++) s -= Lx[p] * y[Li[p]];
      y[j] = s / Lx[Lp[j]];
    }
    for (int k = 0; k < n; k++) x[perm[k]] = y[k];
    std::fill(y.begin(), y.end(), 0.0);
  }

  long long nnz() const { return Lp.empty() ? 0 : Lp[n]; }
};

// Exact minimum-degree ordering on a (node) graph with explicit elimination cliques (bitsets); ties -> lowest index.
// Degrees count weighted neighbours (weight = DOFs per node). Not CSparse's AMD: an independent ordering for tests.
std::vector<int> min_degree(int n, const std::vector<int>& adjp, const std::vector<int>& adj, const std::vector<int>& w) {
  const int words = (n + 63) / 64;
  std::vector<uint64_t> bits((size_t)n * words, 0);
  auto B = [&](int v) { return &bits[(size_t)v * words]; };
  for (int v = 0; v < n; v++)
    for (int p = adjp[v]; p < adjp[v + 1]; p++) {
      int u = adj[p];
      if (u != v) B(v)[u >> 6] |= 1ull << (u & 63);
    }
  std::vector<char> done(n, 0);
  std::vector<long long> deg(n, 0);
  auto degree = [&](int v) {
    long long d = 0;
    const uint64_t* b = B(v);
    for (int k = 0; k < words; k++) {
      uint64_t x = b[k];
      while (x) { int t = __builtin_ctzll(x); d += w[k * 64 + t]; x &= x - 1; }
    }
    return d;
  };
  for (int v = 0; v < n; v++) deg[v] = degree(v);
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> nb;
  for (int it = 0; it < n; it++) {
    int best = -1;
    for (int v = 0; v < n; v++)
      if (!done[v] && (best < 0 || deg[v] < deg[best])) best = v;
    order.push_back(best);
    done[best] = 1;
    nb.clear();
    const uint64_t* bb = B(best);
    for (int k = 0; k < words; k++) {
      uint64_t x = bb[k];
      while (x) { int t = __builtin_ctzll(x); nb.push_back(k * 64 + t); x &= x - 1; }
    }
    for (int u : nb) {
      uint64_t* bu = B(u);
      for (int k = 0; k < words; k++) bu[k] |= bb[k];
      bu[u >> 6] &= ~(1ull << (u & 63));
      bu[best >> 6] &= ~(1ull << (best & 63));
    }
    for (int u : nb) deg[u] = degree(u);
  }
  return order;
}

// ---------------------------------------------------------------------------------------------- the solver
// TetLocator (TetMesh.cs:115-262): bucket grid of the tets' bounding boxes; Find tests a cell's tets in ascending order.
struct OLocator {
  const std::vector<int>* T = nullptr;
  const std::vector<V3>* X = nullptr;
  double cell = 1;
  V3 mn;
  int nx = 0, ny = 0, nz = 0;
  std::vector<int> start, items;
  int Cx(double v) const { return ClampIdx((v - mn.x) / cell, nx); }
  int Cy(double v) const { return ClampIdx((v - mn.y) / cell, ny); }
  int Cz(double v) const { return ClampIdx((v - mn.z) / cell, nz); }
  static int ClampIdx(double q, int n) {
    // Math.Clamp((int)q, 0, n - 1) for the range the grid covers
    if (!(q >= 0)) return 0;
    if (q >= n - 1) return n - 1;
    return (int)q;
  }
  static double Vol(const V3& a, const V3& b, const V3& c, const V3& d) { return (b - a).dot((c - a).cross(d - a)) / 6.0; }
  void Build(const std::vector<int>& tets, const std::vector<V3>& x, double c) {
    T = &tets; X = &x; cell = c;
    double x0 = 1e308, y0 = 1e308, z0 = 1e308, x1 = -1e308, y1 = -1e308, z1 = -1e308;
    for (const V3& p : x) {
      x0 = std::min(x0, p.x); x1 = std::max(x1, p.x);
      y0 = std::min(y0, p.y); y1 = std::max(y1, p.y);
      z0 = std::min(z0, p.z); z1 = std::max(z1, p.z);
    }
    mn = V3(x0 - c, y0 - c, z0 - c);
    nx = (int)((x1 - mn.x) / c) + 2;
    ny = (int)((y1 - mn.y) / c) + 2;
    nz = (int)((z1 - mn.z) / c) + 2;
    int ne = (int)(tets.size() / 4);
    std::vector<int> count((size_t)nx * ny * nz + 1, 0);
    auto each = [&](int e, bool fillPass, std::vector<int>& fill) {
      V3 lo = x[tets[4 * e]], hi = lo;
      for (int a = 1; a < 4; a++) {
        const V3& p = x[tets[4 * e + a]];
        lo = V3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
        hi = V3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
      }
      for (int k = Cz(lo.z); k <= Cz(hi.z); k++)
        for (int j = Cy(lo.y); j <= Cy(hi.y); j++)
          for (int i = Cx(lo.x); i <= Cx(hi.x); i++) {
            int cc = (k * ny + j) * nx + i;
            if (fillPass) items[fill[cc]++] = e; else count[cc + 1]++;
          }
    };
    std::vector<int> dummy;
    for (int e = 0; e < ne; e++) each(e, false, dummy);
    for (size_t i = 0; i + 1 < count.size(); i++) count[i + 1] += count[i];
    start = count;
    items.assign(count.back(), 0);
    std::vector<int> fill = count;
    for (int e = 0; e < ne; e++) each(e, true, fill);
  }
  int Find(const V3& p, double* w) const {
    int c = (Cz(p.z) * ny + Cy(p.y)) * nx + Cx(p.x);
    const std::vector<int>& t = *T;
    const std::vector<V3>& x = *X;
    for (int k = start[c]; k < start[c + 1]; k++) {
      int e = items[k];
      V3 a = x[t[4 * e]], b = x[t[4 * e + 1]], cc = x[t[4 * e + 2]], d = x[t[4 * e + 3]];
      double v = Vol(a, b, cc, d);
      if (v <= 0) continue;
      double b0 = Vol(p, b, cc, d) / v, b1 = Vol(a, p, cc, d) / v, b2 = Vol(a, b, p, d) / v;
      double b3 = 1 - b0 - b1 - b2;
      if (b0 >= -1e-9 && b1 >= -1e-9 && b2 >= -1e-9 && b3 >= -1e-9) {
        if (w) { w[0] = b0; w[1] = b1; w[2] = b2; w[3] = b3; }
        return e;
      }
    }
    return -1;
  }
  bool Contains(const V3& p) const { return Find(p, nullptr) >= 0; }
  V3 Interpolate(const V3& p, const std::vector<V3>& field) const {
    double w[4];
    int e = Find(p, w);
    const std::vector<int>& t = *T;
    if (e >= 0) return w[0] * field[t[4 * e]] + w[1] * field[t[4 * e + 1]] + w[2] * field[t[4 * e + 2]] + w[3] * field[t[4 * e + 3]];
    int best = 0;
    double dmin = 1e308;
    for (size_t n = 0; n < X->size(); n++) {
      V3 d = (*X)[n] - p;
      double dd = d.x * d.x + d.y * d.y + d.z * d.z;
      if (dd < dmin) { dmin = dd; best = (int)n; }
    }
    return field[best];
  }
  int FindNearest(const V3& p) const {
    int e = Find(p, nullptr);
    if (e >= 0) return e;
    int c = (Cz(p.z) * ny + Cy(p.y)) * nx + Cx(p.x);
    const std::vector<int>& t = *T;
    const std::vector<V3>& x = *X;
    double best = 1e308;
    int bestE = -1;
    auto test = [&](int q) {
      V3 cen = 0.25 * (x[t[4 * q]] + x[t[4 * q + 1]] + x[t[4 * q + 2]] + x[t[4 * q + 3]]);
      V3 d = cen - p;
      double dd = d.x * d.x + d.y * d.y + d.z * d.z;
      if (dd < best) { best = dd; bestE = q; }
    };
    for (int k = start[c]; k < start[c + 1]; k++) test(items[k]);
    if (bestE >= 0) return bestE;
    for (int q = 0; q < (int)(t.size() / 4); q++) test(q);
    return bestE;
  }
};

// IsosurfaceStuffing.Stuff / StuffLattice (IsosurfaceStuffing.cs:17-88, :115-294) with the inside test In(p).
template <typename InFn>
void OStuff(InFn In, const V3& mn, const V3& mx, double h, bool symX, bool symY, std::vector<V3>& outX,
            std::vector<int>& outT) {
  const double alphaLong = 0.24999, alphaShort = 0.41189;
  V3 off(0.31 * h, 0.37 * h, 0.17 * h);
  int i0 = (int)std::floor(mn.x / h) - 1, i1 = (int)std::ceil(mx.x / h) + 1;
  int j0 = (int)std::floor(mn.y / h) - 1, j1 = (int)std::ceil(mx.y / h) + 1;
  int k0 = (int)std::floor(mn.z / h) - 1, k1 = (int)std::ceil(mx.z / h) + 1;
  int ni = i1 - i0, nj = j1 - j0, nk = k1 - k0;
  int nCorner = (ni + 1) * (nj + 1) * (nk + 1);
  auto Corner = [&](int i, int j, int k) { return ((k - k0) * (nj + 1) + (j - j0)) * (ni + 1) + (i - i0); };
  auto Center = [&](int i, int j, int k) { return nCorner + ((k - k0) * nj + (j - j0)) * ni + (i - i0); };
  std::vector<V3> pos;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) pos.push_back(V3(i * h, j * h, k * h) + off);
  for (int k = k0; k < k1; k++)
    for (int j = j0; j < j1; j++)
      for (int i = i0; i < i1; i++) pos.push_back(V3((i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h) + off);
  std::vector<int> lat;
  auto AddPair = [&](int c1, int c2, int a, int b, int c, int d) {
    int q[16] = {c1, c2, a, b, c1, c2, b, c, c1, c2, c, d, c1, c2, d, a};
    lat.insert(lat.end(), q, q + 16);
  };
  for (int k = k0; k < k1; k++)
    for (int j = j0; j < j1; j++)
      for (int i = i0; i < i1; i++) {
        int c = Center(i, j, k);
        if (i + 1 < i1) AddPair(c, Center(i + 1, j, k), Corner(i + 1, j, k), Corner(i + 1, j + 1, k), Corner(i + 1, j + 1, k + 1), Corner(i + 1, j, k + 1));
        if (j + 1 < j1) AddPair(c, Center(i, j + 1, k), Corner(i, j + 1, k), Corner(i + 1, j + 1, k), Corner(i + 1, j + 1, k + 1), Corner(i, j + 1, k + 1));
        if (k + 1 < k1) AddPair(c, Center(i, j, k + 1), Corner(i, j, k + 1), Corner(i + 1, j, k + 1), Corner(i + 1, j + 1, k + 1), Corner(i, j + 1, k + 1));
      }
  std::vector<int> sign(pos.size());
  for (size_t v = 0; v < pos.size(); v++) sign[v] = In(pos[v]) ? -1 : 1;
  auto isCorner = [&](int v) { return v < nCorner; };
  auto cross_ = [&](int u, int w) { return (sign[u] < 0 && sign[w] > 0) || (sign[u] > 0 && sign[w] < 0); };
  // cuts: insertion-ordered
  std::map<std::pair<int, int>, int> cutIndex;
  std::vector<std::pair<int, int>> cutKeys;
  std::vector<V3> cutPos;
  const int ea[6] = {0, 0, 0, 1, 1, 2}, eb[6] = {1, 2, 3, 2, 3, 3};
  for (size_t t = 0; t < lat.size(); t += 4)
    for (int q = 0; q < 6; q++) {
      int u = lat[t + ea[q]], w = lat[t + eb[q]];
      if (!cross_(u, w)) continue;
      std::pair<int, int> key = u < w ? std::make_pair(u, w) : std::make_pair(w, u);
      if (cutIndex.count(key)) continue;
      V3 pin = sign[u] < 0 ? pos[u] : pos[w], pout = sign[u] < 0 ? pos[w] : pos[u];
      for (int it = 0; it < 16; it++) {
        V3 m = 0.5 * (pin + pout);
        if (In(m)) pin = m; else pout = m;
      }
      V3 cp = 0.5 * (pin + pout);
      if (symX && std::fabs(cp.x) < 1e-4 * h) cp.x = 0;
      if (symY && std::fabs(cp.y) < 1e-4 * h) cp.y = 0;
      cutIndex[key] = (int)cutPos.size();
      cutKeys.push_back(key);
      cutPos.push_back(cp);
    }
  std::map<int, std::pair<V3, double>> warpTo;
  for (size_t q = 0; q < cutKeys.size(); q++) {
    int u = cutKeys[q].first, w = cutKeys[q].second;
    const V3& c = cutPos[q];
    double len = (pos[w] - pos[u]).len();
    double alpha = isCorner(u) == isCorner(w) ? alphaLong : alphaShort;
    int vv[2] = {u, w};
    for (int v : vv) {
      double d = (c - pos[v]).len();
      if (sign[v] == 0 || d >= alpha * len) continue;
      auto it = warpTo.find(v);
      if (it == warpTo.end() || d < it->second.second) warpTo[v] = std::make_pair(c, d);
    }
  }
  for (auto& kv : warpTo) { pos[kv.first] = kv.second.first; sign[kv.first] = 0; }
  std::map<std::pair<int, int>, int> cutId;
  auto CutVertex = [&](int u, int w) {
    std::pair<int, int> key = u < w ? std::make_pair(u, w) : std::make_pair(w, u);
    auto it = cutId.find(key);
    if (it != cutId.end()) return it->second;
    pos.push_back(cutPos[cutIndex.at(key)]);
    sign.push_back(0);
    int id = (int)pos.size() - 1;
    cutId[key] = id;
    return id;
  };
  double minVol = 1e-6 * h * h * h;
  std::vector<int> outTets;
  auto Emit = [&](int a, int b, int c, int d) {
    double v = OLocator::Vol(pos[a], pos[b], pos[c], pos[d]);
    if (std::fabs(v) < minVol) return;
    if (v < 0) std::swap(c, d);
    int q[4] = {a, b, c, d};
    outTets.insert(outTets.end(), q, q + 4);
  };
  auto ConeFace = [&](const std::vector<int>& f, int apex) {
    if (f.size() < 3 || std::find(f.begin(), f.end(), apex) != f.end()) return;
    int s = 0;
    for (size_t q = 1; q < f.size(); q++) if (f[q] < f[s]) s = (int)q;
    int n = (int)f.size();
    for (int q = 1; q < n - 1; q++) Emit(apex, f[s], f[(s + q) % n], f[(s + q + 1) % n]);
  };
  auto OrderPlanar = [&](std::vector<int>& f, const std::vector<int>& all) {
    V3 c;
    for (int x : f) c = c + pos[x];
    c = (1.0 / f.size()) * c;
    V3 inner;
    int cnt = 0;
    for (int x : all) if (sign[x] < 0) { inner = inner + pos[x]; cnt++; }
    V3 nrm = (c - (1.0 / std::max(cnt, 1)) * inner).normalized();
    V3 u = (std::fabs(nrm.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0)).cross(nrm).normalized();
    V3 w = nrm.cross(u);
    std::vector<std::pair<double, int>> key;
    for (int x : f) key.push_back(std::make_pair(std::atan2((pos[x] - c).dot(w), (pos[x] - c).dot(u)), x));
    // ties keep their order (the reference's List.Sort leaves the order of equal angles unspecified)
    std::stable_sort(key.begin(), key.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
    for (size_t q = 0; q < f.size(); q++) f[q] = key[q].second;
  };
  const int faces[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  for (size_t t = 0; t < lat.size(); t += 4) {
    int v[4] = {lat[t], lat[t + 1], lat[t + 2], lat[t + 3]};
    int neg = 0, posCount = 0;
    for (int x : v) { if (sign[x] < 0) neg++; else if (sign[x] > 0) posCount++; }
    if (neg == 0) continue;
    if (posCount == 0) { Emit(v[0], v[1], v[2], v[3]); continue; }
    std::vector<int> poly, iface;
    for (int x : v) { if (sign[x] <= 0) poly.push_back(x); if (sign[x] == 0) iface.push_back(x); }
    for (int q = 0; q < 6; q++) {
      int u = v[ea[q]], w = v[eb[q]];
      if (cross_(u, w)) { int c = CutVertex(u, w); poly.push_back(c); iface.push_back(c); }
    }
    int apex = *std::min_element(poly.begin(), poly.end());
    for (const auto& f : faces) {
      std::vector<int> face;
      for (int q = 0; q < 3; q++) {
        int a = v[f[q]], b = v[f[(q + 1) % 3]];
        if (sign[a] <= 0) face.push_back(a);
        if (cross_(a, b)) face.push_back(CutVertex(a, b));
      }
      ConeFace(face, apex);
    }
    if (iface.size() >= 3) { OrderPlanar(iface, poly); ConeFace(iface, apex); }
  }
  std::map<int, int> map;
  outX.clear();
  outT.assign(outTets.size(), 0);
  for (size_t q = 0; q < outTets.size(); q++) {
    int old = outTets[q];
    auto it = map.find(old);
    int id;
    if (it == map.end()) { id = (int)outX.size(); map[old] = id; outX.push_back(pos[old]); }
    else id = it->second;
    outT[q] = id;
  }
}

struct Oracle {
  OrcParams P;
  Law law;
  Die bottom, top;
  int nn = 0, ne = 0, ndof = 0;
  std::vector<int> tets;
  std::vector<V3> x, vg, vgPrev;
  std::vector<char> onX0, onY0;
  // TetMesh (TetMesh.cs:9-64)
  std::vector<int> ntStart, ntList, bfaces, bowner;
  std::vector<char> fsym, isSurf;
  // state
  std::vector<double> eps, dmg, elDev, elSig, elRate, elMean, grad, vol, hLocal;
  std::vector<int> contact, nFix, dieAxis;
  std::vector<char> fixedv, release;
  std::vector<double> fixedVal, kFric, aFric;
  std::vector<V3> ax0, ax1, ax2;
  // system
  std::vector<int> colPtr, rowIdx, nbrStart, nbr;
  std::vector<double> val, kvol, v, r, dv, vTrial, rvol, vNode;
  std::vector<int> gStart, gNode;
  std::vector<V3> gCoef;
  std::vector<double> eR, eK, eEnergy, ePower;
  std::vector<int> eNbr;
  std::vector<double> pr, pz, pp, pq;
  // constants (Forging3DSolver.cs:144-150)
  double rateRef, rateCut, sigRef, kPen, u0, gap, gap0, ds, time_ = 0;
  bool warm = false, needAnalyze = true, haveFactor = false;
  int pcgFailures = 0, pcgSkip = 0;
  double lastRel = 0;
  // linear backend
  SparseChol chol;
  std::unique_ptr<PardisoDL> pardiso;
  std::vector<int> userPerm;
  bool useUserPerm = false;
  // stats of the last SolveVelocity
  int statFactor = 0, statPcgIts = 0, statPcgCalls = 0, statSolves = 0, statAssemblies = 0;
  std::string err;

  int Dof(int n, int c) const { return 3 * n + c; }
  V3 Axis(int n, int k) const { return k == 0 ? ax0[n] : k == 1 ? ax1[n] : ax2[n]; }
  V3 ToLocal(int n, const V3& g) const { return V3(g.dot(ax0[n]), g.dot(ax1[n]), g.dot(ax2[n])); }
  V3 ToGlobal(int n, double l0, double l1, double l2) const { return l0 * ax0[n] + l1 * ax1[n] + l2 * ax2[n]; }
  V3 GlobalVelocity(const double* vv, int n) const { return ToGlobal(n, vv[3 * n], vv[3 * n + 1], vv[3 * n + 2]); }
  const Die& Surface(int c) const { return c == TOP ? top : bottom; }
  double Offset(int c, double g) const { return c == TOP ? g : 0.0; }
  V3 DieVelocity(int c) const { return c == TOP ? V3(0, 0, -P.ram) : V3(0, 0, 0); }
  V3 Sym(const V3& p, int n) const { return V3(onX0[n] ? 0 : p.x, onY0[n] ? 0 : p.y, p.z); }

  // TetMesh constructor (TetMesh.cs:9-64)
  void BuildTopology() {
    ntStart.assign(nn + 1, 0);
    for (int q : tets) ntStart[q + 1]++;
    for (int i = 0; i < nn; i++) ntStart[i + 1] += ntStart[i];
    std::vector<int> fill(ntStart.begin(), ntStart.end() - 1);
    ntList.resize(tets.size());
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < 4; a++) ntList[fill[tets[4 * e + a]]++] = e;
    struct F { int A, B, C, owner, count; long long order; };
    std::unordered_map<long long, F> faces;
    faces.reserve(4 * ne);
    long long ord = 0;
    for (int e = 0; e < ne; e++) {
      int a = tets[4 * e], b = tets[4 * e + 1], c = tets[4 * e + 2], d = tets[4 * e + 3];
      int fs[4][3] = {{a, c, b}, {a, b, d}, {a, d, c}, {b, c, d}};
      for (auto& f : fs) {
        int s[3] = {f[0], f[1], f[2]};
        std::sort(s, s + 3);
        long long key = ((long long)s[0] * nn + s[1]) * nn + s[2];
        auto it = faces.find(key);
        if (it == faces.end()) faces.emplace(key, F{f[0], f[1], f[2], e, 1, ord++});
        else it->second.count++;
      }
    }
    std::vector<F> bl;
    for (auto& kv : faces) {
      if (kv.second.count > 2) { err = "non-manifold mesh"; }
      if (kv.second.count == 1) bl.push_back(kv.second);
    }
    std::sort(bl.begin(), bl.end(), [](const F& p, const F& q) { return p.order < q.order; });
    bfaces.clear();
    bowner.clear();
    fsym.clear();
    for (auto& f : bl) {
      bfaces.push_back(f.A); bfaces.push_back(f.B); bfaces.push_back(f.C);
      bowner.push_back(f.owner);
      fsym.push_back((onX0[f.A] && onX0[f.B] && onX0[f.C]) || (onY0[f.A] && onY0[f.B] && onY0[f.C]));
    }
    isSurf.assign(nn, 0);
    for (size_t k = 0; k < fsym.size(); k++)
      if (!fsym[k]) for (int a = 0; a < 3; a++) isSurf[bfaces[3 * k + a]] = 1;
  }

  // LocalLengths (Forging3DSolver.cs:442-466)
  void LocalLengths() {
    std::vector<double> sum(nn, 0.0);
    std::vector<int> cnt(nn, 0);
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) {
          int i = tets[4 * e + a], j = tets[4 * e + b];
          double l = (x[i] - x[j]).len();
          sum[i] += l; sum[j] += l; cnt[i]++; cnt[j]++;
        }
    hLocal.resize(nn);
    for (int n = 0; n < nn; n++) hLocal[n] = cnt[n] > 0 ? std::min(sum[n] / cnt[n], P.h) : P.h;
  }

  // BuildPattern (Forging3DSolver.cs:469-521)
  void BuildPattern() {
    std::vector<std::vector<int>> sets(nn);
    std::vector<int> patch;
    for (int n = 0; n < nn; n++) {
      patch.clear();
      for (int q = ntStart[n]; q < ntStart[n + 1]; q++) {
        int e = ntList[q];
        for (int a = 0; a < 4; a++) patch.push_back(tets[4 * e + a]);
      }
      for (int i : patch) for (int j : patch) sets[i].push_back(j);
      if (n % 64 == 63) for (int i = 0; i < nn; i++) if (sets[i].size() > 4096) {
        std::sort(sets[i].begin(), sets[i].end());
        sets[i].erase(std::unique(sets[i].begin(), sets[i].end()), sets[i].end());
      }
    }
    nbrStart.assign(nn + 1, 0);
    nbr.clear();
    for (int n = 0; n < nn; n++) {
      auto& s = sets[n];
      std::sort(s.begin(), s.end());
      s.erase(std::unique(s.begin(), s.end()), s.end());
      nbr.insert(nbr.end(), s.begin(), s.end());
      nbrStart[n + 1] = (int)nbr.size();
      std::vector<int>().swap(s);
    }
    colPtr.assign(ndof + 1, 0);
    for (int j = 0; j < nn; j++)
      for (int c = 0; c < 3; c++) colPtr[3 * j + c + 1] = colPtr[3 * j + c] + 3 * (nbrStart[j + 1] - nbrStart[j]);
    rowIdx.resize(colPtr[ndof]);
    for (int j = 0; j < nn; j++)
      for (int c = 0; c < 3; c++) {
        int p = colPtr[3 * j + c];
        for (int q = nbrStart[j]; q < nbrStart[j + 1]; q++) {
          int i = nbr[q];
          rowIdx[p++] = 3 * i; rowIdx[p++] = 3 * i + 1; rowIdx[p++] = 3 * i + 2;
        }
      }
    val.assign(rowIdx.size(), 0.0);
    kvol.assign(rowIdx.size(), 0.0);
    needAnalyze = true;
    haveFactor = false;
    // _eNbr (Forging3DSolver.cs:1249-1258)
    eNbr.resize(16 * (size_t)ne);
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) eNbr[16 * e + 4 * a + b] = NbrIndex(tets[4 * e + b], tets[4 * e + a]);
  }

  int NbrIndex(int j, int i) const {
    const int* b = &nbr[nbrStart[j]];
    const int* e = &nbr[0] + nbrStart[j + 1];
    const int* f = std::lower_bound(b, e, i);
    return (f != e && *f == i) ? (int)(f - b) : -1;
  }
  // Pos (Forging3DSolver.cs:524)
  int Pos(int i, int ci, int j, int cj) const { return colPtr[3 * j + cj] + 3 * NbrIndex(j, i) + ci; }

  // PrepareStep (Forging3DSolver.cs:660-758); returns non-zero on an inverted tet (MeshDistortionException)
  int PrepareStep() {
    std::fill(fixedv.begin(), fixedv.end(), 0);
    std::fill(fixedVal.begin(), fixedVal.end(), 0.0);
    std::fill(release.begin(), release.end(), 0);
    for (int n = 0; n < nn; n++) {
      V3 dirs[3], e[3];
      double vals[3], ev[3];
      int k = 0;
      dieAxis[n] = -1;
      if (onX0[n]) { dirs[k] = V3(1, 0, 0); vals[k] = 0; k++; }
      if (onY0[n]) { dirs[k] = V3(0, 1, 0); vals[k] = 0; k++; }
      int dieIdx = -1;
      if (contact[n] != FREE) {
        V3 nrm;
        Surface(contact[n]).project(x[n], Offset(contact[n], gap), &nrm);
        dirs[k] = nrm;
        vals[k] = DieVelocity(contact[n]).dot(nrm);
        dieIdx = k;
        k++;
      }
      int m = 0;
      for (int q = 0; q < k; q++) {
        V3 d = dirs[q];
        double vv = vals[q];
        for (int s = 0; s < m; s++) {
          double proj = d.dot(e[s]);
          d = d - proj * e[s];
          vv -= proj * ev[s];
        }
        double len = d.len();
        if (len < 0.2) continue;
        e[m] = (1 / len) * d;
        ev[m] = vv / len;
        if (q == dieIdx) dieAxis[n] = m;
        m++;
      }
      if (m == 0) { e[0] = V3(1, 0, 0); e[1] = V3(0, 1, 0); e[2] = V3(0, 0, 1); }
      else if (m == 1) {
        V3 helper = std::fabs(e[0].x) < 0.8 ? V3(1, 0, 0) : V3(0, 1, 0);
        e[1] = e[0].cross(helper).normalized();
        e[2] = e[0].cross(e[1]);
      } else if (m == 2) {
        e[2] = e[0].cross(e[1]).normalized();
      }
      ax0[n] = e[0]; ax1[n] = e[1]; ax2[n] = e[2];
      nFix[n] = m;
      for (int q = 0; q < m; q++) { fixedv[Dof(n, q)] = 1; fixedVal[Dof(n, q)] = ev[q]; }
    }
    for (int el = 0; el < ne; el++) {
      V3 p0 = x[tets[4 * el]];
      V3 a = x[tets[4 * el + 1]] - p0, b = x[tets[4 * el + 2]] - p0, c = x[tets[4 * el + 3]] - p0;
      double det = a.dot(b.cross(c));
      if (!(det > 0)) { err = "Tetrahedron " + std::to_string(el) + " is inverted"; return -3; }
      V3 g1 = (1 / det) * b.cross(c), g2 = (1 / det) * c.cross(a), g3 = (1 / det) * a.cross(b);
      V3 g0 = -(g1 + g2 + g3);
      V3 gs[4] = {g0, g1, g2, g3};
      for (int q = 0; q < 4; q++) {
        grad[12 * el + 3 * q] = gs[q].x; grad[12 * el + 3 * q + 1] = gs[q].y; grad[12 * el + 3 * q + 2] = gs[q].z;
      }
      vol[el] = det / 6;
    }
    BuildAnp();
    BuildFriction();
    LoadVelocityGuess();
    return 0;
  }

  // BuildAnp (Forging3DSolver.cs:761-823)
  void BuildAnp() {
    std::fill(vNode.begin(), vNode.end(), 0.0);
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < 4; a++) vNode[tets[4 * e + a]] += vol[e] / 4;
    gStart.assign(nn + 1, 0);
    gNode.clear();
    gCoef.clear();
    std::unordered_map<int, int> map;
    for (int n = 0; n < nn; n++) {
      map.clear();
      for (int q = ntStart[n]; q < ntStart[n + 1]; q++) {
        int e = ntList[q];
        double w = vol[e] / 4 / vNode[n];
        for (int a = 0; a < 4; a++) {
          int m = tets[4 * e + a];
          V3 gc = w * V3(grad[12 * e + 3 * a], grad[12 * e + 3 * a + 1], grad[12 * e + 3 * a + 2]);
          auto it = map.find(m);
          if (it != map.end()) gCoef[it->second] = gCoef[it->second] + gc;
          else { map[m] = (int)gNode.size(); gNode.push_back(m); gCoef.push_back(gc); }
        }
      }
      gStart[n + 1] = (int)gNode.size();
    }
    std::fill(kvol.begin(), kvol.end(), 0.0);
    for (int n = 0; n < nn; n++) {
      double kv = kPen * vNode[n];
      for (int p = gStart[n]; p < gStart[n + 1]; p++) {
        int i = gNode[p];
        V3 ci = ToLocal(i, gCoef[p]);
        double cia[3] = {ci.x, ci.y, ci.z};
        for (int q = gStart[n]; q < gStart[n + 1]; q++) {
          int j = gNode[q];
          V3 cj = ToLocal(j, gCoef[q]);
          double cja[3] = {cj.x, cj.y, cj.z};
          int k = NbrIndex(j, i);
          for (int cc = 0; cc < 3; cc++) {
            int pp0 = colPtr[3 * j + cc] + 3 * k;
            for (int rr = 0; rr < 3; rr++) kvol[pp0 + rr] += kv * (cia[rr] * cja[cc]);
          }
        }
      }
    }
  }

  // BuildFriction (Forging3DSolver.cs:829-850)
  void BuildFriction() {
    std::fill(aFric.begin(), aFric.end(), 0.0);
    std::vector<double> ksum(nn, 0.0);
    std::vector<int> kcnt(nn, 0);
    for (size_t k = 0; k < fsym.size(); k++) {
      if (fsym[k]) continue;
      int a = bfaces[3 * k], b = bfaces[3 * k + 1], c = bfaces[3 * k + 2];
      double area = 0.5 * (x[b] - x[a]).cross(x[c] - x[a]).len();
      double kk = elSig[bowner[k]] / std::sqrt(3.0);
      int ns[3] = {a, b, c};
      for (int n : ns) { aFric[n] += area / 3; ksum[n] += kk; kcnt[n]++; }
    }
    for (int n = 0; n < nn; n++) kFric[n] = kcnt[n] > 0 ? ksum[n] / kcnt[n] : 0;
  }

  // LoadVelocityGuess (Forging3DSolver.cs:854-869)
  void LoadVelocityGuess() {
    bool trend = P.extrapolate_start && warm && (int)vgPrev.size() == nn;
    for (int n = 0; n < nn; n++) {
      V3 g = trend ? (2 * vg[n] - vgPrev[n]) : vg[n];
      V3 l = ToLocal(n, g);
      v[Dof(n, 0)] = l.x; v[Dof(n, 1)] = l.y; v[Dof(n, 2)] = l.z;
    }
    for (int d = 0; d < ndof; d++) if (fixedv[d]) v[d] = fixedVal[d];
  }

  // Kernel (Forging3DSolver.cs:1312-1400): element residual/stiffness in global axes
  void Kernel(int e, bool matrix, bool newton, const double* Ve, double* Re, double* Ke, double& energy, double& power) {
    double B[72] = {0}, MB[72], G[12];
    double volE = vol[e];
    int o = 12 * e;
    for (int a = 0; a < 4; a++) {
      double bx = grad[o + 3 * a], by = grad[o + 3 * a + 1], bz = grad[o + 3 * a + 2];
      int c = 3 * a;
      B[0 * 12 + c] = bx; B[1 * 12 + c + 1] = by; B[2 * 12 + c + 2] = bz;
      B[3 * 12 + c] = by; B[3 * 12 + c + 1] = bx;
      B[4 * 12 + c + 1] = bz; B[4 * 12 + c + 2] = by;
      B[5 * 12 + c] = bz; B[5 * 12 + c + 2] = bx;
    }
    double ep[6];
    for (int i = 0; i < 6; i++) {
      double s = 0;
      for (int k = 0; k < 12; k++) s += B[i * 12 + k] * Ve[k];
      ep[i] = s;
    }
    double ev3 = (ep[0] + ep[1] + ep[2]) / 3;
    double d0 = ep[0] - ev3, d1 = ep[1] - ev3, d2 = ep[2] - ev3;
    double rate = std::sqrt(std::max(2.0 / 3.0 * (d0 * d0 + d1 * d1 + d2 * d2) + (ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5]) / 3.0, 0.0));
    double me[6] = {2.0 / 3.0 * d0, 2.0 / 3.0 * d1, 2.0 / 3.0 * d2, ep[3] / 3, ep[4] / 3, ep[5] / 3};
    for (int k = 0; k < 12; k++) {
      double s = 0;
      for (int i = 0; i < 6; i++) s += B[i * 12 + k] * me[i];
      G[k] = s;
    }
    double strain = eps[e];
    double sigCut = law.stress(strain, rateCut);
    double mu0 = sigCut / rateCut, mExp = law.M;
    double mu, sig, wDens, coefN;
    if (rate > rateCut) {
      sig = law.stress(strain, rate);
      mu = sig / rate;
      wDens = 0.5 * sigCut * rateCut + (sig * rate - sigCut * rateCut) / (mExp + 1);
      coefN = (mExp - 1) * mu / (rate * rate);
    } else {
      mu = mu0;
      sig = mu0 * rate;
      wDens = 0.5 * mu0 * rate * rate;
      coefN = 0;
    }
    energy = volE * wDens;
    power = volE * sig * rate;
    for (int k = 0; k < 12; k++) Re[k] = volE * mu * G[k];
    elSig[e] = sig;
    elRate[e] = rate;
    for (int i = 0; i < 6; i++) elDev[6 * e + i] = mu * me[i];
    if (!matrix) return;
    for (int k = 0; k < 12; k++) {
      double tr = (B[k] + B[12 + k] + B[24 + k]) / 3;
      MB[k] = 2.0 / 3.0 * (B[k] - tr);
      MB[12 + k] = 2.0 / 3.0 * (B[12 + k] - tr);
      MB[24 + k] = 2.0 / 3.0 * (B[24 + k] - tr);
      MB[36 + k] = B[36 + k] / 3;
      MB[48 + k] = B[48 + k] / 3;
      MB[60 + k] = B[60 + k] / 3;
    }
    double wm = volE * mu, wn = newton ? volE * coefN * P.newton_blend : 0;
    for (int p = 0; p < 12; p++)
      for (int q = p; q < 12; q++) {
        double s = 0;
        for (int i = 0; i < 6; i++) s += B[i * 12 + p] * MB[i * 12 + q];
        double vv = wm * s + wn * G[p] * G[q];
        Ke[p * 12 + q] = vv;
        Ke[q * 12 + p] = vv;
      }
  }

  // ElementWork (Forging3DSolver.cs:1265-1310)
  void ElementWork(int e, bool matrix, bool newton, const double* vv) {
    double Ve[12], Re[12], Ke[144];
    for (int a = 0; a < 4; a++) {
      V3 g = GlobalVelocity(vv, tets[4 * e + a]);
      Ve[3 * a] = g.x; Ve[3 * a + 1] = g.y; Ve[3 * a + 2] = g.z;
    }
    Kernel(e, matrix, newton, Ve, Re, Ke, eEnergy[e], ePower[e]);
    for (int a = 0; a < 4; a++) {
      V3 rl = ToLocal(tets[4 * e + a], V3(Re[3 * a], Re[3 * a + 1], Re[3 * a + 2]));
      eR[12 * e + 3 * a] = rl.x; eR[12 * e + 3 * a + 1] = rl.y; eR[12 * e + 3 * a + 2] = rl.z;
    }
    if (!matrix) return;
    double blk[9];
    for (int a = 0; a < 4; a++) {
      int i = tets[4 * e + a];
      for (int b = 0; b < 4; b++) {
        int j = tets[4 * e + b];
        size_t o = 144 * (size_t)e + 9 * (4 * a + b);
        for (int rr = 0; rr < 3; rr++) {
          V3 ar = Axis(i, rr);
          for (int c = 0; c < 3; c++)
            blk[3 * rr + c] = ar.x * Ke[(3 * a) * 12 + 3 * b + c] + ar.y * Ke[(3 * a + 1) * 12 + 3 * b + c] + ar.z * Ke[(3 * a + 2) * 12 + 3 * b + c];
        }
        for (int rr = 0; rr < 3; rr++)
          for (int c = 0; c < 3; c++) {
            V3 bc = Axis(j, c);
            eK[o + 3 * rr + c] = blk[3 * rr] * bc.x + blk[3 * rr + 1] * bc.y + blk[3 * rr + 2] * bc.z;
          }
      }
    }
  }

  // Assemble (Forging3DSolver.cs:1115-1217); returns Pi(v), writes r, val (if matrix), power
  double Assemble(bool matrix, bool newton, const double* vv, double& power) {
    statAssemblies++;
    std::fill(rvol.begin(), rvol.end(), 0.0);
    for (int j = 0; j < ndof; j++) {
      double vj = vv[j];
      if (vj == 0) continue;
      for (int p = colPtr[j]; p < colPtr[j + 1]; p++) rvol[rowIdx[p]] += kvol[p] * vj;
    }
    double energy = 0;
    for (int d = 0; d < ndof; d++) energy += 0.5 * vv[d] * rvol[d];
    power = 2 * energy;
    std::copy(rvol.begin(), rvol.end(), r.begin());
    if (matrix) std::copy(kvol.begin(), kvol.end(), val.begin());
    if (matrix && eK.size() != 144 * (size_t)ne) eK.resize(144 * (size_t)ne);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) ElementWork(e, matrix, newton, vv);
    for (int e = 0; e < ne; e++) {
      energy += eEnergy[e];
      power += ePower[e];
      for (int a = 0; a < 4; a++) {
        int n = tets[4 * e + a];
        r[3 * n] += eR[12 * e + 3 * a];
        r[3 * n + 1] += eR[12 * e + 3 * a + 1];
        r[3 * n + 2] += eR[12 * e + 3 * a + 2];
      }
      if (!matrix) continue;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          int j = tets[4 * e + b], k = eNbr[16 * e + 4 * a + b];
          size_t o = 144 * (size_t)e + 9 * (4 * a + b);
          for (int cj = 0; cj < 3; cj++) {
            int p = colPtr[3 * j + cj] + 3 * k;
            for (int ci = 0; ci < 3; ci++) val[p + ci] += eK[o + 3 * ci + cj];
          }
        }
    }
    double m = P.friction;
    if (m > 0) {
      for (int n = 0; n < nn; n++) {
        if (contact[n] == FREE || nFix[n] >= 3) continue;
        double c = m * kFric[n] * aFric[n] * (2 / M_PI);
        if (c <= 0) continue;
        V3 vd = ToLocal(n, DieVelocity(contact[n]));
        double vda[3] = {vd.x, vd.y, vd.z};
        int k0 = nFix[n];
        double s[3] = {0, 0, 0};
        double s2 = 0;
        for (int q = k0; q < 3; q++) { s[q] = vv[3 * n + q] - vda[q]; s2 += s[q] * s[q]; }
        double sl = std::sqrt(s2), z = sl / u0, at = std::atan(z);
        energy += c * (sl * at - 0.5 * u0 * std::log(1 + z * z));
        power += c * at * sl;
        double gs = sl > 1e-12 * u0 ? at / sl : 1 / u0;
        for (int q = k0; q < 3; q++) r[3 * n + q] += c * gs * s[q];
        if (!matrix) continue;
        double tang = c / (u0 * (1 + z * z));
        for (int q = k0; q < 3; q++)
          for (int p = k0; p < 3; p++) {
            double kqp;
            if (newton && sl > 1e-12 * u0) kqp = c * gs * ((q == p ? 1 : 0) - s[q] * s[p] / s2) + tang * s[q] * s[p] / s2;
            else kqp = q == p ? c * gs : 0;
            val[Pos(n, q, n, p)] += kqp;
          }
      }
    }
    return energy;
  }

  static double Dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
  }
  static double Norm(const std::vector<double>& a) {
    double s = 0;
    for (double q : a) s += q * q;
    return std::sqrt(s);
  }
  static double Norm(const double* a, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * a[i];
    return std::sqrt(s);
  }

  // MatVec (Forging3DSolver.cs:1089-1098)
  void MatVec(const double* xv, double* y) {
    std::fill(y, y + ndof, 0.0);
    for (int j = 0; j < ndof; j++) {
      double vj = xv[j];
      if (vj == 0) continue;
      for (int p = colPtr[j]; p < colPtr[j + 1]; p++) y[rowIdx[p]] += val[p] * vj;
    }
  }

  // ISpdSolver slot: Analyze / Factorize / Solve
  int LinAnalyze() {
    if (P.linear == 1) {
      if (!pardiso) pardiso.reset(new PardisoDL());
      if (!pardiso->ok()) { err = "MKL PARDISO unavailable"; return -10; }
      return pardiso->analyze(ndof, colPtr.data(), rowIdx.data(), useUserPerm ? userPerm.data() : nullptr);
    }
    std::vector<int> perm;
    if (useUserPerm) perm = userPerm;
    else perm = NodeMinDegree();
    return chol.analyze(ndof, colPtr.data(), rowIdx.data(), perm.data(), nullptr);
  }
  int LinFactorize() {
    statFactor++;
    if (P.linear == 1) return pardiso->factorize(val.data());
    return chol.factorize(val.data());
  }
  int LinSolve(const double* b, double* xo) {
    statSolves++;
    if (P.linear == 1) return pardiso->solve(b, xo);
    chol.solve(b, xo);
    for (int i = 0; i < ndof; i++) if (!std::isfinite(xo[i])) return -4;
    return 0;
  }
  long long FactorNonZeros() const { return P.linear == 1 ? (pardiso ? pardiso->nnz_l : 0) : chol.nnz(); }

  std::vector<int> NodeMinDegree() {
    std::vector<int> adjp(nn + 1, 0), adj;
    for (int n = 0; n < nn; n++) {
      for (int q = nbrStart[n]; q < nbrStart[n + 1]; q++) if (nbr[q] != n) adj.push_back(nbr[q]);
      adjp[n + 1] = (int)adj.size();
    }
    std::vector<int> w(nn, 3);
    std::vector<int> ord = min_degree(nn, adjp, adj, w);
    std::vector<int> perm(ndof);
    for (int k = 0; k < nn; k++) for (int c = 0; c < 3; c++) perm[3 * k + c] = 3 * ord[k] + c;
    return perm;
  }

  // Pcg (Forging3DSolver.cs:1051-1087); returns iterations, or -1 if not converged, or < -1 on a solver error
  int Pcg(const double* b, double* xo, double tol, int maxIt) {
    double bnorm = Norm(b, ndof);
    if (bnorm == 0) { std::fill(xo, xo + ndof, 0.0); return 0; }
    int rc = LinSolve(b, xo);
    if (rc) return -100 + rc;
    MatVec(xo, pq.data());
    for (int i = 0; i < ndof; i++) pr[i] = b[i] - pq[i];
    if ((rc = LinSolve(pr.data(), pz.data()))) return -100 + rc;
    pp = pz;
    double rz = Dot(pr, pz);
    for (int it = 0; it < maxIt; it++) {
      statPcgIts++;
      if (Norm(pr) <= tol * bnorm) return it;
      MatVec(pp.data(), pq.data());
      double alpha = rz / Dot(pp, pq);
      for (int i = 0; i < ndof; i++) { xo[i] += alpha * pp[i]; pr[i] -= alpha * pq[i]; }
      if ((rc = LinSolve(pr.data(), pz.data()))) return -100 + rc;
      double rzNew = Dot(pr, pz);
      double beta = rzNew / rz;
      rz = rzNew;
      for (int i = 0; i < ndof; i++) pp[i] = pz[i] + beta * pp[i];
    }
    return Norm(pr) <= tol * bnorm ? maxIt : -1;
  }

  // SolveLinear (Forging3DSolver.cs:965-1019)
  int SolveLinear(const double* b, double* xo) {
    if (haveFactor && pcgSkip > 0) {
      pcgSkip--;
    } else if (haveFactor) {
      statPcgCalls++;
      int its = Pcg(b, xo, 1e-7, P.pcg_budget);
      if (its >= 0) { pcgFailures = 0; return 0; }
      if (its < -1) { haveFactor = false; err = "sparse solve failed"; return -5; }
      if (++pcgFailures >= 2) pcgSkip = std::min(64, 1 << (pcgFailures - 1));
    }
    if (needAnalyze) {
      int rc = LinAnalyze();
      if (rc) { err = "analyze failed"; return rc; }
      needAnalyze = false;
    }
    int rc = LinFactorize();
    if (rc) { haveFactor = false; err = "Sparse factorisation failed (not positive definite)"; return rc; }
    rc = LinSolve(b, xo);
    if (rc) { haveFactor = false; err = "Sparse solve failed"; return rc; }
    haveFactor = true;
    return 0;
  }

  // SolveVelocity (Forging3DSolver.cs:875-948); returns iterations (>= 1) or a negative error code
  int SolveVelocity(int& converged) {
    converged = 0;
    lastRel = INFINITY;
    statFactor = statPcgIts = statPcgCalls = statSolves = statAssemblies = 0;
    int it;
    int halvedInARow = 0;
    const int NewtonLimit = 20;
    for (it = 1; it <= P.max_iterations; it++) {
      bool newton = (warm || it > 2) && it <= NewtonLimit && !(P.adaptive_newton && halvedInARow >= 2);
      double pw;
      double e0 = Assemble(true, newton, v.data(), pw);
      double maxDiag = 0;
      for (int j = 0; j < ndof; j++)
        for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
          int i = rowIdx[p];
          if (fixedv[i] || fixedv[j]) val[p] = i == j ? 1 : 0;
          else if (i == j) maxDiag = std::max(maxDiag, val[p]);
        }
      double reg = 1e-10 * maxDiag;
      for (int j = 0; j < ndof; j++) if (!fixedv[j]) val[Pos(j / 3, j % 3, j / 3, j % 3)] += reg;
      for (int d = 0; d < ndof; d++) if (fixedv[d]) r[d] = 0;
      for (int d = 0; d < ndof; d++) r[d] = -r[d];
      int rc = SolveLinear(r.data(), dv.data());
      if (rc) return rc < 0 ? rc : -rc;
      double slope = 0;
      for (int d = 0; d < ndof; d++) slope -= r[d] * dv[d];
      double beta = 1;
      for (int ls = 0; ls < 8; ls++) {
        for (int d = 0; d < ndof; d++) vTrial[d] = v[d] + beta * dv[d];
        double e1 = Assemble(false, false, vTrial.data(), pw);
        if (e1 <= e0 + 1e-4 * beta * slope + 1e-11 * std::fabs(e0)) break;
        double curv = (e1 - e0 - slope * beta) / (beta * beta);
        beta = P.quadratic_ls && curv > 0 ? std::min(0.7 * beta, std::max(0.1 * beta, -slope / (2 * curv))) : 0.5 * beta;
      }
      v = vTrial;
      halvedInARow = newton && beta < 1 ? halvedInARow + 1 : 0;
      double rel = beta * Norm(dv) / std::max(Norm(v), 1e-30);
      lastRel = rel;
      if (std::isnan(rel)) { err = "Velocity solution became NaN"; return -6; }
      if (rel < P.tolerance) { converged = 1; return it; }
    }
    return it - 1;
  }

  // FinalEvaluation (Forging3DSolver.cs:1402-1451)
  double FinalEvaluation(double& reaction) {
    double power;
    Assemble(false, false, v.data(), power);
    vgPrev = vg;
    reaction = 0;
    double maxB = 0, maxT = 0;
    for (int n = 0; n < nn; n++) {
      vg[n] = GlobalVelocity(v.data(), n);
      if (contact[n] == FREE) continue;
      V3 rg = ToGlobal(n, r[3 * n], r[3 * n + 1], r[3 * n + 2]);
      if (contact[n] == TOP) reaction -= rg.z;
      if (dieAxis[n] < 0) continue;
      double comp = r[3 * n + dieAxis[n]];
      if (contact[n] == TOP) maxT = std::max(maxT, comp); else maxB = std::max(maxB, comp);
    }
    int keptB = 0, keptT = 0;
    for (int n = 0; n < nn; n++) {
      if (contact[n] == FREE || dieAxis[n] < 0) continue;
      double comp = r[3 * n + dieAxis[n]];
      double thr = 0.02 * (contact[n] == TOP ? maxT : maxB);
      if (comp < -thr) release[n] = 1;
      else if (contact[n] == TOP) keptT++;
      else keptB++;
    }
    if (keptB == 0 || keptT == 0)
      for (int n = 0; n < nn; n++)
        if ((keptT == 0 && contact[n] == TOP) || (keptB == 0 && contact[n] == BOTTOM)) release[n] = 0;
    std::vector<double> nodal(nn);
    for (int n = 0; n < nn; n++) {
      double ev = 0;
      for (int p = gStart[n]; p < gStart[n + 1]; p++) ev += gCoef[p].dot(vg[gNode[p]]);
      nodal[n] = kPen * ev;
    }
    for (int e = 0; e < ne; e++)
      elMean[e] = 0.25 * (nodal[tets[4 * e]] + nodal[tets[4 * e + 1]] + nodal[tets[4 * e + 2]] + nodal[tets[4 * e + 3]]);
    return power;
  }

  // AccumulateDamage (Forging3DSolver.cs:1844-1855)
  void AccumulateDamage(double dt) {
    for (int e = 0; e < ne; e++) {
      double sig = elSig[e];
      if (!(sig > 0) || elRate[e] <= rateCut) continue;
      int o = 6 * e;
      double m = elMean[e];
      double s1 = max3(elDev[o] + m, elDev[o + 1] + m, elDev[o + 2] + m, elDev[o + 3], elDev[o + 4], elDev[o + 5]);
      if (s1 > 0) dmg[e] += s1 / sig * elRate[e] * dt;
    }
  }

  static double Quality(const V3& a, const V3& b, const V3& c, const V3& d) {
    double l2 = ((b - a).dot(b - a) + (c - a).dot(c - a) + (d - a).dot(d - a) + (c - b).dot(c - b) + (d - b).dot(d - b) + (d - c).dot(d - c)) / 6.0;
    double V = (b - a).dot((c - a).cross(d - a)) / 6.0;
    return l2 > 0 ? 6 * std::sqrt(2.0) * V / std::pow(l2, 1.5) : -1;
  }

  // SurfaceNormals (Forging3DSolver.cs:1540-1562): area-weighted outward normals of the surface nodes
  std::vector<V3> SurfaceNormals(const std::vector<V3>& xx) const {
    std::vector<V3> nr(nn);
    for (size_t k = 0; k < fsym.size(); k++) {
      if (fsym[k]) continue;
      int a = bfaces[3 * k], b = bfaces[3 * k + 1], c = bfaces[3 * k + 2];
      V3 an = (xx[b] - xx[a]).cross(xx[c] - xx[a]);
      nr[a] = nr[a] + an;
      nr[b] = nr[b] + an;
      nr[c] = nr[c] + an;
    }
    for (int i = 0; i < nn; i++) {
      if (onX0[i]) nr[i].x = 0;
      if (onY0[i]) nr[i].y = 0;
      if (nr[i].dot(nr[i]) > 0) nr[i] = nr[i].normalized();
    }
    return nr;
  }

  // PushOut (:1570-1592): back to the die surface along the workpiece's inward normal (bisection), else the closest point
  V3 PushOut(const Die& die, double off, const V3& p, const V3& outward) const {
    if (outward.dot(outward) == 0 || die.clearance(p, off) >= 0) return die.project(p, off, nullptr);
    V3 d = -outward;
    double step = 0.1 * P.h, lo = 0, hi = -1;
    for (double t = step; t <= 2 * P.h + 1e-12; t += step) {
      if (die.clearance(p + t * d, off) >= 0) { hi = t; break; }
      lo = t;
    }
    if (hi < 0) return die.project(p, off, nullptr);
    for (int i = 0; i < 25; i++) {
      double mid = 0.5 * (lo + hi);
      if (die.clearance(p + mid * d, off) >= 0) hi = mid; else lo = mid;
    }
    return p + hi * d;
  }

  double TetQ(const std::vector<V3>& xx, int e) const {
    return Quality(xx[tets[4 * e]], xx[tets[4 * e + 1]], xx[tets[4 * e + 2]], xx[tets[4 * e + 3]]);
  }

  // WorstAround (:1681-1687)
  double WorstAround(const std::vector<V3>& xx, int n) const {
    double q = 1e308;
    for (int p2 = ntStart[n]; p2 < ntStart[n + 1]; p2++) q = std::min(q, TetQ(xx, ntList[p2]));
    return q;
  }

  // CornerAllowance (:1611-1612) in the normal pass and the first recovery pass (the lenient passes are not restated)
  int recoveryPass = 0;
  double CornerAllowance(double local) const { return (recoveryPass >= 1 ? 0.1 : 0.05) * local; }

  // ImprovePlacement (:1625-1679) with Tol = the local size (the strict passes)
  int ImprovePlacement(std::vector<V3>& xx, std::vector<int>& cc, const std::vector<V3>& pred, const std::vector<V3>& nrm,
                       double gapT, double qMin) {
    int total = 0;
    for (int sweep = 0; sweep < 4; sweep++) {
      int count = 0;
      for (int e = 0; e < ne; e++) {
        if (TetQ(xx, e) >= qMin) continue;
        for (int a = 0; a < 4; a++) {
          int n = tets[4 * e + a];
          if (cc[n] == FREE) continue;
          const Die& die = Surface(cc[n]);
          double off = Offset(cc[n], gapT);
          V3 cand[8];
          int k = 0;
          V3 closest = Sym(die.project(pred[n], off, nullptr), n);
          cand[k++] = closest;
          V3 push = Sym(PushOut(die, off, pred[n], nrm[n]), n);
          if ((push - pred[n]).len() <= 0.5 * hLocal[n]) cand[k++] = push;
          double pen = -die.clearance(pred[n], off);
          const double fs[4] = {0.0, 0.25, 0.5, 0.75};
          for (double f : fs)
            if (pen * (1 - f) <= CornerAllowance(hLocal[n])) cand[k++] = pred[n] + f * (closest - pred[n]);
          V3 old = xx[n];
          double best = WorstAround(xx, n);
          int pick = -1;
          for (int c = 0; c < k; c++) {
            if (cand[c].x == old.x && cand[c].y == old.y && cand[c].z == old.z) continue;
            xx[n] = cand[c];
            double q = WorstAround(xx, n);
            if (q > best + 1e-6) { best = q; pick = c; }
          }
          xx[n] = pick >= 0 ? cand[pick] : old;
          if (pick < 0) continue;
          const V3& cp = cand[pick];
          if (cp.x == pred[n].x && cp.y == pred[n].y && cp.z == pred[n].z && pen < 0) cc[n] = FREE;
          count++;
        }
      }
      total += count;
      if (count == 0) break;
    }
    return total;
  }

  // SmoothInterior (:1776-1838): interior nodes of tets below qTarget towards their neighbours' centroid
  int SmoothInterior(std::vector<V3>& xx, double qTarget, int sweeps) {
    int moves = 0;
    std::vector<char> bad(nn);
    for (int s2 = 0; s2 < sweeps; s2++) {
      std::fill(bad.begin(), bad.end(), 0);
      bool any = false;
      for (int e = 0; e < ne; e++) {
        if (TetQ(xx, e) >= qTarget) continue;
        for (int a = 0; a < 4; a++) bad[tets[4 * e + a]] = 1;
        any = true;
      }
      if (!any) break;
      int sweepMoves = 0;
      for (int n = 0; n < nn; n++) {
        if (!bad[n] || isSurf[n]) continue;
        V3 sum;
        int cnt = 0;
        double worst = 1e308;
        for (int p2 = ntStart[n]; p2 < ntStart[n + 1]; p2++) {
          int e = ntList[p2];
          for (int a = 0; a < 4; a++) {
            int m = tets[4 * e + a];
            if (m == n) continue;
            sum = sum + xx[m];
            cnt++;
          }
          worst = std::min(worst, TetQ(xx, e));
        }
        if (cnt == 0) continue;
        V3 old = xx[n];
        V3 c = (1.0 / cnt) * sum;
        const double lams[3] = {1.0, 0.5, 0.25};
        for (double lambda : lams) {
          V3 p = old + lambda * (c - old);
          xx[n] = V3(onX0[n] ? 0 : p.x, onY0[n] ? 0 : p.y, p.z);
          double q = 1e308;
          for (int p2 = ntStart[n]; p2 < ntStart[n + 1]; p2++) q = std::min(q, TetQ(xx, ntList[p2]));
          if (q > worst + 1e-6) { sweepMoves++; break; }
          xx[n] = old;
        }
      }
      moves += sweepMoves;
      if (sweepMoves == 0) break;
    }
    return moves;
  }

  int lastReverts = 0, lastSmoothMoves = 0;

  // ---------------------------------------------------------------------------------------------- remeshing
  double vol0 = 0;             // Volume() of the initial mesh (times the symmetry factor), :184
  bool symX = false, symY = false;
  int rmImprove = 0, rmReverts = 0, rmSmooth = 0, rmUntangle = 0;
  double rmQOld = 0, rmQNew = 0;

  static double TetVol(const V3& a, const V3& b, const V3& c, const V3& d) { return (b - a).dot((c - a).cross(d - a)) / 6.0; }
  double MinQuality(const std::vector<V3>& xx) const {
    double q = 1e308;
    for (int e = 0; e < ne; e++) q = std::min(q, TetQ(xx, e));
    return q;
  }

  // Allocate (:403-438): a new mesh (its symmetry-plane flags) and positions; every per-node and per-element array is
  // reset
  void Allocate(const std::vector<V3>& nx, const std::vector<int>& nt, const std::vector<char>& ox,
                const std::vector<char>& oy) {
    nn = (int)nx.size(); ne = (int)(nt.size() / 4); ndof = 3 * nn;
    tets = nt; x = nx;
    onX0 = ox; onY0 = oy;
    vg.assign(nn, V3()); vgPrev.clear();
    eps.assign(ne, 0); dmg.assign(ne, 0); elDev.assign(6 * (size_t)ne, 0);
    elSig.assign(ne, 0); elRate.assign(ne, 0); elMean.assign(ne, 0);
    grad.assign(12 * (size_t)ne, 0); vol.assign(ne, 0);
    contact.assign(nn, FREE); nFix.assign(nn, 0); dieAxis.assign(nn, -1);
    fixedv.assign(ndof, 0); fixedVal.assign(ndof, 0); release.assign(nn, 0);
    kFric.assign(nn, 0); aFric.assign(nn, 0);
    ax0.assign(nn, V3()); ax1.assign(nn, V3()); ax2.assign(nn, V3());
    v.assign(ndof, 0); r.assign(ndof, 0); dv.assign(ndof, 0); vTrial.assign(ndof, 0);
    rvol.assign(ndof, 0); vNode.assign(nn, 0);
    pr.assign(ndof, 0); pz.assign(ndof, 0); pp.assign(ndof, 0); pq.assign(ndof, 0);
    eR.assign(12 * (size_t)ne, 0); eEnergy.assign(ne, 0); ePower.assign(ne, 0);
    BuildTopology();
    LocalLengths();
    BuildPattern();
    needAnalyze = true; haveFactor = false; pcgFailures = pcgSkip = 0;
    chol = SparseChol();
    pardiso.reset();
    useUserPerm = false;
  }

  // CompensateVolume (:2108-2145): free-surface nodes move along their normals to restore the initial volume
  void CompensateVolume(std::vector<V3>& xx, const std::vector<int>& cc) const {
    double vv = 0;
    for (int e = 0; e < ne; e++) vv += TetVol(xx[tets[4 * e]], xx[tets[4 * e + 1]], xx[tets[4 * e + 2]], xx[tets[4 * e + 3]]);
    double deficit = vol0 / P.sym_factor - vv;
    std::vector<V3> nr(nn);
    for (size_t k = 0; k < fsym.size(); k++) {
      if (fsym[k]) continue;
      int a = bfaces[3 * k], b = bfaces[3 * k + 1], c = bfaces[3 * k + 2];
      V3 an = 0.5 * (xx[b] - xx[a]).cross(xx[c] - xx[a]);
      nr[a] = nr[a] + an; nr[b] = nr[b] + an; nr[c] = nr[c] + an;
    }
    double rate = 0;
    std::vector<char> use(nn, 0);
    for (int n = 0; n < nn; n++) {
      if (!isSurf[n] || cc[n] != FREE) { nr[n] = V3(); continue; }
      V3 q = nr[n];
      if (onX0[n]) q.x = 0;
      if (onY0[n]) q.y = 0;
      rate += q.len() / 3;
      nr[n] = q.normalized();
    }
    if (rate <= 0) return;
    double shift = std::min(std::max(deficit / rate, -P.h), P.h);
    for (int n = 0; n < nn; n++) {
      if (nr[n].x == 0 && nr[n].y == 0 && nr[n].z == 0) continue;
      V3 moved = xx[n] + shift * nr[n];
      if (bottom.clearance(moved, 0) > 0 && top.clearance(moved, gap) > 0) xx[n] = moved;
    }
  }

  // UntangleSurface (:1698-1772)
  int UntangleSurface(std::vector<V3>& xx, std::vector<int>& cc, const std::vector<V3>& stuffed, double qTarget, int sweeps) const {
    int moves = 0;
    std::vector<char> bad(nn);
    for (int s2 = 0; s2 < sweeps; s2++) {
      std::fill(bad.begin(), bad.end(), 0);
      bool any = false;
      for (int e = 0; e < ne; e++) {
        if (TetQ(xx, e) >= qTarget) continue;
        for (int a = 0; a < 4; a++) bad[tets[4 * e + a]] = 1;
        any = true;
      }
      if (!any) break;
      int sweepMoves = 0;
      for (int n = 0; n < nn; n++) {
        if (!bad[n] || !isSurf[n]) continue;
        V3 sum;
        int cnt = 0;
        for (int p2 = ntStart[n]; p2 < ntStart[n + 1]; p2++) {
          int e = ntList[p2];
          for (int a = 0; a < 4; a++) { int m = tets[4 * e + a]; if (m == n) continue; sum = sum + xx[m]; cnt++; }
        }
        if (cnt == 0) continue;
        V3 old = xx[n];
        double worst = WorstAround(xx, n);
        V3 centroid = (1.0 / cnt) * sum;
        V3 best = old;
        double bestQ = worst;
        const V3 targets[2] = {stuffed[n], centroid};
        const double lams[3] = {1.0, 0.5, 0.25};
        for (const V3& target : targets)
          for (double lambda : lams) {
            V3 p = Sym(old + lambda * (target - old), n);
            if (std::min(bottom.clearance(p, 0), top.clearance(p, gap)) < 0) continue;
            xx[n] = p;
            double q = WorstAround(xx, n);
            if (q > bestQ + 1e-6) { best = p; bestQ = q; }
          }
        xx[n] = best;
        if (!(best.x == old.x && best.y == old.y && best.z == old.z)) { cc[n] = FREE; sweepMoves++; }
      }
      moves += sweepMoves;
      if (sweepMoves == 0) break;
    }
    return moves;
  }

  // RemeshTrigger (:1867-1880) without the graded-mesh case: interior smoothing first, then a remesh when the worst
  // tet is still below q or a surface edge is longer than 2.8 h. 0 none, 1 quality, 2 stretched surface.
  int RemeshTrigger(double remeshQuality) {
    if (MinQuality(x) < remeshQuality) SmoothInterior(x, remeshQuality, 6);
    if (MinQuality(x) < remeshQuality) return 1;
    double maxEdge = 0;
    for (size_t k = 0; k < bfaces.size(); k += 3) {
      const V3 &a = x[bfaces[k]], &b = x[bfaces[k + 1]], &c = x[bfaces[k + 2]];
      maxEdge = std::max(maxEdge, std::max((b - a).len(), std::max((c - b).len(), (a - c).len())));
    }
    return maxEdge > 2.8 * P.h ? 2 : 0;
  }

  // RemeshCore (:1958-2102) with the uniform lattice and no local refinement (RefinementLevels' hot spots and the
  // graded lattice are not restated): true when the new mesh is accepted and the state moved onto it.
  bool RemeshCore() {
    rmImprove = rmReverts = rmSmooth = rmUntangle = 0;
    rmQOld = MinQuality(x);
    OLocator loc;
    loc.Build(tets, x, 2 * P.h);
    V3 mn = x[0], mx = x[0];
    for (const V3& q : x) {
      mn = V3(std::min(mn.x, q.x), std::min(mn.y, q.y), std::min(mn.z, q.z));
      mx = V3(std::max(mx.x, q.x), std::max(mx.y, q.y), std::max(mx.z, q.z));
    }
    const double eps1 = 1e-7 * P.h;
    auto In = [&](const V3& q) {
      if ((symX && q.x < 0) || (symY && q.y < 0)) return false;
      return loc.Contains(V3(symX ? std::max(q.x, eps1) : q.x, symY ? std::max(q.y, eps1) : q.y, q.z));
    };
    std::vector<V3> nx;
    std::vector<int> nt;
    OStuff(In, mn, mx, P.h, symX, symY, nx, nt);
    if (nt.empty()) { rmQNew = -1; return false; }
    // the new mesh's topology and local lengths at the stuffed positions
    Oracle M;
    M.P = P; M.bottom = bottom; M.top = top; M.symX = symX; M.symY = symY; M.gap = gap; M.recoveryPass = recoveryPass;
    M.nn = (int)nx.size(); M.ne = (int)(nt.size() / 4); M.ndof = 3 * M.nn;
    M.tets = nt; M.x = nx;
    M.onX0.assign(M.nn, 0); M.onY0.assign(M.nn, 0);
    for (int n = 0; n < M.nn; n++) { M.onX0[n] = symX && nx[n].x == 0; M.onY0[n] = symY && nx[n].y == 0; }
    M.vol0 = vol0;
    M.BuildTopology();
    if (!M.err.empty()) { rmQNew = -1; return false; }
    M.LocalLengths();
    std::vector<V3> xs = nx, stuffed = nx;
    std::vector<int> cc(M.nn, FREE);
    std::vector<V3> newNrm = M.SurfaceNormals(xs);
    const std::vector<double>& hLoc = M.hLocal;
    for (int n = 0; n < M.nn; n++) {
      if (!M.isSurf[n]) continue;
      double cb = bottom.clearance(xs[n], 0), ctp = top.clearance(xs[n], gap), tolC = 0.12 * hLoc[n];
      if (cb < tolC && cb <= ctp) cc[n] = BOTTOM;
      else if (ctp < tolC) cc[n] = TOP;
      if (cc[n] != FREE) {
        V3 q = Surface(cc[n]).project(xs[n], Offset(cc[n], gap), nullptr);
        xs[n] = V3(M.onX0[n] ? 0 : q.x, M.onY0[n] ? 0 : q.y, q.z);
      }
    }
    M.CompensateVolume(xs, cc);
    rmImprove = M.ImprovePlacement(xs, cc, stuffed, newNrm, gap, 0.1);
    for (int sweep = 0; sweep < 4; sweep++) {
      int reverted = 0;
      for (int e = 0; e < M.ne; e++) {
        if (M.TetQ(xs, e) >= 0.03) continue;
        for (int a = 0; a < 4; a++) {
          int n = M.tets[4 * e + a];
          if (xs[n].x == stuffed[n].x && xs[n].y == stuffed[n].y && xs[n].z == stuffed[n].z) continue;
          if (std::min(bottom.clearance(stuffed[n], 0), top.clearance(stuffed[n], gap)) < -0.05 * hLoc[n]) continue;
          xs[n] = stuffed[n];
          cc[n] = FREE;
          reverted++;
        }
      }
      rmReverts += reverted;
      if (reverted == 0) break;
    }
    rmSmooth = M.SmoothInterior(xs, 0.2, 10);
    rmUntangle = M.UntangleSurface(xs, cc, stuffed, 0.03, 6);
    rmSmooth += M.SmoothInterior(xs, 0.2, 4);
    rmQNew = M.MinQuality(xs);
    if (rmQNew < 0.02 && !(rmQNew > 0.002 && rmQNew > rmQOld + 0.002)) return false;
    // remap: element history from the old tet nearest the new centroid, velocities interpolated
    std::vector<double> nEps(M.ne), nDmg(M.ne), nSig(M.ne), nRate(M.ne), nMean(M.ne);
    for (int e = 0; e < M.ne; e++) {
      V3 c = 0.25 * (xs[nt[4 * e]] + xs[nt[4 * e + 1]] + xs[nt[4 * e + 2]] + xs[nt[4 * e + 3]]);
      int f = loc.FindNearest(c);
      nEps[e] = f >= 0 ? eps[f] : 0;
      nDmg[e] = f >= 0 ? dmg[f] : 0;
      nSig[e] = f >= 0 ? elSig[f] : sigRef;
      nRate[e] = f >= 0 ? elRate[f] : 0;
      nMean[e] = f >= 0 ? elMean[f] : 0;
    }
    std::vector<V3> nVg(M.nn);
    for (int n = 0; n < M.nn; n++) nVg[n] = loc.Interpolate(xs[n], vg);
    Allocate(xs, nt, M.onX0, M.onY0);  // the symmetry flags of the stuffed points (BuildMesh, :205-206)
    eps = nEps; dmg = nDmg; elSig = nSig; elRate = nRate; elMean = nMean;
    vg = nVg;
    contact = cc;
    warm = false;
    return true;
  }

  // TryUpdate (Forging3DSolver.cs:1455-1537): contact and placement, then the placement repair (ImprovePlacement,
  // SmoothInterior) and the collapse check; the recovery ladder's lenient passes are not restated.
  int TryUpdate(double dt) {
    double gapT = std::max(0.0, gap - P.ram * dt);
    if (gapT < 1e-9) gapT = 0;
    std::vector<V3> xT(nn), pred(nn);
    std::vector<int> cT(nn, FREE);
    std::vector<V3> nrm = SurfaceNormals(x);
    double maxPenRel = 0;
    for (int n = 0; n < nn; n++) {
      V3 p = Sym(x[n] + dt * vg[n], n);
      pred[n] = p;
      int c = contact[n];
      if (c != FREE && !release[n]) {
        const Die& s = Surface(c);
        double off = Offset(c, gapT);
        if (s.clearance(p, off) > 0.05 * hLocal[n]) cT[n] = FREE;
        else { p = Sym(s.project(p, off, nullptr), n); cT[n] = c; }
      } else if (isSurf[n]) {
        double cb = bottom.clearance(p, 0), ctp = top.clearance(p, gapT);
        if (cb < 0 || ctp < 0) {
          int hit = cb <= ctp ? BOTTOM : TOP;
          V3 q = Sym(Surface(hit).project(p, Offset(hit, gapT), nullptr), n);
          maxPenRel = std::max(maxPenRel, (q - p).len() / hLocal[n]);
          p = q;
          cT[n] = hit;
        }
      }
      xT[n] = p;
    }
    if (maxPenRel > 0.5) { err = "die penetration exceeds half the local element size"; return 1; }
    lastReverts = ImprovePlacement(xT, cT, pred, nrm, gapT, 0.05);
    lastSmoothMoves = SmoothInterior(xT, 0.05, 6);
    for (int e = 0; e < ne; e++)
      if (Quality(xT[tets[4 * e]], xT[tets[4 * e + 1]], xT[tets[4 * e + 2]], xT[tets[4 * e + 3]]) <= 0.005) {
        err = "tetrahedron collapsed";
        return 1;
      }
    x = xT;
    contact = cT;
    gap = gapT;
    time_ += dt;
    return 0;
  }

  // LayerLimitedIncrement (Forging3DSolver.cs:236-249)
  double LayerLimitedIncrement() const {
    double thinnest = INFINITY;
    for (int n = 0; n < nn; n++) {
      if (contact[n] != TOP) continue;
      double zb = bottom.surface_z(x[n].x, x[n].y), zt = top.surface_z(x[n].x, x[n].y) + gap;
      if (std::isfinite(zb) && std::isfinite(zt) && zt > zb) thinnest = std::min(thinnest, zt - zb);
    }
    return std::isfinite(thinnest) ? std::max(0.1 * thinnest, P.stroke_increment / 32) : INFINITY;
  }

  // One increment (Forging3DSolver.cs:540-638) on a fixed mesh: no remeshing, no recovery ladder.
  // Step (Forging3DSolver.cs:540-638). With remeshing on (remeshQuality > 0): the remesh trigger first, then the
  // recovery ladder's first two passes (as is; after a remesh, with the first pass's corner allowance doubled); the
  // refinement passes are not restated.
  double remeshQuality = 0;
  int remeshCount = 0;
  int Step(double& load, double& reaction, int& iterations) {
    if (remeshQuality > 0 && RemeshTrigger(remeshQuality) && RemeshCore()) remeshCount++;
    const int passes = remeshQuality > 0 ? 2 : 1;
    bool done = false;
    double dt = 0;
    for (int pass = 0; pass < passes && !done; pass++) {
      recoveryPass = pass;
      int rc = PrepareStep();
      if (rc) { recoveryPass = 0; return rc; }
      int converged = 0;
      iterations = SolveVelocity(converged);
      if (iterations < 0) { recoveryPass = 0; return iterations; }
      if (!converged && lastRel >= 1e-2) {
        if (pass < passes - 1 && RemeshCore()) { remeshCount++; continue; }
        err = "Velocity solution did not converge";
        recoveryPass = 0;
        return -7;
      }
      double power = FinalEvaluation(reaction);
      load = power / P.ram * P.sym_factor;
      reaction *= P.sym_factor;
      double dsx = std::min(std::min(ds, P.stroke_increment), LayerLimitedIncrement());
      dt = std::min(dsx, gap) / P.ram;
      if (gap - P.ram * dt < 0.05 * dsx) dt = gap / P.ram;
      for (int attempt = 0; attempt <= 5 && !done; attempt++) {
        double dta = dt / (1 << attempt);
        if (TryUpdate(dta) == 0) { dt = dta; done = true; }
      }
      if (done || pass == passes - 1) break;
      if (RemeshCore()) remeshCount++;  // a rejected remesh still leaves the next pass
    }
    recoveryPass = 0;
    if (!done) return -8;
    for (int e = 0; e < ne; e++) eps[e] += elRate[e] * dt;
    AccumulateDamage(dt);
    warm = true;
    lastDt = dt;
    return 0;
  }
  double lastDt = 0;
};

}  // namespace

extern "C" {

void* orc_create(const OrcParams* prm, int nn, int ne, const int* tets, const double* x, const unsigned char* onX0,
                 const unsigned char* onY0, double gap0) {
  auto* o = new Oracle();
  o->P = *prm;
  o->law = Law{prm->C, prm->N, prm->M, prm->Eps0};
  o->bottom.side = 0; o->bottom.kind = prm->bottom_kind;
  o->top.side = 1; o->top.kind = prm->top_kind;
  for (int k = 0; k < 8; k++) { o->bottom.p[k] = prm->bottom_p[k]; o->top.p[k] = prm->top_p[k]; }
  o->nn = nn; o->ne = ne; o->ndof = 3 * nn;
  o->tets.assign(tets, tets + 4 * (size_t)ne);
  o->x.resize(nn);
  for (int n = 0; n < nn; n++) o->x[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
  o->onX0.assign(onX0, onX0 + nn);
  o->onY0.assign(onY0, onY0 + nn);
  o->vg.assign(nn, V3());
  o->eps.assign(ne, 0); o->dmg.assign(ne, 0); o->elDev.assign(6 * (size_t)ne, 0);
  o->elSig.assign(ne, 0); o->elRate.assign(ne, 0); o->elMean.assign(ne, 0);
  o->grad.assign(12 * (size_t)ne, 0); o->vol.assign(ne, 0);
  o->contact.assign(nn, FREE); o->nFix.assign(nn, 0); o->dieAxis.assign(nn, -1);
  o->fixedv.assign(o->ndof, 0); o->fixedVal.assign(o->ndof, 0); o->release.assign(nn, 0);
  o->kFric.assign(nn, 0); o->aFric.assign(nn, 0);
  o->ax0.assign(nn, V3()); o->ax1.assign(nn, V3()); o->ax2.assign(nn, V3());
  o->v.assign(o->ndof, 0); o->r.assign(o->ndof, 0); o->dv.assign(o->ndof, 0); o->vTrial.assign(o->ndof, 0);
  o->rvol.assign(o->ndof, 0); o->vNode.assign(nn, 0);
  o->pr.assign(o->ndof, 0); o->pz.assign(o->ndof, 0); o->pp.assign(o->ndof, 0); o->pq.assign(o->ndof, 0);
  o->eR.assign(12 * (size_t)ne, 0); o->eEnergy.assign(ne, 0); o->ePower.assign(ne, 0);
  // constants (Forging3DSolver.cs:144-150)
  o->rateRef = prm->ram / prm->size;
  o->rateCut = 1e-3 * o->rateRef;
  o->sigRef = o->law.stress(0, o->rateRef);
  o->kPen = prm->penalty_factor * o->sigRef / o->rateRef;
  o->u0 = 5e-3 * prm->ram;
  o->gap = o->gap0 = gap0;
  o->ds = prm->stroke_increment;
  o->BuildTopology();
  o->LocalLengths();
  o->BuildPattern();
  std::fill(o->elSig.begin(), o->elSig.end(), o->sigRef);
  for (int n = 0; n < nn; n++) { o->symX = o->symX || o->onX0[n]; o->symY = o->symY || o->onY0[n]; }
  for (int e = 0; e < ne; e++)
    o->vol0 += Oracle::TetVol(o->x[tets[4 * e]], o->x[tets[4 * e + 1]], o->x[tets[4 * e + 2]], o->x[tets[4 * e + 3]]);
  o->vol0 *= prm->sym_factor;
  return o;
}

void orc_destroy(void* h) { delete (Oracle*)h; }

const char* orc_error(void* h) { return ((Oracle*)h)->err.c_str(); }

void orc_sizes(void* h, long long* out) {
  auto* o = (Oracle*)h;
  out[0] = o->nn; out[1] = o->ne; out[2] = o->ndof; out[3] = (long long)o->rowIdx.size();
  out[4] = (long long)o->fsym.size(); out[5] = (long long)o->gNode.size();
}

void orc_constants(void* h, double* out) {
  auto* o = (Oracle*)h;
  out[0] = o->rateRef; out[1] = o->rateCut; out[2] = o->sigRef; out[3] = o->kPen; out[4] = o->u0; out[5] = o->gap;
}

void orc_get_pattern(void* h, int* colPtr, int* rowIdx, int* nbrStart, int* nbr) {
  auto* o = (Oracle*)h;
  std::copy(o->colPtr.begin(), o->colPtr.end(), colPtr);
  std::copy(o->rowIdx.begin(), o->rowIdx.end(), rowIdx);
  if (nbrStart) std::copy(o->nbrStart.begin(), o->nbrStart.end(), nbrStart);
  if (nbr) std::copy(o->nbr.begin(), o->nbr.end(), nbr);
}

void orc_get_topology(void* h, int* bfaces, int* owner, unsigned char* fsym, unsigned char* isSurf) {
  auto* o = (Oracle*)h;
  std::copy(o->bfaces.begin(), o->bfaces.end(), bfaces);
  std::copy(o->bowner.begin(), o->bowner.end(), owner);
  for (size_t k = 0; k < o->fsym.size(); k++) fsym[k] = o->fsym[k];
  for (int n = 0; n < o->nn; n++) isSurf[n] = o->isSurf[n];
}

// Initial contact and velocity guess (Forging3DSolver.cs:158-182) for a vertical-cylinder billet.
void orc_initial(void* h, double zmin, double cx, double cy) {
  auto* o = (Oracle*)h;
  double tol = 0.05 * o->P.h;
  int nb = 0, nt = 0;
  for (int n = 0; n < o->nn; n++) {
    if (!o->isSurf[n]) continue;
    if (o->bottom.clearance(o->x[n], 0) < tol) { o->contact[n] = BOTTOM; nb++; }
    else if (o->top.clearance(o->x[n], o->gap) < tol) { o->contact[n] = TOP; nt++; }
  }
  for (int side = BOTTOM; side <= TOP; side++) {
    if ((side == BOTTOM && nb) || (side == TOP && nt)) continue;
    int best = -1;
    double bestC = 1e300;
    for (int n = 0; n < o->nn; n++) {
      if (!o->isSurf[n]) continue;
      double c = side == BOTTOM ? o->bottom.clearance(o->x[n], 0) : o->top.clearance(o->x[n], o->gap);
      if (c < bestC) { bestC = c; best = n; }
    }
    if (best >= 0) o->contact[best] = side;
  }
  for (int n = 0; n < o->nn; n++)
    if (o->contact[n] != FREE) o->x[n] = o->Sym(o->Surface(o->contact[n]).project(o->x[n], o->Offset(o->contact[n], o->gap), nullptr), n);
  for (int n = 0; n < o->nn; n++) {
    V3 p = o->x[n];
    o->vg[n] = V3(0.5 * o->rateRef * (p.x - cx), 0.5 * o->rateRef * (p.y - cy), -o->rateRef * (p.z - zmin));
  }
}

void orc_get_state(void* h, double* x, double* vg, int* contact, double* eps, double* elSig, double* elRate, double* gap,
                   unsigned char* warm) {
  auto* o = (Oracle*)h;
  for (int n = 0; n < o->nn; n++) {
    if (x) { x[3 * n] = o->x[n].x; x[3 * n + 1] = o->x[n].y; x[3 * n + 2] = o->x[n].z; }
    if (vg) { vg[3 * n] = o->vg[n].x; vg[3 * n + 1] = o->vg[n].y; vg[3 * n + 2] = o->vg[n].z; }
    if (contact) contact[n] = o->contact[n];
  }
  if (eps) std::copy(o->eps.begin(), o->eps.end(), eps);
  if (elSig) std::copy(o->elSig.begin(), o->elSig.end(), elSig);
  if (elRate) std::copy(o->elRate.begin(), o->elRate.end(), elRate);
  if (gap) *gap = o->gap;
  if (warm) *warm = o->warm;
}

void orc_set_state(void* h, const double* x, const double* vg, const int* contact, const double* eps, const double* elSig,
                   double gap, int warm) {
  auto* o = (Oracle*)h;
  for (int n = 0; n < o->nn; n++) {
    if (x) o->x[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
    if (vg) o->vg[n] = V3(vg[3 * n], vg[3 * n + 1], vg[3 * n + 2]);
    if (contact) o->contact[n] = contact[n];
  }
  if (eps) std::copy(eps, eps + o->ne, o->eps.begin());
  if (elSig) std::copy(elSig, elSig + o->ne, o->elSig.begin());
  o->gap = gap;
  o->warm = warm != 0;
}

int orc_prepare(void* h) { return ((Oracle*)h)->PrepareStep(); }

void orc_get_frames(void* h, double* ax, int* nFix, int* dieAxis, unsigned char* fixedv, double* fixedVal) {
  auto* o = (Oracle*)h;
  for (int n = 0; n < o->nn; n++) {
    V3 e[3] = {o->ax0[n], o->ax1[n], o->ax2[n]};
    for (int k = 0; k < 3; k++) { ax[9 * n + 3 * k] = e[k].x; ax[9 * n + 3 * k + 1] = e[k].y; ax[9 * n + 3 * k + 2] = e[k].z; }
    nFix[n] = o->nFix[n];
    dieAxis[n] = o->dieAxis[n];
  }
  for (int d = 0; d < o->ndof; d++) { fixedv[d] = o->fixedv[d]; fixedVal[d] = o->fixedVal[d]; }
}

void orc_get_geometry(void* h, double* grad, double* vol, double* vNode, double* kvol, double* aFric, double* kFric) {
  auto* o = (Oracle*)h;
  if (grad) std::copy(o->grad.begin(), o->grad.end(), grad);
  if (vol) std::copy(o->vol.begin(), o->vol.end(), vol);
  if (vNode) std::copy(o->vNode.begin(), o->vNode.end(), vNode);
  if (kvol) std::copy(o->kvol.begin(), o->kvol.end(), kvol);
  if (aFric) std::copy(o->aFric.begin(), o->aFric.end(), aFric);
  if (kFric) std::copy(o->kFric.begin(), o->kFric.end(), kFric);
}

void orc_get_v(void* h, double* v) { auto* o = (Oracle*)h; std::copy(o->v.begin(), o->v.end(), v); }
void orc_set_v(void* h, const double* v) { auto* o = (Oracle*)h; std::copy(v, v + o->ndof, o->v.begin()); }

double orc_assemble(void* h, int matrix, int newton, const double* v, double* r, double* val, double* power) {
  auto* o = (Oracle*)h;
  double pw;
  double e = o->Assemble(matrix != 0, newton != 0, v, pw);
  if (r) std::copy(o->r.begin(), o->r.end(), r);
  if (val && matrix) std::copy(o->val.begin(), o->val.end(), val);
  if (power) *power = pw;
  return e;
}

void orc_get_element_state(void* h, double* elSig, double* elRate, double* elDev, double* elMean, double* dmg) {
  auto* o = (Oracle*)h;
  if (elSig) std::copy(o->elSig.begin(), o->elSig.end(), elSig);
  if (elRate) std::copy(o->elRate.begin(), o->elRate.end(), elRate);
  if (elDev) std::copy(o->elDev.begin(), o->elDev.end(), elDev);
  if (elMean) std::copy(o->elMean.begin(), o->elMean.end(), elMean);
  if (dmg) std::copy(o->dmg.begin(), o->dmg.end(), dmg);
}

// TryUpdate (Forging3DSolver.cs:1455-1537) with the state's velocities over dt: 0 on success (x, contact, gap and time
// advance), else non-zero with the reason in orc_error; moves[2] = nodes moved by ImprovePlacement and SmoothInterior.
int orc_try_update(void* h, double dt, int* moves) {
  auto* o = (Oracle*)h;
  int rc = o->TryUpdate(dt);
  if (moves) { moves[0] = o->lastReverts; moves[1] = o->lastSmoothMoves; }
  return rc;
}

// RemeshTrigger (:1867-1880) on the state (its interior smoothing moves x): 0 none, 1 quality, 2 stretched surface.
int orc_remesh_trigger(void* h, double remeshQuality) { return ((Oracle*)h)->RemeshTrigger(remeshQuality); }

// RemeshCore (:1958-2102) at recovery pass `pass` (0 or 1): 1 when the new mesh was accepted (the state now lives on
// it: orc_sizes, orc_get_mesh, orc_get_state), 0 when rejected. stats[6]: ImprovePlacement moves, reverts, smoothing
// moves, untangling moves, worst quality before and after.
int orc_remesh(void* h, int pass, double* stats) {
  auto* o = (Oracle*)h;
  o->recoveryPass = pass;
  bool ok = o->RemeshCore();
  o->recoveryPass = 0;
  if (stats) {
    stats[0] = o->rmImprove; stats[1] = o->rmReverts; stats[2] = o->rmSmooth; stats[3] = o->rmUntangle;
    stats[4] = o->rmQOld; stats[5] = o->rmQNew;
  }
  return ok ? 1 : 0;
}

// UntangleSurface (:1698-1772) and CompensateVolume (:2108-2145) on positions x[3nn] and contact[nn] (in place), with
// the stuffed positions stuffed[3nn] for UntangleSurface; the first returns its number of moves.
int orc_untangle_surface(void* h, double* x, int* contact, const double* stuffed, double qTarget, int sweeps) {
  auto* o = (Oracle*)h;
  std::vector<V3> xx(o->nn), st(o->nn);
  std::vector<int> cc(contact, contact + o->nn);
  for (int n = 0; n < o->nn; n++) {
    xx[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
    st[n] = V3(stuffed[3 * n], stuffed[3 * n + 1], stuffed[3 * n + 2]);
  }
  int m = o->UntangleSurface(xx, cc, st, qTarget, sweeps);
  for (int n = 0; n < o->nn; n++) { x[3 * n] = xx[n].x; x[3 * n + 1] = xx[n].y; x[3 * n + 2] = xx[n].z; contact[n] = cc[n]; }
  return m;
}
void orc_compensate_volume(void* h, double* x, const int* contact) {
  auto* o = (Oracle*)h;
  std::vector<V3> xx(o->nn);
  std::vector<int> cc(contact, contact + o->nn);
  for (int n = 0; n < o->nn; n++) xx[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
  o->CompensateVolume(xx, cc);
  for (int n = 0; n < o->nn; n++) { x[3 * n] = xx[n].x; x[3 * n + 1] = xx[n].y; x[3 * n + 2] = xx[n].z; }
}

// Overwrite the accumulated Cockcroft-Latham damage dmg[ne].
void orc_set_damage(void* h, const double* dmg) {
  auto* o = (Oracle*)h;
  std::copy(dmg, dmg + o->ne, o->dmg.begin());
}

// Remeshing inside orc_step: the trigger's quality threshold (RemeshQuality, 0.12 by default in the reference), 0 off.
void orc_set_remesh(void* h, double remeshQuality) { ((Oracle*)h)->remeshQuality = remeshQuality; }
int orc_remesh_count(void* h) { return ((Oracle*)h)->remeshCount; }

// The current mesh: tets[4ne], onX0/onY0[nn].
void orc_get_mesh(void* h, int* tets, unsigned char* onX0, unsigned char* onY0) {
  auto* o = (Oracle*)h;
  if (tets) std::copy(o->tets.begin(), o->tets.end(), tets);
  for (int n = 0; n < o->nn; n++) {
    if (onX0) onX0[n] = o->onX0[n];
    if (onY0) onY0[n] = o->onY0[n];
  }
}

// Initial volume (Forging3DSolver.cs:184) and the symmetry planes of the model (set at creation from onX0 / onY0).
void orc_set_remesh_info(void* h, double vol0, int symX, int symY) {
  auto* o = (Oracle*)h;
  o->vol0 = vol0; o->symX = symX != 0; o->symY = symY != 0;
}

// TetLocator (TetMesh.cs:115-262) over (tets, x) with cell size `cell`: per query point Find (tet, weights), FindNearest
// and Interpolate of field[3nn] (any output may be NULL).
void orc_locate(int nn, const double* x, int ne, const int* tets, double cell, int np, const double* p, int* found,
                double* w, int* nearest, const double* field, double* interp) {
  std::vector<V3> xx(nn);
  for (int n = 0; n < nn; n++) xx[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
  std::vector<int> tt(tets, tets + 4 * (size_t)ne);
  OLocator loc;
  loc.Build(tt, xx, cell);
  std::vector<V3> f;
  if (field) { f.resize(nn); for (int n = 0; n < nn; n++) f[n] = V3(field[3 * n], field[3 * n + 1], field[3 * n + 2]); }
  for (int i = 0; i < np; i++) {
    V3 q(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    if (found) found[i] = loc.Find(q, w ? w + 4 * i : nullptr);
    if (nearest) nearest[i] = loc.FindNearest(q);
    if (field && interp) { V3 v = loc.Interpolate(q, f); interp[3 * i] = v.x; interp[3 * i + 1] = v.y; interp[3 * i + 2] = v.z; }
  }
}

// IsosurfaceStuffing.Stuff (IsosurfaceStuffing.cs:17-88) of the interior of the mesh (tets, x) (inside test: its
// TetLocator with cell `cell`) over [bmin, bmax] at spacing h. Call with outX = NULL to get the sizes, then again with
// room for nn_out points and ne_out tets.
void orc_stuff(int nn, const double* x, int ne, const int* tets, double cell, const double* bmin, const double* bmax,
               double h, int symX, int symY, int* nn_out, int* ne_out, double* outX, int* outT) {
  std::vector<V3> xx(nn);
  for (int n = 0; n < nn; n++) xx[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
  std::vector<int> tt(tets, tets + 4 * (size_t)ne);
  OLocator loc;
  loc.Build(tt, xx, cell);
  const double e1 = 1e-7 * h;
  auto In = [&](const V3& q) {
    if ((symX && q.x < 0) || (symY && q.y < 0)) return false;
    return loc.Contains(V3(symX ? std::max(q.x, e1) : q.x, symY ? std::max(q.y, e1) : q.y, q.z));
  };
  std::vector<V3> px;
  std::vector<int> pt;
  OStuff(In, V3(bmin[0], bmin[1], bmin[2]), V3(bmax[0], bmax[1], bmax[2]), h, symX != 0, symY != 0, px, pt);
  *nn_out = (int)px.size();
  *ne_out = (int)(pt.size() / 4);
  if (outX) for (size_t n = 0; n < px.size(); n++) { outX[3 * n] = px[n].x; outX[3 * n + 1] = px[n].y; outX[3 * n + 2] = px[n].z; }
  if (outT) std::copy(pt.begin(), pt.end(), outT);
}

// SmoothInterior (:1776-1838) on the node positions x[3nn] (in place); returns the number of moves.
int orc_smooth_interior(void* h, double* x, double qTarget, int sweeps) {
  auto* o = (Oracle*)h;
  std::vector<V3> xx(o->nn);
  for (int n = 0; n < o->nn; n++) xx[n] = V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]);
  int m = o->SmoothInterior(xx, qTarget, sweeps);
  for (int n = 0; n < o->nn; n++) { x[3 * n] = xx[n].x; x[3 * n + 1] = xx[n].y; x[3 * n + 2] = xx[n].z; }
  return m;
}

// Overwrite the converged element state that AccumulateDamage reads (any pointer may be NULL).
void orc_set_element_state(void* h, const double* elSig, const double* elRate, const double* elDev,
                           const double* elMean) {
  auto* o = (Oracle*)h;
  if (elSig) std::copy(elSig, elSig + o->ne, o->elSig.begin());
  if (elRate) std::copy(elRate, elRate + o->ne, o->elRate.begin());
  if (elDev) std::copy(elDev, elDev + 6 * (size_t)o->ne, o->elDev.begin());
  if (elMean) std::copy(elMean, elMean + o->ne, o->elMean.begin());
}

// AccumulateDamage (Forging3DSolver.cs:1844-1855) on the current element state.
void orc_accumulate_damage(void* h, double dt) { ((Oracle*)h)->AccumulateDamage(dt); }

// Principal.Max3 (Solver/Principal.cs:7-19) of n tensors t6[6n] = (xx, yy, zz, xy, yz, zx).
void orc_max3(int n, const double* t6, double* out) {
  for (int k = 0; k < n; k++) {
    const double* t = t6 + 6 * (size_t)k;
    out[k] = max3(t[0], t[1], t[2], t[3], t[4], t[5]);
  }
}

// stats: [factorizations, pcg iterations, pcg calls, solves, assemblies]
int orc_solve_velocity(void* h, int* converged, double* lastRel, int* stats) {
  auto* o = (Oracle*)h;
  int cv = 0;
  int it = o->SolveVelocity(cv);
  *converged = cv;
  *lastRel = o->lastRel;
  if (stats) { stats[0] = o->statFactor; stats[1] = o->statPcgIts; stats[2] = o->statPcgCalls; stats[3] = o->statSolves; stats[4] = o->statAssemblies; }
  return it;
}

double orc_final(void* h, double* reaction) { return ((Oracle*)h)->FinalEvaluation(*reaction); }

int orc_step(void* h, double* out) {
  auto* o = (Oracle*)h;
  double load = 0, reaction = 0;
  int its = 0;
  int rc = o->Step(load, reaction, its);
  out[0] = load; out[1] = reaction; out[2] = its; out[3] = o->lastDt; out[4] = o->gap;
  out[5] = o->statFactor; out[6] = o->statPcgIts; out[7] = o->statAssemblies;
  return rc;
}

void orc_set_perm(void* h, const int* perm) {
  auto* o = (Oracle*)h;
  if (perm) { o->userPerm.assign(perm, perm + o->ndof); o->useUserPerm = true; }
  else o->useUserPerm = false;
  o->needAnalyze = true;
  o->haveFactor = false;
}

void orc_reset_factor(void* h) { auto* o = (Oracle*)h; o->haveFactor = false; o->pcgFailures = o->pcgSkip = 0; }

long long orc_factor_nnz(void* h) { return ((Oracle*)h)->FactorNonZeros(); }

// TetMesh.Quality (TetMesh.cs:107-112) of every tet at the node positions x[3nn]: the restatement TryUpdate uses
void orc_quality(int ne, const int* tets, const double* x, double* q) {
  auto P = [&](int n) { return V3(x[3 * n], x[3 * n + 1], x[3 * n + 2]); };
  for (int e = 0; e < ne; e++)
    q[e] = Oracle::Quality(P(tets[4 * e]), P(tets[4 * e + 1]), P(tets[4 * e + 2]), P(tets[4 * e + 3]));
}

// ---- stand-alone sparse Cholesky (full symmetric CSC in, FP64) for tests and the CPU baseline
void* orc_chol_create(int n, const int* colPtr, const int* rowIdx, const int* perm, int* parent_out, int* colcount_out, int* rc) {
  auto* c = new SparseChol();
  std::vector<int> cc;
  *rc = c->analyze(n, colPtr, rowIdx, perm, &cc);
  if (parent_out) std::copy(c->parent.begin(), c->parent.end(), parent_out);
  if (colcount_out) std::copy(cc.begin(), cc.end(), colcount_out);
  return c;
}
int orc_chol_factor(void* c, const double* vals) { return ((SparseChol*)c)->factorize(vals); }
void orc_chol_solve(void* c, const double* b, double* x) { ((SparseChol*)c)->solve(b, x); }
long long orc_chol_nnz(void* c) { return ((SparseChol*)c)->nnz(); }
void orc_chol_destroy(void* c) { delete (SparseChol*)c; }

// Elimination tree of P A P^T by the textbook definition (parent(j) = min{i > j : L(i,j) != 0}) from a brute-force
// symbolic elimination on a dense boolean matrix. Small n only: this pins the sparse symbolic paths.
int orc_brute_symbolic(int n, const int* colPtr, const int* rowIdx, const int* perm, int* parent_out, int* colcount_out) {
  if (n > 6000) return -1;
  std::vector<int> iperm(n);
  for (int k = 0; k < n; k++) iperm[perm[k]] = k;
  std::vector<std::vector<char>> M(n, std::vector<char>(n, 0));
  for (int j = 0; j < n; j++)
    for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
      int a = iperm[rowIdx[p]], b = iperm[j];
      M[a][b] = M[b][a] = 1;
    }
  for (int k = 0; k < n; k++) {
    std::vector<int> nz;
    for (int i = k + 1; i < n; i++) if (M[i][k]) nz.push_back(i);
    for (int a : nz) for (int b : nz) M[a][b] = 1;
    parent_out[k] = nz.empty() ? -1 : nz[0];
    colcount_out[k] = 1 + (int)nz.size();
  }
  return 0;
}

// Minimum-degree node ordering of a full symmetric CSC pattern whose 3 consecutive columns form a node.
void orc_min_degree_nodes(int nn, const int* nbrStart, const int* nbr, int* perm_out) {
  std::vector<int> adjp(nn + 1, 0), adj;
  for (int n = 0; n < nn; n++) {
    for (int q = nbrStart[n]; q < nbrStart[n + 1]; q++) if (nbr[q] != n) adj.push_back(nbr[q]);
    adjp[n + 1] = (int)adj.size();
  }
  std::vector<int> w(nn, 3);
  std::vector<int> ord = min_degree(nn, adjp, adj, w);
  for (int k = 0; k < nn; k++) for (int c = 0; c < 3; c++) perm_out[3 * k + c] = 3 * ord[k] + c;
}

// ---- MKL PARDISO on a full symmetric CSC (the reference's MklPardisoSolver settings, MklPardisoSolver.cs:20-99)
void* orc_pardiso_create(int n, const int* colPtr, const int* rowIdx, const int* perm, int want_perm, int* rc) {
  auto* p = new PardisoDL();
  if (!p->ok()) { *rc = -10; return p; }
  p->want_perm = want_perm;
  *rc = p->analyze(n, colPtr, rowIdx, perm);
  return p;
}
int orc_pardiso_factor(void* p, const double* vals) { return ((PardisoDL*)p)->factorize(vals); }
int orc_pardiso_solve(void* p, const double* b, double* x) { return ((PardisoDL*)p)->solve(b, x); }
void orc_pardiso_info(void* p, long long* out) {
  auto* q = (PardisoDL*)p;
  out[0] = q->nnz_l; out[1] = q->mflops; out[2] = q->threads;
}
void orc_pardiso_perm(void* p, int* perm) { auto* q = (PardisoDL*)p; std::copy(q->perm.begin(), q->perm.end(), perm); }
void orc_pardiso_destroy(void* p) { delete (PardisoDL*)p; }
int orc_pardiso_available() { PardisoDL p; return p.ok() ? 1 : 0; }
void orc_set_threads(int n) { PardisoDL::set_threads(n); }

// die surface heights at n points (the height field every contact routine of the oracle reads)
void orc_die_height(int side, int kind, const double* p, int n, const double* x, const double* y, double* out) {
  Die d;
  d.side = side; d.kind = kind;
  std::copy(p, p + 8, d.p);
  for (int i = 0; i < n; ++i) out[i] = d.height(x[i], y[i]);
}

}  // extern "C"
